"""Cell renumbering of a mesh (ins_mesh_renumber): a permutation of the same cells whose
operator is the permuted operator; cells of one shape class are grouped in their old order."""
import numpy as np
import pytest


@pytest.mark.parametrize("dim", [2, 3])
def test_shape_class_renumbering(ip, oracle, dim):
    mesh = ip.Mesh.square(5, jitter=0.1, seed=4) if dim == 2 else ip.Mesh.cube(3, 0.1, 4)
    coords, c0 = mesh.arrays()
    nf0, nb0 = mesh.n_faces, mesh.n_bfaces
    perm = mesh.renumber()
    _, c1 = mesh.arrays()
    assert np.array_equal(np.sort(perm), np.arange(len(perm)))
    assert np.array_equal(c1, c0[perm])            # same cells, same vertex order
    assert (mesh.n_faces, mesh.n_bfaces) == (nf0, nb0)
    # the operator of the renumbered mesh is the permuted operator
    N = 2
    o0 = oracle.Oracle(coords, c0, N)
    o1 = oracle.Oracle(coords, c1, N)
    x = np.random.default_rng(dim).standard_normal((len(perm), o0.Np))
    assert np.abs(o1.mult(x[perm]) - o0.mult(x)[perm]).max() < 1e-12 * np.abs(o0.mult(x)).max()
    # a structured subdivision has 2 (2D) or 6 (3D) classes, each kept in its old order
    assert np.sum(np.diff(perm) < 0) == (1 if dim == 2 else 5)


def test_windowed_shape_class_renumbering(ip, oracle):
    """Windowed grouping: each run of 64 K cells (K = 6 classes on the Kuhn cube) is a
    permutation of the same run, each 64-cell wave holds one class, and the operator of the
    renumbered mesh is the permuted operator."""
    n = 8
    glob = ip.Mesh.cube(n, 0.1, 4)
    pg = glob.renumber(ip.RENUMBER_SHAPE_CLASS)
    E = len(pg)
    cls = np.empty(E, dtype=np.int64)
    cls[pg] = np.arange(E) // (E // 6)               # class of each old cell, from the global grouping
    mesh = ip.Mesh.cube(n, 0.1, 4)
    coords, c0 = mesh.arrays()
    perm = mesh.renumber(ip.RENUMBER_SHAPE_CLASS_WINDOWED)
    _, c1 = mesh.arrays()
    W = 64 * 6
    assert E % W == 0
    for b in range(0, E, W):
        assert np.array_equal(np.sort(perm[b:b + W]), np.arange(b, b + W))
    waves = cls[perm].reshape(-1, 64)
    assert (waves == waves[:, :1]).all()
    assert np.array_equal(c1, c0[perm])
    o0 = oracle.Oracle(coords, c0, 1)
    o1 = oracle.Oracle(coords, c1, 1)
    x = np.random.default_rng(7).standard_normal((E, o0.Np))
    assert np.abs(o1.mult(x[perm]) - o0.mult(x)[perm]).max() < 1e-12 * np.abs(o0.mult(x)).max()


def test_unknown_renumbering_is_rejected(ip):
    mesh = ip.Mesh.square(3)
    with pytest.raises(Exception):
        mesh.renumber(99)


def test_tiled_shape_class_renumbering(ip, oracle):
    """Tiled grouping on the Kuhn cube (8^3 hexes): tiles of 4^3 hexes hold 64 cells of each of
    the 6 classes, so each run of 384 cells lies in one 4h cube, each 64-cell wave holds one
    class, and the operator of the renumbered mesh is the permuted operator."""
    n = 8
    glob = ip.Mesh.cube(n, 0.1, 4)
    pg = glob.renumber(ip.RENUMBER_SHAPE_CLASS)
    E = len(pg)
    cls = np.empty(E, dtype=np.int64)
    cls[pg] = np.arange(E) // (E // 6)
    mesh = ip.Mesh.cube(n, 0.1, 4)
    coords, c0 = mesh.arrays()
    perm = mesh.renumber(ip.RENUMBER_SHAPE_CLASS_TILED)
    _, c1 = mesh.arrays()
    assert np.array_equal(np.sort(perm), np.arange(E))
    assert np.array_equal(c1, c0[perm])
    waves = cls[perm].reshape(-1, 64)
    assert (waves == waves[:, :1]).all()
    cen = coords[c1].mean(axis=1)                      # (E, 3) centroids in the new order
    tiles = np.floor(cen / (4.0 / n)).astype(int).reshape(-1, 384, 3)
    assert (tiles == tiles[:, :1, :]).all()
    assert len({tuple(t) for t in tiles[:, 0, :]}) == E // 384
    o0 = oracle.Oracle(coords, c0, 1)
    o1 = oracle.Oracle(coords, c1, 1)
    x = np.random.default_rng(9).standard_normal((E, o0.Np))
    assert np.abs(o1.mult(x[perm]) - o0.mult(x)[perm]).max() < 1e-12 * np.abs(o0.mult(x)).max()


def test_tiled_renumbering_2d_and_ragged(ip):
    """2D squares (2 classes, tiles of 8x8 squares) and a cube whose side is no multiple of the
    tile: still a permutation of the same cells."""
    for mesh in (ip.Mesh.square(12, jitter=0.1, seed=2), ip.Mesh.cube(5, 0.1, 3)):
        _, c0 = mesh.arrays()
        perm = mesh.renumber(ip.RENUMBER_SHAPE_CLASS_TILED)
        _, c1 = mesh.arrays()
        assert np.array_equal(np.sort(perm), np.arange(len(perm)))
        assert np.array_equal(c1, c0[perm])
