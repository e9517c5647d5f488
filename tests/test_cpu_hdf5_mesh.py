"""OP2 HDF5 meshes (grids/h5 of the reference; three copied under tests/golden/meshes)
and periodic faces.

* The library's reader (ins_mesh_load_op2_hdf5, the stand-in for DGMesh2D/3D(filename),
  src/ins2d.cpp:192) returns the same arrays as an independent numpy parse of the files.
* Periodic faces (cube-20.h5 and vortex_square_small.h5 join opposite sides of their
  boxes) are recognised exactly where the file's periodicFace dat says.
* The oracle's operator with caller-given faces, periodic ones included, equals the
  independent quadrature assembly with the same periodic pairing.
"""
import os

import numpy as np
import pytest

from h5_min import H5
from sipg_textbook import assemble

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
FV = {2: [(0, 1), (1, 2), (2, 0)], 3: [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]}


def path(name):
    return os.path.join(MESHES, name + ".h5")


@pytest.mark.parametrize("name", ["cube-20", "vortex_square_small", "3D-cylinder"])
def test_reader_matches_independent_parse(ip, name):
    m = ip.Mesh.load_hdf5(path(name))
    h = H5(path(name))
    coords, c2n = m.arrays()
    assert np.array_equal(coords, h.read("node_coords"))
    assert np.array_equal(c2n, h.read("cell2nodes"))
    f2c, fn, b2c, bn = m.faces()
    assert np.array_equal(f2c, h.read("face2cells"))
    assert np.array_equal(fn, h.read("faceNum" if "faceNum" in h.items else "edgeNum"))
    if "bface2cells" in h.items:
        assert np.array_equal(b2c, h.read("bface2cells").ravel())
        assert np.array_equal(bn, h.read("bfaceNum").ravel())
    else:
        assert m.n_bfaces == 0 and int(h.read("bfaces")[0]) == 0
    assert m.n_cells == int(h.read("cells")[0]) and m.n_nodes == int(h.read("nodes")[0])
    assert m.n_faces == int(h.read("faces")[0])


def _shared(c2n, dim, f2c, fn):
    return np.array([set(c2n[a][list(FV[dim][x])]) == set(c2n[b][list(FV[dim][y])])
                     for (a, b), (x, y) in zip(f2c, fn)])


@pytest.mark.parametrize("name,expect", [("cube-20", 738), ("vortex_square_small", 120),
                                         ("3D-cylinder", 0)])
def test_periodic_faces(ip, name, expect):
    m = ip.Mesh.load_hdf5(path(name))
    coords, c2n = m.arrays()
    f2c, fn, _, _ = m.faces()
    periodic = ~_shared(c2n, m.dim, f2c, fn)
    assert periodic.sum() == expect
    h = H5(path(name))
    if "periodicFace" in h.items:
        assert np.array_equal(periodic, h.read("periodicFace").ravel() != 0)
    # each periodic face joins two translates: the shift is a period of the box
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    for i in np.nonzero(periodic)[0][:50]:
        (a, b), (x, y) = f2c[i], fn[i]
        t = coords[c2n[b][list(FV[m.dim][y])]].mean(0) - coords[c2n[a][list(FV[m.dim][x])]].mean(0)
        assert np.allclose(np.abs(t), 0) | np.allclose(np.abs(t), hi - lo) or \
            np.all(np.isclose(np.abs(t), 0) | np.isclose(np.abs(t), hi - lo))


def periodic_op2_faces(mesh, axes):
    """The mesh's faces with boundary faces on opposite sides of the unit box paired along
    `axes` (OP2 face sets of a periodic mesh); other boundary faces stay bfaces."""
    coords, c2n = mesh.arrays()
    f2c, fn, b2c, bn = mesh.faces()
    dim = mesh.dim

    def key(c, f, shift):
        return tuple(sorted(tuple(np.round(coords[v] + shift, 9)) for v in c2n[c][list(FV[dim][f])]))

    index = {key(c, f, 0): i for i, (c, f) in enumerate(zip(b2c, bn))}
    pairs, used = [], set()
    for i, (c, f) in enumerate(zip(b2c, bn)):
        for ax in axes:
            sh = np.zeros(dim)
            sh[ax] = 1.0
            j = index.get(key(c, f, sh))
            if j is not None and i not in used and j not in used:
                pairs.append(((c, b2c[j]), (f, bn[j])))
                used.update((i, j))
    keep = [i for i in range(len(b2c)) if i not in used]
    pf2c = np.array([p[0] for p in pairs], dtype=np.int32).reshape(-1, 2)
    pfn = np.array([p[1] for p in pairs], dtype=np.int32).reshape(-1, 2)
    return (np.vstack([f2c, pf2c]), np.vstack([fn, pfn]), b2c[keep].copy(), bn[keep].copy())


@pytest.mark.parametrize("dim,N,axes", [(2, 2, (0, 1)), (2, 3, (0,)), (3, 1, (0, 1, 2)),
                                        (3, 2, (0, 2))])
def test_oracle_periodic_equals_quadrature(ip, oracle, dim, N, axes):
    base = ip.Mesh.square(3, jitter=0.1, seed=5) if dim == 2 else ip.Mesh.cube(2, 0.1, 5)
    faces = periodic_op2_faces(base, axes)
    coords, c2n = base.arrays()
    m = ip.Mesh.from_op2(coords, c2n, *faces)  # the library accepts the periodic pairing
    assert m.n_bfaces == len(faces[2])
    o = oracle.Oracle(coords, c2n, N, faces=faces)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    xyz = o.node_coords()
    periods = []
    for ax in axes:
        t = np.zeros(dim)
        t[ax] = 1.0
        periods.append(t)
    A_q = assemble(coords, c2n, xyz, N, periods=periods)
    A_o = o.dense()
    assert np.abs(A_o - A_q).max() < 1e-12 * np.abs(A_q).max()


@pytest.mark.parametrize("name,N", [("cube-20", 1), ("vortex_square_small", 2)])
def test_oracle_on_periodic_reference_mesh(ip, oracle, name, N):
    m = ip.Mesh.load_hdf5(path(name))
    o = oracle.Oracle(*m.arrays(), N, faces=m.faces())
    one = np.ones((o.n_cells, o.Np))
    assert np.abs(o.mult(one)).max() < 1e-11 * np.abs(o.mult(np.random.default_rng(0).random(one.shape))).max()
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal((2, o.n_cells, o.Np))
    a, b = np.sum(v * o.mult(u)), np.sum(u * o.mult(v))
    assert abs(a - b) < 1e-12 * np.sqrt(np.sum(u * o.mult(u)) * np.sum(v * o.mult(v)))
