"""The C-ABI library loads and exports what include/ins_poisson.h declares;
host-side pieces (reference element, mesh topology) agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ins_poisson.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ins_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol(ip):
    lib = ctypes.CDLL(ip.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(ip.EXPORTED) == names


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_product_reference_element_matches_oracle(ip, oracle, dim, N):
    a, b = oracle.ref_element(dim, N), ip.ref_element(dim, N)
    assert np.array_equal(a["fmask"], b["fmask"])
    for k in ("rst", "D", "M", "Mface"):
        scale = max(1.0, np.abs(a[k]).max())
        assert np.abs(a[k] - b[k]).max() < 1e-12 * scale, k
    for nc in range(1, N):
        assert np.abs(oracle.interp(dim, N, nc) - ip.ref_interp(dim, N, nc)).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_order1_tables_are_the_p1_closed_forms(ip, oracle, dim):
    """The order-1 kernels compile in the P1 basis on the reference simplex (vertices
    (-1,..,-1), e_k): phi_0 = -(1 + sum xi)/2, phi_{k+1} = (1 + xi_k)/2, the mass column sums
    |T|/(dim+1) and the face mass |F|/(dim(dim+1)) (1 + delta_ij) in face coordinates. Both
    the library's tables and the oracle's must be exactly these (a sign, index or factor slip in
    either fails here; the library also checks its tables when it builds them)."""
    np_ = dim + 1
    dphi = np.zeros((dim, np_))
    dphi[:, 0] = -0.5
    for r in range(dim):
        dphi[r, r + 1] = 0.5
    vol = 2.0 if dim == 2 else 4.0 / 3.0       # reference triangle / tetrahedron
    face = 2.0  # the face in face coordinates: the segment [-1, 1] or the reference triangle
    nfp = dim
    mface = face / (nfp * (nfp + 1)) * (np.ones((nfp, nfp)) + np.eye(nfp))
    for R in (ip.ref_element(dim, 1), oracle.ref_element(dim, 1)):
        D = np.asarray(R["D"]).reshape(dim, np_, np_)
        for r in range(dim):
            assert np.abs(D[r] - dphi[r][None, :]).max() < 1e-13  # every row: constant gradient
        M = np.asarray(R["M"]).reshape(np_, np_)
        assert np.abs(M.sum(axis=0) - vol / np_).max() < 1e-13
        Mf = np.asarray(R["Mface"]).reshape(dim + 1, nfp, nfp)
        assert np.abs(Mf - mface[None]).max() < 1e-13
    # the closed forms themselves: gradients of the P1 basis at the vertices reproduce the
    # vertex coordinates' differences, and the face mass integrates 1 to the face size
    verts = np.vstack([-np.ones(dim), -np.ones((dim, dim)) + 2 * np.eye(dim)])
    for r in range(dim):
        assert np.allclose(dphi[r] @ verts[:, r], 1.0)
    assert np.isclose(mface.sum(), face)


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_mesh_topology_matches_oracle(ip, oracle, kind):
    m = ip.Mesh.square(9, 7, jitter=0.1, seed=3) if kind == "square" else ip.Mesh.cube(4, 0.1, 3)
    coords, c2n = m.arrays()
    o = oracle.Oracle(coords, c2n, 2)
    f2c, fnum, b2c, bnum = m.faces()
    of2c, ofnum = o.faces()
    ob2c, obnum = o.bfaces()
    assert np.array_equal(f2c, of2c) and np.array_equal(fnum, ofnum)
    assert np.array_equal(b2c, ob2c) and np.array_equal(bnum, obnum)
    nf = m.dim + 1
    assert 2 * m.n_faces + m.n_bfaces == nf * m.n_cells


def test_benchmark_mesh_sizes(ip):
    m = ip.Mesh.square(32, jitter=0.1, seed=1)
    assert (m.n_cells, m.n_faces, m.n_bfaces) == (2048, 3008, 128)
    m = ip.Mesh.cube(8, jitter=0.1, seed=1)
    assert m.n_cells == 6 * 512 and m.n_bfaces == 6 * 2 * 64


def test_jitter_is_seeded_and_bounded(ip):
    a, _ = ip.Mesh.cube(5, 0.1, 7).arrays()
    b, _ = ip.Mesh.cube(5, 0.1, 7).arrays()
    c, _ = ip.Mesh.cube(5, 0.1, 8).arrays()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    g = np.stack(np.meshgrid(*[np.arange(6) / 5] * 3, indexing="ij"), -1)[..., ::-1]
    g = g.transpose(2, 1, 0, 3).reshape(-1, 3)[:, ::-1]
    assert np.abs(a - g).max() <= 0.1 / 5 + 1e-15


def test_invalid_arguments_are_reported(ip):
    with pytest.raises(ip.InsPoissonError) as e:
        ip.Mesh.square(0)
    assert e.value.code == ip.ERR_INVALID
    with pytest.raises(ip.InsPoissonError):
        ip.ref_element(2, 7)
    coords = np.array([[0.0, 0], [1, 0], [0, 1]])
    with pytest.raises(ip.InsPoissonError):
        ip.Mesh.from_arrays(coords, np.array([[0, 1, 5]]))


def test_user_mesh_orientation_fixed(ip):
    coords = np.array([[0.0, 0], [1, 0], [0, 1], [1, 1]])
    m = ip.Mesh.from_arrays(coords, np.array([[0, 2, 1], [1, 2, 3]]))
    _, c2n = m.arrays()
    for c in c2n:
        a, b, d = coords[c]
        assert (b[0] - a[0]) * (d[1] - a[1]) - (d[0] - a[0]) * (b[1] - a[1]) > 0


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 4), (3, 2), (3, 3)])
def test_mesh_node_coords_match_oracle(ip, oracle, dim, N):
    m = ip.Mesh.square(5, jitter=0.15, seed=4) if dim == 2 else ip.Mesh.cube(2, 0.15, 4)
    o = oracle.Oracle(*m.arrays(), N)
    assert np.abs(m.node_coords(N) - o.node_coords()).max() < 1e-14


def test_batch_host_side_entry_points(ip):
    c = ip.batch_config()
    assert (c.precision, c.cheb_order, c.pre_it, c.post_it, c.max_iter_eig, c.coarse_iters,
            c.use_graph, c.coarse, c.tile) == (ip.PRECISION_MIXED, 3, 1, 1, 15, 40, 1, ip.COARSE_AMG, 0)
    assert abs(c.eig_safety - 1.1) < 1e-15 and abs(c.cheb_min_factor - 0.1) < 1e-15
    # arena bytes grow with cells, order, a coefficient field and FP64 V-cycle storage
    b = lambda *a, **k: ip.problem_bytes(*a, **k)
    assert b(2, 1000, 2) < b(2, 2000, 2) < b(2, 2000, 3)
    assert b(3, 1000, 3) < b(3, 1000, 3, has_factor=True)
    assert b(3, 1000, 3) < b(3, 1000, 3, precision=ip.PRECISION_FP64)
    # cells are padded to whole 256-cell tiles: only the per-bface table grows below a tile
    assert 0 <= b(2, 256, 2) - b(2, 1, 2) <= 256 * 3 * 4 + 256
    with pytest.raises(ip.InsPoissonError):
        b(4, 10, 2)
    with pytest.raises(ip.InsPoissonError):
        b(2, 10, 5)
    # invalid handles and configurations are rejected before any device work
    h = ctypes.c_void_p()
    assert ip.lib.ins_batch_create(5, None, ctypes.byref(h)) == ip.ERR_INVALID
    bad = ip.batch_config(cheb_order=0)
    assert ip.lib.ins_batch_create(2, ctypes.byref(bad), ctypes.byref(h)) == ip.ERR_INVALID
    bad = ip.batch_config(precision=7)
    assert ip.lib.ins_server_create(1 << 30, ctypes.byref(bad), None, ctypes.byref(h)) == ip.ERR_INVALID
    n = ctypes.c_int()
    assert ip.lib.ins_batch_iterate(None, 1, ctypes.byref(n)) == ip.ERR_INVALID
    assert ip.lib.ins_server_poll(None, 0, ctypes.byref(n)) == ip.ERR_INVALID
    with pytest.raises(TypeError):
        ip.batch_config(no_such_option=1)
    bad = ip.batch_config(tile=96)
    assert ip.lib.ins_batch_create(2, ctypes.byref(bad), ctypes.byref(h)) == ip.ERR_INVALID
    bad = ip.batch_config(coarse=5)
    assert ip.lib.ins_batch_create(2, ctypes.byref(bad), ctypes.byref(h)) == ip.ERR_INVALID
    assert ip.amg_bytes(3, 1000, 5000) > ip.amg_bytes(3, 500, 2500) > 0
