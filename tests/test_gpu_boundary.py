"""The drop-in boundary with an OP2 caller's conventions, through the C ABI:

* device SoA dats with the caller's padded stride (n_cells + 37 here), as OP2 GPU
  builds lay them out (-DDG_OP2_SOA, Makefile:40-41); padding entries untouched;
* the caller's face and bface order (ins_mesh_create_op2), which indexes bc_types
  and bc_data (src/solvers/2d/ins_solver.cpp:205,324);
* a negatively oriented cell given to ins_mesh_create (re-oriented, solved as the
  oracle solves it).
All results against the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import bc_values, manufactured_sin

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def _strided(torch, x, stride, fill=-3.0):
    """cells x rows host array -> device SoA (rows x stride), padding set to `fill`."""
    a = np.full((x.shape[1], stride), fill)
    a[:, :x.shape[0]] = x.T
    return torch.from_numpy(a).cuda()


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2), (3, 3)])
def test_caller_stride(ip, oracle, dim, N):
    torch = pytest.importorskip("torch")
    mesh = ip.Mesh.square(8, jitter=0.1, seed=6) if dim == 2 else ip.Mesh.cube(3, 0.1, 6)
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(*mesh.arrays(), N)
    bt = np.zeros(s.n_bfaces, dtype=np.int32)
    bt[::4] = 1
    s.set_bc_types(bt)
    o.set_bc_types(bt)
    s.set_pmg(list(range(N, 0, -1)), coarse_iters=20)
    s.set_tol_and_iter(1e-12, 0.0, 300)
    stride = s.n_cells + 37
    s.set_dat_stride(stride)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((s.n_cells, s.Np))
    dx = _strided(torch, x, stride)
    dy = _strided(torch, np.zeros_like(x), stride, fill=5.0)
    torch.cuda.synchronize()
    s.mult((dx.data_ptr(), dy.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    y = dy.cpu().numpy()
    assert np.all(y[:, s.n_cells:] == 5.0)
    assert rel(y[:, :s.n_cells].T, o.mult(x)) < 1e-12
    # solve: rhs and ans in the caller's stride, ans IN/OUT
    u, f = manufactured_sin(s.node_coords())
    b = o.mass(f)
    db = _strided(torch, b, stride)
    dx0 = _strided(torch, np.zeros_like(b), stride, fill=9.0)
    torch.cuda.synchronize()
    assert s.solve(db.data_ptr(), dx0.data_ptr(), layout=ip.LAYOUT_DEVICE_SOA)
    xs = dx0.cpu().numpy()
    assert np.all(xs[:, s.n_cells:] == 9.0)
    xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-12)
    assert rc == 0
    assert relL2(xs[:, :s.n_cells].T, xo) < 1e-10
    # face dats share the stride: the caller's pressure rhs
    NF = (dim + 1) * s.Nfp
    g0 = rng.standard_normal((s.n_cells, NF))
    g1 = rng.standard_normal((s.n_cells, NF))
    div = rng.standard_normal((s.n_cells, s.Np))
    ref_rhs, ref_g1 = s.pressure_rhs(1.5, -0.5, 0.01, g0, g1, div)
    d0, d1, dd = _strided(torch, g0, stride), _strided(torch, g1, stride), _strided(torch, div, stride)
    torch.cuda.synchronize()
    ip._check(ip.lib.ins_poisson_pressure_rhs(s._h, 1.5, -0.5, 0.01, d0.data_ptr(), d1.data_ptr(),
                                              dd.data_ptr(), ip.LAYOUT_DEVICE_SOA))
    assert np.array_equal(dd.cpu().numpy()[:, :s.n_cells].T, ref_rhs)
    assert np.array_equal(d1.cpu().numpy()[:, :s.n_cells].T, ref_g1)
    assert np.all(dd.cpu().numpy()[:, s.n_cells:] == -3.0)


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_caller_bface_order(ip, oracle, dim, N):
    base = ip.Mesh.square(6, jitter=0.1, seed=7) if dim == 2 else ip.Mesh.cube(3, 0.1, 7)
    coords, c2n = base.arrays()
    f2c, fn, b2c, bn = base.faces()
    perm = np.random.default_rng(2).permutation(len(b2c))
    mesh = ip.Mesh.from_op2(coords, c2n, f2c[::-1], fn[::-1], b2c[perm], bn[perm])
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(coords, c2n, N)
    # boundary types and data given in the caller's bface order
    bt_caller = (np.arange(len(b2c)) % 3 == 0).astype(np.int32)
    s.set_bc_types(bt_caller)
    bt_lib = np.empty_like(bt_caller)
    bt_lib[perm] = bt_caller
    o.set_bc_types(bt_lib)
    xyz = s.node_coords()
    u, f = manufactured_sin(xyz)
    fmask = ip.ref_element(dim, N)["fmask"]
    bc_caller = bc_values(xyz, b2c[perm], bn[perm], fmask, u)
    s.set_bcs(bc_caller)
    bc_lib = np.empty_like(bc_caller)
    bc_lib[perm] = bc_caller
    b = o.mass(f)
    assert rel(s.apply_bc(b), o.apply_bc(bc_lib, b)) < 1e-12
    x = np.random.default_rng(1).standard_normal((s.n_cells, s.Np))
    assert rel(s.mult(x), o.mult(x)) < 1e-12
    s.set_solver(ip.SOLVER_JACOBI)
    s.set_tol_and_iter(1e-12, 0.0, 20000)
    ans = np.zeros_like(b)
    assert s.solve(b, ans)
    xo, rc, _, _ = o.solve_jacobi(o.apply_bc(bc_lib, b), rtol=1e-12)
    assert rc == 0 and relL2(ans, xo) < 1e-10


def test_flipped_cell_reoriented(ip, oracle):
    base = ip.Mesh.cube(3, 0.1, 9)
    coords, c2n = base.arrays()
    c2n = c2n.copy()
    for e in (0, 17, 40):
        c2n[e, [2, 3]] = c2n[e, [3, 2]]
    mesh = ip.Mesh.from_arrays(coords, c2n)
    s = ip.PoissonSolver(mesh, 3)
    o = oracle.Oracle(coords, c2n, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    assert rel(s.node_coords(), o.node_coords()) < 1e-14
    u, f = manufactured_sin(s.node_coords())
    b = s.mass(f)
    s.set_pmg([3, 2, 1], coarse_iters=20)
    s.set_tol_and_iter(1e-12, 0.0, 300)
    x = np.zeros_like(b)
    assert s.solve(b, x)
    xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-12)
    assert rc == 0 and relL2(x, xo) < 1e-10
