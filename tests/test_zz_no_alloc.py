"""The allocation-free apply (ins_poisson_mult_noalloc) against ins_poisson_mult, through the
C ABI, on two calls in a row that share the per-GPU scratch: a 4^3 cube at order 3 in the
solver's own device SoA layout, then a smaller 3^3 cube at order 2 in a caller stride (staged
in the scratch). Results must equal the allocating apply bit for bit (same kernels, same data).
Also: the errors for a handle not yet initialised, a host layout and a mesh whose traces do not
fit the scratch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _strided(torch, x, stride, fill):
    """cells x rows host array -> device SoA (rows x stride), padding set to `fill`."""
    a = np.full((x.shape[1], stride), fill)
    a[:, :x.shape[0]] = x.T
    return torch.from_numpy(a).cuda()


def _solver(ip, n, N, seed):
    mesh = ip.Mesh.cube(n, jitter=0.1, seed=seed)
    s = ip.PoissonSolver(mesh, N)
    bt = np.zeros(s.n_bfaces, dtype=np.int32)
    bt[::3] = 1
    s.set_bc_types(bt)
    rng = np.random.default_rng(seed)
    s.set_factor(1.0 + rng.random((s.n_cells, s.Np)))
    s.set_pmg(list(range(N, 0, -1)), coarse_iters=10)
    return s, mesh


def test_noalloc_matches_mult_two_calls(ip):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    # call 1: 4^3 cube, order 3, the solver's own SoA stride (kernels run on the caller's vectors)
    s1, m1 = _solver(ip, 4, 3, 1)
    s1.init()
    x1 = rng.standard_normal((s1.n_cells, s1.Np))
    stride1 = s1.stride
    dx = _strided(torch, x1, stride1, 0.0)
    dref = torch.zeros_like(dx)
    dout = torch.full_like(dx, 7.0)
    torch.cuda.synchronize()
    s1.mult((dx.data_ptr(), dref.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    rc = ip.lib.ins_poisson_mult_noalloc(s1._h, dx.data_ptr(), dout.data_ptr(), ip.LAYOUT_DEVICE_SOA)
    assert rc == ip.OK, ip.lib.ins_poisson_last_error().decode()
    ref = dref.cpu().numpy()[:, :s1.n_cells]
    out = dout.cpu().numpy()[:, :s1.n_cells]
    assert np.abs(ref).max() > 0
    assert np.array_equal(out, ref)

    # call 2: smaller problem (3^3 cube, order 2) in a caller stride E + 37: staged in the scratch
    s2, m2 = _solver(ip, 3, 2, 2)
    stride2 = s2.n_cells + 37
    s2.set_dat_stride(stride2)
    s2.init()
    x2 = rng.standard_normal((s2.n_cells, s2.Np))
    dx2 = _strided(torch, x2, stride2, -3.0)
    dref2 = _strided(torch, np.zeros_like(x2), stride2, 5.0)
    dout2 = _strided(torch, np.zeros_like(x2), stride2, 5.0)
    torch.cuda.synchronize()
    s2.mult((dx2.data_ptr(), dref2.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    s2.mult_noalloc((dx2.data_ptr(), dout2.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    out2 = dout2.cpu().numpy()
    ref2 = dref2.cpu().numpy()
    assert np.all(out2[:, s2.n_cells:] == 5.0)  # the caller's padding is left as it was
    assert np.array_equal(out2, ref2)

    # call 3: the first problem again, after the smaller one overwrote the shared scratch
    dout.fill_(-1.0)
    torch.cuda.synchronize()
    s1.mult_noalloc((dx.data_ptr(), dout.data_ptr()))
    assert np.array_equal(dout.cpu().numpy()[:, :s1.n_cells], ref)


def test_noalloc_errors(ip):
    torch = pytest.importorskip("torch")
    s, _ = _solver(ip, 2, 2, 3)
    d = torch.zeros((s.Np, s.stride), dtype=torch.float64, device="cuda")
    rc = ip.lib.ins_poisson_mult_noalloc(s._h, d.data_ptr(), d.data_ptr(), ip.LAYOUT_DEVICE_SOA)
    assert rc == ip.ERR_INVALID and "init" in ip.lib.ins_poisson_last_error().decode()
    s.init()
    x = np.zeros((s.n_cells, s.Np))
    y = np.zeros_like(x)
    rc = ip.lib.ins_poisson_mult_noalloc(s._h, x.ctypes.data, y.ctypes.data, ip.LAYOUT_HOST_AOS)
    assert rc == ip.ERR_INVALID
    # 33^3 x 6 cells at order 3 need 40 x stride doubles of traces: beyond 64 MiB
    big = ip.PoissonSolver(ip.Mesh.cube(33, jitter=0.0, seed=0), 3)
    big.set_bc_types(np.zeros(big.n_bfaces, dtype=np.int32))
    big.set_solver(ip.SOLVER_JACOBI)
    big.init()
    db = torch.zeros((big.Np, big.stride), dtype=torch.float64, device="cuda")
    dy = torch.zeros_like(db)
    rc = ip.lib.ins_poisson_mult_noalloc(big._h, db.data_ptr(), dy.data_ptr(), ip.LAYOUT_DEVICE_SOA)
    assert rc == ip.ERR_INVALID
    assert "too large" in ip.lib.ins_poisson_last_error().decode()
