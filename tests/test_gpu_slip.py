"""The vector (velocity) viscous solve with slip walls on the GPU, through the C ABI
(ins_viscous_vec_*), against the FP64 oracle: apply, diagonal, boundary terms, block-Jacobi
blocks and solves with every ViscousSolver preconditioner; device SoA with a caller stride.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FV2 = [(0, 1), (1, 2), (2, 0)]
FV3 = [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def _case(ip, oracle, dim, N, n=None, seed=1, factor=True, mm="var"):
    mesh = (ip.Mesh.square(n or 4, jitter=0.1, seed=seed) if dim == 2
            else ip.Mesh.cube(n or 2, 0.1, seed + 1))
    coords, c2n = mesh.arrays()
    s = ip.ViscousVecSolver(mesh, N)
    o = oracle.Oracle(coords, c2n, N)
    b2c, bn = o.bfaces()
    _, _, mb2c, mbn = mesh.faces()
    assert np.array_equal(b2c, mb2c) and np.array_equal(bn, mbn)
    fv = FV2 if dim == 2 else FV3
    bct = np.zeros((dim, o.n_bfaces), dtype=np.int32)
    for b in range(o.n_bfaces):
        cen = coords[c2n[b2c[b]][list(fv[bn[b]])]].mean(axis=0)
        if abs(cen[0]) < 1e-9:
            bct[:, b] = 0
        elif abs(cen[0] - 1) < 1e-9 or abs(cen[1] - 1) < 1e-9:
            bct[:, b] = 2
        elif abs(cen[1]) < 1e-9:
            bct[:, b] = (2, 1) + ((0,) if dim == 3 else ())
        else:
            bct[:, b] = 1
    s.set_bc_types(list(bct))
    rng = np.random.default_rng(seed + 10 * N)
    if factor:
        k = 1 + rng.random((len(c2n), o.Np))
        s.set_factor(k)
        o.set_factor(k)
    if mm == "var":
        m = 50 + 10 * rng.random((len(c2n), o.Np))
        s.set_mm_factor(m)
        o.set_mm_factor(m)
    elif mm:
        s.set_mm_const(float(mm))
        o.set_mm_const(float(mm))
    return s, o, bct, rng


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_vec_mult_matches_oracle(ip, oracle, dim, N):
    s, o, bct, rng = _case(ip, oracle, dim, N)
    u = rng.standard_normal((dim, s.n_cells, s.Np))
    tol = 1e-11 if N == 4 else 1e-12
    assert rel(s.mult(u), o.mult_vec(bct, u)) < tol


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_vec_constant_factor_matches_oracle(ip, oracle, dim, N):
    """ViscousMatrix2D: constant mass factor, no variable coefficient."""
    s, o, bct, rng = _case(ip, oracle, dim, N, factor=False, mm=123.0)
    u = rng.standard_normal((dim, s.n_cells, s.Np))
    assert rel(s.mult(u), o.mult_vec(bct, u)) < 1e-12


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 4), (3, 1), (3, 3)])
def test_vec_diag_and_bc_match_oracle(ip, oracle, dim, N):
    s, o, bct, rng = _case(ip, oracle, dim, N)
    assert rel(s.diag(), o.diag_vec(bct, consistent=True)) < 1e-12
    g = rng.standard_normal((dim, s.n_bfaces, s.Nfp))
    s.set_bcs(list(g))
    f = rng.standard_normal((dim, s.n_cells, s.Np))
    assert rel(s.apply_bc(f), o.apply_bc_vec(bct, g, f)) < 1e-12


@pytest.mark.parametrize("pre", ["NONE", "INV_MASS", "FACTOR_INV_MASS", "RECP_FACTOR_DAT_INV_MASS",
                                 "JACOBI", "BLOCK_JACOBI"])
@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_vec_solve_matches_oracle(ip, oracle, dim, N, pre):
    s, o, bct, rng = _case(ip, oracle, dim, N, n=6 if dim == 2 else 3)
    g = rng.standard_normal((dim, s.n_bfaces, s.Nfp))
    f = rng.standard_normal((dim, s.n_cells, s.Np))
    s.set_bcs(list(g))
    recp = 0.5 + rng.random((s.n_cells, s.Np))
    s.set_preconditioner(getattr(ip, "VEC_PRE_" + pre))
    s.set_inv_mass_factor(0.02)
    s.set_inv_mass_recp_factor(recp)
    s.set_tol_and_iter(1e-12, 0.0, 5000)
    x = np.zeros((dim, s.n_cells, s.Np))
    assert s.solve(f, x), s.last_stats
    okind = {"NONE": 0, "INV_MASS": 1, "FACTOR_INV_MASS": 2, "RECP_FACTOR_DAT_INV_MASS": 3,
             "JACOBI": 4, "BLOCK_JACOBI": 4}[pre]
    xo, rc, it, rr = o.solve_vec(bct, f, bc=g, rtol=1e-12, precond=okind, factor=0.02, recp_factor=recp)
    assert rc == 0
    assert relL2(x, xo) < 1e-9
    if pre != "BLOCK_JACOBI":
        assert abs(s.last_stats["iterations"] - it) <= 2


def test_vec_block_jacobi_blocks(ip, oracle):
    s, o, bct, rng = _case(ip, oracle, 2, 2, n=3)
    s.set_preconditioner(ip.VEC_PRE_BLOCK_JACOBI)
    inv = s.inv_block_diag()
    A = o.dense_vec(bct)
    n, Np = s.n_cells * s.Np, s.Np
    for c in range(2):
        for e in range(s.n_cells):
            i0 = c * n + e * Np
            blk = A[i0:i0 + Np, i0:i0 + Np]
            assert rel(inv[c, e], np.linalg.inv(blk)) < 1e-10


def test_vec_device_soa_caller_stride(ip, oracle):
    torch = pytest.importorskip("torch")
    s, o, bct, rng = _case(ip, oracle, 3, 2)
    stride = s.n_cells + 37
    s.set_dat_stride(stride)
    u = rng.standard_normal((3, s.n_cells, s.Np))

    def dev(x, fill):
        a = np.full((s.Np, stride), fill)
        a[:, :s.n_cells] = x.T
        return torch.from_numpy(a).cuda()

    du = [dev(u[c], -3.0) for c in range(3)]
    dy = [dev(np.zeros_like(u[c]), 5.0) for c in range(3)]
    torch.cuda.synchronize()
    s.mult(([t.data_ptr() for t in du], [t.data_ptr() for t in dy]), layout=ip.LAYOUT_DEVICE_SOA)
    y = np.stack([t.cpu().numpy() for t in dy])
    assert np.all(y[:, :, s.n_cells:] == 5.0)
    assert rel(np.transpose(y[:, :, :s.n_cells], (0, 2, 1)), o.mult_vec(bct, u)) < 1e-12
    # solve in the caller's layout, ans IN/OUT
    g = rng.standard_normal((3, s.n_bfaces, s.Nfp))
    s.set_bcs(list(g))
    s.set_preconditioner(ip.VEC_PRE_JACOBI)
    s.set_tol_and_iter(1e-12, 0.0, 2000)
    f = rng.standard_normal((3, s.n_cells, s.Np))
    df = [dev(f[c], 0.0) for c in range(3)]
    dx = [dev(np.zeros_like(f[c]), 9.0) for c in range(3)]
    torch.cuda.synchronize()
    assert s.solve([t.data_ptr() for t in df], [t.data_ptr() for t in dx], layout=ip.LAYOUT_DEVICE_SOA)
    x = np.stack([t.cpu().numpy() for t in dx])
    assert np.all(x[:, :, s.n_cells:] == 9.0)
    xo, rc, _, _ = o.solve_vec(bct, f, bc=g, rtol=1e-12, precond=4)
    assert rc == 0 and relL2(np.transpose(x[:, :, :s.n_cells], (0, 2, 1)), xo) < 1e-9
