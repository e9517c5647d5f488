"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Tolerances: FP64 kernels within 1e-12 relative (max norm) of the oracle;
FP32 kernels within 2e-5; FP64 solves within 1e-10 relative L2 of the oracle's
FP64 solve; mixed-precision solves within 1e-8 after refinement.
"""
import numpy as np
import pytest

from helpers import bc_values, density_field, manufactured_sin, poly_solution

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def tol64(N, base=1e-12):
    """FP64 parity bound; the oracle's order-4 monomial basis carries ~5e-13 of its own error."""
    return base if N < 4 else 10 * base


def make_pair(ip, oracle, dim, N, n=None, seed=3, bc="dirichlet", factor=False):
    mesh = ip.Mesh.square(n or 6, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(n or 3, 0.1, seed)
    coords, c2n = mesh.arrays()
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(coords, c2n, N)
    nb = s.n_bfaces
    if bc == "dirichlet":
        bt = np.zeros(nb, dtype=np.int32)
    elif bc == "neumann":
        bt = np.ones(nb, dtype=np.int32)
    else:
        bt = np.zeros(nb, dtype=np.int32)
        bt[1::3] = 1
    s.set_bc_types(bt)
    o.set_bc_types(bt)
    k = None
    if factor:
        xyz = o.node_coords()
        h = 1.0 / (n or (6 if dim == 2 else 3))
        k = 1.0 / density_field(xyz, h, rho_in=10.0)
        s.set_factor(k)
        o.set_factor(k)
    return mesh, s, o, k


CASES = [(2, N) for N in (1, 2, 3, 4)] + [(3, N) for N in (1, 2, 3, 4)]


@pytest.mark.parametrize("dim,N", CASES)
@pytest.mark.parametrize("bc", ["dirichlet", "mixed"])
def test_apply_fp64_matches_oracle(ip, oracle, dim, N, bc):
    _, s, o, _ = make_pair(ip, oracle, dim, N, bc=bc)
    rng = np.random.default_rng(10 * dim + N)
    x = rng.standard_normal((s.n_cells, s.Np))
    assert rel(s.mult(x), o.mult(x)) < tol64(N)


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2), (3, 3), (3, 4)])
def test_apply_variable_factor_matches_oracle(ip, oracle, dim, N):
    _, s, o, _ = make_pair(ip, oracle, dim, N, bc="mixed", factor=True)
    x = np.random.default_rng(7).standard_normal((s.n_cells, s.Np))
    assert rel(s.mult(x), o.mult(x)) < tol64(N)
    assert rel(s.diag(), o.diag(consistent=True)) < tol64(N)


@pytest.mark.parametrize("dim,N", CASES)
def test_apply_fp32_matches_oracle(ip, oracle, dim, N):
    _, s, o, _ = make_pair(ip, oracle, dim, N, bc="mixed")
    x = np.random.default_rng(5).standard_normal((s.n_cells, s.Np))
    assert rel(s.mult_fp32(x), o.mult(x)) < 2e-5


@pytest.mark.parametrize("dim,N", CASES)
def test_diag_bc_mass_match_oracle(ip, oracle, dim, N):
    mesh, s, o, _ = make_pair(ip, oracle, dim, N, bc="mixed")
    assert rel(s.diag(), o.diag(consistent=False)) < tol64(N)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((s.n_cells, s.Np))
    assert rel(s.mass(x), o.mass(x)) < tol64(N, 1e-13)
    bc = rng.standard_normal((s.n_bfaces, s.Nfp))
    s.set_bcs(bc)
    assert rel(s.apply_bc(x), o.apply_bc(bc, x)) < tol64(N)
    xyz = s.node_coords()
    assert rel(xyz, o.node_coords()) < 1e-14


def test_polynomial_consistency_gpu(ip, oracle):
    """A u = M f + BC terms for u in P_N, through the device kernels."""
    for dim, N in ((2, 3), (3, 3)):
        _, s, o, _ = make_pair(ip, oracle, dim, N)
        xyz = s.node_coords()
        u, f = poly_solution(xyz, N)
        R = ip.ref_element(dim, N)
        _, _, b2c, bn = s.mesh.faces()
        s.set_bcs(bc_values(xyz, b2c, bn, R["fmask"], u))
        rhs = s.apply_bc(s.mass(f))
        assert rel(s.mult(u), rhs) < 1e-10


def test_config0_jacobi_solve(ip, oracle):
    """2D 32x32x2 jittered triangles, N=3, Dirichlet, Jacobi-PCG (the CPU-runnable workload)."""
    G = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "oracle_golden.npz"), allow_pickle=False)
    mesh = ip.Mesh.square(32, jitter=0.1, seed=int(G["cfg0_seed"]))
    s = ip.PoissonSolver(mesh, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_JACOBI)
    s.set_tol_and_iter(1e-12, 0.0, 20000)
    o = oracle.Oracle(*mesh.arrays(), 3)
    xyz = s.node_coords()
    u, f = manufactured_sin(xyz)
    b = s.mass(f)
    x = np.zeros((s.n_cells, s.Np))
    assert s.solve(b, x)
    assert s.last_stats["rel_residual"] <= 1e-12
    xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-12)
    assert rc == 0
    assert relL2(x, xo) < 1e-10
    err = o.mass_norm(x - u)
    assert abs(err - float(G["cfg0_l2err"])) < 0.005 * float(G["cfg0_l2err"])


# coarse_iters 20 and 21: the coarse CG brings x up to date every second iteration, so an odd
# count ends on a single update
@pytest.mark.parametrize("dim,N,orders,coarse_iters", [(3, 3, [3, 2, 1], 20), (2, 4, [4, 2, 1], 20),
                                                       (3, 2, [2, 1], 20), (3, 3, [3, 2, 1], 21)])
def test_pmg_solve_matches_oracle(ip, oracle, dim, N, orders, coarse_iters):
    mesh = ip.Mesh.cube(4, 0.1, 8) if dim == 3 else ip.Mesh.square(12, jitter=0.1, seed=8)
    s = ip.PoissonSolver(mesh, N)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_PMULTIGRID)
    s.set_pmg(orders, coarse_iters=coarse_iters)
    s.set_tol_and_iter(1e-12, 0.0, 300)
    o = oracle.Oracle(*mesh.arrays(), N)
    u, f = manufactured_sin(s.node_coords())
    b = s.mass(f)
    x = np.zeros((s.n_cells, s.Np))
    assert s.solve(b, x)
    it_gpu = s.last_stats["iterations"]
    xo, rc, it_o, _, lmax = o.solve_pmg(b, orders, coarse_iters=coarse_iters, rtol=1e-12, max_iter=300)
    assert rc == 0
    assert relL2(x, xo) < 1e-10
    assert abs(it_gpu - it_o) <= 2
    for l in range(len(orders) - 1):
        assert abs(s.lambda_max(l) - lmax[l]) < 1e-8 * lmax[l]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_variable_coefficient_neumann(ip, oracle, precision):
    """Config-2 form at small size: rho contrast 1000 sphere, pure Neumann, mean-zero."""
    n, N = 4, 3
    mesh = ip.Mesh.cube(n, 0.1, 21)
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(*mesh.arrays(), N)
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    o.set_bc_types(np.ones(o.n_bfaces, dtype=np.int32))
    k = 1.0 / density_field(s.node_coords(), 1.0 / n)
    s.set_factor(k)
    o.set_factor(k)
    s.set_nullspace(True)
    s.set_pmg([3, 2, 1], coarse_iters=30)
    if precision == "mixed":
        s.set_precision(ip.PRECISION_MIXED)
    s.set_tol_and_iter(1e-10, 0.0, 400)
    rng = np.random.default_rng(4)
    f = rng.standard_normal((s.n_cells, s.Np))
    b = s.mass(f)
    b -= b.mean()
    x = np.zeros_like(b)
    assert s.solve(b, x)
    assert abs(x.mean()) < 1e-12 * np.abs(x).max()
    xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-12, max_iter=50000, nullspace=True)
    assert rc == 0
    assert relL2(x, xo) < (1e-8 if precision == "mixed" else 1e-9)


def test_device_soa_layout_matches_host(ip, oracle):
    torch = pytest.importorskip("torch")
    _, s, o, _ = make_pair(ip, oracle, 3, 3, bc="mixed")
    x = np.random.default_rng(2).standard_normal((s.n_cells, s.Np))
    soa = np.zeros((s.Np, s.stride))
    soa[:, :s.n_cells] = x.T
    dx = torch.from_numpy(soa).cuda()
    dy = torch.zeros_like(dx)
    torch.cuda.synchronize()
    s.mult((dx.data_ptr(), dy.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    y = dy.cpu().numpy()[:, :s.n_cells].T
    assert rel(y, o.mult(x)) < 1e-12


def test_not_converged_and_breakdown(ip, oracle):
    _, s, _, _ = make_pair(ip, oracle, 2, 2)
    s.set_solver(ip.SOLVER_JACOBI)
    s.set_tol_and_iter(1e-12, 0.0, 1)
    b = np.random.default_rng(1).standard_normal((s.n_cells, s.Np))
    x = np.zeros_like(b)
    assert s.solve(b, x) is False
    assert s.last_stats["iterations"] == 1
    b[3, 1] = np.nan
    with pytest.raises(ip.InsPoissonError) as e:
        s.solve(b, np.zeros_like(b))
    assert e.value.code == ip.ERR_BREAKDOWN


def test_solve_is_deterministic(ip, oracle):
    _, s, _, _ = make_pair(ip, oracle, 3, 3, n=3)
    s.set_pmg([3, 2, 1])
    b = np.random.default_rng(9).standard_normal((s.n_cells, s.Np))
    x1, x2 = np.zeros_like(b), np.zeros_like(b)
    s.solve(b, x1)
    s.solve(b, x2)
    assert np.array_equal(x1, x2)


def test_profiling_timers(ip, oracle):
    _, s, _, _ = make_pair(ip, oracle, 3, 3, n=4)
    s.set_pmg([3, 2, 1])
    s.set_profiling(True)
    b = np.random.default_rng(9).standard_normal((s.n_cells, s.Np))
    s.solve(b, np.zeros_like(b))
    t = s.take_timing()
    assert t["apply_launches"] >= s.last_stats["fine_applies"] > 0
    assert t["apply_ms"] > 0 and t["trace_ms"] > 0
