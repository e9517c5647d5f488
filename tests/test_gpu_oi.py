"""Over-integrated variable-coefficient operator on the GPU (ins_poisson_set_factor_oi /
ins_poisson_set_factor_surf_oi), through the C ABI, against the oracle's restatement with the
same rules: FP64 apply 1e-12 (1e-11 at N = 4), FP32 apply 2e-5, the diagonal against the
oracle's assembled matrix, and p-multigrid solves (FP64 and mixed precision) whose true
residual under the oracle's operator is below the tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FV2 = [(0, 1), (1, 2), (2, 0)]
FV3 = [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _case(ip, oracle, dim, N, n=None, seed=3):
    mesh = (ip.Mesh.square(n or 4, jitter=0.1, seed=seed) if dim == 2
            else ip.Mesh.cube(n or 2, 0.1, seed))
    coords, c2n = mesh.arrays()
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(coords, c2n, N)
    rng = np.random.default_rng(seed + N)
    bt = (rng.random(s.n_bfaces) < 0.6).astype(np.int32)
    s.set_bc_types(bt)
    o.set_bc_types(bt)
    vx, vw, sb, sw = rule = ip.oi_rule(dim, N)
    # a density-like coefficient, sampled at the nodes and at the points of each cell
    kn = 0.5 + rng.random((s.n_cells, s.Np))
    kv = 0.5 + rng.random((s.n_cells, len(vw)))
    ks = 0.5 + rng.random((s.n_cells, (dim + 1) * len(sw)))
    s.set_factor(kn)
    o.set_factor(kn)
    s.set_factor_oi(kv)
    s.set_factor_surf_oi(ks)
    return s, o, rule, kv, ks, rng


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_oi_apply_matches_oracle(ip, oracle, dim, N):
    s, o, rule, kv, ks, rng = _case(ip, oracle, dim, N)
    x = rng.standard_normal((s.n_cells, s.Np))
    ref = o.mult_oi(rule, kv, ks, x)
    assert rel(s.mult(x), ref) < (1e-12 if N < 4 else 1e-11)
    assert rel(s.mult_fp32(x), ref) < 2e-5
    # clearing one part turns over-integration off: the nodal operator again
    s.set_factor_surf_oi(None)
    assert rel(s.mult(x), o.mult(x).reshape(s.n_cells, s.Np)) < (1e-12 if N < 4 else 1e-11)


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_oi_diag_matches_oracle(ip, oracle, dim, N):
    s, o, rule, kv, ks, rng = _case(ip, oracle, dim, N, n=3 if dim == 2 else 2)
    n = s.n_cells * s.Np
    d = np.zeros(n)
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        d[j] = o.mult_oi(rule, kv, ks, e).ravel()[j]
    assert rel(s.diag().ravel(), d) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2), (3, 4)])
def test_oi_pmg_solve(ip, oracle, dim, N, prec):
    s, o, rule, kv, ks, rng = _case(ip, oracle, dim, N, n=8 if dim == 2 else 3)
    s.set_solver(ip.SOLVER_PMULTIGRID)
    s.set_pmg(list(range(N, 0, -1)), coarse_iters=30)
    s.set_precision(ip.PRECISION_FP64 if prec == "fp64" else ip.PRECISION_MIXED)
    s.set_tol_and_iter(1e-10, 0.0, 300)
    b = rng.standard_normal((s.n_cells, s.Np))
    x = np.zeros_like(b)
    assert s.solve(b, x), s.last_stats
    r = b - o.mult_oi(rule, kv, ks, x)
    assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
