"""GPU parity of the viscous (Helmholtz) operator and its solvers against the FP64 oracle.

Operator A u + J M (mm o u) (per-node factor) or A u + mm J M u (constant): apply within 1e-12
(FP64) and 2e-5 (FP32) relative, diagonal within 1e-12; solves with the inverse-mass
preconditioner (scalar or per-node factor), Jacobi, block-Jacobi and p-multigrid within 1e-10
relative L2 of the oracle's inverse-mass solve.
"""
import numpy as np
import pytest

from helpers import density_field, manufactured_sin

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def pair(ip, oracle, dim, N, n, mode, factor=False, seed=7):
    mesh = ip.Mesh.square(n, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(n, 0.1, seed)
    coords, c2n = mesh.arrays()
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(coords, c2n, N)
    bt = np.zeros(s.n_bfaces, dtype=np.int32)
    bt[1::4] = 1
    s.set_bc_types(bt)
    o.set_bc_types(bt)
    mm = None
    if mode == "const":
        s.set_mm_const(40.0)
        o.set_mm_const(40.0)
    else:
        mm = 20.0 + 10.0 * np.random.default_rng(seed).random((s.n_cells, s.Np))
        s.set_mm_factor(mm)
        o.set_mm_factor(mm)
    if factor:
        k = 1.0 + density_field(o.node_coords(), 1.0 / n, rho_in=5.0)
        s.set_factor(k)
        o.set_factor(k)
    return s, o, mm


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 3), (3, 1), (3, 2), (3, 4)])
@pytest.mark.parametrize("mode", ["const", "node"])
def test_viscous_apply_and_diag(ip, oracle, dim, N, mode):
    print(f"viscous apply dim={dim} N={N} {mode}", flush=True)
    s, o, _ = pair(ip, oracle, dim, N, 5 if dim == 2 else 2, mode, factor=(N == 2))
    x = np.random.default_rng(N).standard_normal((s.n_cells, s.Np))
    y = o.mult(x)
    tol = 1e-12 if N < 4 else 1e-11
    assert rel(s.mult(x), y) < tol
    assert rel(s.mult_fp32(x), y) < 2e-5
    assert rel(s.diag(), o.diag(consistent=True)) < tol


@pytest.mark.parametrize("mode", ["const", "node"])
def test_inv_mass_solve_matches_oracle(ip, oracle, mode):
    print(f"viscous inv-mass solve {mode}", flush=True)
    s, o, mm = pair(ip, oracle, 2, 3, 8, mode)
    b = np.random.default_rng(3).standard_normal((s.n_cells, s.Np))
    if mode == "const":
        xo, rc, it_o, _ = o.solve_inv_mass(b, rtol=1e-12, factor=1.0 / 40.0)
    else:
        xo, rc, it_o, _ = o.solve_inv_mass(b, rtol=1e-12, recp_factor=mm)
    assert rc == 0
    s.set_solver(ip.SOLVER_INV_MASS)
    s.set_tol_and_iter(1e-12, 0.0, 2000)
    if mode == "const":
        s.set_inv_mass_factor(1.0 / 40.0)
    else:
        s.set_inv_mass_recp_factor(mm)
    x = np.zeros_like(b)
    assert s.solve(b, x)
    assert relL2(x, xo) < 1e-10
    assert abs(s.last_stats["iterations"] - it_o) <= 2


@pytest.mark.parametrize("solver", ["jacobi", "block", "pmg"])
def test_viscous_solvers_agree(ip, oracle, solver):
    print(f"viscous solve {solver}", flush=True)
    s, o, mm = pair(ip, oracle, 3, 2, 3, "node", factor=True)
    b = np.random.default_rng(4).standard_normal((s.n_cells, s.Np))
    xo, rc, _, _ = o.solve_inv_mass(b, rtol=1e-12, recp_factor=mm)
    assert rc == 0
    s.set_solver({"jacobi": ip.SOLVER_JACOBI, "block": ip.SOLVER_BLOCK_JACOBI,
                  "pmg": ip.SOLVER_PMULTIGRID}[solver])
    if solver == "pmg":
        s.set_pmg([2, 1])
    s.set_tol_and_iter(1e-12, 0.0, 3000)
    x = np.zeros_like(b)
    assert s.solve(b, x)
    assert relL2(x, xo) < 1e-10


def test_helmholtz_discretisation_error_matches_oracle(ip, oracle):
    """-lap u + lam u = f, u = sin(pi x) sin(pi y), Dirichlet, N=3: same L2 error to 2 digits."""
    print("helmholtz manufactured N=3", flush=True)
    lam = 100.0
    mesh = ip.Mesh.square(16, jitter=0.1, seed=12)
    coords, c2n = mesh.arrays()
    s = ip.PoissonSolver(mesh, 3)
    o = oracle.Oracle(coords, c2n, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_mm_const(lam)
    o.set_mm_const(lam)
    xyz = s.node_coords()
    u, f = manufactured_sin(xyz)
    b = s.mass(f + lam * u)
    s.set_solver(ip.SOLVER_INV_MASS)
    s.set_inv_mass_factor(1.0 / lam)
    s.set_tol_and_iter(1e-12, 0.0, 2000)
    x = np.zeros_like(b)
    assert s.solve(b, x)
    xo, rc, _, _ = o.solve_inv_mass(b, rtol=1e-12, factor=1.0 / lam)
    assert rc == 0
    e_gpu, e_ora = o.mass_norm(x - u), o.mass_norm(xo - u)
    assert abs(e_gpu - e_ora) <= 0.005 * e_ora
    assert relL2(x, xo) < 1e-10
