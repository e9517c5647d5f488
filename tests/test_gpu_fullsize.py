"""The configurations at their full sizes, through size-independent properties (the oracle cannot
run these sizes in seconds; the bench compares every form with it at oracle-feasible sizes):

* 3D cube 64^3 x 6 (1,572,864 tets), N = 3, FP64: the operator is symmetric (x.Ay = y.Ax), linear,
  positive on Dirichlet and annihilates constants under pure Neumann, to rounding; the p-multigrid
  solve of the headline configuration reaches 1e-10 in the true FP64 residual.
* 3D cube 48^3 x 6 (663,552 tets), N = 4, rho 1000 sphere, pure Neumann: the mixed-precision solve
  on the matrix cores with the AMG coarse solve reaches 1e-8 in the true FP64 residual, mean zero.
"""
import numpy as np
import pytest

from helpers import density_field, manufactured_sin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube64(ip):
    mesh = ip.Mesh.cube(64, 0.1, 20261202)
    mesh.renumber(ip.RENUMBER_SHAPE_CLASS_WINDOWED)
    return mesh


def test_cube64_operator_properties(ip, cube64):
    s = ip.PoissonSolver(cube64, 3)
    assert s.n_cells == 1572864
    bt = np.zeros(s.n_bfaces, dtype=np.int32)
    bt[::3] = 1
    s.set_bc_types(bt)
    s.set_solver(ip.SOLVER_JACOBI)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((s.n_cells, s.Np))
    y = rng.standard_normal((s.n_cells, s.Np))
    Ax, Ay = s.mult(x), s.mult(y)
    xAy, yAx = float(np.vdot(x, Ay)), float(np.vdot(y, Ax))
    assert abs(xAy - yAx) <= 1e-11 * abs(xAy)
    assert float(np.vdot(x, Ax)) > 0
    Axy = s.mult(x + 2 * y)
    assert np.abs(Axy - (Ax + 2 * Ay)).max() <= 1e-12 * np.abs(Axy).max()
    # pure Neumann: constants are in the null space
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    A1 = s.mult(np.ones((s.n_cells, s.Np)))
    assert np.abs(A1).max() <= 1e-11 * np.abs(s.mult(x)).max()


def test_cube64_headline_solve(ip, cube64):
    s = ip.PoissonSolver(cube64, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_PMULTIGRID)
    s.set_pmg([3, 2, 1], pre_it=[1] * 3, post_it=[1] * 3, cheb_orders=[3] * 3, coarse_iters=150)
    s.set_tol_and_iter(1e-10, 0.0, 500)
    u, f = manufactured_sin(s.node_coords())
    b = s.mass(f)
    x = np.zeros_like(b)
    assert s.solve(b, x), s.last_stats
    r = b - s.mult(x)
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
    # discretisation error of the manufactured solution: O(h^4) at h = 1/64
    e = x - u
    assert np.sqrt(np.vdot(e, s.mass(e)) / np.vdot(u, s.mass(u))) < 1e-6


def test_cube48_precision_form_solve(ip):
    mesh = ip.Mesh.cube(48, 0.1, 20261206)
    mesh.renumber(ip.RENUMBER_SHAPE_CLASS_WINDOWED)
    s = ip.PoissonSolver(mesh, 4)
    assert s.n_cells == 663552
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    s.set_nullspace(True)
    k = 1.0 / density_field(s.node_coords(), 1.0 / 48)
    s.set_factor(k)
    s.set_solver(ip.SOLVER_PMULTIGRID)
    s.set_pmg([4, 3, 2, 1], pre_it=[1] * 4, post_it=[1] * 4, cheb_orders=[3] * 4, coarse_iters=150)
    s.set_coarse(ip.COARSE_AMG)
    s.set_precision(ip.PRECISION_MIXED)
    s.set_tol_and_iter(1e-8, 0.0, 1000)
    b = s.mass(np.random.default_rng(20261206).standard_normal((s.n_cells, s.Np)))
    b -= b.mean()
    x = np.zeros_like(b)
    assert s.solve(b, x), s.last_stats
    r = b - s.mult(x)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b)
    assert abs(x.mean()) <= 1e-10 * np.abs(x).max()
