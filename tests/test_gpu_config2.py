"""The precision-form solve at oracle size: 3D, N = 4, rho = 1000 inside a sphere.

div(1/rho grad p) = f on a 5^3 x 6 Kuhn cube, rho = 1000 inside the sphere of
radius 0.25 at the centre and 1 outside, smoothed over a band of 2h
(ls_step.h:5-6), f Gaussian projected to zero mean, pure Neumann with the
mean-zero constraint (src/solvers/3d/mp_ins_solver.cpp:513-632 sets up this
solve). The solver is the one the precision form runs at full size: p-multigrid
4 -> 3 -> 2 -> 1 with Chebyshev(3), the vertex-space AMG coarse solve, and the
FP32 V-cycle whose order-4 element pass runs on the matrix cores inside the FP64
flexible CG.

Tolerances: mixed precision within 1e-8 relative L2 of the oracle's FP64
Jacobi-PCG solution (rtol 1e-12), the plain FP64 path within 1e-9.
"""
import numpy as np
import pytest

from helpers import density_field

pytestmark = pytest.mark.gpu

N_CUBE, ORDER, SEED = 5, 4, 22


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


@pytest.fixture(scope="module")
def case(ip, oracle):
    mesh = ip.Mesh.cube(N_CUBE, 0.1, SEED)
    o = oracle.Oracle(*mesh.arrays(), ORDER)
    o.set_bc_types(np.ones(o.n_bfaces, dtype=np.int32))
    k = 1.0 / density_field(o.node_coords(), 1.0 / N_CUBE, rho_in=1000.0)
    o.set_factor(k)
    f = np.random.default_rng(4).standard_normal((o.n_cells, o.Np))
    b = o.mass(f)
    b -= b.mean()
    xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-12, max_iter=200000, nullspace=True)
    assert rc == 0
    return mesh, k, b, xo


def solve(ip, mesh, k, b, precision, coarse, epass, orders=(4, 3, 2, 1)):
    s = ip.PoissonSolver(mesh, ORDER)
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    s.set_nullspace(True)
    s.set_factor(k)
    s.set_solver(ip.SOLVER_PMULTIGRID)
    nl = len(orders)
    s.set_pmg(list(orders), pre_it=[1] * nl, post_it=[1] * nl, cheb_orders=[3] * nl, coarse_iters=30)
    s.set_coarse(coarse)
    s.set_precision(precision)
    s.set_element_pass(epass)
    s.set_tol_and_iter(1e-10, 0.0, 1000)
    x = np.zeros_like(b)
    assert s.solve(b, x), s.last_stats
    assert abs(x.mean()) < 1e-12 * np.abs(x).max()
    return x, s.last_stats


def test_config2_mixed_matrix_cores_matches_oracle(ip, case):
    mesh, k, b, xo = case
    x_auto, st = solve(ip, mesh, k, b, ip.PRECISION_MIXED, ip.COARSE_AMG, ip.ELEMENT_PASS_AUTO)
    assert relL2(x_auto, xo) < 1e-8, (relL2(x_auto, xo), st)
    # AUTO takes the matrix cores at order 4: not the same bits as the vector FMA element pass;
    # forcing the matrix cores on the order 3 and 2 levels too, or the vector pass everywhere,
    # solves to the same tolerance
    x_mat, _ = solve(ip, mesh, k, b, ip.PRECISION_MIXED, ip.COARSE_AMG, ip.ELEMENT_PASS_MATRIX)
    x_vec, _ = solve(ip, mesh, k, b, ip.PRECISION_MIXED, ip.COARSE_AMG, ip.ELEMENT_PASS_VECTOR)
    assert not np.array_equal(x_auto, x_vec)
    assert relL2(x_mat, xo) < 1e-8
    assert relL2(x_vec, xo) < 1e-8
    # with order 4 the only level above order 1, AUTO and MATRIX run the same kernels: same bits
    x_a41, _ = solve(ip, mesh, k, b, ip.PRECISION_MIXED, ip.COARSE_AMG, ip.ELEMENT_PASS_AUTO, (4, 1))
    x_m41, _ = solve(ip, mesh, k, b, ip.PRECISION_MIXED, ip.COARSE_AMG, ip.ELEMENT_PASS_MATRIX, (4, 1))
    assert np.array_equal(x_a41, x_m41)
    assert relL2(x_a41, xo) < 1e-8


@pytest.mark.parametrize("coarse", ["amg", "jacobi_cg"])
def test_config2_fp64_matches_oracle(ip, case, coarse):
    mesh, k, b, xo = case
    kind = ip.COARSE_AMG if coarse == "amg" else ip.COARSE_JACOBI_CG
    x, st = solve(ip, mesh, k, b, ip.PRECISION_FP64, kind, ip.ELEMENT_PASS_AUTO)
    assert relL2(x, xo) < 1e-9, (relL2(x, xo), st)
