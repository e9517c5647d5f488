"""GPU parity of the block-Jacobi preconditioner (reference: calc_inv_block_diag,
src/solvers/2d/slip_matrix/factor_viscous_matrix.cpp:258-326, and cg_block_jacobi.h),
through the C ABI.

  - The inverse per-cell blocks equal numpy inverses of the diagonal blocks of the
    oracle's assembled operator, within 1e-10 of the largest entry.
  - Block-Jacobi PCG reaches the oracle's FP64 solution within 1e-10 relative L2.
"""
import numpy as np
import pytest

from helpers import density_field, manufactured_sin

pytestmark = pytest.mark.gpu


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 3), (3, 2), (3, 4)])
def test_inverse_blocks_match_oracle(ip, oracle, dim, N):
    mesh = ip.Mesh.square(4, jitter=0.1, seed=2) if dim == 2 else ip.Mesh.cube(2, 0.1, 2)
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(*mesh.arrays(), N)
    bt = np.zeros(s.n_bfaces, dtype=np.int32)
    bt[1::3] = 1
    s.set_bc_types(bt)
    o.set_bc_types(bt)
    k = 1.0 / density_field(s.node_coords(), 0.5, rho_in=10.0)
    s.set_factor(k)
    o.set_factor(k)
    s.set_solver(ip.SOLVER_BLOCK_JACOBI)
    inv = s.inv_block_diag()
    A = o.dense()
    Np = s.Np
    for e in range(s.n_cells):
        blk = A[e * Np:(e + 1) * Np, e * Np:(e + 1) * Np]
        ref = np.linalg.inv(blk)
        assert np.abs(inv[e] - ref).max() < 1e-10 * np.abs(ref).max(), e


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_block_jacobi_solve_matches_oracle(ip, oracle, dim, N):
    mesh = ip.Mesh.square(8, jitter=0.1, seed=6) if dim == 2 else ip.Mesh.cube(3, 0.1, 6)
    its = {}
    for kind in (ip.SOLVER_BLOCK_JACOBI, ip.SOLVER_JACOBI):
        s = ip.PoissonSolver(mesh, N)
        s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
        s.set_solver(kind)
        s.set_tol_and_iter(1e-12, 0.0, 5000)
        _, f = manufactured_sin(s.node_coords())
        b = s.mass(f)
        x = np.zeros_like(b)
        assert s.solve(b, x)
        its[kind] = s.last_stats["iterations"]
        if kind == ip.SOLVER_BLOCK_JACOBI:
            o = oracle.Oracle(*mesh.arrays(), N)
            xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-13, max_iter=100000)
            assert rc == 0
            assert relL2(x, xo) < 1e-10
    assert its[ip.SOLVER_BLOCK_JACOBI] < its[ip.SOLVER_JACOBI], its
