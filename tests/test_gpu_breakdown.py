"""Indefinite systems end in INS_ERR_BREAKDOWN, through the C ABI, as the oracle's solve does
(tests/test_cpu_oracle_pins.py): a density that changes sign makes p.Ap <= 0 (Jacobi-PCG and
p-multigrid FCG), and a nonfinite rhs is reported the same way."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,solver", [(2, "jacobi"), (3, "jacobi"), (3, "pmg")])
def test_indefinite_factor_reports_breakdown(ip, dim, solver):
    mesh = ip.Mesh.square(6, jitter=0.1, seed=2) if dim == 2 else ip.Mesh.cube(3, 0.1, 2)
    s = ip.PoissonSolver(mesh, 2)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    xyz = s.node_coords()
    s.set_factor(np.where(xyz[0] < 0.5, 1.0, -1.0))
    if solver == "jacobi":
        s.set_solver(ip.SOLVER_JACOBI)
    else:
        s.set_solver(ip.SOLVER_PMULTIGRID)
        s.set_pmg([2, 1], coarse_iters=10)
    s.set_tol_and_iter(1e-10, 0.0, 500)
    b = s.mass(np.ones((s.n_cells, s.Np)))
    x = np.zeros_like(b)
    with pytest.raises(ip.InsPoissonError) as e:
        s.solve(b, x)
    assert e.value.code == ip.ERR_BREAKDOWN


def test_nonfinite_rhs_reports_breakdown(ip):
    mesh = ip.Mesh.square(4, jitter=0.1, seed=3)
    s = ip.PoissonSolver(mesh, 2)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_JACOBI)
    b = np.ones((s.n_cells, s.Np))
    b[1, 2] = np.nan
    x = np.zeros_like(b)
    with pytest.raises(ip.InsPoissonError) as e:
        s.solve(b, x)
    assert e.value.code == ip.ERR_BREAKDOWN
