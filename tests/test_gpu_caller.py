"""GPU parity of the caller-side pieces around the pressure solve, through the C ABI.

  - pressure rhs (ins_pressure_rhs_2d.h:1-12 + op2_gemv(LIFT) + mesh->mass,
    src/solvers/2d/ins_solver.cpp:310-319) equals J M (-div/dt + LIFT g), g = b0 dPdN +
    b1 dPdNOld, formed with the oracle's independently derived mass and face-mass
    matrices (LIFT = M^-1 EMAT), within 1e-12 relative;
  - initial-guess extrapolation is exact for fields polynomial in time of degree below the
    history size, and shortens the next solve.
"""
import numpy as np
import pytest

from helpers import manufactured_sin

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 3), (2, 4), (3, 2), (3, 3), (3, 4)])
def test_pressure_rhs_matches_oracle(ip, oracle, dim, N):
    mesh = ip.Mesh.square(5, jitter=0.1, seed=1) if dim == 2 else ip.Mesh.cube(2, 0.1, 1)
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(*mesh.arrays(), N)
    rng = np.random.default_rng(dim * 10 + N)
    NF = (dim + 1) * s.Nfp
    div = rng.standard_normal((s.n_cells, s.Np))
    g_new = rng.standard_normal((s.n_cells, NF))
    g_old = rng.standard_normal((s.n_cells, NF))
    b0, b1, dt = 2.0, -1.0, 0.013
    rhs, g_out = s.pressure_rhs(b0, b1, dt, g_new, g_old, div)
    g = b0 * g_new + b1 * g_old
    assert rel(g_out, g) < 1e-15
    R = oracle.ref_element(dim, N)
    Np, Nfp = s.Np, s.Nfp
    EMAT = np.zeros((Np, NF))
    for f in range(dim + 1):
        EMAT[R["fmask"][f][:, None], f * Nfp + np.arange(Nfp)[None, :]] += R["Mface"][f]
    LIFT = np.linalg.solve(R["M"], EMAT)
    ref = o.mass(-div / dt + g @ LIFT.T)
    assert rel(rhs, ref) < 1e-12


def test_initial_guess_extrapolation(ip):
    mesh = ip.Mesh.square(6, jitter=0.1, seed=2)
    s = ip.PoissonSolver(mesh, 2)
    rng = np.random.default_rng(0)
    A, B, Cq = (rng.standard_normal((s.n_cells, s.Np)) for _ in range(3))
    field = lambda t: A + B * t + Cq * t * t
    guess = np.full((s.n_cells, s.Np), 7.0)
    s.initial_guess(0.4, guess)
    assert np.all(guess == 7.0)                    # empty history keeps the caller's guess
    for t in (0.1, 0.2, 0.3):
        s.history_push(t, field(t))
    s.initial_guess(0.4, guess)
    assert rel(guess, field(0.4)) < 1e-12          # quadratic in time, 3 points: exact
    with pytest.raises(ip.InsPoissonError):
        s.history_push(0.3, field(0.3))            # duplicate time
    s.set_history_size(2)                          # keeps 0.2, 0.3: linear extrapolation
    lin = lambda t: A + B * t
    s2 = ip.PoissonSolver(mesh, 2)
    s2.set_history_size(2)
    for t in (0.1, 0.2, 0.3):
        s2.history_push(t, lin(t))
    s2.initial_guess(0.5, guess)
    assert rel(guess, lin(0.5)) < 1e-12


def test_extrapolated_guess_shortens_solve(ip):
    mesh = ip.Mesh.square(12, jitter=0.1, seed=3)
    s = ip.PoissonSolver(mesh, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_pmg([3, 2, 1])
    s.set_tol_and_iter(1e-10, 0.0, 500)
    _, f = manufactured_sin(s.node_coords())
    its = []
    for step, t in enumerate((0.1, 0.2, 0.3, 0.4)):
        b = s.mass(f * (1.0 + t))
        x = np.zeros_like(b)
        if step == 3:
            s.initial_guess(t, x)
        assert s.solve(b, x)
        its.append(s.last_stats["iterations"])
        s.history_push(t, x)
    assert its[3] < its[2], its
