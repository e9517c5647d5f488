"""GPU parity of the pressure projection (gradient with central flux, velocity update) against
the FP64 oracle: within 1e-12 relative (max norm) for 2D and 3D, orders 1-3, and 5e-11 at order 4,
where the oracle's monomial-basis reference element carries ~5e-13 of its own error and the
derivative and lift matrices amplify it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_grad_central_and_projection(ip, oracle, dim, N):
    print(f"projection dim={dim} N={N}", flush=True)
    mesh = ip.Mesh.square(6, jitter=0.1, seed=21) if dim == 2 else ip.Mesh.cube(3, 0.1, 21)
    coords, c2n = mesh.arrays()
    s = ip.PoissonSolver(mesh, N)
    o = oracle.Oracle(coords, c2n, N)
    rng = np.random.default_rng(dim * 10 + N)
    p = rng.standard_normal((s.n_cells, s.Np))
    go = o.grad_central(p)
    tol = 1e-12 if N < 4 else 5e-11
    assert rel(s.grad_central(p), go) < tol
    uT = rng.standard_normal((dim, s.n_cells, s.Np))
    dt = 0.0137
    assert rel(s.project_velocity(dt, p, uT), uT - dt * go) < tol
