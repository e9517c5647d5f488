"""Pins the oracle's gradient with central flux (the pressure projection's grad_with_central_flux).

The toolkit routine is not vendored with the reference, so the restatement follows the textbook
strong form and is pinned by what the mathematics fixes: exactness on global polynomials of
degree <= N, a discrete Gauss identity for discontinuous fields that vanish on the boundary
(the central flux makes the interior face terms cancel: sum over cells of J M grad_h p = 0), and
O(h^N) convergence of the gradient of a smooth field.
"""
import numpy as np
import pytest

from helpers import poly_solution


def make(ip, oracle, dim, N, n, seed=3):
    mesh = ip.Mesh.square(n, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(n, 0.1, seed)
    coords, c2n = mesh.arrays()
    return oracle.Oracle(coords, c2n, N)


def poly_grad(xyz, N):
    """Gradient of helpers.poly_solution's polynomial."""
    dim = xyz.shape[0]
    x, y = xyz[0], xyz[1]
    if dim == 2:
        if N == 1:
            return np.stack([2 + 0 * x, -3 + 0 * x])
        gx, gy = 2 * x - 0.5 * y + 1, -0.5 * x + 4 * y
        if N >= 3:
            gx, gy = gx + 3 * x * x - y * y, gy - 2 * x * y
        return np.stack([gx, gy])
    z = xyz[2]
    if N == 1:
        return np.stack([2 + 0 * x, -3 + 0 * x, 1 + 0 * x])
    gx, gy, gz = 2 * x - 0.5 * y + 1, -0.5 * x + 4 * y - z, 2 * z - y
    if N >= 3:
        gx, gz = gx + 3 * x * x - z * z, gz - 2 * x * z
    return np.stack([gx, gy, gz])


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_grad_exact_on_polynomials(ip, oracle, dim, N):
    o = make(ip, oracle, dim, N, 4 if dim == 2 else 2)
    xyz = o.node_coords()
    u, _ = poly_solution(xyz, N)
    g = o.grad_central(u)
    ge = poly_grad(xyz, N)
    assert np.abs(g - ge).max() < 1e-10 * max(1.0, np.abs(ge).max())


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 3), (3, 2)])
def test_discrete_gauss_identity(ip, oracle, dim, N):
    """For discontinuous p vanishing at boundary face nodes, sum_e 1^T J M (grad_h p) = 0."""
    o = make(ip, oracle, dim, N, 4 if dim == 2 else 2, seed=5)
    p = np.random.default_rng(1).standard_normal((o.n_cells, o.Np))
    b2c, bn = o.bfaces()
    fmask = oracle.ref_element(dim, N)["fmask"]
    for c, f in zip(b2c, bn):
        p[c, fmask[f]] = 0.0
    g = o.grad_central(p)
    for x in range(dim):
        total = o.mass(g[x]).sum()
        assert abs(total) < 1e-11 * np.abs(o.mass(np.abs(g[x]))).sum()
    # a wrong flux sign or weight breaks the identity
    assert np.abs(o.mass(g[0])).sum() > 1e-3


@pytest.mark.parametrize("N", [1, 2, 3])
def test_grad_h_convergence(ip, oracle, N):
    errs = []
    for n in (8, 16, 32):
        o = make(ip, oracle, 2, N, n, seed=11)
        xyz = o.node_coords()
        u = np.sin(np.pi * xyz[0]) * np.cos(np.pi * xyz[1])
        gx = np.pi * np.cos(np.pi * xyz[0]) * np.cos(np.pi * xyz[1])
        g = o.grad_central(u)
        errs.append(o.mass_norm(g[0] - gx))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > N - 0.3, rates
