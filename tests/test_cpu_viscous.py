"""Pins the oracle's viscous (Helmholtz) operator: the SIPG operator plus a mass term.

The reference's viscous solve applies  A u + J M (mm o u)  (fvmf_2d_mult_mm_geof.h) or
A u + mm J M u  (vmf_2d_mult_mm_geof.h), preconditioned by the inverse mass matrix
(viscous_solver.cpp:334-366) or by Jacobi / block-Jacobi. Pins: under pure Neumann
conditions the SIPG part annihilates constants, so the operator maps 1 to exactly the
(independently pinned) mass of the factor; symmetry and the exact diagonal; Helmholtz
manufactured solutions converge at O(h^{N+1}); inverse-mass PCG reaches the dense solve.
"""
import numpy as np
import pytest

from helpers import manufactured_sin


def make(ip, oracle, dim, N, n, seed=3):
    mesh = ip.Mesh.square(n, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(n, 0.1, seed)
    coords, c2n = mesh.arrays()
    return oracle.Oracle(coords, c2n, N)


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 3), (3, 2)])
def test_mass_term_maps_constants_to_mass(ip, oracle, dim, N):
    o = make(ip, oracle, dim, N, 4 if dim == 2 else 2)
    o.set_bc_types(np.ones(o.n_bfaces, dtype=np.int32))
    one = np.ones((o.n_cells, o.Np))
    assert np.abs(o.mult(one)).max() < 1e-10          # SIPG . 1 = 0 under Neumann
    o.set_mm_const(7.5)
    y = o.mult(one)
    assert np.abs(y - 7.5 * o.mass(one)).max() < 1e-11 * np.abs(y).max()
    rng = np.random.default_rng(1)
    mm = 0.5 + rng.random((o.n_cells, o.Np))
    o.set_mm_const(0.0)
    o.set_mm_factor(mm)
    y = o.mult(one)
    assert np.abs(y - o.mass(mm)).max() < 1e-11 * np.abs(y).max()


def test_mass_term_symmetric_and_diagonal(ip, oracle):
    o = make(ip, oracle, 2, 2, 3)
    bt = np.zeros(o.n_bfaces, dtype=np.int32)
    bt[::2] = 1
    o.set_bc_types(bt)
    A0 = o.dense()
    o.set_mm_const(3.0)
    A = o.dense()
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    # the added term is block diagonal: 3 J M per cell
    B = A - A0
    Np = o.Np
    for e in range(o.n_cells):
        blk = B[e * Np:(e + 1) * Np, e * Np:(e + 1) * Np]
        unit = np.zeros((o.n_cells, Np))
        cols = []
        for j in range(Np):
            unit[:] = 0
            unit[e, j] = 1
            cols.append(3.0 * o.mass(unit)[e])
        assert np.abs(blk - np.array(cols).T).max() < 1e-12 * np.abs(B).max()
    off = B.copy()
    for e in range(o.n_cells):
        off[e * Np:(e + 1) * Np, e * Np:(e + 1) * Np] = 0
    assert np.abs(off).max() < 1e-12 * np.abs(B).max()
    rng = np.random.default_rng(2)
    o.set_mm_factor(0.5 + rng.random((o.n_cells, Np)))
    o.set_factor(0.5 + rng.random((o.n_cells, Np)))
    A = o.dense()
    assert np.abs(o.diag(consistent=True).ravel() - np.diag(A)).max() < 1e-12 * np.abs(A).max()


@pytest.mark.parametrize("N", [1, 2, 3])
def test_helmholtz_h_convergence(ip, oracle, N):
    """-lap u + lam u = f with u = sin(pi x) sin(pi y), Dirichlet: O(h^{N+1}) in L2."""
    lam = 50.0
    errs = []
    for n in (8, 16, 32):
        o = make(ip, oracle, 2, N, n, seed=11)
        o.set_mm_const(lam)
        xyz = o.node_coords()
        u, f = manufactured_sin(xyz)
        x, rc, it, rr = o.solve_inv_mass(o.mass(f + lam * u), rtol=1e-12, max_iter=20000,
                                         factor=1.0 / lam)
        assert rc == 0
        errs.append(o.mass_norm(x - u))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > N + 1 - 0.35, rates


def test_inv_mass_pcg_matches_dense(ip, oracle):
    o = make(ip, oracle, 2, 3, 3, seed=2)
    rng = np.random.default_rng(0)
    mm = 20.0 + rng.random((o.n_cells, o.Np))
    o.set_mm_factor(mm)
    A = o.dense()
    b = rng.standard_normal(A.shape[0])
    xd = np.linalg.solve(A, b)
    x1, rc1, it1, _ = o.solve_inv_mass(b, rtol=1e-13, factor=1.0 / 20.0)
    x2, rc2, it2, _ = o.solve_inv_mass(b, rtol=1e-13, recp_factor=mm)
    xj, rcj, itj, _ = o.solve_jacobi(b, rtol=1e-13)
    assert rc1 == 0 and rc2 == 0 and rcj == 0
    for x in (x1, x2, xj):
        assert np.linalg.norm(x.ravel() - xd) <= 1e-11 * np.linalg.norm(xd)
