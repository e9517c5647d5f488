"""Pins of the oracle's operator and eigenvalue estimate by independent computations.

1. The oracle's dense operator equals an SIPG bilinear form assembled here by
   quadrature from scratch (tests/sipg_textbook.py: its own basis, rules, face
   search and penalty), for 2D/3D, N = 1..4, Dirichlet / Neumann / mixed
   boundaries and constant / variable coefficients. The same assembly with the
   penalty or the symmetry term scaled by 2 is shown to differ, so a slip of that
   size in the oracle's tau (fvmf_2d_calc_tau_faces.h:9) or in its 1/2 jump
   bookkeeping (fvmf_2d_mult_flux.h:9-14, vmf_2d_mult_avg_jump.h:32-38) fails here.
2. The oracle's power iteration on D^-1 A (set_rand_vec / rayleigh_quotient /
   vec_normalise; rayleigh_quotient.h:2) converges to the largest eigenvalue of
   D^-1 A from a dense eigen-solve, and its 15-iteration estimate used by the
   Chebyshev smoother lies below that value.
"""
import numpy as np
import pytest

from helpers import density_field
from sipg_textbook import assemble


def _mesh(ip, dim, n, seed):
    return ip.Mesh.square(n, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(n, 0.1, seed)


def _bc_rule(kind):
    if kind == "dirichlet":
        return lambda c: 0
    if kind == "neumann":
        return lambda c: 1
    # Neumann on the faces x = 1 and y = 0, Dirichlet elsewhere
    return lambda c: int(c[0] > 1 - 1e-9 or c[1] < 1e-9)


def _oracle_bc_types(o, oracle, rule):
    xyz = o.node_coords()
    fmask = oracle.ref_element(o.dim, o.order)["fmask"]
    b2c, bn = o.bfaces()
    t = np.zeros(o.n_bfaces, dtype=np.int32)
    for b in range(o.n_bfaces):
        c = xyz[:, b2c[b], fmask[bn[b]]].mean(axis=1)
        t[b] = rule(c)
    return t


def _pair(ip, oracle, dim, N, bc, var):
    n = 3 if dim == 2 else 2
    mesh = _mesh(ip, dim, n, 11 + N)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    rule = _bc_rule(bc)
    o.set_bc_types(_oracle_bc_types(o, oracle, rule))
    xyz = o.node_coords()
    k = None
    if var:
        k = 1.0 / density_field(xyz, 1.0 / n, rho_in=10.0, radius=0.3)
        o.set_factor(k)
    return o, coords, c2n, xyz, rule, k


CASES = ([(d, N, "dirichlet", False) for d in (2, 3) for N in (1, 2, 3, 4)]
         + [(d, N, "mixed", True) for d in (2, 3) for N in (1, 2, 3, 4)]
         + [(d, N, "neumann", False) for d in (2, 3) for N in (2, 3)]
         + [(2, 3, "neumann", True), (3, 2, "dirichlet", True)])


@pytest.mark.parametrize("dim,N,bc,var", CASES)
def test_oracle_operator_equals_quadrature_sipg(ip, oracle, dim, N, bc, var):
    o, coords, c2n, xyz, rule, k = _pair(ip, oracle, dim, N, bc, var)
    A_o = o.dense()
    A_q = assemble(coords, c2n, xyz, N, bc_type_of=rule, k=k)
    scale = np.abs(A_q).max()
    assert np.abs(A_o - A_q).max() < 1e-12 * scale * (10 if N == 4 else 1)


@pytest.mark.parametrize("dim,N", [(2, 2), (3, 3)])
def test_quadrature_sipg_detects_factor_two_slips(ip, oracle, dim, N):
    o, coords, c2n, xyz, rule, k = _pair(ip, oracle, dim, N, "mixed", True)
    A_o = o.dense()
    scale = np.abs(A_o).max()
    for kw in ({"tau_scale": 2.0}, {"tau_scale": 0.5}, {"sym_scale": 2.0}, {"sym_scale": 0.0}):
        A_bad = assemble(coords, c2n, xyz, N, bc_type_of=rule, k=k, **kw)
        assert np.abs(A_o - A_bad).max() > 1e-3 * scale, kw


@pytest.mark.parametrize("dim,N,var", [(2, 2, False), (2, 4, True), (3, 2, False), (3, 3, True)])
def test_power_iteration_matches_dense_eigenvalue(ip, oracle, dim, N, var):
    o, *_ = _pair(ip, oracle, dim, N, "dirichlet", var)
    A = o.dense()
    D = o.diag(consistent=True).ravel()
    lam = np.linalg.eigvals(A / D[:, None])
    top = lam[np.argmax(lam.real)]
    # D^-1 A is similar to a symmetric matrix for constant k; with a variable k the
    # reference's flux makes A slightly nonsymmetric, but the top eigenvalue stays real
    assert abs(top.imag) < 1e-8 * abs(top)
    lmax = top.real
    b = np.zeros((o.n_cells, o.Np))
    b[0, 0] = 1.0
    # the oracle's estimate after 15 iterations (the solver default) and after many
    est15 = o.solve_pmg(b, [N, 1], max_iter_eig=15, max_iter=1)[4][0]
    est = o.solve_pmg(b, [N, 1], max_iter_eig=3000, max_iter=1)[4][0]
    assert abs(est - lmax) < 1e-6 * lmax
    assert 0.5 * lmax < est15 <= lmax * (1 + 1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_indefinite_factor_breaks_down(ip, oracle, dim):
    """A density that changes sign (an indefinite operator) ends the flexible CG with a
    breakdown (p.Ap <= 0), status -3, as the GPU solve reports INS_ERR_BREAKDOWN."""
    mesh = ip.Mesh.square(6, jitter=0.1, seed=2) if dim == 2 else ip.Mesh.cube(3, 0.1, 2)
    o = oracle.Oracle(*mesh.arrays(), 2)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    xyz = o.node_coords()
    o.set_factor(np.where(xyz[0] < 0.5, 1.0, -1.0))
    b = o.mass(np.ones((o.n_cells, o.Np)))
    x, rc, it, rr = o.solve_jacobi(b, rtol=1e-10, max_iter=500)
    assert rc == -3
