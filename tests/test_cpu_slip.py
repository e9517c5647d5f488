"""The oracle's vector (velocity) operator with slip walls (Matrix2Vec / Matrix3Vec), pinned:

* against an independent quadrature assembly (tests/sipg_textbook.py assemble_vec): the textbook
  SIPG form per component plus, on slip walls, the normal projection n n^T of the Dirichlet face
  form; variable coefficient and per-node mass factor included;
* by frame invariance: on a rotated mesh with slip walls, A_rot (R u) = R (A u) node by node
  (a wrong sign or index in the projection breaks it);
* reduction to the scalar operator of each component when no face is a slip wall, and a lone
  type 2 adds nothing;
* diagonal = diag of the assembled matrix; apply_bc per component; PCG = dense solve.
"""
import numpy as np
import pytest

import sipg_textbook as tb

FV2 = [(0, 1), (1, 2), (2, 0)]
FV3 = [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]


def _mesh(ip, dim, seed=1):
    return ip.Mesh.square(3, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(2, 0.1, seed + 1)


def _types_of(dim):
    def f(cen):
        if abs(cen[0]) < 1e-9:
            return (0,) * dim                       # no-slip wall
        if abs(cen[0] - 1) < 1e-9:
            return (2,) * dim                       # slip wall
        if abs(cen[1]) < 1e-9:
            return (2, 1) + ((0,) if dim == 3 else ())  # lone 2: nothing for that component
        return (1,) * dim                           # outflow
    return f


def _bct(o, coords, c2n, types_of):
    dim = coords.shape[1]
    b2c, bn = o.bfaces()
    fv = FV2 if dim == 2 else FV3
    bct = np.zeros((dim, o.n_bfaces), dtype=np.int32)
    for b in range(o.n_bfaces):
        cen = coords[c2n[b2c[b]][list(fv[bn[b]])]].mean(axis=0)
        bct[:, b] = types_of(cen)
    return bct


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_oracle_vector_operator_matches_textbook(ip, oracle, dim, N):
    mesh = _mesh(ip, dim)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    tof = _types_of(dim)
    bct = _bct(o, coords, c2n, tof)
    assert (bct == 2).all(axis=0).any()  # at least one slip wall
    rng = np.random.default_rng(N)
    k = 1 + rng.random((len(c2n), o.Np))
    mm = 2 + rng.random((len(c2n), o.Np))
    o.set_factor(k)
    o.set_mm_factor(mm)
    A = o.dense_vec(bct)
    T = tb.assemble_vec(coords, c2n, o.node_coords(), N, tof, k=k, mm=mm)
    assert np.abs(A - T).max() <= 1e-11 * np.abs(T).max()


def test_textbook_slip_detects_a_projection_slip(ip, oracle):
    """The pin fails for a plausible mistake: the slip ghost taken without the n_a factor
    (jump 2 n.U on every component) differs from the oracle's operator."""
    mesh = _mesh(ip, 2)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, 2)
    bct = np.full((2, o.n_bfaces), 2, dtype=np.int32)
    A = o.dense_vec(bct)
    blocks = []
    A0 = tb.assemble(coords, c2n, o.node_coords(), 2, bc_type_of=lambda c: 1, bface_blocks=blocks)
    n = A0.shape[0]
    Np = o.Np
    W = np.zeros_like(A)
    W[:n, :n] = A0
    W[n:, n:] = A0
    for e, nrm, FV, Db in blocks:
        for a in range(2):
            for b in range(2):
                W[a * n + e * Np:a * n + (e + 1) * Np, b * n + e * Np:b * n + (e + 1) * Np] += nrm[b] * Db
    assert np.abs(A - W).max() > 1e-3 * np.abs(A).max()


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 3), (3, 2)])
def test_slip_frame_invariance(ip, oracle, dim, N):
    mesh = _mesh(ip, dim, seed=4)
    coords, c2n = mesh.arrays()
    if dim == 2:
        th = 0.7
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    else:
        q, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))
        R = q if np.linalg.det(q) > 0 else -q
    o = oracle.Oracle(coords, c2n, N)
    orot = oracle.Oracle(coords @ R.T, c2n, N)
    bct = np.full((dim, o.n_bfaces), 2, dtype=np.int32)
    bct[:, ::5] = 0  # some no-slip faces too
    rng = np.random.default_rng(6)
    k = 1 + rng.random((len(c2n), o.Np))
    for oo in (o, orot):
        oo.set_factor(k)
        oo.set_mm_const(3.0)
    assert np.array_equal(o.bfaces()[0], orot.bfaces()[0])
    assert np.allclose(orot.node_coords(), np.einsum("ij,jen->ien", R, o.node_coords()))
    u = rng.standard_normal((dim, len(c2n), o.Np))
    Au = o.mult_vec(bct, u)
    ARu = orot.mult_vec(bct, np.einsum("ij,jen->ien", R, u))
    assert np.abs(ARu - np.einsum("ij,jen->ien", R, Au)).max() <= 1e-11 * np.abs(Au).max()


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_reduces_to_scalar_operator(ip, oracle, dim, N):
    mesh = _mesh(ip, dim)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    rng = np.random.default_rng(7)
    bct = rng.integers(0, 2, size=(dim, o.n_bfaces)).astype(np.int32)
    lone = bct.copy()
    lone[0, ::3] = 2  # a lone 2 on component 0: no terms, as Neumann in the operator
    u = rng.standard_normal((dim, len(c2n), o.Np))
    Au = o.mult_vec(bct, u)
    Al = o.mult_vec(lone, u)
    for c in range(dim):
        o.set_bc_types(bct[c])
        assert np.abs(Au[c] - o.mult(u[c]).reshape(len(c2n), o.Np)).max() <= 1e-13 * np.abs(Au[c]).max()
    t0 = lone[0].copy()
    t0[t0 == 2] = 1
    o.set_bc_types(t0)
    assert np.abs(Al[0] - o.mult(u[0]).reshape(len(c2n), o.Np)).max() <= 1e-13 * np.abs(Al[0]).max()
    # boundary terms: a lone 2 adds nothing (not the Neumann terms)
    g = rng.standard_normal((dim, o.n_bfaces, o.Nfp))
    rl = o.apply_bc_vec(lone, g, np.zeros_like(u))
    t0 = lone[0].copy()
    t0[t0 == 2] = 7  # a type the scalar kernel treats as neither Dirichlet nor Neumann
    o.set_bc_types(t0)
    assert np.abs(rl[0] - o.apply_bc(g[0], np.zeros_like(u[0])).reshape(len(c2n), o.Np)).max() <= 1e-13 * (np.abs(rl[0]).max() + 1)


@pytest.mark.parametrize("dim,N", [(2, 2), (3, 1)])
def test_diag_and_solve(ip, oracle, dim, N):
    mesh = _mesh(ip, dim)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    bct = _bct(o, coords, c2n, _types_of(dim))
    rng = np.random.default_rng(8)
    k = 1 + rng.random((len(c2n), o.Np))
    o.set_factor(k)
    o.set_mm_const(5.0)
    A = o.dense_vec(bct)
    d = o.diag_vec(bct, consistent=True)
    assert np.abs(d.ravel() - np.diag(A)).max() <= 1e-12 * np.abs(np.diag(A)).max()
    f = rng.standard_normal((dim, len(c2n), o.Np))
    g = rng.standard_normal((dim, o.n_bfaces, o.Nfp))
    b = o.apply_bc_vec(bct, g, f)
    xd = np.linalg.solve(A, b.ravel())
    for pre in (0, 1, 2, 4):
        x, rc, it, rr = o.solve_vec(bct, f, bc=g, rtol=1e-13, precond=pre, factor=0.5)
        assert rc == 0, (pre, rr)
        assert np.linalg.norm(x.ravel() - xd) <= 1e-10 * np.linalg.norm(xd)
