"""Pins the FP64 oracle (CPU, no GPU): closed forms, invariants, brute force.

The reference ships no golden vectors for this path, so the oracle is pinned
by what the mathematics fixes and by a second, independently derived copy of
the reference element (the product's orthonormal-basis construction).
"""
import numpy as np
import pytest

from helpers import manufactured_sin, poly_solution, bc_values


def gll(N):
    from numpy.polynomial import legendre as L
    c = np.zeros(N + 1)
    c[N] = 1
    inner = L.legroots(L.legder(c))
    return np.concatenate([[-1.0], np.sort(inner), [1.0]])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_reference_element_closed_forms(oracle, dim, N):
    R = oracle.ref_element(dim, N)
    rst, D, M = R["rst"], R["D"], R["M"]
    # derivative matrices are exact on monomials of degree <= N
    for a in range(N + 1):
        for b in range(N + 1 - a):
            f = rst[0] ** a * rst[1] ** b
            fr = a * rst[0] ** max(a - 1, 0) * rst[1] ** b if a else 0 * f
            fs = b * rst[0] ** a * rst[1] ** max(b - 1, 0) if b else 0 * f
            assert np.abs(D[0] @ f - fr).max() < 1e-11
            assert np.abs(D[1] @ f - fs).max() < 1e-11
    vol = 2.0 if dim == 2 else 4.0 / 3.0
    assert abs(M.sum() - vol) < 1e-12
    # mass integrates r^2 exactly: int_T r^2 (2D) = 2/3*... computed via unit simplex
    if dim == 2:
        exact = 4 * (4 * 2 / 24 - 2 * 2 / 6 + 1 / 2)  # int (2x-1)^2 over unit tri * 4
        assert abs(np.ones(len(M)) @ M @ rst[0] ** 2 - exact) < 1e-12 or N < 2
    # face masses integrate 1 to the reference face measure (2)
    for f in range(dim + 1):
        assert abs(R["Mface"][f].sum() - 2.0) < 1e-12
    # N = 1 nodes are the vertices
    if N == 1:
        verts = {(-1, -1), (1, -1), (-1, 1)} if dim == 2 else {
            (-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)}
        got = {tuple(np.round(rst[:, i]).astype(int)) for i in range(rst.shape[1])}
        assert got == verts


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("N", [2, 3, 4])
def test_edge_nodes_are_gauss_lobatto(oracle, dim, N):
    """Warp & blend places the nodes of every edge on the 1D GLL points."""
    rst = oracle.ref_element(dim, N)["rst"]
    on_edge = np.abs(rst[1] + 1) < 1e-10
    if dim == 3:
        on_edge &= np.abs(rst[2] + 1) < 1e-10
    r = np.sort(rst[0][on_edge])
    assert np.abs(r - gll(N)).max() < 1e-12


def test_2d_order2_nodes_equidistant(oracle):
    rst = oracle.ref_element(2, 2)["rst"]
    vals = set(np.round(rst.ravel(), 12))
    assert vals <= {-1.0, 0.0, 1.0}


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_operator_symmetric_spd_and_nullspace(ip, oracle, dim, N):
    mesh = ip.Mesh.square(3, jitter=0.1, seed=5) if dim == 2 else ip.Mesh.cube(2, 0.1, 5)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    A = o.dense()
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    o.set_bc_types(np.ones(o.n_bfaces, dtype=np.int32))
    A = o.dense()
    assert np.abs(A @ np.ones(len(A))).max() < 1e-10 * np.abs(A).max()
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev[0] > -1e-10 * ev[-1] and ev[1] > 1e-8 * ev[-1]


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_polynomial_consistency(ip, oracle, dim, N):
    """SIPG is consistent: for u in P_N, A u = M(-lap u) + BC terms to round-off."""
    mesh = ip.Mesh.square(4, jitter=0.1, seed=9) if dim == 2 else ip.Mesh.cube(2, 0.1, 9)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    xyz = o.node_coords()
    u, f = poly_solution(xyz, N)
    R = oracle.ref_element(dim, N)
    b2c, bn = o.bfaces()
    g = bc_values(xyz, b2c, bn, R["fmask"], u)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    lhs = o.mult(u)
    rhs = o.apply_bc(g, o.mass(f))
    assert np.abs(lhs.ravel() - rhs.ravel()).max() < 1e-9 * np.abs(rhs).max()
    # pure Neumann with g = n.grad u is consistent too (checked through A u - BC = M f)


@pytest.mark.parametrize("N", [1, 2, 3])
def test_h_convergence_2d(ip, oracle, N):
    errs = []
    for n in (8, 16, 32):
        mesh = ip.Mesh.square(n, jitter=0.1, seed=11)
        coords, c2n = mesh.arrays()
        o = oracle.Oracle(coords, c2n, N)
        xyz = o.node_coords()
        u, f = manufactured_sin(xyz)
        x, rc, it, rr = o.solve_jacobi(o.mass(f), rtol=1e-12, max_iter=20000)
        assert rc == 0
        errs.append(o.mass_norm(x - u))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > N + 1 - 0.35, rates


def test_h_convergence_3d(ip, oracle):
    errs = []
    for n in (2, 4, 8):
        mesh = ip.Mesh.cube(n, jitter=0.1, seed=11)
        coords, c2n = mesh.arrays()
        o = oracle.Oracle(coords, c2n, 2)
        xyz = o.node_coords()
        u, f = manufactured_sin(xyz)
        x, rc, it, rr = o.solve_jacobi(o.mass(f), rtol=1e-12, max_iter=20000)
        assert rc == 0
        errs.append(o.mass_norm(x - u))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > 2.5, rates


def test_fcg_matches_dense_cholesky(ip, oracle):
    mesh = ip.Mesh.square(3, jitter=0.1, seed=2)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, 3)
    A = o.dense()
    rng = np.random.default_rng(0)
    b = rng.standard_normal(A.shape[0])
    xd = np.linalg.solve(A, b)
    x, rc, it, rr = o.solve_jacobi(b, rtol=1e-13)
    assert rc == 0
    assert np.linalg.norm(x.ravel() - xd) <= 1e-11 * np.linalg.norm(xd)


def test_diag_is_exact_diagonal(ip, oracle):
    mesh = ip.Mesh.square(3, jitter=0.1, seed=4)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, 2)
    rng = np.random.default_rng(1)
    o.set_factor(0.5 + rng.random((o.n_cells, o.Np)))
    bt = np.zeros(o.n_bfaces, dtype=np.int32)
    bt[::3] = 1
    o.set_bc_types(bt)
    A = o.dense()
    d = o.diag(consistent=True).ravel()
    assert np.abs(d - np.diag(A)).max() < 1e-12 * np.abs(A).max()
    # the reference's own formula (fvmf_2d_op2_diag.h) equals it for constant factor
    o.set_factor(None)
    A = o.dense()
    assert np.abs(o.diag(consistent=False).ravel() - np.diag(A)).max() < 1e-12 * np.abs(A).max()


def test_pmg_and_jacobi_agree(ip, oracle):
    mesh = ip.Mesh.cube(3, jitter=0.1, seed=6)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, 3)
    xyz = o.node_coords()
    u, f = manufactured_sin(xyz)
    b = o.mass(f)
    xj, rc1, itj, _ = o.solve_jacobi(b, rtol=1e-11)
    xm, rc2, itm, _, lmax = o.solve_pmg(b, [3, 2, 1], rtol=1e-11, coarse_iters=30)
    assert rc1 == 0 and rc2 == 0
    assert itm < itj / 4
    assert np.linalg.norm(xm - xj) <= 1e-9 * np.linalg.norm(xj)
    assert np.all(lmax[:2] > 1.0) and np.all(lmax[:2] < 10.0)


def test_golden_fixtures(ip, oracle):
    """Regression pin of the oracle against committed fixtures (tests/golden/make_golden.py)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.npz")
    G = np.load(path, allow_pickle=False)
    mesh = ip.Mesh.square(32, jitter=0.1, seed=int(G["cfg0_seed"]))
    coords, c2n = mesh.arrays()
    assert np.array_equal(coords, G["cfg0_coords"])
    o = oracle.Oracle(coords, c2n, 3)
    xyz = o.node_coords()
    u, f = manufactured_sin(xyz)
    x, rc, it, rr = o.solve_jacobi(o.mass(f), rtol=1e-10)
    assert rc == 0
    assert abs(o.mass_norm(x - u) - float(G["cfg0_l2err"])) < 1e-3 * float(G["cfg0_l2err"])
    assert abs(np.sqrt(o.l2sq(x - u)) - float(G["cfg0_l2err_ref_pattern"])) < 1e-3 * float(G["cfg0_l2err"])
    assert np.linalg.norm(x - G["cfg0_solution"]) < 1e-9 * np.linalg.norm(G["cfg0_solution"])
    m2 = ip.Mesh.square(2, jitter=0.1, seed=1)
    o2 = oracle.Oracle(*m2.arrays(), 2)
    assert np.abs(o2.dense() - G["dense_2d_n2"]).max() < 1e-12 * np.abs(G["dense_2d_n2"]).max()
