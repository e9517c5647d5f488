"""Over-integration (FactorPoissonMatrixFreeDiagOI): the library's cubature rules and the oracle's
over-integrated operator, pinned without the toolkit (whose class is not vendored: parity with
its own rules and point order is unpinned, the caller evaluates its coefficient at the library's
points):

* the rules integrate every monomial of degree <= 2N exactly (barycentric moments), the face
  rules are symmetric (the neighbour's point is a permutation), the edge rule has 2N+1 points;
* the interpolation to the points reproduces polynomials of degree N;
* the oracle's operator equals an independent exact assembly of the SIPG form with a linear
  coefficient k(x) (integrands of degree <= 2N: the rules are exact), and fails it when the
  volume weights or the face weights slip;
* a constant coefficient reproduces the nodal operator.
"""
import itertools
from math import factorial

import numpy as np
import pytest

import sipg_textbook as tb

FV2 = [(0, 1), (1, 2), (2, 0)]
FV3 = [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]
RV2 = np.array([[-1, -1], [1, -1], [-1, 1]], dtype=float)
RV3 = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)


def _bary(dim, pts):
    RV = RV2 if dim == 2 else RV3
    A = np.vstack([RV.T, np.ones(dim + 1)])
    return np.linalg.solve(A, np.vstack([pts.T, np.ones(len(pts))]))  # (dim+1, n)


@pytest.mark.parametrize("dim,N", [(d, n) for d in (2, 3) for n in (1, 2, 3, 4)])
def test_rules_exact_to_degree_2N(ip, dim, N):
    vx, vw, sb, sw = ip.oi_rule(dim, N)
    lam = _bary(dim, vx)
    vol = 2.0 if dim == 2 else 4.0 / 3.0
    assert lam.min() > -1e-14 and sb.min() > 0  # interior points
    for e in itertools.product(range(2 * N + 1), repeat=dim + 1):
        if sum(e) > 2 * N:
            continue
        exact = vol * factorial(dim) * np.prod([factorial(x) for x in e]) / factorial(sum(e) + dim)
        assert abs(np.sum(vw * np.prod(lam ** np.array(e)[:, None], axis=0)) - exact) < 1e-14
    for e in itertools.product(range(2 * N + 1), repeat=dim):
        if sum(e) > 2 * N:
            continue
        exact = 2.0 * factorial(dim - 1) * np.prod([factorial(x) for x in e]) / factorial(sum(e) + dim - 1)
        assert abs(np.sum(sw * np.prod(sb.T ** np.array(e)[:, None], axis=0)) - exact) < 1e-14
    if dim == 2:
        assert len(sw) == 2 * N + 1  # the reference's DG_CUB_SURF_2D_NP (preprocessor.py:4-30)
        assert len(vw) == [3, 6, 12, 16][N - 1]  # DG_CUB_2D_NP
    else:
        assert len(sw) == [3, 6, 12, 16][N - 1]  # DG_CUB_SURF_3D_NP (preprocessor.py:32-60)
    # symmetric face rule: every permutation of a point's barycentrics is a point of the rule
    for p in itertools.permutations(range(dim)):
        perm = sb[:, list(p)]
        d = np.abs(perm[:, None, :] - sb[None, :, :]).sum(axis=2).min(axis=1)
        assert d.max() < 1e-13


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 4), (3, 1), (3, 3)])
def test_interp_reproduces_degree_N(ip, dim, N):
    Iv, Is = ip.oi_interp(dim, N)
    vx, vw, sb, sw = ip.oi_rule(dim, N)
    nodes = ip.ref_element(dim, N)["rst"].T  # (Np, dim)
    rng = np.random.default_rng(N)
    c = rng.standard_normal(dim)

    def poly(x):
        return (1.0 + np.atleast_2d(x) @ c) ** N

    assert np.abs(Iv @ poly(nodes) - poly(vx)).max() < 1e-12
    RV = RV2 if dim == 2 else RV3
    fv = FV2 if dim == 2 else FV3
    xs = np.concatenate([sb @ RV[list(fv[f])] for f in range(dim + 1)])
    assert np.abs(Is @ poly(nodes) - poly(xs)).max() < 1e-12


def _oi_inputs(ip, oracle, dim, N, kfn, seed=1):
    mesh = ip.Mesh.square(3, jitter=0.1, seed=seed) if dim == 2 else ip.Mesh.cube(2, 0.1, seed + 1)
    coords, c2n = mesh.arrays()
    o = oracle.Oracle(coords, c2n, N)
    E = len(c2n)
    fv = FV2 if dim == 2 else FV3

    def tof(cen):
        return 0 if cen[0] < 1e-9 or cen[1] > 1 - 1e-9 else 1

    b2c, bn = o.bfaces()
    bt = np.array([tof(coords[c2n[b2c[b]][list(fv[bn[b]])]].mean(axis=0)) for b in range(o.n_bfaces)],
                  dtype=np.int32)
    o.set_bc_types(bt)
    xyz = o.node_coords()
    kn = kfn(xyz.reshape(dim, -1).T).reshape(E, o.Np)
    o.set_factor(kn)
    rule = ip.oi_rule(dim, N)
    vx, vw, sb, sw = rule
    lam = _bary(dim, vx)
    kv = np.zeros((E, len(vw)))
    ks = np.zeros((E, (dim + 1) * len(sw)))
    for e in range(E):
        X = coords[c2n[e]]
        kv[e] = kfn(lam.T @ X)
        for f in range(dim + 1):
            ks[e, f * len(sw):(f + 1) * len(sw)] = kfn(sb @ X[list(fv[f])])
    return mesh, coords, c2n, o, rule, kv, ks, kn, tof, xyz


def _dense_oi(o, rule, kv, ks):
    n = o.n_cells * o.Np
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        A[:, j] = o.mult_oi(rule, kv, ks, e).ravel()
    return A


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_oracle_oi_matches_exact_form(ip, oracle, dim, N):
    a = np.array([0.7, -0.4, 0.3][:dim])

    def kfn(x):
        return 2.0 + np.atleast_2d(x) @ a

    mesh, coords, c2n, o, rule, kv, ks, kn, tof, xyz = _oi_inputs(ip, oracle, dim, N, kfn)
    A = _dense_oi(o, rule, kv, ks)
    T = tb.assemble(coords, c2n, xyz, N, bc_type_of=tof, k=kn, k_fn=kfn)
    assert np.abs(A - T).max() <= 1e-12 * np.abs(T).max()
    assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()
    # a slip in the volume weights or in the face weights is seen
    vx, vw, sb, sw = rule
    for bad in ((vx, vw * 1.01, sb, sw), (vx, vw, sb, sw * 1.01)):
        B = _dense_oi(o, bad, kv, ks)
        assert np.abs(B - T).max() > 1e-4 * np.abs(T).max()


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_oracle_oi_constant_factor_is_nodal_operator(ip, oracle, dim, N):
    mesh, coords, c2n, o, rule, kv, ks, kn, tof, xyz = _oi_inputs(ip, oracle, dim, N, lambda x: 0 * x[:, 0] + 2.5)
    x = np.random.default_rng(3).standard_normal((o.n_cells, o.Np))
    a = o.mult_oi(rule, kv, ks, x)
    b = o.mult(x).reshape(o.n_cells, o.Np)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
