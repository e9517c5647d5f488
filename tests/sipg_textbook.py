"""Independent SIPG assembly by quadrature: a checker for the oracle's operator.

Nothing here shares code or tables with the oracle or the product. Each cell's
Lagrange basis is built from its own nodal points (a monomial Vandermonde in
scaled physical coordinates); volume and face integrals use collapsed
Gauss-Legendre rules on the physical simplices; faces, normals, measures and
face nodes are found from the vertex coordinates alone. The bilinear form is
the textbook symmetric interior penalty form written with the reference's
conventions (Appendix A of the reference kernels; each term cites its kernel):

  a(u, v) = sum_K int_K Pi(k grad u) . grad v                      fvmf_2d_grad.h:8-11
          + sum_F int_F (tau/2 [u] - n . {Pi(k grad u)}) v          fvmf_2d_mult_flux.h:9-10
          + sum_F int_F Pi_F(-1/2 k [u]) n . grad v                 fvmf_2d_mult_flux.h:11-14
  tau = 2(N+1)(N+2) max(fscale_L max_F k_L, fscale_R max_F k_R)    fvmf_2d_calc_tau_faces.h:9
  Dirichlet faces: [u] = 2u, {q} = q                               vmf_2d_mult_avg_jump.h:32-38
  tau_b = 2(N+1)(N+2) fscale max_F k                               fvmf_2d_calc_tau_bfaces.h:2-10

with fscale = sJ / J, sJ = |F| / |F_ref| (|F_ref| = 2 in 2D and 3D) and
J = |K| / |K_ref| (|K_ref| = 2 in 2D, 4/3 in 3D): the Hesthaven-Warburton
surface-to-volume ratio. Pi(.) is the nodal interpolant of the product at the
cell's nodes (the reference forms k * grad u node by node), Pi_F the
interpolant through the face's nodes.
"""
import itertools
import math

import numpy as np


def _gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


def unit_simplex_rule(d, n):
    """Collapsed Gauss-Legendre rule on the unit d-simplex (vertex 0 at the origin).

    Returns points (Q, d) and weights summing to 1/d!. Exact for polynomials of
    total degree <= 2n - 1 - (d - 1)."""
    if d == 1:
        x, w = _gl01(n)
        return x[:, None], w
    x, w = _gl01(n)
    pts, wts = [], []
    for idx in itertools.product(range(n), repeat=d):
        a = [x[i] for i in idx]
        wt = 1.0
        for i in idx:
            wt *= w[i]
        # Duffy map: y0 = a0, y1 = a1 (1 - a0), y2 = a2 (1 - a0)(1 - a1)
        y, s = [], 1.0
        for j in range(d):
            y.append(a[j] * s)
            if j + 1 < d:
                wt *= (1 - a[j]) ** (d - 1 - j)
            s *= (1 - a[j])
        pts.append(y)
        wts.append(wt)
    return np.array(pts), np.array(wts)


def monomial_exponents(dim, N):
    return [e for e in itertools.product(range(N + 1), repeat=dim) if sum(e) <= N]


class CellBasis:
    """Lagrange basis of P_N through a cell's nodal points."""

    def __init__(self, nodes, N):
        self.dim = nodes.shape[1]
        self.c = nodes.mean(axis=0)
        self.h = np.ptp(nodes, axis=0).max()
        self.ex = monomial_exponents(self.dim, N)
        V = self._phi(nodes)
        self.C = np.linalg.inv(V)

    def _phi(self, x):
        z = (np.atleast_2d(x) - self.c) / self.h
        return np.stack([np.prod(z ** np.array(e), axis=1) for e in self.ex], axis=1)

    def _dphi(self, x, d):
        z = (np.atleast_2d(x) - self.c) / self.h
        cols = []
        for e in self.ex:
            if e[d] == 0:
                cols.append(np.zeros(z.shape[0]))
                continue
            e2 = list(e)
            e2[d] -= 1
            cols.append(e[d] * np.prod(z ** np.array(e2), axis=1) / self.h)
        return np.stack(cols, axis=1)

    def val(self, x):
        """(Q, Np): basis values at points x."""
        return self._phi(x) @ self.C

    def grad(self, x):
        """(dim, Q, Np): basis gradients at points x."""
        return np.stack([self._dphi(x, d) @ self.C for d in range(self.dim)])


def simplex_measure(V):
    d = V.shape[1]
    m = V.shape[0] - 1
    E = (V[1:] - V[0]).T  # d x m
    return math.sqrt(abs(np.linalg.det(E.T @ E))) / math.factorial(m)


def assemble(coords, c2n, xyz, N, bc_type_of=None, k=None, nq=None, tau_scale=1.0, sym_scale=1.0,
             periods=(), bface_blocks=None, k_fn=None):
    """Dense SIPG matrix (E*Np)^2 in the cells x Np node order of `xyz`.

    coords (nn, dim), c2n (E, dim+1), xyz (dim, E, Np) nodal points;
    bc_type_of(centroid) -> 0 Dirichlet / 1 Neumann (default Dirichlet);
    k (E, Np) nodal coefficient or None (k = 1); tau_scale and sym_scale multiply the
    penalty and the symmetry term (1 for the reference's form; other values let a test
    show that a slip in either is detected); periods: translation vectors that join
    boundary faces into periodic faces (a face and its translate are one interior face);
    bface_blocks: a list that receives (cell, outward unit normal, face vertices, Np x Np block
    of the Dirichlet face terms) for every boundary face, whatever its type;
    k_fn: a coefficient k(x) (points (Q, dim) -> (Q,)) integrated exactly, in place of the nodal
    interpolants Pi(k grad u) and Pi_F(k [u]) (the over-integrated form); the penalty still
    takes the nodal k."""
    dim = coords.shape[1]
    E, Np = xyz.shape[1], xyz.shape[2]
    nq = nq or N + 3
    kk = np.ones((E, Np)) if k is None else np.asarray(k, dtype=float).reshape(E, Np)
    pen = 2.0 * (N + 1) * (N + 2) * tau_scale
    Kref = 2.0 if dim == 2 else 4.0 / 3.0
    Fref = 2.0
    qv, wv = unit_simplex_rule(dim, nq)
    qf, wf = unit_simplex_rule(dim - 1, nq)
    nodes = [xyz[:, e, :].T for e in range(E)]
    bases = [CellBasis(nodes[e], N) for e in range(E)]
    G = []  # nodal gradient matrices G[e][d][i, j] = d_d l_j(x_i)
    for e in range(E):
        G.append(bases[e].grad(nodes[e]))
    A = np.zeros((E * Np, E * Np))
    vol = np.zeros(E)

    def blk(a, b):
        return slice(a * Np, (a + 1) * Np), slice(b * Np, (b + 1) * Np)

    for e in range(E):
        V = coords[c2n[e]]
        vol[e] = simplex_measure(V)
        xq = V[0] + qv @ (V[1:] - V[0])
        w = wv * vol[e] * math.factorial(dim)
        L = bases[e].val(xq)
        dL = bases[e].grad(xq)
        for d in range(dim):
            if k_fn is None:
                A[blk(e, e)] += dL[d].T @ (w[:, None] * (L @ (kk[e][:, None] * G[e][d])))
            else:
                A[blk(e, e)] += dL[d].T @ ((w * k_fn(xq))[:, None] * dL[d])

    faces = {}
    for e in range(E):
        for vs in itertools.combinations(range(dim + 1), dim):
            key = tuple(sorted(int(c2n[e][v]) for v in vs))
            faces.setdefault(key, []).append(e)

    def face_nodes(e, FV):
        # nodes of cell e on the plane/line through the face vertices FV
        n = outward_normal(coords[c2n[e]], FV)
        dist = (nodes[e] - FV[0]) @ n
        return np.nonzero(np.abs(dist) < 1e-10 * bases[e].h)[0]

    def outward_normal(Vc, FV):
        if dim == 2:
            t = FV[1] - FV[0]
            n = np.array([t[1], -t[0]])
        else:
            n = np.cross(FV[1] - FV[0], FV[2] - FV[0])
        n = n / np.linalg.norm(n)
        if (Vc.mean(axis=0) - FV[0]) @ n > 0:
            n = -n
        return n

    # face records: [(cell, face vertex coordinates), ...] with one or two sides
    records = []
    lone = []
    for key, cells in faces.items():
        if len(cells) == 2:
            records.append([(cells[0], coords[list(key)]), (cells[1], coords[list(key)])])
        else:
            lone.append((cells[0], coords[list(key)]))
    if periods:
        def ckey(FV):
            return tuple(sorted(tuple(np.round(v, 9)) for v in FV))
        index = {ckey(FV): i for i, (_, FV) in enumerate(lone)}
        paired = set()
        for i, (e, FV) in enumerate(lone):
            if i in paired:
                continue
            for t in list(periods) + [-np.asarray(t) for t in periods]:
                j = index.get(ckey(FV + np.asarray(t)))
                if j is not None and j not in paired and j != i:
                    records.append([(e, FV), lone[j]])
                    paired.update((i, j))
                    break
        lone = [lone[i] for i in range(len(lone)) if i not in paired]
    records += [[r] for r in lone]

    for rec in records:
        FV = rec[0][1]
        cells = [c for c, _ in rec]
        area = simplex_measure(FV)
        sJ = area / Fref
        fs = {e: sJ / (vol[e] / Kref) for e in cells}
        fn = {e: face_nodes(e, F) for e, F in rec}
        kmax = {e: kk[e][fn[e]].max() for e in cells}
        if len(rec) == 2:
            tau = pen * max(fs[e] * kmax[e] for e in cells)
            for (e, FVe), (o, FVo) in (rec, rec[::-1]):
                t = FVo.mean(axis=0) - FVe.mean(axis=0)  # zero unless periodic
                xq = FVe[0] + qf @ (FVe[1:] - FVe[0])
                w = wf * area * math.factorial(dim - 1)
                n = outward_normal(coords[c2n[e]], FVe)
                Le, Lo = bases[e].val(xq), bases[o].val(xq + t)
                ndL = np.tensordot(n, bases[e].grad(xq), axes=1)  # (Q, Np)
                if k_fn is None:
                    qe = sum(n[d] * (Le @ (kk[e][:, None] * G[e][d])) for d in range(dim))
                    qo = sum(n[d] * (Lo @ (kk[o][:, None] * G[o][d])) for d in range(dim))
                else:
                    kq = k_fn(xq)[:, None]
                    qe = kq * ndL
                    qo = kq * np.tensordot(n, bases[o].grad(xq + t), axes=1)
                A[blk(e, e)] += Le.T @ (w[:, None] * (0.5 * tau * Le - 0.5 * qe))
                A[blk(e, o)] += Le.T @ (w[:, None] * (-0.5 * tau * Lo - 0.5 * qo))
                if k_fn is None:
                    # symmetric term: Pi_F(-1/2 k_e (u_e - u_o)) through e's face nodes
                    S = Le[:, fn[e]] * kk[e][fn[e]][None, :]
                    Ie = np.eye(Np)[fn[e]]
                    Uo = bases[o].val(nodes[e][fn[e]] + t)  # u_o at e's face nodes
                    A[blk(e, e)] += -0.5 * sym_scale * ndL.T @ (w[:, None] * (S @ Ie))
                    A[blk(e, o)] += 0.5 * sym_scale * ndL.T @ (w[:, None] * (S @ Uo))
                else:
                    kw = (w * k_fn(xq))[:, None]
                    A[blk(e, e)] += -0.5 * sym_scale * ndL.T @ (kw * Le)
                    A[blk(e, o)] += 0.5 * sym_scale * ndL.T @ (kw * Lo)
        else:
            e = cells[0]
            xq = FV[0] + qf @ (FV[1:] - FV[0])
            w = wf * area * math.factorial(dim - 1)
            centroid = FV.mean(axis=0)
            bt = 0 if bc_type_of is None else bc_type_of(centroid)
            if bt != 0 and bface_blocks is None:
                continue
            tau = pen * fs[e] * kmax[e]
            n = outward_normal(coords[c2n[e]], FV)
            Le = bases[e].val(xq)
            ndL = np.tensordot(n, bases[e].grad(xq), axes=1)
            if k_fn is None:
                qe = sum(n[d] * (Le @ (kk[e][:, None] * G[e][d])) for d in range(dim))
                S = Le[:, fn[e]] * kk[e][fn[e]][None, :]
                Db = (Le.T @ (w[:, None] * (tau * Le - qe))
                      - sym_scale * ndL.T @ (w[:, None] * (S @ np.eye(Np)[fn[e]])))
            else:
                kq = k_fn(xq)[:, None]
                Db = (Le.T @ (w[:, None] * (tau * Le - kq * ndL))
                      - sym_scale * ndL.T @ ((w[:, None] * kq) * Le))
            if bface_blocks is not None:
                bface_blocks.append((e, n, FV, Db))
            if bt == 0:
                A[blk(e, e)] += Db
    return A


def assemble_vec(coords, c2n, xyz, N, bc_types_of, k=None, mm=None, nq=None):
    """Dense matrix of the vector (velocity) operator with slip walls, (d E Np)^2 with the
    components as d blocks of E*Np, from the textbook scalar form above:

      component blocks   a_c(u, v) with the Dirichlet terms on faces where that component is 0
      slip walls         sum_{a,b} n_a n_b D_F(u_b, v_a) on faces where every component is 2
                         (D_F: the Dirichlet face form; the wall holds n.u = 0 weakly and
                         leaves the tangential part free: the normal projection n n^T of the
                         reference's slip ghost, vmf_2d_mult_avg_jump.h:7-27)
      mass term          mm(x) u.v integrated exactly (mm a nodal field, or a constant)

    bc_types_of(centroid) -> tuple of d bc types of that boundary face."""
    dim = coords.shape[1]
    E, Np = xyz.shape[1], xyz.shape[2]
    blocks = []
    A0 = assemble(coords, c2n, xyz, N, bc_type_of=lambda c: 1, k=k, nq=nq, bface_blocks=blocks)
    n = E * Np
    A = np.zeros((dim * n, dim * n))
    for a in range(dim):
        A[a * n:(a + 1) * n, a * n:(a + 1) * n] = A0
    for e, nrm, FV, Db in blocks:
        t = bc_types_of(FV.mean(axis=0))
        sl = slice(e * Np, (e + 1) * Np)
        if all(x == 2 for x in t):
            for a in range(dim):
                for b in range(dim):
                    A[a * n + e * Np:a * n + (e + 1) * Np, b * n + e * Np:b * n + (e + 1) * Np] += nrm[a] * nrm[b] * Db
            continue
        for a in range(dim):
            if t[a] == 0:
                A[a * n:(a + 1) * n, a * n:(a + 1) * n][sl, sl] += Db
    if mm is not None:
        Mm = mass_matrix(coords, c2n, xyz, N, mm, nq=nq)
        for a in range(dim):
            A[a * n:(a + 1) * n, a * n:(a + 1) * n] += Mm
    return A


def mass_matrix(coords, c2n, xyz, N, mm, nq=None):
    """Block-diagonal matrix of int_K Pi(mm u) v: the nodal interpolant of the product, as the
    reference forms mm * u node by node (fvmf_2d_mult_mm_geof.h); mm a nodal (E, Np) field or a
    constant. By quadrature."""
    dim = coords.shape[1]
    E, Np = xyz.shape[1], xyz.shape[2]
    nq = nq or N + 3
    qv, wv = unit_simplex_rule(dim, nq)
    A = np.zeros((E * Np, E * Np))
    for e in range(E):
        V = coords[c2n[e]]
        vol = simplex_measure(V)
        xq = V[0] + qv @ (V[1:] - V[0])
        w = wv * vol * math.factorial(dim)
        L = CellBasis(xyz[:, e, :].T, N).val(xq)
        Me = L.T @ (w[:, None] * L)
        A[e * Np:(e + 1) * Np, e * Np:(e + 1) * Np] = Me * (float(mm) if np.isscalar(mm) else np.asarray(mm)[e][None, :])
    return A
