"""ctypes bindings of the FP64 CPU oracle - TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker. The product (ins-solver_amd/) never imports this package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def _load():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp, i64, d, ci = C.c_void_p, C.c_int64, C.c_double, C.c_int
    lib.oracle_create.argtypes = [ci, ci, i64, vp, i64, vp]
    lib.oracle_create.restype = vp
    lib.oracle_create_faces.argtypes = [ci, ci, i64, vp, i64, vp, i64, vp, vp, i64, vp, vp]
    lib.oracle_create_faces.restype = vp
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_destroy.restype = None
    for name in ("oracle_np", "oracle_nfp"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = ci
    for name in ("oracle_n_faces", "oracle_n_bfaces"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = i64
    sigs = {
        "oracle_bfaces": [vp, vp, vp], "oracle_faces": [vp, vp, vp],
        "oracle_node_coords": [vp, vp], "oracle_set_bc_types": [vp, vp],
        "oracle_set_factor": [vp, vp], "oracle_mult": [vp, vp, vp],
        "oracle_diag": [vp, ci, vp], "oracle_apply_bc": [vp, vp, vp],
        "oracle_mass": [vp, vp, vp],
    }
    for name, a in sigs.items():
        getattr(lib, name).argtypes = a
        getattr(lib, name).restype = None
    lib.oracle_set_penalty_order.argtypes = [vp, ci]
    lib.oracle_set_penalty_order.restype = None
    lib.oracle_set_threads.argtypes = [ci]
    lib.oracle_set_threads.restype = None
    lib.oracle_l2sq.argtypes = [vp, vp]
    lib.oracle_l2sq.restype = d
    lib.oracle_solve_jacobi.argtypes = [vp, vp, vp, d, d, ci, ci, C.POINTER(ci), C.POINTER(d)]
    lib.oracle_solve_jacobi.restype = ci
    lib.oracle_solve_pmg.argtypes = [vp, vp, ci, vp, vp, vp, d, d, ci, ci, vp, vp, d, d, ci, ci,
                                     C.POINTER(ci), C.POINTER(d), vp]
    lib.oracle_solve_pmg.restype = ci
    lib.oracle_ref_element.argtypes = [ci, ci, vp, vp, vp, vp, vp]
    lib.oracle_ref_element.restype = ci
    lib.oracle_interp.argtypes = [ci, ci, ci, vp]
    lib.oracle_interp.restype = ci
    lib.oracle_grad_central.argtypes = [vp, vp, vp, vp, vp]
    lib.oracle_grad_central.restype = None
    lib.oracle_set_mm_factor.argtypes = [vp, vp]
    lib.oracle_set_mm_factor.restype = None
    lib.oracle_set_mm_const.argtypes = [vp, d]
    lib.oracle_set_mm_const.restype = None
    lib.oracle_solve_inv_mass.argtypes = [vp, vp, vp, d, d, ci, d, vp, C.POINTER(ci), C.POINTER(d)]
    lib.oracle_solve_inv_mass.restype = ci
    lib.oracle_mult_oi.argtypes = [vp, ci, vp, vp, ci, vp, vp, vp, vp, vp, vp]
    lib.oracle_mult_oi.restype = None
    lib.oracle_mult_vec.argtypes = [vp, ci, vp, vp, vp]
    lib.oracle_mult_vec.restype = None
    lib.oracle_diag_vec.argtypes = [vp, ci, vp, ci, vp]
    lib.oracle_diag_vec.restype = None
    lib.oracle_apply_bc_vec.argtypes = [vp, ci, vp, vp, vp]
    lib.oracle_apply_bc_vec.restype = None
    lib.oracle_solve_vec.argtypes = [vp, ci, vp, vp, vp, vp, d, d, ci, ci, d, vp, C.POINTER(ci),
                                     C.POINTER(d)]
    lib.oracle_solve_vec.restype = ci
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def np_of(dim, N):
    return (N + 1) * (N + 2) // 2 if dim == 2 else (N + 1) * (N + 2) * (N + 3) // 6


def nfp_of(dim, N):
    return N + 1 if dim == 2 else (N + 1) * (N + 2) // 2


def set_threads(n):
    lib.oracle_set_threads(int(n))


def ref_element(dim, N):
    Np, Nfp, Nf = np_of(dim, N), nfp_of(dim, N), dim + 1
    rst = np.zeros((dim, Np))
    D = np.zeros((dim, Np, Np))
    M = np.zeros((Np, Np))
    fmask = np.zeros((Nf, Nfp), dtype=np.int32)
    Mf = np.zeros((Nf, Nfp, Nfp))
    if lib.oracle_ref_element(dim, N, _p(rst), _p(D), _p(M), _p(fmask), _p(Mf)) != 0:
        raise RuntimeError("oracle_ref_element failed")
    return {"rst": rst, "D": D, "M": M, "fmask": fmask, "Mface": Mf}


def interp(dim, nf, nc):
    P = np.zeros((np_of(dim, nf), np_of(dim, nc)))
    if lib.oracle_interp(dim, nf, nc, _p(P)) != 0:
        raise RuntimeError("oracle_interp failed")
    return P


class Oracle:
    def __init__(self, coords, cell2nodes, order, faces=None):
        """faces: optional (face2cells, faceNum, bface2cells, bfaceNum) of the caller (an OP2
        mesh file's face sets, periodic faces included) instead of the oracle's own search."""
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        c2n = np.ascontiguousarray(cell2nodes, dtype=np.int32)
        self.dim = coords.shape[1]
        self.order = order
        if faces is None:
            self._h = lib.oracle_create(self.dim, order, coords.shape[0], _p(coords), c2n.shape[0],
                                        _p(c2n))
        else:
            f2c, fn, b2c, bn = (np.ascontiguousarray(a, dtype=np.int32) for a in faces)
            self._h = lib.oracle_create_faces(self.dim, order, coords.shape[0], _p(coords),
                                              c2n.shape[0], _p(c2n), f2c.shape[0], _p(f2c), _p(fn),
                                              b2c.size, _p(b2c), _p(bn))
        if not self._h:
            raise RuntimeError("oracle_create failed")
        self.n_cells = c2n.shape[0]
        self.Np = lib.oracle_np(self._h)
        self.Nfp = lib.oracle_nfp(self._h)
        self.n_faces = lib.oracle_n_faces(self._h)
        self.n_bfaces = lib.oracle_n_bfaces(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            lib.oracle_destroy(self._h)
            self._h = None

    def bfaces(self):
        b2c = np.zeros(self.n_bfaces, dtype=np.int32)
        bn = np.zeros(self.n_bfaces, dtype=np.int32)
        lib.oracle_bfaces(self._h, _p(b2c), _p(bn))
        return b2c, bn

    def faces(self):
        f2c = np.zeros((self.n_faces, 2), dtype=np.int32)
        fn = np.zeros((self.n_faces, 2), dtype=np.int32)
        lib.oracle_faces(self._h, _p(f2c), _p(fn))
        return f2c, fn

    def node_coords(self):
        xyz = np.zeros((self.dim, self.n_cells, self.Np))
        lib.oracle_node_coords(self._h, _p(xyz))
        return xyz

    def set_bc_types(self, t):
        t = np.ascontiguousarray(t, dtype=np.int32)
        assert t.size == self.n_bfaces
        lib.oracle_set_bc_types(self._h, _p(t))

    def set_penalty_order(self, n):
        lib.oracle_set_penalty_order(self._h, int(n))

    def set_factor(self, k):
        k = None if k is None else np.ascontiguousarray(k, dtype=np.float64)
        lib.oracle_set_factor(self._h, _p(k))

    def grad_central(self, p):
        """Gradient with central flux: (dim, cells, Np)."""
        p = self._vec(p)
        g = np.zeros((3, self.n_cells, self.Np))
        lib.oracle_grad_central(self._h, _p(p), _p(g[0]), _p(g[1]), _p(g[2]))
        return g[:self.dim]

    def set_mm_factor(self, mm):
        """Per-node mass-term factor of the viscous operator (cells x Np), None = none."""
        mm = None if mm is None else np.ascontiguousarray(mm, dtype=np.float64)
        lib.oracle_set_mm_factor(self._h, _p(mm))

    def set_mm_const(self, mm):
        """Constant mass-term factor of the viscous operator, 0 = none."""
        lib.oracle_set_mm_const(self._h, float(mm))

    def solve_inv_mass(self, rhs, x0=None, rtol=1e-10, atol=0.0, max_iter=5000, factor=1.0,
                       recp_factor=None):
        """Flexible PCG with the inverse-mass preconditioner (times factor, or divided by
        the per-node recp_factor)."""
        x = np.zeros(self.n_cells * self.Np) if x0 is None else self._vec(x0).copy()
        rf = None if recp_factor is None else self._vec(recp_factor)
        it, rr = C.c_int(), C.c_double()
        rc = lib.oracle_solve_inv_mass(self._h, _p(self._vec(rhs)), _p(x), rtol, atol, max_iter,
                                       float(factor), _p(rf), C.byref(it), C.byref(rr))
        return x.reshape(self.n_cells, self.Np), rc, it.value, rr.value

    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.n_cells * self.Np
        return x

    def mult(self, x):
        x = self._vec(x)
        y = np.zeros_like(x)
        lib.oracle_mult(self._h, _p(x), _p(y))
        return y

    def diag(self, consistent=True):
        d = np.zeros((self.n_cells, self.Np))
        lib.oracle_diag(self._h, int(consistent), _p(d))
        return d

    def apply_bc(self, bc, rhs):
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        r = self._vec(rhs).copy()
        lib.oracle_apply_bc(self._h, _p(bc), _p(r))
        return r

    def mass(self, x):
        x = self._vec(x)
        y = np.zeros_like(x)
        lib.oracle_mass(self._h, _p(x), _p(y))
        return y

    def l2sq(self, e):
        """The reference's error pattern sum_e 1^T J M (e o e) (l2_error_vortex.cpp:25-44)."""
        return lib.oracle_l2sq(self._h, _p(self._vec(e)))

    def mass_norm(self, e):
        """sqrt(e^T (J M) e): the L2 norm of the nodal interpolant."""
        e = self._vec(e)
        return float(np.sqrt(e.ravel() @ self.mass(e).ravel()))

    def mult_oi(self, rule, kv, ks, x):
        """Over-integrated operator: rule = (vol_rst, vol_w, surf_bary, surf_w) (the library's
        rules), kv (cells, n_vol), ks (cells, Nfaces * n_surf)."""
        vx, vw, sb, sw = (np.ascontiguousarray(a, dtype=np.float64) for a in rule)
        kv = np.ascontiguousarray(kv, dtype=np.float64)
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        assert kv.size == self.n_cells * vw.size and ks.size == self.n_cells * (self.dim + 1) * sw.size
        x = self._vec(x)
        y = np.zeros_like(x)
        lib.oracle_mult_oi(self._h, vw.size, _p(vx), _p(vw), sw.size, _p(sb), _p(sw), _p(kv), _p(ks),
                           _p(x), _p(y))
        return y.reshape(self.n_cells, self.Np)

    # ---- vector (velocity) operator with slip walls: dim components, (dim, cells, Np) arrays,
    # bc types (dim, n_bfaces): 0 Dirichlet, 1 Neumann, 2 slip ----
    def _vvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.dim * self.n_cells * self.Np
        return x

    def _vbct(self, bct):
        t = np.ascontiguousarray(bct, dtype=np.int32)
        assert t.size == self.dim * self.n_bfaces
        return t

    def mult_vec(self, bct, x):
        x = self._vvec(x)
        y = np.zeros_like(x)
        lib.oracle_mult_vec(self._h, self.dim, _p(self._vbct(bct)), _p(x), _p(y))
        return y.reshape(self.dim, self.n_cells, self.Np)

    def diag_vec(self, bct, consistent=True):
        d = np.zeros((self.dim, self.n_cells, self.Np))
        lib.oracle_diag_vec(self._h, self.dim, _p(self._vbct(bct)), int(consistent), _p(d))
        return d

    def apply_bc_vec(self, bct, bc, rhs):
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        assert bc.size == self.dim * self.n_bfaces * self.Nfp
        r = self._vvec(rhs).copy()
        lib.oracle_apply_bc_vec(self._h, self.dim, _p(self._vbct(bct)), _p(bc), _p(r))
        return r.reshape(self.dim, self.n_cells, self.Np)

    def solve_vec(self, bct, rhs, bc=None, x0=None, rtol=1e-10, atol=0.0, max_iter=5000, precond=4,
                  factor=1.0, recp_factor=None):
        """ViscousSolver: boundary terms of bc added to the rhs, then PCG (precond 0 none,
        1 inverse mass, 2 inverse mass times factor, 3 inverse mass / recp_factor, 4 Jacobi)."""
        n = self.dim * self.n_cells * self.Np
        x = np.zeros(n) if x0 is None else self._vvec(x0).copy()
        bcv = None if bc is None else np.ascontiguousarray(bc, dtype=np.float64)
        rf = None if recp_factor is None else self._vec(recp_factor)
        it, rr = C.c_int(), C.c_double()
        rc = lib.oracle_solve_vec(self._h, self.dim, _p(self._vbct(bct)), _p(bcv), _p(self._vvec(rhs)), _p(x),
                                  rtol, atol, max_iter, int(precond), float(factor), _p(rf),
                                  C.byref(it), C.byref(rr))
        return x.reshape(self.dim, self.n_cells, self.Np), rc, it.value, rr.value

    def dense_vec(self, bct):
        """Dense vector operator by unit vectors (tiny meshes only)."""
        n = self.dim * self.n_cells * self.Np
        A = np.zeros((n, n))
        e = np.zeros(n)
        for j in range(n):
            e[:] = 0
            e[j] = 1
            A[:, j] = self.mult_vec(bct, e).ravel()
        return A

    def dense(self):
        """Dense operator by applying mult to unit vectors (tiny meshes only)."""
        n = self.n_cells * self.Np
        A = np.zeros((n, n))
        e = np.zeros(n)
        for j in range(n):
            e[:] = 0
            e[j] = 1
            A[:, j] = self.mult(e)
        return A

    def solve_jacobi(self, rhs, x0=None, rtol=1e-10, atol=0.0, max_iter=5000, nullspace=False):
        x = np.zeros(self.n_cells * self.Np) if x0 is None else self._vec(x0).copy()
        it, rr = C.c_int(), C.c_double()
        rc = lib.oracle_solve_jacobi(self._h, _p(self._vec(rhs)), _p(x), rtol, atol, max_iter,
                                     int(nullspace), C.byref(it), C.byref(rr))
        return x.reshape(self.n_cells, self.Np), rc, it.value, rr.value

    def solve_pmg(self, rhs, orders, pre_it=None, post_it=None, cheb_orders=None, eig_safety=1.1,
                  cheb_min_factor=0.1, max_iter_eig=15, coarse_iters=40, x0=None, rtol=1e-10,
                  atol=0.0, max_iter=500, nullspace=False):
        nl = len(orders)

        def arr(v, dflt):
            return np.ascontiguousarray(v if v is not None else [dflt] * nl, dtype=np.int32)

        o, pre, post, ch = arr(orders, 0), arr(pre_it, 1), arr(post_it, 1), arr(cheb_orders, 3)
        x = np.zeros(self.n_cells * self.Np) if x0 is None else self._vec(x0).copy()
        it, rr = C.c_int(), C.c_double()
        lmax = np.zeros(nl)
        rc = lib.oracle_solve_pmg(self._h, _p(o), nl, _p(pre), _p(post), _p(ch), eig_safety,
                                  cheb_min_factor, max_iter_eig, coarse_iters,
                                  _p(self._vec(rhs)), _p(x), rtol, atol, max_iter, int(nullspace),
                                  C.byref(it), C.byref(rr), _p(lmax))
        return x.reshape(self.n_cells, self.Np), rc, it.value, rr.value, lmax
