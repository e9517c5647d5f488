"""Solves on the reference's own meshes (grids/h5 copies under tests/golden/meshes), read by
ins_mesh_load_op2_hdf5, against the oracle built on the same face sets.

cube-20.h5 and vortex_square_small.h5 are periodic boxes [-10, 10]^d (every face interior,
some joining opposite sides); 3D-cylinder.h5 has 2,278 boundary faces.
Tolerances: FP64 applies 5e-12 relative in the max norm (1e-11 at N = 4): these meshes mix cell
sizes and shapes more than the generated ones, and the oracle's monomial basis loses a few
more digits on them (1.8e-12 seen on vortex_square_small at N = 3); FP64 solves 1e-10 relative L2.
"""
import os

import numpy as np
import pytest

from helpers import bc_values

pytestmark = pytest.mark.gpu
MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def load(ip, oracle, name, N):
    m = ip.Mesh.load_hdf5(os.path.join(MESHES, name + ".h5"))
    s = ip.PoissonSolver(m, N)
    o = oracle.Oracle(*m.arrays(), N, faces=m.faces())
    return m, s, o


@pytest.mark.parametrize("name,N", [("cube-20", 1), ("cube-20", 2), ("cube-20", 3),
                                    ("vortex_square_small", 1), ("vortex_square_small", 2),
                                    ("vortex_square_small", 3), ("vortex_square_small", 4),
                                    ("3D-cylinder", 2), ("3D-cylinder", 3)])
def test_apply_on_reference_mesh(ip, oracle, name, N):
    m, s, o = load(ip, oracle, name, N)
    if s.n_bfaces:
        bt = (np.arange(s.n_bfaces) % 2).astype(np.int32)
        s.set_bc_types(bt)
        o.set_bc_types(bt)
    x = np.random.default_rng(N).standard_normal((s.n_cells, s.Np))
    assert rel(s.mult(x), o.mult(x)) < (5e-12 if N < 4 else 1e-11)
    assert rel(s.diag(), o.diag(consistent=False)) < (5e-12 if N < 4 else 1e-11)


def test_periodic_solve_vortex_square(ip, oracle):
    """Fully periodic square [-10, 10]^2: u = sin(pi x/10) sin(pi y/10), pure Neumann-like
    (no boundary), mean-zero; p-multigrid FCG against the oracle's Jacobi-PCG."""
    m, s, o = load(ip, oracle, "vortex_square_small", 3)
    xyz = s.node_coords()
    k = np.pi / 10
    u = np.sin(k * xyz[0]) * np.sin(k * xyz[1])
    b = s.mass(2 * k * k * u)
    b -= b.mean()
    s.set_nullspace(True)
    s.set_pmg([3, 2, 1], coarse_iters=30)
    s.set_tol_and_iter(1e-12, 0.0, 500)
    x = np.zeros_like(b)
    assert s.solve(b, x), s.last_stats
    xo, rc, _, _ = o.solve_jacobi(b, rtol=1e-12, max_iter=100000, nullspace=True)
    assert rc == 0
    assert relL2(x, xo) < 1e-10
    e = x - u
    e -= e.mean()
    assert o.mass_norm(e) < 1e-3 * o.mass_norm(u)


def test_dirichlet_solve_cylinder(ip, oracle):
    """3D-cylinder.h5, N = 2, u = sin(x) cos(y) exp(z/2) with Dirichlet data on every bface."""
    m, s, o = load(ip, oracle, "3D-cylinder", 2)
    xyz = s.node_coords()
    u = np.sin(xyz[0]) * np.cos(xyz[1]) * np.exp(0.5 * xyz[2])
    f = (1 + 1 - 0.25) * u
    _, _, b2c, bn = m.faces()
    bc = bc_values(xyz, b2c, bn, ip.ref_element(3, 2)["fmask"], u)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    s.set_bcs(bc)
    b = s.mass(f)
    s.set_pmg([2, 1], coarse_iters=30)
    s.set_tol_and_iter(1e-12, 0.0, 500)
    x = np.zeros_like(b)
    assert s.solve(b, x), s.last_stats
    xo, rc, _, _ = o.solve_jacobi(o.apply_bc(bc, b), rtol=1e-12, max_iter=100000)
    assert rc == 0
    assert relL2(x, xo) < 1e-10
    assert o.mass_norm(x - u) < 1e-2 * o.mass_norm(u)
