"""GPU: solves on renumbered meshes (ins_mesh_renumber, both shape-class orders).

The renumbered mesh holds the same cells in another order, so its operator is the permuted
operator: the apply of the renumbered mesh on x[perm] equals the oracle's apply on the original
mesh, permuted (FP64 1e-12, FP32 2e-5), and a p-multigrid solve returns the permuted solution
(1e-10 relative L2). The cube of 6^3 x 6 = 1296 cells ends in a partial run of 384 cells, and its
packed neighbour words (neighbour << 5 | code) take the element passes' packed path.
"""
import numpy as np
import pytest

from helpers import manufactured_sin

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def renumbered(ip, method, n=6, seed=13):
    mesh = ip.Mesh.cube(n, 0.1, seed)
    coords, c0 = mesh.arrays()
    perm = mesh.renumber(method)
    return mesh, coords, c0, perm


@pytest.mark.parametrize("method", ["global", "windowed"])
@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_apply_on_renumbered_mesh(ip, oracle, method, N):
    m = ip.RENUMBER_SHAPE_CLASS if method == "global" else ip.RENUMBER_SHAPE_CLASS_WINDOWED
    mesh, coords, c0, perm = renumbered(ip, m)
    tol = 1e-12 if N < 4 else 1e-11
    # mixed Dirichlet / Neumann faces, against the oracle on the renumbered arrays
    s = ip.PoissonSolver(mesh, N)
    bt = np.zeros(s.n_bfaces, dtype=np.int32)
    bt[1::3] = 1
    s.set_bc_types(bt)
    o1 = oracle.Oracle(*mesh.arrays(), N)
    o1.set_bc_types(bt)
    x = np.random.default_rng(N).standard_normal((s.n_cells, s.Np))
    ref = o1.mult(x)
    assert rel(s.mult(x), ref) < tol
    assert rel(s.mult_fp32(x), ref) < 2e-5
    # all-Dirichlet: the renumbered operator is the permuted operator of the original mesh
    # (new cell i = old cell perm[i], so x_old = x[argsort(perm)] and y_new = y_old[perm])
    sd = ip.PoissonSolver(mesh, N)
    sd.set_bc_types(np.zeros(sd.n_bfaces, dtype=np.int32))
    o = oracle.Oracle(coords, c0, N)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    y0 = o.mult(np.ascontiguousarray(x[np.argsort(perm)]))
    assert rel(sd.mult(x), y0[perm]) < tol


def test_pmg_solve_on_renumbered_mesh_is_the_permuted_solve(ip):
    base = ip.Mesh.cube(6, 0.1, 13)
    win, _, _, perm = renumbered(ip, ip.RENUMBER_SHAPE_CLASS_WINDOWED)
    xs = []
    for mesh in (base, win):
        s = ip.PoissonSolver(mesh, 3)
        s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
        s.set_solver(ip.SOLVER_PMULTIGRID)
        s.set_pmg([3, 2, 1], coarse_iters=30)
        s.set_tol_and_iter(1e-12, 0.0, 300)
        _, f = manufactured_sin(s.node_coords())
        b = s.mass(f)
        x = np.zeros((s.n_cells, s.Np))
        assert s.solve(b, x)
        xs.append(x)
    assert relL2(xs[1], xs[0][perm]) < 1e-10
