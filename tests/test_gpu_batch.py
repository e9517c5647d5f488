"""GPU parity of the batched and serving forms, through the C ABI.

Each problem in a batch runs the flexible CG of src/solvers/2d/viscous_solver.cpp:158-266
preconditioned by a two-level p-multigrid V-cycle (orders N,1). Bars:
  - FP64 V-cycle solves to rtol 1e-12 within 1e-9 relative L2 of the oracle's
    two-level p-multigrid solve, with outer iteration counts within 3 of it;
  - mixed precision (FP32 V-cycle) within 1e-8 of the oracle after the FP64 outer CG;
  - a problem's result is bitwise the same alone, in any batch mix and through the server.
"""
import threading
import time

import numpy as np
import pytest

from helpers import density_field, manufactured_sin

pytestmark = pytest.mark.gpu


def relL2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def sq(ip, n, seed):
    return ip.Mesh.square(n, jitter=0.1, seed=seed)


def cube(ip, n, seed):
    return ip.Mesh.cube(n, 0.1, seed)


def sin_problem(ip, mesh, N, k=(1, 2, 1), rtol=1e-12, max_iter=400, **kw):
    """Dirichlet problem with u = prod sin(k_i pi x_i), rhs as nodal f (b = J M f)."""
    _, f = manufactured_sin(mesh.node_coords(N), k)
    return dict(mesh=mesh, order=N, rhs=f, rhs_nodal=True, rtol=rtol, max_iter=max_iter, **kw)


def mass_rhs(ip, mesh, N, f):
    s = ip.PoissonSolver(mesh, N)
    return s.mass(f)


def run_batch(ip, dim, problems, **cfg):
    b = ip.Batch(dim, **cfg)
    ids = b.add(problems)
    b.run()
    done = set(b.take_finished())
    assert done == set(ids)
    out = [b.result(i) for i in ids]
    return b, ids, out


MIX2 = [(5, 1), (7, 2), (6, 3), (4, 4), (9, 2)]
MIX3 = [(2, 1), (3, 2), (2, 3), (2, 4)]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("coarse", ["jacobi_cg", "amg"])
def test_batch_fp64_matches_oracle(ip, oracle, dim, coarse):
    mix = MIX2 if dim == 2 else MIX3
    meshes = [(sq if dim == 2 else cube)(ip, n, 30 + i) for i, (n, _) in enumerate(mix)]
    probs = [sin_problem(ip, m, N) for m, (_, N) in zip(meshes, mix)]
    kind = ip.COARSE_AMG if coarse == "amg" else ip.COARSE_JACOBI_CG
    _, _, res = run_batch(ip, dim, probs, precision=ip.PRECISION_FP64, coarse=kind)
    for m, (_, N), p, (rc, x, st) in zip(meshes, mix, probs, res):
        assert rc == ip.OK, st
        assert st["converged"] == 1 and st["rel_residual"] <= 1e-12
        o = oracle.Oracle(*m.arrays(), N)
        b = o.mass(p["rhs"])
        xo, orc, it_o, _, _ = o.solve_pmg(b, [N, 1] if N > 1 else [1], rtol=1e-12, max_iter=400)
        assert orc == 0
        assert relL2(x, xo) < 1e-9, (N, relL2(x, xo))
        if N > 1 and coarse == "jacobi_cg":  # the oracle's p-multigrid has the Jacobi-CG coarse solve
            assert abs(st["iterations"] - it_o) <= 3, (N, st["iterations"], it_o)


def test_batch_mixed_variable_neumann(ip, oracle):
    """rho contrast 10 circle, pure Neumann with the mean-zero nullspace, FP32 V-cycle with the
    AMG coarse solve (the default)."""
    probs, meshes = [], []
    for i, (n, N) in enumerate([(6, 2), (8, 3), (5, 4)]):
        m = sq(ip, n, 50 + i)
        xyz = m.node_coords(N)
        k = 1.0 / density_field(xyz, 1.0 / n, rho_in=10.0)
        rng = np.random.default_rng(i)
        f = rng.standard_normal(xyz[0].shape)
        b = mass_rhs(ip, m, N, f)
        b -= b.mean()
        probs.append(dict(mesh=m, order=N, rhs=b, factor=k, nullspace=True,
                          bc_types=np.ones(m.n_bfaces, dtype=np.int32), rtol=1e-11, max_iter=400))
        meshes.append(m)
    _, _, res = run_batch(ip, 2, probs, precision=ip.PRECISION_MIXED)
    for m, p, (rc, x, st) in zip(meshes, probs, res):
        assert rc == ip.OK, st
        assert abs(x.mean()) < 1e-12 * np.abs(x).max()
        o = oracle.Oracle(*m.arrays(), p["order"])
        o.set_bc_types(p["bc_types"])
        o.set_factor(p["factor"])
        xo, orc, _, _ = o.solve_jacobi(p["rhs"], rtol=1e-13, max_iter=100000, nullspace=True)
        assert orc == 0
        assert relL2(x, xo) < 1e-8, relL2(x, xo)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_result_independent_of_mix(ip, precision):
    """The same problem alone, in a 2D mix, and added later into a running batch: bitwise equal."""
    prec = ip.PRECISION_FP64 if precision == "fp64" else ip.PRECISION_MIXED
    target_mesh = sq(ip, 8, 77)
    target = sin_problem(ip, target_mesh, 3, rtol=1e-10)
    _, _, alone = run_batch(ip, 2, [target], precision=prec)
    others = [sin_problem(ip, sq(ip, n, 60 + n), N, rtol=1e-10) for n, N in MIX2]
    _, ids, mixed = run_batch(ip, 2, others[:2] + [target] + others[2:], precision=prec)
    rc_a, x_a, st_a = alone[0]
    rc_m, x_m, st_m = mixed[2]
    assert rc_a == rc_m == ip.OK
    assert np.array_equal(x_a, x_m)
    assert st_a["iterations"] == st_m["iterations"]
    # joins a batch that is already iterating
    b = ip.Batch(2, precision=prec)
    first = b.add(others)
    b.iterate(3)
    late = b.add([target])[0]
    b.run()
    got = set(b.take_finished())
    assert got == set(first) | {late}
    rc_l, x_l, st_l = b.result(late)
    assert rc_l == ip.OK and np.array_equal(x_a, x_l)
    assert st_l["iterations"] == st_a["iterations"]


def test_batch_nodal_rhs_equals_mass_weighted(ip):
    m = sq(ip, 7, 5)
    _, f = manufactured_sin(m.node_coords(2), (2, 1))
    b_mass = mass_rhs(ip, m, 2, f)
    _, _, r1 = run_batch(ip, 2, [dict(mesh=m, order=2, rhs=f, rhs_nodal=True, rtol=1e-10)])
    _, _, r2 = run_batch(ip, 2, [dict(mesh=m, order=2, rhs=b_mass, rtol=1e-10)])
    assert np.array_equal(r1[0][1], r2[0][1])


def test_batch_graph_replay_matches_eager(ip):
    probs = [sin_problem(ip, sq(ip, n, 90 + n), N, rtol=1e-10) for n, N in MIX2]
    _, _, g = run_batch(ip, 2, probs, use_graph=1)
    _, _, e = run_batch(ip, 2, probs, use_graph=0)
    for (rg, xg, sg), (re_, xe, se) in zip(g, e):
        assert rg == re_ == ip.OK
        assert np.array_equal(xg, xe) and sg["iterations"] == se["iterations"]


def test_batch_chunked_replays_match_single_iterations(ip):
    """Replays of several lockstep iterations with the device-side convergence mask give the
    same bits and iteration counts as one iteration per host round trip."""
    probs = [sin_problem(ip, sq(ip, n, 70 + n), N, rtol=1e-10) for n, N in MIX2]
    ref = run_batch(ip, 2, probs, chunk=1)[2]
    for chunk in (3, 8, 64):
        got = run_batch(ip, 2, probs, chunk=chunk)[2]
        for (r1, x1, s1), (r2, x2, s2) in zip(ref, got):
            assert r1 == r2 == ip.OK
            assert np.array_equal(x1, x2)
            assert s1["iterations"] == s2["iterations"]


def test_batch_not_converged_release_and_reuse(ip):
    b = ip.Batch(2, precision=ip.PRECISION_FP64)
    m = sq(ip, 6, 1)
    ids = b.add([sin_problem(ip, m, 2, rtol=1e-14, max_iter=2)])
    b.run()
    assert b.take_finished() == ids
    rc, x, st = b.result(ids[0])
    assert rc == ip.NOT_CONVERGED and st["iterations"] == 2 and st["converged"] == 0
    used = b.stats()["bytes_in_use"]
    arena = ip.problem_bytes(2, m.n_cells, 2, precision=ip.PRECISION_FP64)
    assert arena < used <= arena + ip.amg_bytes(2, m.n_nodes, m.n_cells, ip.PRECISION_FP64)
    b.release(ids[0])
    assert b.stats()["bytes_in_use"] == 0
    with pytest.raises(ip.InsPoissonError):
        b.result(ids[0])
    ids2 = b.add([sin_problem(ip, m, 2, rtol=1e-10)])
    b.run()
    assert b.result(ids2[0])[0] == ip.OK


def test_server_matches_batch_and_budget(ip):
    """Requests from a client thread with staggered arrivals; every result equals the
    batched one bitwise, queueing happens under a tight budget, oversize is rejected."""
    cfg = dict(precision=ip.PRECISION_MIXED)
    reqs2 = [sin_problem(ip, sq(ip, n, 200 + i), N, rtol=1e-10) for i, (n, N) in enumerate(MIX2)]
    reqs3 = [sin_problem(ip, cube(ip, n, 300 + i), N, rtol=1e-10) for i, (n, N) in enumerate(MIX3)]
    _, _, ref2 = run_batch(ip, 2, reqs2, **cfg)
    _, _, ref3 = run_batch(ip, 3, reqs3, **cfg)
    sizes = [ip.problem_bytes(r["mesh"].dim, r["mesh"].n_cells, r["order"]) +
             ip.amg_bytes(r["mesh"].dim, r["mesh"].n_nodes, r["mesh"].n_cells) for r in reqs2 + reqs3]
    budget = max(sizes) * 2
    srv = ip.Server(budget, cfg2d=cfg, cfg3d=cfg)
    ids = []

    def client():
        for r2, r3 in zip(reqs2, reqs3 + [None]):
            ids.append(("2", srv.submit(**r2)))
            if r3 is not None:
                ids.append(("3", srv.submit(**r3)))
            time.sleep(0.01)

    t = threading.Thread(target=client)
    t.start()
    t.join()
    big = ip.Mesh.square(40, jitter=0.1, seed=1)
    assert ip.problem_bytes(2, big.n_cells, 4) + ip.amg_bytes(2, big.n_nodes, big.n_cells) > budget
    rej = srv.submit(**sin_problem(ip, big, 4))
    srv.wait_all()
    rc, _, _, _, _ = srv.wait(rej)
    assert rc == ip.ERR_NOMEM
    k2 = k3 = 0
    for kind, rid in ids:
        rc, x, st, lat, q = srv.wait(rid)
        ref = ref2[k2] if kind == "2" else ref3[k3]
        if kind == "2":
            k2 += 1
        else:
            k3 += 1
        assert rc == ip.OK and np.array_equal(x, ref[1])
        assert st["iterations"] == ref[2]["iterations"]
        assert lat >= q >= 0.0
    s = srv.stats()
    assert s["completed"] == len(ids)
    assert 0 < s["peak_bytes"] <= budget
    srv.close()


def test_batch_amg_high_contrast(ip, oracle):
    """rho contrast 1000 circle/sphere, pure Neumann: the AMG coarse solve reaches the oracle's
    solution, in fewer lockstep iterations than Jacobi-CG(40) once the mesh is large enough for
    the fixed Jacobi-CG count to stop being an exact coarse solve."""
    cases = [(2, sq(ip, 12, 71), 2), (3, cube(ip, 4, 72), 3)]
    for dim, m, N in cases:
        xyz = m.node_coords(N)
        n = 12 if dim == 2 else 4
        k = 1.0 / density_field(xyz, 1.0 / n, rho_in=1000.0)
        f = np.random.default_rng(dim).standard_normal(xyz[0].shape)
        b = mass_rhs(ip, m, N, f)
        b -= b.mean()
        prob = dict(mesh=m, order=N, rhs=b, factor=k, nullspace=True,
                    bc_types=np.ones(m.n_bfaces, dtype=np.int32), rtol=1e-10, max_iter=500)
        its = {}
        for coarse in (ip.COARSE_AMG, ip.COARSE_JACOBI_CG):
            _, _, res = run_batch(ip, dim, [prob], precision=ip.PRECISION_MIXED, coarse=coarse)
            rc, x, st = res[0]
            assert rc == ip.OK, (dim, coarse, st)
            its[coarse] = st["iterations"]
            if coarse == ip.COARSE_AMG:
                o = oracle.Oracle(*m.arrays(), N)
                o.set_bc_types(prob["bc_types"])
                o.set_factor(k)
                xo, orc, _, _ = o.solve_jacobi(b, rtol=1e-12, max_iter=200000, nullspace=True)
                assert orc == 0
                assert relL2(x, xo) < 1e-7, relL2(x, xo)
        if dim == 2:  # on the 384-tet cube Jacobi-CG(40) is still an exact coarse solve
            assert its[ip.COARSE_AMG] < its[ip.COARSE_JACOBI_CG], (dim, its)


def test_batch_tile_sizes_and_autotuner(ip):
    """Every tile size gives the same solution to rounding; the autotuner times all three."""
    probs = [sin_problem(ip, sq(ip, n, 120 + n), N, rtol=1e-11) for n, N in MIX2]
    xs = {}
    for tile in (64, 128, 256):
        b, ids, res = run_batch(ip, 2, probs, precision=ip.PRECISION_FP64, tile=tile)
        assert b.tuning() == (tile, {})
        assert all(r[0] == ip.OK for r in res)
        xs[tile] = [r[1] for r in res]
    for q in range(len(probs)):
        for tile in (64, 128):
            assert relL2(xs[tile][q], xs[256][q]) < 1e-9
    b, _, res = run_batch(ip, 3, [sin_problem(ip, cube(ip, 3, 5), 3, rtol=1e-8)])
    tile, times = b.tuning()
    assert tile in (64, 128, 256) and sorted(times) == [64, 128, 256]
    assert times[tile] == min(times.values())
