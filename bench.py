"""Benchmark of the DG Poisson solve on one MI355X.

Default workload (the 3D plain-FP64 configuration): unit cube, 64^3 hexes split
into 6 Kuhn tetrahedra each (1,572,864 tets), vertices jittered +-0.1h, order
N=3 (Np=20, 31,457,280 DoF), Dirichlet, manufactured u = sin(pi x) sin(pi y)
sin(pi z); flexible CG preconditioned by a p-multigrid V-cycle 3->2->1 with
Chebyshev(3) smoothing and Jacobi-CG on order 1; outer tolerance 1e-10 (true
relative residual). A step is one solve from a zero initial guess with rhs and
solution resident in HBM (device SoA layout).

Prints one JSON line. `value` is GDoF/s = DoF x solves / wall time.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "ins-solver_amd")]

METRIC = "Poisson GDoF/s & solves/s to 1e-8 rel res, mixed vs FP64; p99 solve latency"

# Canonical algorithmic work of one SIPG operator apply per cell (3D, N=3, FP64):
# flops 6*d*Np^2 + 2(d+1)*Np*Nf*Nfp + 4d^2*Np + Nf*Nfp*(4d+6), bytes 2*Np*8 + 10*8 + Nf*5*8 + Nf*8.
FLOPS_PER_CELL = {(3, 3): 15040, (3, 2): 4512, (3, 1): 1032, (3, 4): 41190}
BYTES_PER_CELL_F64 = {(3, 3): 592, (3, 2): 432, (3, 1): 336, (3, 4): 832}
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def manufactured_sin(xyz):
    u = np.sin(np.pi * xyz[0]) * np.sin(np.pi * xyz[1]) * np.sin(np.pi * xyz[2])
    return u, 3 * np.pi ** 2 * u


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="hexes per cube edge")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--coarse-iters", type=int, default=40)
    ap.add_argument("--cpu-n", type=int, default=20, help="cube size of the CPU baseline sample")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-mixed-leg", action="store_true")
    return ap.parse_args()


def cpu_baseline(n, order, rtol):
    import oracle as O
    import ins_poisson as ip
    mesh = ip.Mesh.cube(n, 0.1, 20261203)
    coords, c2n = mesh.arrays()
    o = O.Oracle(coords, c2n, order)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    u, f = manufactured_sin(o.node_coords())
    b = o.mass(f)
    orders = list(range(order, 0, -1))
    t0 = time.perf_counter()
    x, rc, it, rr, _ = o.solve_pmg(b, orders, rtol=rtol, coarse_iters=40)
    dt = time.perf_counter() - t0
    dof = o.n_cells * o.Np
    return {"value": dof / dt * 1e-9, "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"oracle p-MG FP64 solve, cube {n}^3x6 tets N={order} ({dof} DoF), "
                      f"rtol {rtol:g}, {it} its, {dt:.1f} s, converged={rc == 0}"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import ins_poisson as ip

    mesh = ip.Mesh.cube(a.n, 0.1, 20261202)
    s = ip.PoissonSolver(mesh, a.order)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_PMULTIGRID)
    orders = list(range(a.order, 0, -1))
    s.set_pmg(orders, pre_it=[1] * len(orders), post_it=[1] * len(orders),
              cheb_orders=[3] * len(orders), coarse_iters=a.coarse_iters)
    s.set_tol_and_iter(a.rtol, 0.0, 500)
    if a.precision == "mixed":
        s.set_precision(ip.PRECISION_MIXED)
    s.init()
    xyz = s.node_coords()
    u, f = manufactured_sin(xyz)
    dof = s.n_cells * s.Np

    def to_soa(x):
        soa = np.zeros((s.Np, s.stride))
        soa[:, :s.n_cells] = x.T
        return torch.from_numpy(soa).cuda()

    f_d = to_soa(f)
    b_d = torch.zeros_like(f_d)
    x_d = torch.zeros_like(f_d)
    torch.cuda.synchronize()
    s.mass((f_d.data_ptr(), b_d.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    del f, xyz

    def one_solve():
        x_d.zero_()
        ok = s.solve(b_d.data_ptr(), x_d.data_ptr(), layout=ip.LAYOUT_DEVICE_SOA)
        return ok, dict(s.last_stats)

    for _ in range(a.warmup):
        one_solve()
    torch.cuda.synchronize()
    s.take_timing()
    s.set_profiling(True)
    lat, stats = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        ok, st = one_solve()
        lat.append(time.perf_counter() - ts)
        stats.append(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s.set_profiling(False)
    timing = s.take_timing()
    wall = t1 - t0
    gdofs = dof * a.steps / wall * 1e-9

    # discretisation error of the last solution (nodal mass norm)
    x = x_d.cpu().numpy()[:, :s.n_cells].T
    e = x - u
    err = float(np.sqrt(np.sum(e * s.mass(e))))
    unorm = float(np.sqrt(np.sum(u * s.mass(u))))

    key = (3, a.order)
    cells = s.n_cells
    t_apply = (timing["trace_ms"] + timing["apply_ms"]) / max(timing["apply_launches"], 1) * 1e-3
    flops = FLOPS_PER_CELL.get(key, 0) * cells
    bytes_ = BYTES_PER_CELL_F64.get(key, 0) * cells
    achieved_tf = flops / t_apply * 1e-12 if t_apply > 0 else None
    achieved_gbs = bytes_ / t_apply * 1e-9 if t_apply > 0 else None
    roofline = {
        "bound": "mfma",
        "achieved": achieved_tf,
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved_tf / PEAK_FP64_TFLOPS if achieved_tf else None,
        "traffic": None,
        "kernel": f"SIPG apply order {a.order} FP64 (k_trace + k_apply), {cells} cells",
        "avg_launch_ms": t_apply * 1e3,
        "launches": timing["apply_launches"],
        "algorithmic_flops_per_launch": flops,
        "algorithmic_bytes_per_launch": bytes_,
        "hbm_achieved_gbs": achieved_gbs,
        "hbm_frac": achieved_gbs / PEAK_HBM_GBS if achieved_gbs else None,
    }
    out = {
        "metric": METRIC,
        "value": gdofs,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64+f32",
        "data": "synthetic: seeded jittered Kuhn-tet cube, manufactured u=sin(pi x)sin(pi y)sin(pi z)",
        "config": {
            "workload": f"3D cube {a.n}^3x6 Kuhn tets, N={a.order}, Dirichlet, p-MG "
                        f"{'->'.join(map(str, orders))} Chebyshev(3) + Jacobi-CG({a.coarse_iters}) "
                        f"coarse, flexible CG to rel res {a.rtol:g}, {a.precision}",
            "cells": cells, "dof": dof, "order": a.order, "rtol": a.rtol,
            "precision": a.precision,
        },
        "solves_per_s": a.steps / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "iterations": [st["iterations"] for st in stats],
        "rel_residual": max(st["rel_residual"] for st in stats),
        "converged": all(st["converged"] for st in stats),
        "discretisation_error_l2": err,
        "relative_discretisation_error": err / unorm,
        "fine_applies_per_solve": stats[-1]["fine_applies"],
        "smoother_applies_per_solve": stats[-1]["smoother_applies"],
        "roofline": roofline,
    }
    if not a.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.order, a.rtol)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
