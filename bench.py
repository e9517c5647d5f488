"""Benchmark of the DG Poisson solve on one MI355X.

Default workload (the 3D plain-FP64 configuration): unit cube, 64^3 hexes split
into 6 Kuhn tetrahedra each (1,572,864 tets), vertices jittered +-0.1h, cells
grouped by shape class within runs of 384 cells (ins_mesh_renumber, as OP2's
op_renumber would), order
N=3 (Np=20, 31,457,280 DoF), Dirichlet, manufactured u = sin(pi x) sin(pi y)
sin(pi z); flexible CG preconditioned by a p-multigrid V-cycle 3->2->1 with
Chebyshev(3) smoothing and Jacobi-CG on order 1; outer tolerance 1e-10 (true
relative residual). A step is one solve from a zero initial guess with rhs and
solution resident in HBM (device SoA layout).

Prints one JSON line. `value` is GDoF/s = DoF x solves / wall time.

Beside it the line carries the other forms of the same solve:
  mixed_vs_fp64  the same cube at rel res 1e-8, FP64 V-cycle vs FP32 V-cycle;
  coarse_amg     the same cube with the vertex-space AMG coarse solve;
  viscous        the velocity step's Helmholtz solve on the same cube (mass term
                 g0 Re/dt, inverse-mass PCG, rel res 1e-8);
  precision_form 48^3x6 tets, N=4, div(1/rho grad p) = f with rho = 1000 inside a
                 sphere, pure Neumann, rel res 1e-8: plain FP64 path, FP64 with the
                 AMG coarse solve, FP32 V-cycle inside the FP64 flexible CG;
  batch          256 independent 2D problems (log-uniform 500..50k triangles,
                 N in {2,3,4}, u = sin(k pi x) sin(m pi y), k,m in 1..4) solved
                 to 1e-8 one at a time on the plain FP64 path, then all at once
                 by the batch engine with the mixed V-cycle;
  serving        a request stream (exponential inter-arrival times, 70% 2D /
                 30% 3D, 1k..200k elements log-uniform, N in 1..4, rho contrast
                 in {1,10,1000} around a circle/sphere, pure Neumann with the
                 mean-zero constraint, rel res 1e-8) through the serving loop,
                 with throughput and p50/p99 latency per arrival rate.
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
# HBM bytes of one trace + element pass on the 64^3x6 cube at N=3, FP64 (cells grouped by shape
# class within runs of 384), measured with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of
# this bench, calibrated by the 512 MiB device copy of ins_device_peaks:
# traffic = 2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE.
MEASURED_TRAFFIC = {(64, 3, "fp64"): 3.077e9}
# order-1 apply of the coarse Jacobi-CG (k_apply_p1<3, double, false, true, true, true>: cell-by-cell
# values and metric), same calibration, counters over the 5250 launches of one headline solve: bytes
# per launch
MEASURED_TRAFFIC_P1 = {(64, "fp64"): 0.3095e9}
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_MATRIX_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    ap.add_argument("--coarse-iters", type=int, default=150,
                    help="Jacobi-CG iterations of the coarse solve (150: fastest on the 64^3 cube)")
    ap.add_argument("--cpu-n", type=int, default=20, help="cube size of the CPU baseline sample")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-mixed-leg", action="store_true")
    ap.add_argument("--skip-amg", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 256 MiB device copy 3 times first (counter calibration)")
    ap.add_argument("--precision-n", type=int, default=48, help="cube size of the precision form")
    ap.add_argument("--precision-order", type=int, default=4)
    ap.add_argument("--skip-precision", action="store_true")
    ap.add_argument("--precision-variants", type=str, default=",".join(PRECISION_VARIANTS),
                    help="with --precision-only: the solvers to run (profiling)")
    ap.add_argument("--skip-viscous", action="store_true")
    ap.add_argument("--precision-only", action="store_true",
                    help="run only the precision-form leg and print its JSON (a profiling aid)")
    ap.add_argument("--apply-only", type=int, default=0, metavar="ORDER",
                    help="run only --steps FP64 operator applies of this order on the --n cube and "
                         "print their mean time (a profiling aid for counter runs)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batch leg and print its JSON (a profiling aid)")
    ap.add_argument("--skip-oracle-check", action="store_true")
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="with --batch-only: leave out the one-at-a-time FP64 path (engine profiles)")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-rates", type=str, default="10,25,50,100,200,500,1000,2000",
                    help="req/s, comma list")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-viscous-vec", action="store_true")
    ap.add_argument("--skip-over-int", action="store_true")
    return ap.parse_args()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def config0_leg(ip, reps=5):
    """The CPU-runnable configuration: unit square 32x32 cells x 2 triangles, jitter 0.1h, N=3,
    u = sin(pi x) sin(pi y), Dirichlet, Jacobi-PCG to 1e-10; GPU (FP64) and oracle on the host."""
    import oracle as O
    mesh = ip.Mesh.square(32, jitter=0.1, seed=20261202)
    s = ip.PoissonSolver(mesh, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_JACOBI)
    s.set_tol_and_iter(1e-10, 0.0, 5000)
    s.init()
    xyz = s.node_coords()
    u = np.sin(np.pi * xyz[0]) * np.sin(np.pi * xyz[1])
    b = s.mass(2 * np.pi ** 2 * u)
    x = np.zeros_like(b)
    s.solve(b, x)
    t = []
    for _ in range(reps):
        x = np.zeros_like(b)
        t0 = time.perf_counter()
        s.solve(b, x)
        t.append(time.perf_counter() - t0)
    gpu_its = s.last_stats["iterations"]
    e = x - u
    err_gpu = float(np.sqrt(np.sum(e * s.mass(e))))
    coords, c2n = mesh.arrays()
    o = O.Oracle(coords, c2n, 3)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    t0 = time.perf_counter()
    xo, rc, its, rr = o.solve_jacobi(b, rtol=1e-10, max_iter=5000)
    t_cpu = time.perf_counter() - t0
    eo = xo - u
    err_cpu = float(np.sqrt(np.sum(eo * s.mass(eo))))
    dof = s.n_cells * s.Np
    return {"workload": "unit square 32x32x2 tris (jitter 0.1h), N=3, Dirichlet, Jacobi-PCG to 1e-10",
            "dof": dof, "gpu_ms": min(t) * 1e3, "gpu_iterations": gpu_its,
            "cpu_ms": t_cpu * 1e3, "cpu_iterations": its, "cpu_threads": threads,
            "l2_error_gpu": err_gpu, "l2_error_cpu": err_cpu,
            "rel_l2_gpu_vs_oracle": float(np.linalg.norm(x - xo) / np.linalg.norm(xo))}


def cpu_baseline(n, order, rtol, coarse_iters):
    import oracle as O
    import ins_poisson as ip
    mesh = cube(ip, n, 20261203)
    coords, c2n = mesh.arrays()
    o = O.Oracle(coords, c2n, order)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    u, f = manufactured_sin(o.node_coords())
    b = o.mass(f)
    orders = list(range(order, 0, -1))
    t0 = time.perf_counter()
    x, rc, it, rr, _ = o.solve_pmg(b, orders, rtol=rtol, coarse_iters=coarse_iters)
    dt = time.perf_counter() - t0
    dof = o.n_cells * o.Np
    return {"value": dof / dt * 1e-9, "unit": "GDoF/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"oracle p-MG FP64 solve (Jacobi-CG({coarse_iters}) coarse), cube {n}^3x6 tets "
                      f"N={order} ({dof} DoF), "
                      f"rtol {rtol:g}, {it} its, {dt:.1f} s, converged={rc == 0}"}


def cube(ip, n, seed):
    """Jittered Kuhn-tet cube, cells grouped by shape class within runs of 64 x 6 cells."""
    mesh = ip.Mesh.cube(n, 0.1, seed)
    mesh.renumber(ip.RENUMBER_SHAPE_CLASS_WINDOWED)
    return mesh


def cube_solver(ip, n, order, precision, rtol, coarse_iters):
    mesh = cube(ip, n, 20261202)
    s = ip.PoissonSolver(mesh, order)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_PMULTIGRID)
    orders = list(range(order, 0, -1))
    s.set_pmg(orders, pre_it=[1] * len(orders), post_it=[1] * len(orders),
              cheb_orders=[3] * len(orders), coarse_iters=coarse_iters)
    s.set_tol_and_iter(rtol, 0.0, 500)
    s.set_precision(ip.PRECISION_MIXED if precision == "mixed" else ip.PRECISION_FP64)
    s.init()
    return mesh, s, orders


def apply_leg(ip, torch, n, order, steps):
    """steps FP64 operator applies (device SoA, Dirichlet) on the bench cube, timed apart with HIP
    events on the solver's stream: the workload of counter runs on one kernel."""
    mesh = cube(ip, n, 20261202)
    s = ip.PoissonSolver(mesh, order)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_solver(ip.SOLVER_JACOBI)
    s.init()
    x = torch.rand((s.Np, s.stride), dtype=torch.float64, device="cuda")
    x[:, s.n_cells:] = 0
    y = torch.zeros_like(x)
    s.take_timing()
    s.set_profiling(True)
    for _ in range(max(steps, 1)):
        s.mult((x.data_ptr(), y.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
    torch.cuda.synchronize()
    s.set_profiling(False)
    t = s.take_timing()
    k = max(t["apply_launches"], 1)
    return {"workload": f"cube {n}^3x6 Kuhn tets, order {order}, FP64 apply", "cells": s.n_cells,
            "launches": t["apply_launches"], "trace_ms": t["trace_ms"] / k, "apply_ms": t["apply_ms"] / k}


def timed_solves(s, ip, torch, b_d, x_d, steps, warmup, profile=False, tag=None):
    """steps solves from a zero guess, rhs/solution resident in HBM; returns wall, latencies, stats."""
    def one():
        x_d.zero_()
        s.solve(b_d.data_ptr(), x_d.data_ptr(), layout=ip.LAYOUT_DEVICE_SOA)
        st = dict(s.last_stats)
        if tag:
            log(f"{tag}: solve done, {st['iterations']} its, rel res {st['rel_residual']:.2e}")
        return st

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    if profile:
        s.take_timing()
        s.set_profiling(True)
    lat, stats = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        stats.append(one())
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    timing = None
    if profile:
        s.set_profiling(False)
        timing = s.take_timing()
    return wall, lat, stats, timing


def mixed_vs_fp64_leg(ip, torch, s, b_d, x_d, dof, steps):
    """The cube at rel res 1e-8: FP64 V-cycle, then FP32 V-cycle, same solver and rhs."""
    out = {}
    for prec in ("fp64", "mixed"):
        s.set_precision(ip.PRECISION_MIXED if prec == "mixed" else ip.PRECISION_FP64)
        s.set_tol_and_iter(1e-8, 0.0, 500)
        s.init()
        wall, lat, stats, _ = timed_solves(s, ip, torch, b_d, x_d, steps, 1)
        out[prec] = {"gdofs": dof * steps / wall * 1e-9, "ms_per_solve": wall / steps * 1e3,
                     "iterations": [st["iterations"] for st in stats],
                     "fp64_applies": [st["fine_applies"] for st in stats],
                     "rel_residual": max(st["rel_residual"] for st in stats),
                     "converged": all(st["converged"] for st in stats)}
    out["speedup_mixed_over_fp64"] = out["mixed"]["gdofs"] / out["fp64"]["gdofs"]
    out["rtol"] = 1e-8
    return out


def coarse_amg_leg(ip, torch, s, b_d, x_d, dof, steps):
    """The cube with the AMG coarse solve (vertex space below order 1) instead of Jacobi-CG:
    FP64 and mixed at rel res 1e-8 (the like-for-like pair), FP64 at 1e-10 as well; set-up
    time reported apart."""
    out = {}
    for name, prec, rtol in (("fp64", "fp64", 1e-8), ("mixed", "mixed", 1e-8),
                             ("fp64_rtol_1e-10", "fp64", 1e-10)):
        s.set_precision(ip.PRECISION_MIXED if prec == "mixed" else ip.PRECISION_FP64)
        s.set_coarse(ip.COARSE_AMG)
        s.set_tol_and_iter(rtol, 0.0, 500)
        t0 = time.perf_counter()
        s.init()
        setup = time.perf_counter() - t0
        wall, lat, stats, _ = timed_solves(s, ip, torch, b_d, x_d, steps, 1)
        out[name] = {"rtol": rtol, "gdofs": dof * steps / wall * 1e-9, "ms_per_solve": wall / steps * 1e3,
                     "setup_s": setup, "iterations": [st["iterations"] for st in stats],
                     "fp64_applies": [st["fine_applies"] for st in stats],
                     "rel_residual": max(st["rel_residual"] for st in stats),
                     "converged": all(st["converged"] for st in stats),
                     "amg_levels": s.amg_levels()}
    out["speedup_mixed_over_fp64"] = out["mixed"]["gdofs"] / out["fp64"]["gdofs"]
    s.set_coarse(ip.COARSE_JACOBI_CG)
    return out


PRECISION_VARIANTS = ("plain_fp64", "mixed_jacobi_cg", "fp64_amg", "mixed_amg")


def precision_leg(ip, torch, n, order, steps, contrast=1000.0, seed=20261206, peaks=None,
                  variants=PRECISION_VARIANTS):
    """Variable-coefficient pure-Neumann cube: div(1/rho grad p) = f, rho = contrast inside a
    sphere of radius 0.25 at the centre (smoothed over 2h), f Gaussian projected to zero mean,
    rel res 1e-8. Runs the plain FP64 path (Jacobi-CG coarse), FP64 with the AMG coarse solve and
    the FP32 V-cycle with the AMG coarse solve inside the FP64 flexible CG; reports GDoF/s of each
    and the relative L2 (mass norm, mean removed) of each solution against the FP64 AMG one."""
    mesh = cube(ip, n, seed)
    s = ip.PoissonSolver(mesh, order)
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    s.set_nullspace(True)
    xyz = s.node_coords()
    k = 1.0 / density(xyz, 1.0 / n, contrast)
    s.set_factor(k)
    s.set_solver(ip.SOLVER_PMULTIGRID)
    orders = list(range(order, 0, -1))
    s.set_pmg(orders, pre_it=[1] * len(orders), post_it=[1] * len(orders),
              cheb_orders=[3] * len(orders), coarse_iters=150)
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((s.n_cells, s.Np))
    b = s.mass(f)
    b -= b.sum() / b.size
    dof = s.n_cells * s.Np

    def to_soa(x):
        soa = np.zeros((s.Np, s.stride))
        soa[:, :s.n_cells] = x.T
        return torch.from_numpy(soa).cuda()

    b_d = to_soa(b)
    x_d = torch.zeros_like(b_d)
    out = {"workload": f"cube {n}^3x6 Kuhn tets, N={order}, rho {contrast:g} in a sphere r=0.25 "
                       f"(band 2h), pure Neumann, mean-zero, rel res 1e-8", "dof": dof}
    sols = {}
    for name, prec, coarse in (("plain_fp64", ip.PRECISION_FP64, ip.COARSE_JACOBI_CG),
                               ("mixed_jacobi_cg", ip.PRECISION_MIXED, ip.COARSE_JACOBI_CG),
                               ("fp64_amg", ip.PRECISION_FP64, ip.COARSE_AMG),
                               ("mixed_amg", ip.PRECISION_MIXED, ip.COARSE_AMG)):
        if name not in variants:
            continue
        log(f"precision form: {name}")
        s.set_precision(prec)
        s.set_coarse(coarse)
        s.set_tol_and_iter(1e-8, 0.0, 1000)
        t0 = time.perf_counter()
        s.init()
        setup = time.perf_counter() - t0
        wall, lat, stats, _ = timed_solves(s, ip, torch, b_d, x_d, steps, 1, tag=name)
        x = x_d.cpu().numpy()[:, :s.n_cells].T.copy()
        x -= np.sum(x * s.mass(np.ones_like(x))) / np.sum(s.mass(np.ones_like(x)))
        sols[name] = x
        out[name] = {"gdofs": dof * steps / wall * 1e-9, "ms_per_solve": wall / steps * 1e3,
                     "setup_s": setup, "iterations": [st["iterations"] for st in stats],
                     "fp64_applies": [st["fine_applies"] for st in stats],
                     "rel_residual": max(st["rel_residual"] for st in stats),
                     "converged": all(st["converged"] for st in stats)}
    if tuple(variants) != PRECISION_VARIANTS:
        return out  # a profiling run of some of the solvers
    ref = sols["fp64_amg"]
    nref = float(np.sqrt(np.sum(ref * s.mass(ref))))
    for name in ("plain_fp64", "mixed_jacobi_cg", "mixed_amg"):
        d = sols[name] - ref
        out[name]["rel_l2_vs_fp64_amg"] = float(np.sqrt(np.sum(d * s.mass(d)))) / nref
    out["fp32_apply"] = fp32_apply_kernels(ip, torch, s, order,
                                           peak_f32=peaks["mfma_f32_tflops"] if peaks else None)
    # like-for-like ratios: the same solver and tolerance, FP32 inner iterations vs FP64
    out["speedup_mixed_over_fp64_jacobi_cg"] = out["mixed_jacobi_cg"]["gdofs"] / out["plain_fp64"]["gdofs"]
    out["speedup_mixed_over_fp64_amg"] = out["mixed_amg"]["gdofs"] / out["fp64_amg"]["gdofs"]
    # the fastest mixed solver over the plain FP64 path as configured
    out["speedup_mixed_amg_over_plain_fp64"] = out["mixed_amg"]["gdofs"] / out["plain_fp64"]["gdofs"]
    return out


def viscous_leg(ip, torch, n, order, steps, reynolds=100.0, g0=1.5, seed=20261207):
    """The viscous (Helmholtz) solve of the velocity step: -lap u + (g0 Re/dt) u = f on the
    Kuhn cube, Dirichlet, dt = h/(N+1)^2 (CFL 1 at unit velocity), inverse-mass PCG (factor
    dt/(g0 Re)) to rel res 1e-8 (the reference's viscous-solve r_tol),
    u = sin(pi x) sin(pi y) sin(pi z)."""
    mesh = cube(ip, n, seed)
    s = ip.PoissonSolver(mesh, order)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    dt = (1.0 / n) / (order + 1) ** 2
    mm = g0 * reynolds / dt
    s.set_mm_const(mm)
    s.set_solver(ip.SOLVER_INV_MASS)
    s.set_inv_mass_factor(1.0 / mm)
    s.set_tol_and_iter(1e-8, 0.0, 1000)
    s.init()
    xyz = s.node_coords()
    u, f = manufactured_sin(xyz)
    b = s.mass(f + mm * u)
    dof = s.n_cells * s.Np

    def to_soa(x):
        soa = np.zeros((s.Np, s.stride))
        soa[:, :s.n_cells] = x.T
        return torch.from_numpy(soa).cuda()

    b_d = to_soa(b)
    x_d = torch.zeros_like(b_d)
    wall, lat, stats, _ = timed_solves(s, ip, torch, b_d, x_d, steps, 1)
    x = x_d.cpu().numpy()[:, :s.n_cells].T
    e = x - u
    err = float(np.sqrt(np.sum(e * s.mass(e))) / np.sqrt(np.sum(u * s.mass(u))))
    return {"workload": f"cube {n}^3x6 Kuhn tets, N={order}, -lap u + (g0 Re/dt) u = f, Re={reynolds:g}, "
                        f"dt={dt:.3g}, Dirichlet, inverse-mass PCG, FP64, rel res 1e-8",
            "dof": dof, "mass_factor": mm, "gdofs": dof * steps / wall * 1e-9,
            "ms_per_solve": wall / steps * 1e3, "iterations": [st["iterations"] for st in stats],
            "rel_residual": max(st["rel_residual"] for st in stats),
            "converged": all(st["converged"] for st in stats),
            "relative_discretisation_error": err}


def mfma_issued_flops_per_cell(order):
    """Flops issued on the matrix cores per cell by the 3D element pass (sipg_mfma.hip):
    per 16-cell tile Q^2 16x16x4 products for G = D4 U and, per k-step of the second
    product, one 16x16x4 per full 16-row output tile and one 4x4x1 (16 blocks) per 4-row group."""
    np_ = (order + 1) * (order + 2) * (order + 3) // 6
    nfp = (order + 1) * (order + 2) // 2
    q, mtf, r4, j4 = (np_ + 3) // 4, np_ // 16, (np_ % 16 + 3) // 4, (nfp + 3) // 4
    ks2 = 3 * q + 4 * 4 * j4
    per_tile = q * q * 2048 + ks2 * (mtf * 2048 + r4 * 512)
    return per_tile / 16.0


def fp32_apply_kernels(ip, torch, s, order, reps=20, peak_f32=None):
    """FP32 operator apply of the solver's fine level with each element-pass kernel, timed
    with HIP events on the solver's stream (trace and element pass apart)."""
    res = {}
    cells = s.n_cells
    x = torch.rand((s.Np, s.stride), dtype=torch.float64, device="cuda")
    x[:, cells:] = 0
    y = torch.zeros_like(x)
    for name, kind in (("matrix_cores", ip.ELEMENT_PASS_MATRIX), ("vector", ip.ELEMENT_PASS_VECTOR)):
        s.set_element_pass(kind)
        s.init()
        for _ in range(3):
            s.mult_fp32((x.data_ptr(), y.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
        torch.cuda.synchronize()
        s.take_timing()
        s.set_profiling(True)
        for _ in range(reps):
            s.mult_fp32((x.data_ptr(), y.data_ptr()), layout=ip.LAYOUT_DEVICE_SOA)
        torch.cuda.synchronize()
        t = s.take_timing()
        s.set_profiling(False)
        apply_ms = t["apply_ms"] / max(t["apply_launches"], 1)
        trace_ms = t["trace_ms"] / max(t["trace_launches"], 1)
        useful = FLOPS_PER_CELL.get((3, order), 0) * cells
        r = {"element_pass_ms": apply_ms, "trace_ms": trace_ms,
             "useful_tflops": useful / (apply_ms + trace_ms) * 1e-9 if apply_ms > 0 else None}
        if name == "matrix_cores":
            issued = mfma_issued_flops_per_cell(order) * cells
            r["mfma_issued_tflops"] = issued / apply_ms * 1e-9
            r["mfma_frac_of_fp32_matrix_peak"] = issued / apply_ms * 1e-9 / PEAK_FP32_MATRIX_TFLOPS
            if peak_f32:
                r["mfma_frac_of_measured_fp32_matrix_peak"] = issued / apply_ms * 1e-9 / peak_f32
        res[name] = r
    s.set_element_pass(ip.ELEMENT_PASS_AUTO)
    res["speedup_matrix_over_vector"] = res["vector"]["element_pass_ms"] / res["matrix_cores"]["element_pass_ms"]
    del x, y
    return res


def batch_problems(ip, count, seed=20261204, wavenumbers=None):
    """2D batch set: log-uniform 500..50k triangles, N in {2,3,4}, sin(k pi x) sin(m pi y)
    (the (k, m) of each problem are appended to `wavenumbers` when a list is given)."""
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        tris = float(np.exp(rng.uniform(np.log(500), np.log(50000))))
        n = max(2, int(round(np.sqrt(tris / 2))))
        N = int(rng.integers(2, 5))
        k, m = (int(v) for v in rng.integers(1, 5, size=2))
        mesh = ip.Mesh.square(n, jitter=0.1, seed=seed + i)
        mesh.renumber(ip.RENUMBER_SHAPE_CLASS_WINDOWED)
        xyz = mesh.node_coords(N)
        f = ((k * k + m * m) * np.pi ** 2 * np.sin(k * np.pi * xyz[0]) * np.sin(m * np.pi * xyz[1]))
        probs.append(dict(mesh=mesh, order=N, rhs=f, rhs_nodal=True, rtol=1e-8, max_iter=500))
        if wavenumbers is not None:
            wavenumbers.append((k, m))
    return probs


def batch_leg(ip, count, plain=True):
    """Same problems solved one at a time on the plain FP64 path, then batched (mixed V-cycle).
    plain=False leaves the one-at-a-time path out (profiles of the batch engine alone)."""
    probs = batch_problems(ip, count)
    dofs = sum(p["mesh"].n_cells * ip.np_of(2, p["order"]) for p in probs)
    # plain FP64 path, one problem at a time (set-up untimed, solve timed)
    plain_t, plain_its, plain_ok = 0.0, [], True
    for i, p in enumerate(probs if plain else []):
        if i % 64 == 0:
            log(f"plain FP64 problem {i}/{len(probs)}")
        s = ip.PoissonSolver(p["mesh"], p["order"])
        s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
        s.set_solver(ip.SOLVER_PMULTIGRID)
        s.set_pmg(list(range(p["order"], 0, -1)), coarse_iters=40)
        s.set_tol_and_iter(1e-8, 0.0, 500)
        s.init()
        b = s.mass(p["rhs"])
        x = np.zeros_like(b)
        t0 = time.perf_counter()
        plain_ok &= s.solve(b, x)
        plain_t += time.perf_counter() - t0
        plain_its.append(s.last_stats["iterations"])
        del s
    out = {"problems": count, "dof": dofs}
    if plain:
        out["plain_fp64"] = {"gdofs": dofs / plain_t * 1e-9, "seconds": plain_t,
                             "solves_per_s": count / plain_t, "mean_iterations": float(np.mean(plain_its)),
                             "converged": bool(plain_ok)}
    # the batch engine in each precision: one warm-up batch of a few problems, then the whole set
    for name, prec in (("batched_mixed", ip.PRECISION_MIXED), ("batched_fp64", ip.PRECISION_FP64)):
        log(f"batch engine, {name}")
        warm = ip.Batch(2, precision=prec)
        warm.add(probs[:4])
        warm.run()
        del warm
        bt = ip.Batch(2, precision=prec)
        t0 = time.perf_counter()
        ids = bt.add(probs)
        t1 = time.perf_counter()
        bt.run()
        t2 = time.perf_counter()
        res = [bt.result(i) for i in ids]
        setup_parts = bt.setup_seconds()
        tile, tile_ms = bt.tuning()
        its = [r[2]["iterations"] for r in res]
        ok = all(r[0] == ip.OK for r in res)
        worst = max(r[2]["rel_residual"] for r in res)
        st = bt.stats()
        out[name] = {"gdofs": dofs / (t2 - t1) * 1e-9, "seconds": t2 - t1,
                     "setup_seconds": t1 - t0, "setup_parts_s": setup_parts,
                     "tile": tile, "tile_ms": tile_ms, "solves_per_s": count / (t2 - t1),
                     "gdofs_incl_setup": dofs / (t2 - t0) * 1e-9,
                     "lockstep_iterations": st["iterations_run"],
                     "mean_iterations": float(np.mean(its)), "max_rel_residual": worst,
                     "converged": bool(ok)}
        del bt
    out["speedup_batched_mixed_over_batched_fp64"] = (out["batched_mixed"]["gdofs"] /
                                                      out["batched_fp64"]["gdofs"])
    if plain:
        out["speedup_batched_mixed_over_plain_fp64"] = out["batched_mixed"]["gdofs"] / out["plain_fp64"]["gdofs"]
    return out


def density(xyz, h, contrast):
    """rho = contrast inside a circle/sphere of radius 0.25 at the centre, 1 outside, smoothed
    over a band of 2h (the level-set smoothing of src/kernels/ls_step.h:5-6)."""
    r = np.sqrt(sum((xyz[d] - 0.5) ** 2 for d in range(xyz.shape[0]))) - 0.25
    step = np.tanh(np.pi * r / (2.0 * h))
    return 0.5 * (1 + step) + 0.5 * contrast * (1 - step)


def serve_requests(ip, count, seed=20261205):
    """Request set: 70% 2D / 30% 3D, 1k..200k elements, N in 1..4, rho contrast {1,10,1000},
    pure Neumann with the mean-zero constraint, Gaussian f projected to zero mean."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        dim = 2 if rng.uniform() < 0.7 else 3
        cells = float(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        N = int(rng.integers(1, 5))
        contrast = float(rng.choice([1.0, 10.0, 1000.0]))
        if dim == 2:
            n = max(2, int(round(np.sqrt(cells / 2))))
            mesh = ip.Mesh.square(n, jitter=0.1, seed=seed + i)
        else:
            n = max(2, int(round((cells / 6) ** (1 / 3))))
            mesh = ip.Mesh.cube(n, 0.1, seed + i)
        mesh.renumber(ip.RENUMBER_SHAPE_CLASS_WINDOWED)
        xyz = mesh.node_coords(N)
        k = None if contrast == 1.0 else 1.0 / density(xyz, 1.0 / n, contrast)
        f = rng.standard_normal(xyz[0].shape)
        f -= f.mean()
        reqs.append(dict(mesh=mesh, order=N, rhs=f, rhs_nodal=True, factor=k, nullspace=True,
                         bc_types=np.ones(mesh.n_bfaces, dtype=np.int32), rtol=1e-8, max_iter=500))
    return reqs


def serve_leg(ip, count, rates, budget_gib=64):
    import threading
    reqs = serve_requests(ip, count)
    dofs = [r["mesh"].n_cells * ip.np_of(r["mesh"].dim, r["order"]) for r in reqs]
    out = {"requests": count, "dof_total": int(sum(dofs)), "hbm_budget_gib": budget_gib, "rates": []}
    # warm-up: graphs and code objects of both engines
    srv = ip.Server(budget_gib << 30)
    for r in reqs[:4]:
        srv.submit(**r)
    srv.wait_all()
    srv.close()
    for rate in rates:
        log(f"serving at {rate:g} req/s")
        srv = ip.Server(budget_gib << 30)
        rng = np.random.default_rng(int(rate))
        gaps = rng.exponential(1.0 / rate, size=count)
        ids = []

        def client():
            t_next = time.perf_counter()
            for r, g in zip(reqs, gaps):
                t_next += g
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids.append(srv.submit(**r))

        t0 = time.perf_counter()
        th = threading.Thread(target=client)
        th.start()
        th.join()
        srv.wait_all()
        wall = time.perf_counter() - t0
        lat, its, ok = [], [], True
        for rid in ids:
            rc, _, st, l, q = srv.wait(rid)
            lat.append(l)
            its.append(st["iterations"])
            ok &= rc == ip.OK
        peak = srv.stats()["peak_bytes"]
        srv.close()
        out["rates"].append({
            "offered_req_per_s": rate, "achieved_req_per_s": count / wall,
            "gdofs": sum(dofs) / wall * 1e-9,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                           "p99": float(np.percentile(lat, 99) * 1e3)},
            "mean_iterations": float(np.mean(its)), "all_converged": bool(ok),
            "peak_resident_gib": peak / 2 ** 30})
    return out


def oracle_check_leg(ip, torch):
    """Checker, run after every timed region: the bench's own solvers against the FP64 oracle
    on inputs the oracle solves in seconds (same settings, smaller meshes). Relative L2 in the
    mass norm; for manufactured solutions also both discretisation errors."""
    import oracle as O

    def mnorm(o, v):
        return float(np.sqrt(np.sum(v * o.mass(v))))

    def compare(o, x, xo, u=None):
        r = {"rel_l2_vs_oracle": mnorm(o, x - xo) / mnorm(o, xo)}
        if u is not None:
            eg, eo = mnorm(o, x - u), mnorm(o, xo - u)
            r["discretisation_error_gpu"] = eg
            r["discretisation_error_oracle"] = eo
            r["discretisation_error_match_2_digits"] = bool(abs(eg - eo) <= 0.005 * eo)
        return r

    out = {}
    # configuration 1 at 12^3 x 6: FP64 to 1e-10 (the headline solver) and mixed to 1e-8
    n = 12
    mesh = cube(ip, n, 20261202)
    s = ip.PoissonSolver(mesh, 3)
    s.set_bc_types(np.zeros(s.n_bfaces, dtype=np.int32))
    s.set_pmg([3, 2, 1], pre_it=[1] * 3, post_it=[1] * 3, cheb_orders=[3] * 3, coarse_iters=150)
    o = O.Oracle(*mesh.arrays(), 3)
    o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
    u, f = manufactured_sin(o.node_coords())
    b = o.mass(f)
    xo, rc, _, _, _ = o.solve_pmg(b, [3, 2, 1], coarse_iters=20, rtol=1e-12, max_iter=500)
    c1 = {"workload": f"cube {n}^3x6, N=3, Dirichlet, p-MG 3->2->1 Chebyshev(3) + Jacobi-CG(150)",
          "oracle_converged": rc == 0}
    for name, prec, rtol in (("fp64_rtol_1e-10", ip.PRECISION_FP64, 1e-10),
                             ("mixed_rtol_1e-8", ip.PRECISION_MIXED, 1e-8)):
        s.set_precision(prec)
        s.set_tol_and_iter(rtol, 0.0, 500)
        x = np.zeros_like(b)
        s.solve(b, x)
        c1[name] = compare(o, x, xo, u)
        c1[name]["iterations"] = s.last_stats["iterations"]
    out["config1_cube12"] = c1
    del s, o
    # configuration 2 at 8^3 x 6: N=4, rho 1000, pure Neumann; FP64 and mixed with the AMG coarse
    n, N = 8, 4
    mesh = cube(ip, n, 20261206)
    s = ip.PoissonSolver(mesh, N)
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    s.set_nullspace(True)
    o = O.Oracle(*mesh.arrays(), N)
    o.set_bc_types(np.ones(o.n_bfaces, dtype=np.int32))
    k = 1.0 / density(o.node_coords(), 1.0 / n, 1000.0)
    s.set_factor(k)
    o.set_factor(k)
    s.set_pmg([4, 3, 2, 1], pre_it=[1] * 4, post_it=[1] * 4, cheb_orders=[3] * 4, coarse_iters=150)
    s.set_coarse(ip.COARSE_AMG)
    f = np.random.default_rng(20261206).standard_normal((s.n_cells, s.Np))
    b = o.mass(f)
    b -= b.mean()
    xo, rc, _, _, _ = o.solve_pmg(b, [4, 3, 2, 1], coarse_iters=30, rtol=1e-12, max_iter=3000,
                                  nullspace=True)
    c2 = {"workload": f"cube {n}^3x6, N=4, rho 1000 sphere, pure Neumann, p-MG 4->1 + AMG coarse, "
                      "rel res 1e-8", "oracle_converged": rc == 0}
    for name, prec in (("fp64_amg", ip.PRECISION_FP64), ("mixed_amg", ip.PRECISION_MIXED)):
        s.set_precision(prec)
        s.set_tol_and_iter(1e-8, 0.0, 2000)
        x = np.zeros_like(b)
        s.solve(b, x)
        c2[name] = compare(o, x, xo)
        c2[name]["iterations"] = s.last_stats["iterations"]
    out["config2_cube8"] = c2
    del s, o
    # batch: the four smallest problems of the batch set, batched in each precision
    kms = []
    allp = batch_problems(ip, 256, wavenumbers=kms)
    pick = sorted(range(len(allp)), key=lambda i: allp[i]["mesh"].n_cells)[:4]
    probs = [allp[i] for i in pick]
    cb = {"workload": "the 4 smallest problems of the batch set (2D, Dirichlet, manufactured), rel res 1e-8"}
    refs = []
    for i, p in zip(pick, probs):
        o = O.Oracle(*p["mesh"].arrays(), p["order"])
        o.set_bc_types(np.zeros(o.n_bfaces, dtype=np.int32))
        xo, rc, _, _ = o.solve_jacobi(o.mass(p["rhs"]), rtol=1e-12, max_iter=100000)
        xyz = o.node_coords()
        k_, m_ = kms[i]
        refs.append((o, xo, np.sin(k_ * np.pi * xyz[0]) * np.sin(m_ * np.pi * xyz[1])))
    for name, prec in (("batched_mixed", ip.PRECISION_MIXED), ("batched_fp64", ip.PRECISION_FP64)):
        bt = ip.Batch(2, precision=prec)
        ids = bt.add(probs)
        bt.run()
        res = []
        for i, (o, xo, u) in zip(ids, refs):
            _, x, _ = bt.result(i)
            res.append(compare(o, x, xo, u))
        cb[name] = {"max_rel_l2_vs_oracle": float(max(r["rel_l2_vs_oracle"] for r in res)),
                    "discretisation_errors_match_2_digits": all(r["discretisation_error_match_2_digits"]
                                                                for r in res),
                    "discretisation_error_gpu": [r["discretisation_error_gpu"] for r in res],
                    "discretisation_error_oracle": [r["discretisation_error_oracle"] for r in res]}
        del bt
    out["batch"] = cb
    # serving: four small requests through the server (pure Neumann, rho contrast, mean-zero)
    reqs = sorted(serve_requests(ip, 40), key=lambda r: r["mesh"].n_cells)[:4]
    srv = ip.Server(8 << 30)
    rids = [srv.submit(**r) for r in reqs]
    srv.wait_all()
    errs = []
    for rid, r in zip(rids, reqs):
        _, x, _, _, _ = srv.wait(rid)
        o = O.Oracle(*r["mesh"].arrays(), r["order"])
        o.set_bc_types(r["bc_types"])
        if r["factor"] is not None:
            o.set_factor(r["factor"])
        b = o.mass(r["rhs"])
        b -= b.mean()
        xo, rc, _, _, _ = o.solve_pmg(b, list(range(r["order"], 0, -1)), coarse_iters=30, rtol=1e-12,
                                      max_iter=5000, nullspace=True)
        errs.append(mnorm(o, x - xo) / mnorm(o, xo))
    srv.close()
    out["serving"] = {"workload": "the 4 smallest of 40 serving requests", "max_rel_l2_vs_oracle": float(max(errs))}
    # viscous solve with slip walls at 3^3 x 6, N=3: the oracle's ViscousSolver restatement
    n = 3
    mesh = cube(ip, n, 20261208)
    s = ip.ViscousVecSolver(mesh, 3)
    bct = vec_bc_types(mesh)
    s.set_bc_types(bct)
    o = O.Oracle(*mesh.arrays(), 3, faces=mesh.faces())
    mm = 1.5 * 100.0 / ((1.0 / n) / 16)
    s.set_mm_const(mm)
    o.set_mm_const(mm)
    s.set_preconditioner(ip.VEC_PRE_FACTOR_INV_MASS)
    s.set_inv_mass_factor(1.0 / mm)
    s.set_tol_and_iter(1e-8, 0.0, 1000)
    rng = np.random.default_rng(20261208)
    g = [np.ascontiguousarray(rng.standard_normal((s.n_bfaces, s.Nfp))) for _ in range(3)]
    s.set_bcs(g)
    f = mm * rng.standard_normal((3, s.n_cells, s.Np))
    x = np.zeros_like(f)
    s.solve(f, x)
    xo, rc, _, _ = o.solve_vec(np.stack(bct), f, bc=np.stack(g), rtol=1e-12, precond=2, factor=1.0 / mm)
    out["viscous_vec_cube3"] = {"workload": "cube 3^3x6, N=3, 3 components, slip walls, FACTOR_INV_MASS PCG, rel res 1e-8",
                                "oracle_converged": rc == 0,
                                "rel_l2_vs_oracle": float(np.linalg.norm(x - xo) / np.linalg.norm(xo)),
                                "iterations": s.last_stats["iterations"]}
    del s, o
    # over-integration at 4^3 x 6, N=4: the true residual of the GPU solution under the oracle's
    # over-integrated operator (the oracle restates the operator, not its solver)
    n, N = 4, 4
    mesh = cube(ip, n, 20261206)
    s = ip.PoissonSolver(mesh, N)
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    s.set_nullspace(True)
    o = O.Oracle(*mesh.arrays(), N, faces=mesh.faces())
    o.set_bc_types(np.ones(o.n_bfaces, dtype=np.int32))
    kn = 1.0 / density(o.node_coords(), 1.0 / n, 1000.0)
    s.set_factor(kn)
    o.set_factor(kn)
    pv, ps = oi_points(ip, mesh, N)
    kv = 1.0 / density(np.moveaxis(pv, 2, 0), 1.0 / n, 1000.0)
    ks = 1.0 / density(np.moveaxis(ps, 2, 0), 1.0 / n, 1000.0)
    s.set_factor_oi(kv)
    s.set_factor_surf_oi(ks)
    s.set_pmg([4, 3, 2, 1], pre_it=[1] * 4, post_it=[1] * 4, cheb_orders=[3] * 4, coarse_iters=150)
    s.set_coarse(ip.COARSE_AMG)
    b = o.mass(np.random.default_rng(20261206).standard_normal((s.n_cells, s.Np)))
    b -= b.mean()
    coi = {"workload": f"cube {n}^3x6, N=4, rho 1000 sphere, over-integrated, pure Neumann, AMG coarse, rel res 1e-8"}
    rule = ip.oi_rule(3, N)
    for name, prec in (("fp64_amg", ip.PRECISION_FP64), ("mixed_amg", ip.PRECISION_MIXED)):
        s.set_precision(prec)
        s.set_tol_and_iter(1e-8, 0.0, 2000)
        x = np.zeros_like(b)
        s.solve(b, x)
        r = b - o.mult_oi(rule, kv, ks, x)
        coi[name] = {"rel_residual_under_oracle_operator": float(np.linalg.norm(r) / np.linalg.norm(b)),
                     "iterations": s.last_stats["iterations"]}
    out["over_int_cube4"] = coi
    return out


FV3 = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]])


def vec_bc_types(mesh):
    """Per-component bc types of the channel set-up on the unit cube: inflow x=0 no-slip
    (Dirichlet), outflow x=1 Neumann, slip walls on y=0, y=1, z=0, z=1."""
    coords, c2n = mesh.arrays()
    _, _, b2c, bnum = mesh.faces()
    cen = coords[c2n[b2c[:, None], FV3[bnum]]].mean(axis=1)
    t = np.full(len(b2c), 2, dtype=np.int32)
    t[cen[:, 0] < 1e-9] = 0
    t[cen[:, 0] > 1 - 1e-9] = 1
    return [t.copy(), t.copy(), t.copy()]


def viscous_vec_leg(ip, torch, n, order, steps, reynolds=100.0, g0=1.5, seed=20261208):
    """The velocity step's viscous solve with slip walls (Matrix3Vec: ViscousMatrix3D /
    ViscousSolver3D): three components, -lap u + (g0 Re/dt) u = f with a smooth f, inflow no-slip,
    outflow Neumann, slip on the four side walls, the reference's single-phase preconditioner
    (FACTOR_INV_MASS, factor dt/(g0 Re)), FP64, rel res 1e-8."""
    mesh = cube(ip, n, seed)
    s = ip.ViscousVecSolver(mesh, order)
    s.set_bc_types(vec_bc_types(mesh))
    dt = (1.0 / n) / (order + 1) ** 2
    mm = g0 * reynolds / dt
    s.set_mm_const(mm)
    s.set_preconditioner(ip.VEC_PRE_FACTOR_INV_MASS)
    s.set_inv_mass_factor(1.0 / mm)
    s.set_tol_and_iter(1e-8, 0.0, 1000)
    # a smooth forcing (the velocity step's rhs is mass times a smooth field) and homogeneous data
    g = [np.zeros((s.n_bfaces, s.Nfp)) for _ in range(3)]
    s.set_bcs(g)
    dof = 3 * s.n_cells * s.Np
    xyz = ip.PoissonSolver(mesh, order).node_coords()
    smooth, _ = manufactured_sin(xyz)

    def to_soa(x):
        soa = np.zeros((s.Np, s.stride))
        soa[:, :s.n_cells] = x.T
        return torch.from_numpy(soa).cuda()

    mass = ip.PoissonSolver(mesh, order)
    b_d = [to_soa(mass.mass(mm * (1.0 + 0.3 * c) * smooth)) for c in range(3)]
    del mass
    x_d = [torch.zeros_like(b_d[0]) for _ in range(3)]
    bp, xp = [t.data_ptr() for t in b_d], [t.data_ptr() for t in x_d]
    t0 = time.perf_counter()
    s.solve(bp, xp, layout=ip.LAYOUT_DEVICE_SOA)  # warm-up (set-up included)
    setup_and_first = time.perf_counter() - t0
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for t in x_d:
            t.zero_()
        s.solve(bp, xp, layout=ip.LAYOUT_DEVICE_SOA)
        stats.append(s.last_stats)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"workload": f"cube {n}^3x6 Kuhn tets, N={order}, 3 components, -lap u + (g0 Re/dt) u = f, "
                        f"Re={reynolds:g}, dt={dt:.3g}, inflow no-slip / outflow Neumann / 4 slip walls, "
                        f"FACTOR_INV_MASS PCG, FP64, rel res 1e-8",
            "dof": dof, "gdofs": dof * steps / wall * 1e-9, "ms_per_solve": wall / steps * 1e3,
            "first_solve_with_setup_s": setup_and_first,
            "iterations": [st["iterations"] for st in stats],
            "rel_residual": max(st["rel_residual"] for st in stats),
            "converged": all(st["converged"] for st in stats)}


def oi_points(ip, mesh, order):
    """Physical coordinates of the over-integration points: volume (cells, nv, 3) and faces
    (cells, 4*ns, 3)."""
    coords, c2n = mesh.arrays()
    vx, vw, sb, sw = ip.oi_rule(3, order)
    RV = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)
    lam = np.linalg.solve(np.vstack([RV.T, np.ones(4)]), np.vstack([vx.T, np.ones(len(vw))])).T  # (nv, 4)
    X = coords[c2n]  # (E, 4, 3)
    pv = np.einsum("qv,evx->eqx", lam, X)
    ps = np.concatenate([np.einsum("qk,ekx->eqx", sb, X[:, FV3[f]]) for f in range(4)], axis=1)
    return pv, ps


def over_int_leg(ip, torch, n, order, steps, contrast=1000.0, seed=20261206):
    """The precision form's problem with the coefficient over-integrated (the reference's
    [pressure-solve] over_int = 1: FactorPoissonMatrixFreeDiagOI3D): 1/rho sampled at the
    cubature points of the cells and faces, the nodal 1/rho for the penalty and the coarser
    levels; p-MG with the AMG coarse solve, FP64 and mixed precision, rel res 1e-8."""
    mesh = cube(ip, n, seed)
    s = ip.PoissonSolver(mesh, order)
    s.set_bc_types(np.ones(s.n_bfaces, dtype=np.int32))
    s.set_nullspace(True)
    h = 1.0 / n
    s.set_factor(1.0 / density(s.node_coords(), h, contrast))
    pv, ps = oi_points(ip, mesh, order)
    s.set_factor_oi(1.0 / density(np.moveaxis(pv, 2, 0), h, contrast))
    s.set_factor_surf_oi(1.0 / density(np.moveaxis(ps, 2, 0), h, contrast))
    del pv, ps
    s.set_solver(ip.SOLVER_PMULTIGRID)
    orders = list(range(order, 0, -1))
    s.set_pmg(orders, pre_it=[1] * len(orders), post_it=[1] * len(orders),
              cheb_orders=[3] * len(orders), coarse_iters=150)
    s.set_coarse(ip.COARSE_AMG)
    rng = np.random.default_rng(seed)
    b = s.mass(rng.standard_normal((s.n_cells, s.Np)))
    b -= b.sum() / b.size
    dof = s.n_cells * s.Np
    soa = np.zeros((s.Np, s.stride))
    soa[:, :s.n_cells] = b.T
    b_d = torch.from_numpy(soa).cuda()
    x_d = torch.zeros_like(b_d)
    out = {"workload": f"cube {n}^3x6 Kuhn tets, N={order}, rho {contrast:g} in a sphere r=0.25 (band 2h), "
                       f"coefficient over-integrated ({ip.oi_rule(3, order)[1].size} volume, "
                       f"{ip.oi_rule(3, order)[3].size} per-face points), pure Neumann, AMG coarse, rel res 1e-8",
           "dof": dof}
    for name, prec in (("fp64_amg", ip.PRECISION_FP64), ("mixed_amg", ip.PRECISION_MIXED)):
        log(f"over-integration: {name}")
        s.set_precision(prec)
        s.set_tol_and_iter(1e-8, 0.0, 1000)
        s.init()
        wall, lat, stats, _ = timed_solves(s, ip, torch, b_d, x_d, steps, 1, tag="oi_" + name)
        out[name] = {"gdofs": dof * steps / wall * 1e-9, "ms_per_solve": wall / steps * 1e3,
                     "iterations": [st["iterations"] for st in stats],
                     "rel_residual": max(st["rel_residual"] for st in stats),
                     "converged": all(st["converged"] for st in stats)}
    out["speedup_mixed_over_fp64_amg"] = out["mixed_amg"]["gdofs"] / out["fp64_amg"]["gdofs"]
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import ins_poisson as ip

    peaks = ip.device_peaks()
    log(f"device peaks: {peaks}")
    if a.precision_only:
        res = precision_leg(ip, torch, a.precision_n, a.precision_order, a.steps, peaks=peaks,
                            variants=tuple(v for v in a.precision_variants.split(",") if v))
        print(json.dumps({"precision_form": res}), flush=True)
        return
    if a.batch_only:
        print(json.dumps({"batch": batch_leg(ip, a.batch_problems, plain=not a.batch_no_plain)}), flush=True)
        return
    if a.apply_only:
        print(json.dumps({"apply_only": apply_leg(ip, torch, a.n, a.apply_only, a.steps)}), flush=True)
        return
    if a.copy_probe:
        # a plain copy of known size: 256 MiB read + 256 MiB written per launch
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mesh, s, orders = cube_solver(ip, a.n, a.order, a.precision, a.rtol, a.coarse_iters)
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
    del f, xyz, f_d

    wall, lat, stats, timing = timed_solves(s, ip, torch, b_d, x_d, a.steps, a.warmup, profile=True)
    gdofs = dof * a.steps / wall * 1e-9

    # discretisation error of the last solution (nodal mass norm)
    x = x_d.cpu().numpy()[:, :s.n_cells].T
    e = x - u
    err = float(np.sqrt(np.sum(e * s.mass(e))))
    unorm = float(np.sqrt(np.sum(u * s.mass(u))))
    del x, e, u

    key = (3, a.order)
    cells = s.n_cells
    t_apply = (timing["trace_ms"] + timing["apply_ms"]) / max(timing["apply_launches"], 1) * 1e-3
    flops = FLOPS_PER_CELL.get(key, 0) * cells
    bytes_ = BYTES_PER_CELL_F64.get(key, 0) * cells
    achieved_tf = flops / t_apply * 1e-12 if t_apply > 0 else None
    achieved_gbs = bytes_ / t_apply * 1e-9 if t_apply > 0 else None
    fine = {
        "bound": "mfma",
        "achieved": achieved_tf,
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved_tf / PEAK_FP64_TFLOPS if achieved_tf else None,
        "traffic": MEASURED_TRAFFIC.get((a.n, a.order, a.precision)),
        "kernel": f"SIPG apply order {a.order} FP64 (k_trace + k_apply), {cells} cells",
        "avg_launch_ms": t_apply * 1e3,
        "launches": timing["apply_launches"],
        "algorithmic_flops_per_launch": flops,
        "algorithmic_bytes_per_launch": bytes_,
        "hbm_achieved_gbs": achieved_gbs,
        "hbm_frac": achieved_gbs / PEAK_HBM_GBS if achieved_gbs else None,
    }
    tr = fine["traffic"]
    if tr and t_apply > 0:
        fine["traffic_gbs"] = tr / t_apply * 1e-9
        fine["traffic_over_algorithmic"] = tr / bytes_ if bytes_ else None
    # The kernel with the most device time in this solve is the order-1 apply of the coarse
    # Jacobi-CG (HBM-bound); one launch in four is timed with HIP events on the solver's stream.
    t_p1 = timing["coarse_apply_ms"] / max(timing["coarse_apply_launches"], 1) * 1e-3
    bytes_p1 = BYTES_PER_CELL_F64[(3, 1)] * cells
    ach_p1 = bytes_p1 / t_p1 * 1e-9 if t_p1 > 0 else None
    roofline = {
        "bound": "hbm",
        "achieved": ach_p1,
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": ach_p1 / PEAK_HBM_GBS if ach_p1 else None,
        "traffic": MEASURED_TRAFFIC_P1.get((a.n, a.precision)),
        "kernel": f"order-1 SIPG apply of the coarse Jacobi-CG, FP64 (k_apply_p1), {cells} cells",
        "avg_launch_ms": t_p1 * 1e3,
        "launches_timed": timing["coarse_apply_launches"],
        "algorithmic_bytes_per_launch": bytes_p1,
    }
    roofline["measured_peak"] = peaks["copy_gbs"]
    roofline["frac_of_measured_peak"] = ach_p1 / peaks["copy_gbs"] if ach_p1 else None
    fine["measured_peak"] = peaks["mfma_f64_tflops"]
    fine["frac_of_measured_peak"] = achieved_tf / peaks["mfma_f64_tflops"] if achieved_tf else None
    if roofline["traffic"] and t_p1 > 0:
        roofline["traffic_gbs"] = roofline["traffic"] / t_p1 * 1e-9
        roofline["traffic_over_algorithmic"] = roofline["traffic"] / bytes_p1
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
            "workload": f"3D cube {a.n}^3x6 Kuhn tets (cells grouped by shape class in runs of 384), N={a.order}, "
                        f"Dirichlet, p-MG "
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
        "roofline_fine_apply": fine,
        "device_peaks_measured": peaks,
    }
    log(f"plain leg: {gdofs:.4f} GDoF/s, its {out['iterations']}")
    if not a.no_mixed_leg:
        log("mixed vs fp64 leg")
        out["mixed_vs_fp64"] = mixed_vs_fp64_leg(ip, torch, s, b_d, x_d, dof, 2)
    if not a.skip_amg:
        log("AMG coarse leg")
        out["coarse_amg"] = coarse_amg_leg(ip, torch, s, b_d, x_d, dof, 2)
    del s, b_d, x_d, mesh
    torch.cuda.empty_cache()
    if not a.skip_viscous:
        log("viscous (Helmholtz) leg")
        out["viscous"] = viscous_leg(ip, torch, a.n, a.order, 3)
        torch.cuda.empty_cache()
    if not a.skip_viscous_vec:
        log("viscous solve with slip walls (vector) leg")
        out["viscous_vec"] = viscous_vec_leg(ip, torch, a.n, a.order, 3)
        torch.cuda.empty_cache()
    if not a.skip_over_int:
        log("over-integration leg")
        out["over_int"] = over_int_leg(ip, torch, a.precision_n, a.precision_order, 2)
        torch.cuda.empty_cache()
    if not a.skip_precision:
        log("precision form leg")
        out["precision_form"] = precision_leg(ip, torch, a.precision_n, a.precision_order, 2,
                                              peaks=peaks)
        torch.cuda.empty_cache()
    if not a.skip_batch:
        log("batch leg")
        out["batch"] = batch_leg(ip, a.batch_problems)
    if not a.skip_serve:
        rates = [float(r) for r in a.serve_rates.split(",") if r]
        log("serving leg")
        out["serving"] = serve_leg(ip, a.serve_requests, rates)
    if not a.skip_cpu:
        log("configuration 0 (2D, GPU and oracle)")
        out["config0"] = config0_leg(ip)
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.order, a.rtol, a.coarse_iters)
    if not a.skip_oracle_check:
        log("oracle check (every form at oracle-feasible sizes)")
        out["oracle_check"] = oracle_check_leg(ip, torch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
