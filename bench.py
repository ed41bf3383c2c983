"""Benchmark of the projection-method step on one MI355X.

Workload: the Karman vortex street on a 2048x1024 MAC grid (inflow u=1 with a
seeded 1e-3 inlet perturbation, Air outlet column, free-slip walls, cylinder
of radius 0.05H at (0.25L, 0.5H)), MacCormack advection of velocity and 3 dye
channels, mixed-precision geometric multigrid V(2,2) with red-black
Gauss-Seidel down to 16x8 and an FP64 direct coarsest solve, every pressure
solve to 1e-6 relative residual, dt at CFL 0.5.

A step is one fused fh_step over the resident fields (inputs already in HBM);
the timed region is K steps bracketed by device synchronisation.  Prints one
JSON line.  `roofline` is measured live: the dominant kernel's launches are
bracketed by HIP events on the engine's own stream during the timed steps.
`cpu_baseline` times the FP64 C oracle (one thread) on a bounded sample of the
same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s + ms to 1e-6 pressure residual, learned vs plain MG, 1 GPU"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per cell of one launch of each traceable
# kernel: every array element read or written once, neighbour reuse perfect.
#   rb_sweep_fine: one colour half-sweep = half of a full red-black sweep's
#     13 B/cell (x 4 read + 4 write, rhs 4, label 1).
#   maccormack_faces: u^n, v^n 8 + forward-pass u, v 8 + out u, v 8 + label 1
#     (u^n is both the carrier and the limiter's stencil: read once).
#   advect_faces: carrier = source u, v 8 + out u, v 8 + label 1.
#   dc_update: e 4 + p 8 read + 8 write + div 8 + label 1 + r32 4; under the
#     flexible-CG solver the traced launch (the solve's first cycle) also
#     writes the FP64 residual r, 8 more (algo_bytes below).
#   smooth_restrict_fine: b 4 + label 1 + x 4 written + coarse b 1 (x starts
#     at zero, so it is not read).
#   prolong_smooth_fine: x 4 + b 4 + label 1 + coarse e 1 + x 4 written.
#   fcg_dots: z 4 + previous direction 4 + r 8 + info 1.
#   fcg_update: z 4 + previous direction 4 + p 8 + r 8 + info 1 read;
#     direction 4 + p 8 + r 8 + r32 4 written.
SOLVER_NAMES = {
    "mg-dc": "MG-DC V(2,2) to 16x8",
    "mg-fcg": "one MG-DC V(2,2) cycle (to 16x8), then flexible CG preconditioned by it",
}

def algo_bytes(kernel, solver):
    """Algorithmic bytes per cell of one traced launch of `kernel`."""
    if kernel == "dc_update" and solver == "mg-fcg":
        return ALGO_BYTES_PER_CELL[kernel] + 8.0
    return ALGO_BYTES_PER_CELL.get(kernel)


ALGO_BYTES_PER_CELL = {
    "smooth_restrict_fine": 10.0,
    "prolong_smooth_fine": 14.0,
    "rb_sweep_fine": 6.5,
    "maccormack_faces": 25.0,
    "advect_faces": 17.0,
    "maccormack_dye": 45.0,
    "advect_dye": 33.0,
    "dc_update": 33.0,
    "fcg_dots": 17.0,
    "fcg_update": 49.0,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--kernel", default=None,
                    help="kernel traced for the roofline (default: the solver's fine-level "
                         "update, dc_update or fcg_update)")
    ap.add_argument("--solver", default="mg-fcg", choices=sorted(SOLVER_NAMES),
                    help="pressure solver of the GPU legs (tolerance 1e-6 either way)")
    ap.add_argument("--cpu-steps", type=int, default=4, help="oracle steps timed (0 = skip)")
    ap.add_argument("--batch-scenes", type=int, default=64,
                    help="scenes of the batched decaying-turbulence run (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--large-n", type=int, default=8192,
                    help="side of the large single-domain plume (0 = skip)")
    ap.add_argument("--large-steps", type=int, default=4)
    ap.add_argument("--large-learned", type=int, default=1,
                    help="also step the large domain from the learned warm start (0 = skip)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (the known-size "
                         "copy kernel that scales hardware traffic counters)")
    args = ap.parse_args()
    if args.kernel is None:
        args.kernel = "fcg_update" if args.solver == "mg-fcg" else "dc_update"
    return args


PMC_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic_karman_2048x1024.json")


def measured_traffic(kernel, nx, ny):
    """HBM bytes per launch of `kernel` from the committed counter passes
    (FETCH_SIZE and WRITE_SIZE, scaled by a known-size copy), for the
    2048x1024 workload they were taken on; None otherwise."""
    if (nx, ny) != (2048, 1024) or not os.path.exists(PMC_TRAFFIC):
        return None
    with open(PMC_TRAFFIC) as f:
        k = json.load(f)["kernels"].get(kernel)
    return None if k is None else k["traffic_bytes"]


def run_gpu(args, unet=None, trace=True, tiles=None):
    """The Karman workload: W warm-up steps, then K timed steps.  unet: solve
    every pressure from the learned warm start instead of the previous p."""
    from flowlab_amd import scenes as S
    from flowlab_amd.simulation import Simulation

    nx, ny = args.nx, args.ny
    g, scene = S.karman_street(nx, ny)
    params = dict(scheme="macCormack", cfl_target=0.5, solver=args.solver, pressure_iters=100,
                  tol=1e-6, residual_every=0)
    sim = Simulation(nx, ny, params, readback=False)
    sim.reset(scene)
    e = sim.stepper.engine
    if unet is not None:
        e.set_warm_start(unet)
    for _ in range(args.warmup):
        sim.step()
    if tiles is None:
        tiles = e.tune_tiles(reps=10)["tiles"]  # autotuned tile shape per level
    else:
        for level, name in enumerate(tiles):
            e.set_tile(name, level)
    # Timed pass: K steps with nothing else in the loop (per-step phase
    # timing would make the host wait for every step's end).
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    # Instrumented pass over the same K steps (the run is deterministic:
    # reset, the same warm-up, then K steps with solve times, cycle counts,
    # residuals and the roofline kernel's launch times recorded).
    sim.reset(scene)
    for _ in range(args.warmup):
        sim.step()
    e.set_timing(True)
    if trace:
        e.trace_kernel(args.kernel)
    solve_ms, iters, hist_max = [], [], 0.0
    for _ in range(args.steps):
        sim.step()
        ph = e.timings()
        solve_ms.append(ph["solve"])
        iters.append(sim.iters)
        hist_max = max(hist_max, float(e.solve_history()[-1]))
    e.synchronize()
    n_launch, k_ms, bracket_ms = 0, 0.0, 0.0
    if trace:
        n_launch, k_ms = e.trace_stats()
        bracket_ms = e.trace_overhead(200)
        e.trace_kernel(None)
    e.set_timing(False)
    e.set_warm_start(None)
    return dict(sim=sim, elapsed=elapsed, solve_ms=solve_ms, iters=iters, hist_max=hist_max,
                n_launch=n_launch, k_ms=k_ms, bracket_ms=bracket_ms, g=g, tiles=tiles)


def run_learned(args, plain):
    """The same workload with the trained U-Net warm start (weights in
    flowlab_amd/weights), against the plain run's solves."""
    from flowlab_amd import learned as L
    if not os.path.exists(L.WEIGHTS):
        return None
    u = L.UNet()
    u.load(L.WEIGHTS)
    r = run_gpu(args, unet=u, trace=False, tiles=plain["tiles"])
    K = args.steps
    ms_l = sum(r["solve_ms"]) / K
    ms_p = sum(plain["solve_ms"]) / K
    return {
        "ms_per_solve": ms_l,
        "vcycles_per_solve": sum(r["iters"]) / K,
        "max_relative_residual": r["hist_max"],
        "solves_per_s_vs_plain": ms_p / ms_l,
        "cell_steps_per_s": args.nx * args.ny * K / r["elapsed"],
        "weights": os.path.relpath(L.WEIGHTS, ROOT),
    }


def run_batched(args):
    """64 decaying-turbulence scenes of uneven size (256^2 .. 1024^2) in one
    batch: one engine per scene on its own stream, stepped from a host thread
    pool so the scenes' kernels interleave; tiles autotuned per scene; dt per
    scene at CFL 0.5; pressure by mg-fcg to 1e-6."""
    from flowlab_amd import scenes as S
    from flowlab_amd.batch import HipBatch

    scenes = S.turbulence_batch(args.batch_scenes)
    b = HipBatch([g for g, _ in scenes])
    for s, (_, sc) in enumerate(scenes):
        b.reset(s, sc)
    kw = dict(scheme="macCormack", cfl=0.5, dt_max=1.0, tol=1e-6, max_cycles=100)
    b.step(**kw)
    tune = b.tune(reps=5)
    for _ in range(max(args.warmup - 1, 1)):
        b.step(**kw)
    iters, res_max = [], 0.0
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        b.step(**kw)
        _, it, res = b.results()
        iters.append(float(it.mean()))
        res_max = max(res_max, float(res.max()))
    elapsed = time.perf_counter() - t0
    cells = b.cells
    sizes = sorted({(g.nx, g.ny) for g, _ in scenes})
    return {
        "workload": f"{len(scenes)} decaying-turbulence scenes, sizes {sizes} drawn by "
                    "SplitMix64(20240612), square-obstacle lattices, k^-3 velocity RMS 1, "
                    "MacCormack, one MG-DC cycle then flexible CG to 1e-6, CFL 0.5",
        "cells": cells,
        "steps": args.batch_steps,
        "value": cells * args.batch_steps / elapsed,
        "unit": "cell-steps/s",
        "ms_per_step": elapsed * 1e3 / args.batch_steps,
        "vcycles_per_solve_mean": sum(iters) / len(iters),
        "max_relative_residual": res_max,
        "largest_scene_tiles": tune["tiles"],
        "execution": "one engine and stream per scene, 16 host threads taking scenes largest "
                     "first",
    }


LARGE_KERNELS = ("dc_update", "smooth_restrict_fine", "prolong_smooth_fine", "maccormack_faces",
                 "advect_faces", "maccormack_dye", "advect_dye")


def run_large(args):
    """One n x n domain (default 8192^2, 67 M cells): closed box, 200 seeded
    circular obstacles, 16 buoyant heat sources (buoyancy 1), MacCormack,
    MG-DC to 1e-6.  Times K steps, replays them instrumented (solve times,
    cycles), then traces each large kernel for one step and reports its
    algorithmic GB/s against the HBM peak."""
    from flowlab_amd import scenes as S
    from flowlab_amd.simulation import Simulation

    n = args.large_n
    g, scene = S.large_plume(n)
    params = dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 * g.h, solver=args.solver,
                  pressure_iters=100, tol=1e-6, residual_every=0, buoyancy=1.0)
    sim = Simulation(n, n, params, readback=False)
    sim.reset(scene)
    e = sim.stepper.engine
    for _ in range(2):
        sim.step()
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.large_steps):
        sim.step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    sim.reset(scene)  # the same steps again, instrumented
    for _ in range(2):
        sim.step()
    e.set_timing(True)
    solve_ms, iters, res = [], [], 0.0
    for _ in range(args.large_steps):
        sim.step()
        solve_ms.append(e.timings()["solve"])
        iters.append(sim.iters)
        res = max(res, float(e.solve_history()[-1]))
    e.synchronize()
    bracket = e.trace_overhead(200)
    kernels = {}
    for k in LARGE_KERNELS:
        e.trace_kernel(k)
        sim.step()
        cnt, ms = e.trace_stats()
        e.trace_kernel(None)
        if cnt:
            avg = ms / cnt - bracket
            gbs = algo_bytes(k, args.solver) * n * n / (avg * 1e-3) / 1e9
            kernels[k] = {"avg_launch_ms": avg, "launches": cnt, "achieved_GBs": gbs,
                          "frac": gbs / HBM_PEAK_GBS}
    e.set_timing(False)
    K = args.large_steps
    learned = run_large_learned(args, sim, scene, K) if args.large_learned else None
    if learned:
        learned["solves_per_s_vs_plain"] = sum(solve_ms) / K / learned["ms_per_solve"]
    return {
        "workload": f"thermal plume {n}x{n}, closed box, 200 circular obstacles, 16 buoyant "
                    f"sources, MacCormack, {SOLVER_NAMES[args.solver]}, tol 1e-6, "
                    "SplitMix64(20240613)",
        "cells": n * n,
        "steps": K,
        "value": n * n * K / elapsed,
        "unit": "cell-steps/s",
        "ms_per_step": elapsed * 1e3 / K,
        "ms_per_solve": sum(solve_ms) / K,
        "vcycles_per_solve": sum(iters) / K,
        "max_relative_residual": res,
        "kernels": kernels,
        "learned_warm_start": learned,
    }


def run_large_learned(args, sim, scene, K):
    """The large domain's steps again from the same reset, every pressure
    solved from the trained U-Net warm start (inference at the full 8192^2
    resolution, then the same solver to 1e-6): K timed steps, then K more
    with per-solve times and cycle counts."""
    from flowlab_amd import learned as L
    if not os.path.exists(L.WEIGHTS):
        return None
    n = args.large_n
    u = L.UNet()
    u.load(L.WEIGHTS)
    e = sim.stepper.engine
    sim.reset(scene)
    e.set_warm_start(u)
    for _ in range(2):
        sim.step()
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        sim.step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    e.set_timing(True)
    solve_ms, iters, res = [], [], 0.0
    for _ in range(K):
        sim.step()
        solve_ms.append(e.timings()["solve"])
        iters.append(sim.iters)
        res = max(res, float(e.solve_history()[-1]))
    e.set_timing(False)
    e.set_warm_start(None)
    return {
        "value": n * n * K / elapsed,
        "unit": "cell-steps/s",
        "ms_per_step": elapsed * 1e3 / K,
        "ms_per_solve": sum(solve_ms) / K,
        "vcycles_per_solve": sum(iters) / K,
        "max_relative_residual": res,
        "weights": os.path.relpath(L.WEIGHTS, ROOT),
    }


def run_cpu_baseline(args):
    from flowlab_amd import scenes as S
    from oracle import pyoracle as O

    nx, ny = args.nx, args.ny
    g, scene = S.karman_street(nx, ny)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=100, tol=1e-6,
                        cfl_target=0.5, residual_every=0)
    osim.step(prm)  # warm-up (the impulsive start's first solve)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        osim.step(prm)
    dt = time.perf_counter() - t0
    return nx * ny * args.cpu_steps / dt, dt


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not a measured path)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_mb > 0:
        import torch
        n = args.copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    r = run_gpu(args)
    learned = run_learned(args, r)
    batched = run_batched(args) if args.batch_scenes > 0 else None
    large = run_large(args) if args.large_n > 0 else None
    nx, ny, K = args.nx, args.ny, args.steps
    cells = nx * ny
    value = cells * K / r["elapsed"]
    # per-launch device time: event bracket around each launch, less the
    # bracket's own cost (an empty event pair on the same stream)
    k_avg_ms = r["k_ms"] / max(r["n_launch"], 1) - r["bracket_ms"]
    # kernels without a bandwidth figure (the latency-bound coarse sub-cycle)
    # report their launch time only
    bpc = algo_bytes(args.kernel, args.solver)
    bytes_per_launch = bpc * cells if bpc else None
    achieved = (bytes_per_launch / (k_avg_ms * 1e-3) / 1e9
                if r["n_launch"] and bytes_per_launch else None)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": r["elapsed"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (advection, smoothing, transfers) + f64 (pressure, residual, norms)",
        "data": "synthetic",
        "config": {
            "workload": f"karman vortex street {nx}x{ny}, MacCormack, "
                        f"{SOLVER_NAMES[args.solver]}, tol 1e-6, CFL 0.5",
            "grid": [nx, ny],
            "solver": args.solver,
            "tol": 1e-6,
            "mg_tiles": r["tiles"],
        },
        "pressure": {
            "ms_per_solve": sum(r["solve_ms"]) / K,
            "vcycles_per_solve": sum(r["iters"]) / K,
            "max_relative_residual": r["hist_max"],
            "learned_warm_start": learned,
        },
        "roofline": {
            "kernel": args.kernel,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.kernel, nx, ny),
            "traffic_source": os.path.relpath(PMC_TRAFFIC, ROOT),
            "avg_launch_ms": k_avg_ms,
            "event_bracket_ms": r["bracket_ms"],
            "launches": r["n_launch"],
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    if batched is not None:
        out["batched"] = batched
    if large is not None:
        out["large"] = large
    if args.cpu_steps > 0:
        v, secs = run_cpu_baseline(args)
        out["cpu_baseline"] = {
            "value": v,
            "unit": "cell-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same {nx}x{ny} workload after 1 warm-up "
                      f"step, FP64 C oracle, one thread ({secs:.1f} s)",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
