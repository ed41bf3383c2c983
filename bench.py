"""Benchmark of the projection-method step on one MI355X.

Headline workload: the 8192x8192 thermal plume (67 M cells; closed box, 200
seeded circular obstacles, 16 buoyant heat sources, buoyancy 1, MacCormack
advection of velocity and 3 dye channels, dt = 0.5 h capped at CFL 0.5,
every pressure solve to 1e-6 relative residual by mixed-precision
defect-correction multigrid, V(2,2) cycles to 16x8, started from the last two
solutions extrapolated linearly in time), 200 steps from rest.  (The Karman
legs run flexible CG around the same V-cycle, faster there; `--solver` and
`--c2-solver` choose.)

A step is one fused fh_step over the resident fields (inputs already in
HBM); the timed region is `--warmup` untimed steps after a reset, then
exactly `--steps` steps bracketed by device synchronisation.  The whole
200-step workload is also run from its reset with per-step phase timing
(`workload_run`: solve times, V-cycles and residuals over every step).

`roofline` is the dominant kernel (the largest device time per step at
8192^2) measured live: each of its launches in the timed steps is bracketed
by HIP events on the engine's stream; achieved = algorithmic bytes per launch
/ mean bracketed duration (nothing subtracted).  `traffic` comes from the
committed FETCH_SIZE/WRITE_SIZE passes at this size.  `legs` run the other
configurations at their stated step counts: the 128^2 plume on red-black SOR
(C1), the 2048x1024 Karman street (C2), the 64-scene decaying-turbulence
batch (C3) and the learned warm-start evaluation at 1024^2 and 2048^2 (C4).
`cpu_baseline` times the FP64 C oracle (one thread) on one step of the
headline workload, continued from the device's state after its 200 steps.
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
WORKLOAD_STEPS = 200

# Algorithmic (compulsory) bytes per cell of one launch of each traceable
# kernel: every array element read or written once, neighbour reuse perfect.
#   maccormack_faces: as one pass (the engine's default: the forward-pass
#     samples computed in LDS, FACES_ONE_PASS) u^n, v^n 8 + out u, v 8 +
#     label 1 (17); as the correction pass of two, u^n, v^n 8 + forward-pass
#     u, v 8 + out u, v 8 + label 1 + the limiter stencils the forward pass
#     recorded 1 (26); with buoyancy alone as the body force either applies
#     it too, reading the heat (+4).
#   advect_faces: carrier = source u, v 8 + out u, v 8 + label 1 + the
#     recorded limiter stencils 1 (MacCormack's forward pass).
#   maccormack_dye: as one pass (the engine's default at dye scale 1: the
#     forward-pass samples computed in LDS, DYE_ONE_PASS) u, v 8 + 3 channels
#     of q^n 12 + label 1 + 3 out 12 (33); as the correction pass of two,
#     u, v 8 + 3 channels of q^n and of the forward pass 24 + 3 out 12 +
#     label 1 + recorded stencils 1 (46); advect_dye (the forward pass of two):
#     u, v 8 + 3 in 12 + 3 out 12 + label 1 + recorded stencils 1.
#   dc_update: a defect-correction cycle's update, fused with the finest
#     prolongation + post-smoothing (the engine's default): the pre-smoothed
#     x 4 + b 4 + label 1 + the coarse correction 1 + p 8 + div 8 read, p 8 +
#     r32 4 written (38); as its own pass (FUSED_UPDATE False) e 4 + p 8 +
#     div 8 + label 1 read, p 8 + r32 4 written (33); under the flexible-CG
#     solver either also writes the FP64 residual (+8).  (The solve's initial
#     defect pass is the divergence kernel below.)
#   smooth_restrict_fine: b 4 + label 1 + x 4 written + coarse b 1.
#   prolong_smooth_fine: x 4 + b 4 + label 1 + coarse e 1 + x 4 written.
#   fcg_dots: z 4 + previous direction 4 + r 8 + info 1.
#   fcg_update: z 4 + previous direction 4 + p 8 + r 8 + info 1 read;
#     direction 4 + p 8 + r 8 + r32 4 written.
#   divergence: under the FP64 solvers formed by the solve's initial defect
#     pass (the traced launch): u, v 8 + p 8 + label 1 read, div 8 + r32 4
#     written (the first cycle's update stores the FP64 residual flexible CG
#     keeps); otherwise its own pass, u, v 8 read + div 8 written (16).
#   forces (buoyancy alone, where the face pass does not apply it): v 4 read
#     + 4 written + heat 4 + label 1.
#   project (gradient subtraction + outflow): u, v 8 read + 8 written +
#     FP64 p 8 + label 1.
#   rb_sweep_fine (one colour): half of a sweep's 13 (x 4 + 4, rhs 4, label 1).
ALGO_BYTES_PER_CELL = {
    "maccormack_faces": 26.0, "advect_faces": 18.0, "maccormack_dye": 46.0,
    "advect_dye": 34.0, "dc_update": 33.0, "smooth_restrict_fine": 10.0,
    "prolong_smooth_fine": 14.0, "fcg_dots": 17.0, "fcg_update": 49.0, "divergence": 16.0,
    "forces": 13.0, "project": 25.0, "rb_sweep_fine": 6.5,
}
LARGE_KERNELS = ("maccormack_dye", "maccormack_faces", "advect_dye", "advect_faces", "dc_update",
                 "prolong_smooth_fine", "smooth_restrict_fine", "project", "divergence",
                 "forces", "fcg_dots", "fcg_update")
SOLVER_NAMES = {
    "mg-dc": "MG-DC V(2,2) to 16x8",
    "mg-fcg": "one MG-DC V(2,2) cycle (to 16x8), then flexible CG preconditioned by it",
}
PMC_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic_plume_8192.json")
# the engines fuse each defect-correction cycle's finest prolongation +
# smoothing with its update (fh_set_fused_update), traced as dc_update
FUSED_UPDATE = True
# the engines run the MacCormack dye as one pass (fh_set_dye_one_pass), traced
# as maccormack_dye
DYE_ONE_PASS = True
# and the MacCormack velocity as one pass (fh_set_faces_one_pass), traced as
# maccormack_faces
FACES_ONE_PASS = True
# start of each step's solve: the previous pressure (the reference's), or the
# last two solutions extrapolated linearly in time (fh_set_warm_extrapolation)
WARM_STARTS = {"previous": (0, "the previous pressure"),
               "extrapolated": (1, "the last two solutions extrapolated linearly in time")}


def algo_bytes(kernel, solver, buoyancy=True, warm="previous"):
    """Algorithmic bytes per cell of one traced launch of `kernel` (buoyancy:
    the workload's forces are buoyancy alone, which the MacCormack face pass
    applies, reading the heat, +4; warm "extrapolated": the solve's initial
    pass also reads the previous solution and writes the start it forms,
    +16)."""
    b = ALGO_BYTES_PER_CELL.get(kernel)
    if kernel == "maccormack_faces" and FACES_ONE_PASS:
        b = 17.0
    if kernel == "maccormack_faces" and buoyancy:
        b += 4.0
    if kernel == "maccormack_dye" and DYE_ONE_PASS:
        b = 33.0
    if kernel == "dc_update" and FUSED_UPDATE:
        b = 38.0
    if kernel == "dc_update" and solver == "mg-fcg":
        b += 8.0
    if kernel == "divergence" and solver in ("mg-dc", "mg-fcg"):
        b = 29.0
        if warm == "extrapolated":
            b += 16.0
    return b




def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=WORKLOAD_STEPS,
                    help="timed steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed steps after the reset, before the timed ones")
    ap.add_argument("--n", type=int, default=8192, help="side of the headline plume")
    ap.add_argument("--solver", default="mg-dc", choices=sorted(SOLVER_NAMES),
                    help="pressure solver of the headline plume")
    ap.add_argument("--c2-solver", default="mg-fcg", choices=sorted(SOLVER_NAMES),
                    help="pressure solver of the Karman street legs (C2 and through Node)")
    ap.add_argument("--warm", default="extrapolated", choices=sorted(WARM_STARTS),
                    help="start of each step's pressure solve (headline and C2)")
    ap.add_argument("--kernel", default=None,
                    help="kernel traced for the roofline (default: the largest device time "
                         "per step)")
    ap.add_argument("--legs", default="c1,c2,c3,c4,learned,node",
                    help="comma list of extra legs: c1, c2, c3, c4, learned, node (none: '')")
    ap.add_argument("--node-steps", type=int, default=200)
    ap.add_argument("--c2-steps", type=int, default=1000)
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--c1-steps", type=int, default=200)
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0 = skip the oracle timing")
    ap.add_argument("--workload-run", type=int, default=1,
                    help="0 = skip the instrumented 200-step run (profiler passes)")
    ap.add_argument("--table", type=int, default=1,
                    help="0 = skip the per-kernel table (profiler passes)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (the known-size "
                         "copy kernel that scales hardware traffic counters)")
    ap.add_argument("--pmc-summary", default=None, metavar="FETCH_CSV,WRITE_CSV",
                    help="no GPU: fold the counter_collection.csv files of a FETCH_SIZE and a "
                         "WRITE_SIZE pass (run with --copy-mb) into the traffic table")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed counter passes
    (FETCH_SIZE and WRITE_SIZE, scaled by a known-size copy) at this size."""
    if not os.path.exists(PMC_TRAFFIC):
        return None
    with open(PMC_TRAFFIC) as f:
        d = json.load(f)
    if d.get("grid") != [n, n]:
        return None
    k = d["kernels"].get(kernel)
    return None if k is None else k["traffic_bytes"]


# rocprofv3 kernel names (template arguments and parameter lists dropped) of
# the kernels whose traffic the table reports
PMC_KERNELS = {
    "k_maccormack_dye_fused": "maccormack_dye", "k_maccormack_dye_lds": "maccormack_dye",
    "k_maccormack_faces_fused": "maccormack_faces", "k_maccormack_faces_lds": "maccormack_faces",
    "k_advect_dye_lds": "advect_dye", "k_advect_faces_lds": "advect_faces",
    "k_prolong_smooth_dc": "dc_update", "k_fcg_update": "fcg_update", "k_fcg_dots": "fcg_dots",
    "k_smooth_restrict": "smooth_restrict_fine", "k_prolong_smooth": "prolong_smooth_fine",
    "k_subtract_outflow": "project", "k_buoyancy": "forces", "k_div_defect": "divergence",
}
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def pmc_summary(args):
    """HBM traffic per launch from a FETCH_SIZE pass and a WRITE_SIZE pass
    (KiB per dispatch) of `bench.py --copy-mb M ...`.  Each counter is scaled
    by what it reads for the M MiB device-to-device copies (M MiB moved each
    way); each kernel's row is its dispatches at its largest grid (the fine
    level for the multigrid kernels), of the instantiation dispatched most.
    Writes the JSON the roofline's `traffic` reads and prints the table as
    markdown."""
    import csv
    import re
    from collections import defaultdict
    fetch_csv, write_csv = args.pmc_summary.split(",")

    def load(path, counter):
        out = defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    raise SystemExit(f"{path}: counter {r['Counter_Name']}, expected {counter}")
                out[(r["Kernel_Name"], int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
        return out

    F, W = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_kib = args.copy_mb * 1024
    if copy_kib <= 0:
        raise SystemExit("--pmc-summary needs the --copy-mb the passes ran with")
    # the copies of the known size: every copy dispatch of that many work-items or more
    big = max(g for (k, g) in F if k.startswith(COPY_KERNEL))
    cf = [v for (k, g), vs in F.items() if k.startswith(COPY_KERNEL) and g == big for v in vs]
    cw = [v for (k, g), vs in W.items() if k.startswith(COPY_KERNEL) and g == big for v in vs]
    sf, sw = copy_kib / (sum(cf) / len(cf)), copy_kib / (sum(cw) / len(cw))

    def short(name):
        return re.sub(r"\(.*", "", name.replace("void ", "").replace("fh::", ""))

    # per kernel: each instantiation's finest-level dispatches (its largest
    # grid), then the instantiation the run used most (the tuned tile shape;
    # the tuner runs every other shape only while timing them)
    finest = {}
    for (kn, g) in F:
        if kn not in finest or finest[kn] < g:
            finest[kn] = g
    pick = {}
    # with the fused update the defect-correction solve runs no unfused
    # prolongation on the finest level: its coarser dispatches are no row
    skip = {"prolong_smooth_fine"} if FUSED_UPDATE and args.solver == "mg-dc" else set()
    for kn, g in finest.items():
        key = PMC_KERNELS.get(re.sub(r"<.*", "", short(kn)))
        if key in skip:
            continue
        if key and (key not in pick or len(F[pick[key]]) < len(F[(kn, g)])):
            pick[key] = (kn, g)
    n = args.n
    cells = n * n
    out = {"grid": [n, n], "solver": args.solver, "copy_mb": args.copy_mb,
           "copy_scale_fetch": sf, "copy_scale_write": sw, "kernels": {}}
    rows = []
    for key, (kn, g) in pick.items():
        fv, wv = F[(kn, g)], W.get((kn, g), [0.0])
        fb = sum(fv) / len(fv) * 1024 * sf
        wb = sum(wv) / len(wv) * 1024 * sw
        a = algo_bytes(key, args.solver, warm=args.warm) * cells
        out["kernels"][key] = {"rocprof_name": short(kn), "dispatches": len(fv),
                               "fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb,
                               "algorithmic_bytes": a, "ratio": (fb + wb) / a}
        rows.append((short(kn), len(fv), fb / 1e6, wb / 1e6, (fb + wb) / 1e6, a / 1e6,
                     (fb + wb) / a))
    with open(PMC_TRAFFIC, "w") as f:
        json.dump(out, f, indent=1)
    print(f"copy {args.copy_mb} MiB: fetch x{sf:.4f}, write x{sw:.4f}")
    print("| kernel | dispatches | fetched MB | written MB | traffic MB | algorithmic MB | "
          "traffic / algorithmic |")
    print("|---|---|---|---|---|---|---|")
    for r in sorted(rows, key=lambda r: -r[4]):
        print(f"| `{r[0]}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} | {r[4]:.1f} | {r[5]:.1f} | "
              f"{r[6]:.3f} |")


# ---- single-grid legs ---------------------------------------------------------

def make_sim(nx, ny, scene, params):
    from flowlab_amd.simulation import Simulation
    sim = Simulation(nx, ny, params, readback=False)
    sim.reset(scene)
    return sim


def instrumented(sim, scene, steps, warmup=0):
    """`steps` steps after a reset (and `warmup` more first) with per-step
    phase timing: solve ms, V-cycles and the final residual of each solve."""
    e = sim.stepper.engine
    sim.reset(scene)
    sim.stepper.upload(sim.f)
    for _ in range(warmup):
        sim.step()
    e.set_timing(True)
    solve_ms, iters, res, phases = [], [], 0.0, {}
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        ph = e.timings()
        solve_ms.append(ph["solve"])
        for k, v in ph.items():
            phases[k] = phases.get(k, 0.0) + v
        iters.append(sim.iters)
        h = e.solve_history()
        if len(h):
            res = max(res, float(h[-1]))
    e.synchronize()
    elapsed = time.perf_counter() - t0
    e.set_timing(False)
    return dict(elapsed=elapsed, solve_ms=solve_ms, iters=iters, max_res=res,
                phases={k: v / steps for k, v in phases.items()})


def kernel_table(sim, cells, solver, steps=2, buoyancy=False, warm="previous"):
    """Each kernel of LARGE_KERNELS bracketed by events over `steps` steps:
    mean launch ms, launches per step and algorithmic GB/s."""
    e = sim.stepper.engine
    out = {}
    for k in LARGE_KERNELS:
        e.trace_kernel(k)
        for _ in range(steps):
            sim.step()
        n, ms = e.trace_stats()
        e.trace_kernel(None)
        if not n:
            continue
        avg = ms / n
        bpc = algo_bytes(k, solver, buoyancy, warm)
        gbs = bpc * cells / (avg * 1e-3) / 1e9
        out[k] = {"avg_launch_ms": avg, "launches_per_step": n / steps,
                  "ms_per_step": ms / steps, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                  "algorithmic_bytes_per_cell": bpc}
    return out


def run_headline(args):
    from flowlab_amd import scenes as S
    n = args.n
    log(f"headline: plume {n}^2 scene")
    g, scene = S.large_plume(n)
    params = dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 * g.h, solver=args.solver,
                  pressure_iters=100, tol=1e-6, residual_every=0, buoyancy=1.0,
                  warm_extrapolation=WARM_STARTS[args.warm][0])
    sim = make_sim(n, n, scene, params)
    e = sim.stepper.engine
    sim.step()  # uploads the fields and builds the multigrid hierarchy
    tiles = e.tune_tiles(reps=5)["tiles"]
    log(f"headline: tiles {tiles}; per-kernel table")
    kernels = (kernel_table(sim, n * n, args.solver, buoyancy=True, warm=args.warm)
               if args.table else {})
    dominant = args.kernel or (max(kernels, key=lambda k: kernels[k]["ms_per_step"])
                               if kernels else "maccormack_dye")
    # timed pass: reset and upload (the fields are resident before the
    # clock starts), W untimed steps, K timed steps with the dominant
    # kernel's launches bracketed on the engine stream
    sim.reset(scene)
    sim.stepper.upload(sim.f)
    for _ in range(args.warmup):
        sim.step()
    e.trace_kernel(dominant)
    e.synchronize()
    log(f"headline: timed pass ({args.warmup} + {args.steps} steps), tracing {dominant}")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    n_launch, k_ms = e.trace_stats()
    e.trace_kernel(None)
    log("headline: the 200-step workload, instrumented")
    wl = instrumented(sim, scene, WORKLOAD_STEPS if args.workload_run else args.steps)
    other = other_start(sim, scene, len(wl["iters"]), wl, args.warm)
    state = None
    if args.cpu_baseline:
        u, v = e.read_velocity()
        state = dict(u=u, v=v, p=e.read_pressure(), dye=e.read_dye(), label=sim.f.label.copy(),
                     g=g, scene=scene)
    learned = None
    if "learned" in args.legs:
        learned = run_learned(sim, scene, WORKLOAD_STEPS, wl)
    sim.stepper.engine.close()
    return dict(n=n, elapsed=elapsed, steps=args.steps, tiles=tiles, kernels=kernels,
                dominant=dominant, n_launch=n_launch, k_ms=k_ms, workload=wl,
                learned=learned, state=state, other_start=other)


def other_start(sim, scene, steps, wl, warm):
    """The same steps from the same reset with the other start of the solve
    (previous pressure against the extrapolation): ms and V-cycles per solve."""
    e = sim.stepper.engine
    name = "previous" if warm == "extrapolated" else "extrapolated"
    e.set_warm_extrapolation(WARM_STARTS[name][0])
    r = instrumented(sim, scene, steps)
    e.set_warm_extrapolation(WARM_STARTS[warm][0])
    K = len(r["iters"])
    ms = sum(r["solve_ms"]) / K
    return {"start": WARM_STARTS[name][1], "steps": K, "ms_per_solve": ms,
            "vcycles_per_solve": sum(r["iters"]) / K, "max_relative_residual": r["max_res"],
            "ms_per_step_with_phase_timing": r["elapsed"] * 1e3 / K,
            "solves_per_s_of_the_run_vs_this": ms / (sum(wl["solve_ms"]) / len(wl["iters"]))}


def run_learned(sim, scene, steps, plain):
    """The same steps from the same reset, every pressure solved from the
    trained U-Net warm start applied to the run's own start (then the same
    solver to 1e-6)."""
    from flowlab_amd import learned as L
    if not os.path.exists(L.WEIGHTS):
        return None
    u = L.UNet()
    u.load(L.WEIGHTS)
    e = sim.stepper.engine
    e.set_warm_start(u)
    r = instrumented(sim, scene, steps)
    e.set_warm_start(None)
    cells = sim.g.nx * sim.g.ny
    ms_l = sum(r["solve_ms"]) / steps
    ms_p = sum(plain["solve_ms"]) / steps
    return {"steps": steps, "ms_per_solve": ms_l, "vcycles_per_solve": sum(r["iters"]) / steps,
            "max_relative_residual": r["max_res"], "solves_per_s_vs_plain": ms_p / ms_l,
            "cell_steps_per_s_with_phase_timing": cells * steps / r["elapsed"],
            "weights": os.path.relpath(L.WEIGHTS, ROOT)}


def leg_summary(cells, steps, elapsed, wl, tiles=None):
    K = len(wl["iters"])
    out = {"cells": cells, "steps": steps, "value": cells * steps / elapsed,
           "unit": "cell-steps/s", "ms_per_step": elapsed * 1e3 / steps,
           "ms_per_solve": sum(wl["solve_ms"]) / K, "vcycles_per_solve": sum(wl["iters"]) / K,
           "max_relative_residual": wl["max_res"],
           "phase_ms_per_step": {k: round(v, 4) for k, v in wl["phases"].items()}}
    if tiles is not None:
        out["mg_tiles"] = tiles
    return out


def run_c2(args):
    """Karman vortex street 2048x1024 at its 1000 steps (cache-resident: the
    working set is below the 256 MB Infinity Cache)."""
    from flowlab_amd import scenes as S
    nx, ny, K = 2048, 1024, args.c2_steps
    g, scene = S.karman_street(nx, ny)
    params = dict(scheme="macCormack", cfl_target=0.5, solver=args.c2_solver, pressure_iters=100,
                  tol=1e-6, residual_every=0, warm_extrapolation=WARM_STARTS[args.warm][0])
    sim = make_sim(nx, ny, scene, params)
    e = sim.stepper.engine
    sim.step()
    tiles = e.tune_tiles(reps=10)["tiles"]
    kernels = kernel_table(sim, nx * ny, args.c2_solver, buoyancy=False, warm=args.warm)
    sim.reset(scene)
    sim.stepper.upload(sim.f)
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        sim.step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    wl = instrumented(sim, scene, K)
    out = leg_summary(nx * ny, K, elapsed, wl, tiles)
    out["workload"] = ("karman vortex street 2048x1024, MacCormack, "
                       f"{SOLVER_NAMES[args.c2_solver]}, tol 1e-6, CFL 0.5, {K} steps, each solve "
                       f"from {WARM_STARTS[args.warm][1]}")
    out["other_start"] = other_start(sim, scene, K, wl, args.warm)
    out["kernels"] = kernels
    out["note"] = "working set inside the 256 MB Infinity Cache: GB/s are cache-effective"
    if "learned" in args.legs:
        out["learned_warm_start"] = run_learned(sim, scene, K, wl)
    e.close()
    return out


def run_c1(args):
    """128^2 closed-box thermal plume: MacCormack, buoyancy 1, vorticity
    confinement 0.1, red-black SOR omega 1.9 x 200 sweeps, dt = 0.5 h, 200
    steps."""
    from flowlab_amd import scenes as S
    g, scene = S.thermal_plume_128()
    K = args.c1_steps
    params = dict(scheme="macCormack", cfl_target=1e30, dt_max=0.5 * g.h, solver="rbsor",
                  pressure_iters=200, omega=1.9, tol=0.0, buoyancy=1.0, confinement=0.1,
                  residual_every=0)
    sim = make_sim(128, 128, scene, params)
    e = sim.stepper.engine
    for _ in range(3):
        sim.step()
    sim.reset(scene)
    sim.stepper.upload(sim.f)
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        sim.step()
    e.synchronize()
    elapsed = time.perf_counter() - t0
    wl = instrumented(sim, scene, K)
    out = leg_summary(128 * 128, K, elapsed, wl)
    out["workload"] = ("thermal plume 128x128, MacCormack, buoyancy 1, confinement 0.1, "
                       f"red-black SOR omega 1.9 x 200, dt 0.5 h, {K} steps")
    e.close()
    return out


def run_c3(args):
    """64 decaying-turbulence scenes of uneven size (256^2 .. 1024^2), each
    phase one launch over every scene's tiles; dt per scene at CFL 0.5;
    pressure by mg-fcg to 1e-6 per scene; 500 steps."""
    from flowlab_amd import scenes as S
    from flowlab_amd.batch import HipBatch
    scenes = S.turbulence_batch(64)
    b = HipBatch([g for g, _ in scenes])
    b.set_warm_extrapolation(WARM_STARTS[args.warm][0])
    for s, (_, sc) in enumerate(scenes):
        b.reset(s, sc)
    kw = dict(scheme="macCormack", cfl=0.5, dt_max=1.0, tol=1e-6, max_cycles=100)
    b.step(**kw)
    tune = b.tune(reps=5)
    for s, (_, sc) in enumerate(scenes):
        b.reset(s, sc)
    K = args.c3_steps
    iters, res_max, t_steps = [], 0.0, 0.0
    t0 = time.perf_counter()
    for _ in range(K):
        b.step(**kw)
        _, it, res = b.results()
        iters.append(float(it.mean()))
        res_max = max(res_max, float(res.max()))
    elapsed = time.perf_counter() - t0
    cells = b.cells
    sizes = sorted({(g.nx, g.ny) for g, _ in scenes})
    b.close()
    return {
        "workload": f"{len(scenes)} decaying-turbulence scenes, sizes {sizes} drawn by "
                    "SplitMix64(20240612), square-obstacle lattices, k^-3 velocity RMS 1, "
                    f"MacCormack, {SOLVER_NAMES['mg-fcg']}, tol 1e-6, CFL 0.5, {K} steps, each "
                    f"solve from {WARM_STARTS[args.warm][1]}",
        "cells": cells, "steps": K, "value": cells * K / elapsed, "unit": "cell-steps/s",
        "ms_per_step": elapsed * 1e3 / K, "vcycles_per_solve_mean": sum(iters) / len(iters),
        "max_relative_residual": res_max, "mg_tiles": tune["tiles"],
        "execution": "one launch per phase over a (scene, block) work list of all scenes, one "
                     "stream; per-scene convergence flags on the device, one host read per "
                     "cycle",
    }


def run_c4(args):
    """The learned warm start at 1024^2 and 2048^2 (3 seeded scenes each, the
    next step's solve to 1e-6), V-cycles and device ms per solve from six
    starts: zero, previous pressure, linear extrapolation in time, and the
    network on each (`network`, `learned`, `learned_extrapolated`)."""
    from flowlab_amd import learned as L
    if not os.path.exists(L.WEIGHTS):
        return None
    u = L.UNet()
    u.load(L.WEIGHTS)
    out = {}
    for n in (1024, 2048):
        r = L.evaluate(u, n, scenes=3)
        for net, base in (("network", "zero"), ("learned", "previous"),
                          ("learned_extrapolated", "extrapolated")):
            r[f"{net}_vcycles_vs_{base}"] = r[net]["vcycles"] / r[base]["vcycles"]
            r[f"{net}_solves_per_s_vs_{base}"] = r[base]["ms"] / r[net]["ms"]
        r["learned_extrapolated_solves_per_s_vs_previous"] = (
            r["previous"]["ms"] / r["learned_extrapolated"]["ms"])
        out[f"{n}x{n}"] = r
    return out


def run_node_c2(args):
    """The Karman street 2048x1024 stepped through the N-API HipStepper from
    Node (flowlab_amd/node/bench_step.js): the host upload, the one-step-stale
    velocity readback into the host fields every step, the host CFL scan and
    the profiler marks are all inside the timed loop, as a TypeScript host
    pays them."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    from flowlab_amd import scenes as S
    node = shutil.which("node")
    script = os.path.join(ROOT, "flowlab_amd", "node", "bench_step.js")
    if node is None or not os.path.exists(os.path.join(ROOT, "flowlab_amd", "node",
                                                       "flowhip.node")):
        return None
    nx, ny = 2048, 1024
    g, scene = S.karman_street(nx, ny)
    f = S_fields(g, scene)
    with tempfile.TemporaryDirectory() as d:
        for name in ("u", "v", "p"):
            np.asarray(getattr(f, name), "<f8").tofile(os.path.join(d, f"{name}.f64"))
        np.concatenate(f.dye).astype("<f8").tofile(os.path.join(d, "dye.f64"))
        f.label.tofile(os.path.join(d, "label.u8"))
        case = dict(nx=nx, ny=ny, h=g.h, scheme="macCormack", solver=args.c2_solver, tol=1e-6,
                    cfl=0.5, dtMax=1 / 30, warm=WARM_STARTS[args.warm][0])
        with open(os.path.join(d, "case.json"), "w") as fh:
            json.dump(case, fh)
        r = subprocess.run([node, script, d, "10", str(args.node_steps)], capture_output=True,
                           text=True, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["workload"] = (f"karman vortex street {nx}x{ny} through the N-API HipStepper from Node "
                       f"{out.get('node')}, {SOLVER_NAMES[args.c2_solver]}, tol 1e-6, CFL 0.5 from "
                       "the host's one-step-stale velocity, each solve from "
                       f"{WARM_STARTS[args.warm][1]}; upload, readback and host CFL scan inside "
                       "the timed loop")
    return out


def run_cpu_baseline_js(args):
    """The reference's CPU path restated in plain JavaScript (FP64
    Float64Array fields, one thread, oracle/js/cpuStep.js) on this host's
    Node: the 2048x1024 Karman workload, one warm-up step, then two timed
    steps, MG-DC to 1e-6."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    from flowlab_amd import scenes as S
    node = shutil.which("node")
    if node is None:
        return None
    nx, ny, warm, steps = 2048, 1024, 1, 2
    g, scene = S.karman_street(nx, ny)
    f = S_fields(g, scene)
    with tempfile.TemporaryDirectory() as d:
        for name in ("u", "v", "p"):
            np.asarray(getattr(f, name), "<f8").tofile(os.path.join(d, f"{name}.f64"))
        np.concatenate(f.dye).astype("<f8").tofile(os.path.join(d, "dye.f64"))
        f.label.tofile(os.path.join(d, "label.u8"))
        case = dict(nx=nx, ny=ny, h=g.h, dyeScale=1, scheme="macCormack", solver="mg-dc",
                    pressureIters=100, tol=1e-6, cflTarget=0.5, dtMax=1 / 30, patch=None)
        with open(os.path.join(d, "case.json"), "w") as fh:
            json.dump(case, fh)
        r = subprocess.run([node, os.path.join(ROOT, "oracle", "js", "cpuStep.js"), d, str(warm),
                            str(steps)], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    o = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": o["cell_steps_per_s"], "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of the {nx}x{ny} Karman workload after {warm} warm-up "
                      f"step, the reference CPU path restated in JavaScript (Float64Array, one "
                      f"thread) on Node {o['node']}, MG-DC to 1e-6 ({o['iters']} V-cycles, "
                      f"{o['seconds']:.1f} s)",
            "host_cpus": o["cpus"], "cpu_model": o["cpu_model"]}


def S_fields(g, scene):
    from flowlab_amd.simulation import scene_fields
    return scene_fields(g, scene)


def run_cpu_baseline(state, args):
    """One step of the headline workload by the FP64 C oracle (one thread),
    continued from the device's fields after the 200-step run, with the
    oracle's defect-correction multigrid to 1e-6."""
    import numpy as np
    from oracle import pyoracle as O
    n = args.n
    g, scene = state["g"], state["scene"]
    osim = O.OracleSim(n, n)
    osim.reset(g, scene)  # labels and the inlet patch
    osim.u[:] = state["u"]
    osim.v[:] = state["v"]
    osim.p[:] = state["p"]
    osim.dye[:] = state["dye"]
    assert np.array_equal(osim.label, state["label"])
    pt = osim.patch
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=100, tol=1e-6,
                        cfl_target=0.5, dt_max=0.5 * g.h, buoyancy=1.0, residual_every=0,
                        patch=(pt.i0, pt.j0, pt.nx, pt.ny))
    t0 = time.perf_counter()
    osim.step(prm, pt.data)
    secs = time.perf_counter() - t0
    return {"value": n * n / secs, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"1 step of the {n}x{n} plume (step {WORKLOAD_STEPS + 1}, continued from "
                      f"the device's fields), FP64 C oracle, one thread, MG-DC to 1e-6 "
                      f"({prm.iters} V-cycles, {secs:.1f} s)"}


def main():
    args = parse()
    if args.pmc_summary:
        pmc_summary(args)
        return
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note, not a measured path)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_mb > 0:
        import torch
        nb = args.copy_mb * (1 << 20) // 4
        a = torch.ones(nb, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    legs = {k for k in args.legs.split(",") if k}
    h = run_headline(args)
    n, K = h["n"], args.steps
    cells = n * n
    wl = h["workload"]
    out_legs = {}
    if "c1" in legs:
        log("leg C1")
        out_legs["C1_plume_128_rbsor"] = run_c1(args)
    if "c2" in legs:
        log("leg C2")
        out_legs["C2_karman_2048x1024"] = run_c2(args)
    if "c3" in legs:
        log("leg C3")
        out_legs["C3_turbulence_batch_64"] = run_c3(args)
    if "c4" in legs:
        log("leg C4")
        out_legs["C4_learned_warm_start_eval"] = run_c4(args)
    if "node" in legs:
        log("leg C2 through Node")
        out_legs["C2_karman_2048x1024_node_boundary"] = run_node_c2(args)
    dom = h["dominant"]
    bpc = algo_bytes(dom, args.solver, warm=args.warm)
    k_avg_ms = h["k_ms"] / max(h["n_launch"], 1)
    bytes_per_launch = bpc * cells
    achieved = bytes_per_launch / (k_avg_ms * 1e-3) / 1e9 if h["n_launch"] else None
    wsum = leg_summary(cells, len(wl["iters"]), wl["elapsed"], wl)
    wsum["note"] = "per-step phase timing on (the host waits for every step's end)"
    out = {
        "metric": METRIC,
        "value": cells * K / h["elapsed"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": h["elapsed"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (advection, smoothing, transfers) + f64 (pressure, residual, norms)",
        "data": "synthetic",
        "config": {
            "workload": f"thermal plume {n}x{n}, closed box, 200 circular obstacles, 16 "
                        f"buoyant sources, buoyancy 1, MacCormack, {SOLVER_NAMES[args.solver]}, "
                        "tol 1e-6, dt 0.5 h capped at CFL 0.5, SplitMix64(20240613), each "
                        f"solve from {WARM_STARTS[args.warm][1]}",
            "warm_start": args.warm,
            "grid": [n, n],
            "workload_steps": WORKLOAD_STEPS,
            "timed_steps": f"{args.warmup + 1}..{args.warmup + K} after the reset",
            "solver": args.solver,
            "tol": 1e-6,
            "mg_tiles": h["tiles"],
        },
        "pressure": {
            "ms_per_solve": wsum["ms_per_solve"],
            "vcycles_per_solve": wsum["vcycles_per_solve"],
            "max_relative_residual": wsum["max_relative_residual"],
            "over": f"all {WORKLOAD_STEPS} solves of the workload",
            "start": WARM_STARTS[args.warm][1],
            "other_start": h["other_start"],
            "learned_warm_start": h["learned"],
        },
        "workload_run": wsum,
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, n),
            "traffic_source": os.path.relpath(PMC_TRAFFIC, ROOT),
            "avg_launch_ms": k_avg_ms,
            "launches": h["n_launch"],
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "algorithmic_bytes_per_cell": bpc,
        },
        "kernels": h["kernels"],
        "legs": out_legs,
    }
    if args.cpu_baseline and h["state"] is not None:
        log("cpu baseline")
        out["cpu_baseline"] = run_cpu_baseline(h["state"], args)
        log("cpu baseline, JavaScript")
        out["cpu_baseline_js"] = run_cpu_baseline_js(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
