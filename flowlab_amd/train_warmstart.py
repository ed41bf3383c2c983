"""Train and evaluate the learned warm start on the GPU.

    python -m flowlab_amd.train_warmstart [--samples 20000] [--minutes 6] \
        [--out flowlab_amd/weights/unet_warmstart.npz] [--report report.json]

1. Data: plain multigrid solves (tol 1e-8) of seeded random 512x512 scenes
   (random circles/rectangles, 1-3 buoyant sources, vortex blobs), captured
   on the device by the engine itself on the network's 128x128 coarse grid:
   each step's solve from the previous pressure (even scenes) or from the
   linear extrapolation in time (odd scenes), and every fifth step the same
   right-hand side solved from zero.
2. Training: Adam on the masked MSE of the normalised coarse correction, by
   the library's own MFMA forward and backward kernels.
3. Evaluation at 1024^2 and 2048^2: V-cycles and device ms to 1e-6 from
   zero, the previous pressure and the extrapolation, each with and without
   the network.
Progress is printed at least once a minute.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import learned as L


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps-per-scene", type=int, default=50)
    ap.add_argument("--minutes", type=float, default=6.0, help="training time budget")
    ap.add_argument("--max-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-sizes", default="1024,2048")
    ap.add_argument("--eval-scenes", type=int, default=3)
    ap.add_argument("--out", default=L.WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--eval-only", action="store_true",
                    help="load --out and only evaluate it")
    a = ap.parse_args(argv)

    if a.eval_only:
        u = L.UNet()
        u.load(a.out)
        evals = {}
        for n in [int(s) for s in a.eval_sizes.split(",") if s]:
            evals[str(n)] = L.evaluate(u, n, scenes=a.eval_scenes)
            print(f"eval {n}^2: " + ", ".join(f"{k}: {v['vcycles']:.2f} cycles {v['ms']:.3f} ms"
                                            for k, v in evals[str(n)].items()), flush=True)
        if a.report:
            with open(a.report, "w") as f:
                json.dump(dict(weights=a.out, eval=evals), f, indent=1)
        return

    t0 = time.perf_counter()
    u = L.UNet(seed=a.seed)
    scenes = L.generate(u, a.samples, n=a.n, steps_per_scene=a.steps_per_scene, seed=100,
                        log=lambda s: print(s, flush=True))
    t_gen = time.perf_counter() - t0
    print(f"generated {u.samples} samples from {scenes} scenes in {t_gen:.1f} s", flush=True)
    u.calibrate(128)
    print(f"output scale c0 = {u.scale:.4g}", flush=True)

    losses = []
    t1 = time.perf_counter()
    chunk = 100
    lr = a.lr
    while len(losses) < a.max_steps and time.perf_counter() - t1 < 60 * a.minutes:
        # the last fifth of the budget (steps or time) at a tenth of the rate
        frac = max(len(losses) / max(1, a.max_steps),
                   (time.perf_counter() - t1) / (60 * a.minutes))
        lr_now = lr * (0.1 if frac > 0.8 else 1.0)
        losses += list(u.train(chunk, batch=a.batch, lr=lr_now, seed=a.seed + len(losses)))
        el = time.perf_counter() - t1
        print(f"train: step {len(losses)} loss(last {chunk}) {np.mean(losses[-chunk:]):.4g} "
              f"({el:.0f} s, {1e3 * el / len(losses):.1f} ms/step)", flush=True)
    t_train = time.perf_counter() - t1
    meta = dict(samples=u.samples, scenes=scenes, n=a.n, steps=len(losses), batch=a.batch,
                lr=a.lr, first_loss=float(np.mean(losses[:chunk])),
                last_loss=float(np.mean(losses[-chunk:])), train_s=t_train, gen_s=t_gen)
    u.save(a.out, meta)
    print(f"saved {a.out}", flush=True)

    evals = {}
    for n in [int(s) for s in a.eval_sizes.split(",") if s]:
        r = L.evaluate(u, n, scenes=a.eval_scenes)
        evals[str(n)] = r
        print(f"eval {n}^2: " + ", ".join(f"{k}: {v['vcycles']:.2f} cycles {v['ms']:.3f} ms"
                                        for k, v in r.items()), flush=True)
    report = dict(meta=meta, eval=evals, loss_curve=[float(x) for x in losses[::max(1, len(losses) // 200)]])
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(dict(meta=meta, eval=evals)), flush=True)


if __name__ == "__main__":
    main()
