"""Learned warm start for the pressure solve (no reference counterpart).

A 4-level convolutional U-Net (3x3 convs, channels 16-32-64-64, BF16 weights
and activations, FP32 accumulation on MFMA) runs inside libflowhip.  It sees
the defect of the previous pressure for the new right-hand side, the solid
and Air masks and the previous pressure, and predicts a correction; the
defect-correction multigrid then runs its V-cycles to the tolerance exactly
as without it, so plain multigrid stays both the fallback and the arbiter
of convergence.

This module: the network handle (parameters, save/load), the seeded scene
generator of the training workload, data generation from the engine's own
multigrid solves, training, and the zero / previous-pressure / learned
comparison.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import hip, scenes as S
from .hip import fptr

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "unet_warmstart.npz")


def _ucheck(status):
    if status != hip.FH_OK:
        raise hip.HipError(status, (hip.lib().fh_unet_last_error() or b"").decode())


class UNet:
    def __init__(self, seed: int = 0):
        self.lib = hip.lib()
        ptr = self.lib.fh_unet_create(seed)
        if not ptr:
            raise hip.HipError(hip.FH_ERR_NO_DEVICE, (self.lib.fh_unet_last_error() or b"").decode())
        self._u = C.c_void_p(ptr)
        self.n = self.lib.fh_unet_num_params()

    def close(self):
        if getattr(self, "_u", None):
            self.lib.fh_unet_destroy(self._u)
            self._u = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ------------------------------------------------------
    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.n, np.float32)
        _ucheck(self.lib.fh_unet_get_params(self._u, fptr(out), self.n))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _ucheck(self.lib.fh_unet_set_params(self._u, fptr(p), self.n))

    @property
    def scale(self):
        c0, a0 = C.c_double(), C.c_double()
        _ucheck(self.lib.fh_unet_get_scale(self._u, C.byref(c0), C.byref(a0)))
        return c0.value, a0.value

    @scale.setter
    def scale(self, v):
        _ucheck(self.lib.fh_unet_set_scale(self._u, float(v[0]), float(v[1])))

    def save(self, path: str = WEIGHTS, meta: dict | None = None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        c0, a0 = self.scale
        np.savez(path, params=self.params, c0=c0, area0=a0,
                 meta=json.dumps(meta or {}))

    def load(self, path: str = WEIGHTS):
        d = np.load(path, allow_pickle=False)
        self.params = d["params"]
        self.scale = (float(d["c0"]), float(d["area0"]))
        return json.loads(str(d["meta"]))

    # ---- raw network (tests) ---------------------------------------------
    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        N, H, W, _ = x.shape
        out = np.empty((N, H, W), np.float32)
        _ucheck(self.lib.fh_unet_forward(self._u, N, H, W, fptr(x), fptr(out)))
        return out

    def backward(self, x: np.ndarray, dout: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        dout = np.ascontiguousarray(dout, np.float32)
        N, H, W, _ = x.shape
        g = np.empty(self.n, np.float32)
        _ucheck(self.lib.fh_unet_backward(self._u, N, H, W, fptr(x), fptr(dout), fptr(g)))
        return g

    # ---- training ----------------------------------------------------------
    def reserve(self, capacity: int, H: int, W: int):
        _ucheck(self.lib.fh_unet_dataset_reserve(self._u, capacity, H, W))

    @property
    def samples(self) -> int:
        return self.lib.fh_unet_dataset_size(self._u)

    def calibrate(self, samples: int = 64):
        _ucheck(self.lib.fh_unet_calibrate(self._u, samples))

    def train(self, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        loss = np.zeros(steps)
        _ucheck(self.lib.fh_unet_train(self._u, steps, batch, lr, seed,
                                       loss.ctypes.data_as(C.POINTER(C.c_double))))
        return loss


# ---- the training / evaluation workload ----------------------------------

def random_scene(n: int, seed: int):
    """Seeded scene of the learned-warm-start workload on an n x n grid:
    3-8 random solid circles and 2-5 rectangles, 1-3 buoyant heat sources,
    a few Gaussian vortex blobs; closed box, or an Air outlet column with
    probability 0.3.  Returns (grid, scene, buoyancy weight)."""
    rng = S.SplitMix64(seed)
    g = S.Grid(n, n, 1.0 / n)
    seeds = []
    if rng.uniform() < 0.3:
        seeds.append(S.open_right())
    for _ in range(rng.randint(3, 8)):
        seeds.append(S.solid_disk(rng.uniform(0.1, 0.9), rng.uniform(0.3, 0.9),
                                  rng.uniform(0.02, 0.07)))
    for _ in range(rng.randint(2, 5)):
        x0, y0 = rng.uniform(0.05, 0.85), rng.uniform(0.3, 0.85)
        seeds.append(S.solid_rect(x0, y0, x0 + rng.uniform(0.02, 0.12),
                                  y0 + rng.uniform(0.02, 0.12)))
    blobs = [(rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.04, 0.1),
              rng.uniform(-1.0, 1.0)) for _ in range(rng.randint(1, 4))]

    def seed_fn(g, u, v):
        x = (np.arange(g.nx + 1) * g.h)[None, :]
        y = (np.arange(g.ny + 1) * g.h)[:, None]
        psi = np.zeros((g.ny + 1, g.nx + 1))
        for bx, by, sg, amp in blobs:
            psi += amp * sg * np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (sg * sg))
        S.add_curl_of_stream(g, u, v, psi, 1.0)

    sources = [(rng.uniform(0.15, 0.85), rng.uniform(0.05, 0.2), rng.uniform(0.02, 0.05))
               for _ in range(rng.randint(1, 3))]

    def dye_source(dg, gg):
        parts = [S.heat_disk_source(cx, cy, r)(dg, gg) for cx, cy, r in sources]
        # one full-width patch holding every source (coverage carries shape)
        cov = np.zeros((dg.ny, dg.nx), np.float32)
        val = np.zeros((3, dg.ny, dg.nx), np.float32)
        for p in parts:
            c = p.data[3 * p.nx * p.ny:].reshape(p.ny, p.nx)
            sl = (slice(p.j0, p.j0 + p.ny), slice(p.i0, p.i0 + p.nx))
            m = c > cov[sl]
            cov[sl][m] = c[m]
            for k in range(3):
                val[k][sl][m] = p.data[k * p.nx * p.ny:(k + 1) * p.nx * p.ny].reshape(p.ny, p.nx)[m]
        j1 = int(np.nonzero(cov.any(axis=1))[0].max()) + 1 if cov.any() else 1
        data = np.concatenate([val[:, :j1].reshape(3, -1).ravel(), cov[:j1].ravel()])
        return S.DyePatch(0, 0, dg.nx, j1, data.astype(np.float32))

    return g, S.SceneSpec(labels=S.all_labels(*seeds), seed=seed_fn,
                          dye_source=dye_source), rng.uniform(0.5, 2.0)


def _sim(n, seed, tol=1e-6):
    from .simulation import Simulation
    g, scene, w = random_scene(n, seed)
    sim = Simulation(n, n, dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 / n * 4,
                                solver="mg-dc", pressure_iters=100, tol=tol, buoyancy=w,
                                residual_every=0), readback=False)
    sim.reset(scene)
    return sim


def generate(unet: UNet, samples: int, n: int = 512, steps_per_scene: int = 50, seed: int = 100,
             log=print):
    """Fill unet's training set from plain multigrid solves (tol 1e-8) of
    seeded random n x n scenes, `steps_per_scene` steps each."""
    unet.reserve(samples, n, n)
    k = 0
    t0 = time.perf_counter()
    while unet.samples < samples:
        sim = _sim(n, seed + k, tol=1e-8)
        sim.stepper.engine.set_capture(unet)
        for _ in range(steps_per_scene):
            if unet.samples >= samples:
                break
            sim.step()
        sim.stepper.engine.set_capture(None)
        k += 1
        if log and k % 20 == 0:
            log(f"generate: {unet.samples}/{samples} samples from {k} scenes "
                f"({time.perf_counter() - t0:.1f} s)")
    return k


def evaluate(unet: UNet, n: int, scenes: int = 4, warm_steps: int = 8, seed: int = 9000,
             tol: float = 1e-6):
    """For `scenes` seeded n x n scenes advanced `warm_steps` steps, solve the
    next step's pressure from (a) zero, (b) the previous pressure, (c) the
    learned warm start on the previous pressure, (d) the learned warm start on
    a zero pressure (the network's own contribution, without the previous
    pressure), each to `tol`; returns V-cycles and device ms."""
    from .engine import HipPressureSolver
    res = {"zero": [], "previous": [], "learned": [], "network": []}
    for s in range(scenes):
        sim = _sim(n, seed + s, tol=tol)
        for _ in range(warm_steps):
            sim.step()
        e = sim.stepper.engine
        # the next step's right-hand side: advance once more with the
        # solve's input captured from the engine's own buffers
        p_prev = e.read_pressure()
        sim.step()
        div = e.read_divergence()
        label = sim.f.label.copy()
        scale = sim.g.h * sim.g.h / sim.dt
        solver = HipPressureSolver(n, n, "mg-dc")
        for name in ("zero", "previous", "learned", "network"):
            solver.engine.set_warm_start(unet if name in ("learned", "network") else None)
            from_zero = name in ("zero", "network")
            p = np.zeros(n * n) if from_zero else p_prev.copy()
            solver.solve(None, p, div, label, scale, 1, 1.0, 0)  # warm-up: labels, buffers
            p = np.zeros(n * n) if from_zero else p_prev.copy()
            it = solver.solve(None, p, div, label, scale, 200, 1.0, tol)
            res[name].append((it, solver.engine.last_solve_ms()))
        solver.engine.set_warm_start(None)
    out = {}
    for k, v in res.items():
        a = np.array(v)
        out[k] = {"vcycles": float(a[:, 0].mean()), "ms": float(a[:, 1].mean())}
    return out
