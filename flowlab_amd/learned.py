"""Learned warm start for the pressure solve (no reference counterpart).

A 4-level convolutional U-Net (3x3 convs, channels 16-32-64-64, BF16 weights
and activations, FP32 accumulation on MFMA) runs inside libflowhip on a
coarse view of the grid: the defect of the solve's start summed over F x F
blocks (coarse sides at most 128), the blocks' solid and Air fractions and
their fluid-mean start pressure.  Its output, interpolated onto the fine
fluid cells, is added to the start with the energy-optimal step; the
defect-correction multigrid then runs its V-cycles to the tolerance exactly
as without it, so plain multigrid stays both the fallback and the arbiter of
convergence.

This module: the network handle (parameters, save/load), the seeded scene
generator of the training workload, data generation from the engine's own
multigrid solves, training, and the comparison of starts.
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
    def scale(self) -> float:
        """c0: the coarse correction is s_r * c0 * the network's output."""
        c0 = C.c_double()
        _ucheck(self.lib.fh_unet_get_scale(self._u, C.byref(c0)))
        return c0.value

    @scale.setter
    def scale(self, c0: float):
        _ucheck(self.lib.fh_unet_set_scale(self._u, float(c0)))

    def save(self, path: str = WEIGHTS, meta: dict | None = None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, c0=self.scale, meta=json.dumps(meta or {}))

    def load(self, path: str = WEIGHTS):
        d = np.load(path, allow_pickle=False)
        self.params = d["params"]
        self.scale = float(d["c0"])
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
        """Room for `capacity` samples of an H x W coarse grid (coarse_grid)."""
        _ucheck(self.lib.fh_unet_dataset_reserve(self._u, capacity, H, W))
        self._shape = (H, W)

    @property
    def samples(self) -> int:
        return self.lib.fh_unet_dataset_size(self._u)

    def sample(self, index: int):
        """Captured sample `index`: (x [H][W][4], t [H][W], (s_r, s_p))."""
        n = self.lib.fh_unet_dataset_size(self._u)
        if not 0 <= index < n:
            raise IndexError(index)
        H, W = self._shape
        x = np.empty((H, W, 4), np.float32)
        t = np.empty((H, W), np.float32)
        sc = np.empty(2, np.float32)
        _ucheck(self.lib.fh_unet_dataset_get(self._u, index, fptr(x), fptr(t), fptr(sc)))
        return x, t, (float(sc[0]), float(sc[1]))

    def calibrate(self, samples: int = 64):
        _ucheck(self.lib.fh_unet_calibrate(self._u, samples))

    def train(self, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        loss = np.zeros(steps)
        _ucheck(self.lib.fh_unet_train(self._u, steps, batch, lr, seed,
                                       loss.ctypes.data_as(C.POINTER(C.c_double))))
        return loss


def coarse_grid(nx: int, ny: int):
    """(F, coarse width, coarse height) of the grid the network runs on for
    an nx x ny grid; F = 0 when the network cannot take it."""
    cw, ch = C.c_int(), C.c_int()
    F = hip.lib().fh_unet_coarse_grid(nx, ny, C.byref(cw), C.byref(ch))
    return F, cw.value, ch.value


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


def _sim(n, seed, tol=1e-6, extrapolation=0):
    from .simulation import Simulation
    g, scene, w = random_scene(n, seed)
    sim = Simulation(n, n, dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 / n * 4,
                                solver="mg-dc", pressure_iters=100, tol=tol, buoyancy=w,
                                residual_every=0, warm_extrapolation=extrapolation),
                     readback=False)
    sim.reset(scene)
    return sim


def generate(unet: UNet, samples: int, n: int = 512, steps_per_scene: int = 50, seed: int = 100,
             zero_every: int = 5, log=print):
    """Fill unet's training set from plain multigrid solves (tol 1e-8) of
    seeded random n x n scenes, `steps_per_scene` steps each: each step's
    solve, from the previous pressure in even-numbered scenes and from the
    linear extrapolation in time in odd-numbered ones, and every
    `zero_every`-th step also the same right-hand side solved from zero."""
    from .engine import HipPressureSolver
    F, cw, ch = coarse_grid(n, n)
    if not F:
        raise ValueError(f"no coarse grid for {n} x {n}")
    unet.reserve(samples, ch, cw)
    zero = HipPressureSolver(n, n, "mg-dc")
    zero.engine.set_capture(unet)
    k = 0
    t0 = time.perf_counter()
    while unet.samples < samples:
        sim = _sim(n, seed + k, tol=1e-8, extrapolation=k % 2)
        sim.stepper.engine.set_capture(unet)
        for st in range(steps_per_scene):
            if unet.samples >= samples:
                break
            sim.step()
            if zero_every and st % zero_every == zero_every - 1 and unet.samples < samples:
                e = sim.stepper.engine
                zero.solve(None, np.zeros(n * n), e.read_divergence(), sim.f.label,
                           sim.g.h * sim.g.h / sim.dt, 100, 1.0, 1e-8)
        sim.stepper.engine.set_capture(None)
        k += 1
        if log and k % 20 == 0:
            log(f"generate: {unet.samples}/{samples} samples from {k} scenes "
                f"({time.perf_counter() - t0:.1f} s)")
    zero.engine.set_capture(None)
    return k


START_NAMES = ("zero", "network", "previous", "learned", "extrapolated", "learned_extrapolated")


def evaluate(unet: UNet, n: int, scenes: int = 4, warm_steps: int = 8, seed: int = 9000,
             tol: float = 1e-6):
    """For `scenes` seeded n x n scenes advanced `warm_steps` steps, solve the
    next step's pressure to `tol` from six starts: zero; the network on a
    zero pressure (its own contribution, without the previous pressure); the
    previous pressure (the reference's warm start); the network on the
    previous pressure; the last two pressures extrapolated linearly in time
    at the steps' own dt; the network on that extrapolation.  Returns
    V-cycles and device ms per start."""
    from .engine import HipPressureSolver
    res = {k: [] for k in START_NAMES}
    for s in range(scenes):
        sim = _sim(n, seed + s, tol=tol)
        for _ in range(warm_steps - 1):
            sim.step()
        e = sim.stepper.engine
        p_a = e.read_pressure()
        sim.step()
        dt_a = sim.dt
        p_b = e.read_pressure()
        # the next step's right-hand side, captured from the engine's buffers
        sim.step()
        dt_b = sim.dt
        div = e.read_divergence()
        label = sim.f.label.copy()
        scale = sim.g.h * sim.g.h / sim.dt
        starts = {"zero": np.zeros(n * n), "previous": p_b,
                  "extrapolated": p_b + (dt_b / dt_a) * (p_b - p_a)}
        solver = HipPressureSolver(n, n, "mg-dc")
        for name in START_NAMES:
            base = {"network": "zero", "learned": "previous",
                    "learned_extrapolated": "extrapolated"}.get(name, name)
            solver.engine.set_warm_start(unet if base != name else None)
            p = starts[base].copy()
            solver.solve(None, p, div, label, scale, 1, 1.0, 0)  # warm-up: labels, buffers
            p = starts[base].copy()
            it = solver.solve(None, p, div, label, scale, 200, 1.0, tol)
            res[name].append((it, solver.engine.last_solve_ms()))
        solver.engine.set_warm_start(None)
    out = {}
    for k, v in res.items():
        a = np.array(v)
        out[k] = {"vcycles": float(a[:, 0].mean()), "ms": float(a[:, 1].mean())}
    return out
