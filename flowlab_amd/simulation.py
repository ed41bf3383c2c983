"""Simulation: the reference's orchestrator (src/core/simulation.ts:259-593)
driving the HIP fused step.

Owns the host fields, dt/CFL and the step order; every kernel of the step
runs on the GPU through HipStepper (the reference's production path, where
`stepper` is set and `advector`/`solver` sit unused, simulation.ts:307-312).
There is no CPU path here: without libflowhip.so construction fails.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence, Union

import numpy as np

from .engine import HipStepper
from .scenes import AIR, FLUID, SOLID, DyePatch, Grid, SceneSpec

PerChannel = Sequence[float]


@dataclass
class SimulationParams:
    """simulation.ts:134-207 plus the build's knobs (solver, confinement)."""
    cfl_target: float = 2.0
    scheme: str = "semiLagrangian"
    dt_max: float = 1 / 30
    pressure_iters: int = 100
    tol: float = 5e-3
    omega: float = 0.0  # 0 -> optimal_omega(nx, ny)
    rho: float = 1.0
    dye_decay: Union[float, PerChannel] = 0.0
    buoyancy: float = 0.0
    residual_every: int = 1
    confinement: float = 0.0
    solver: str = "mg-dc"
    # build-defined: start each solve from the previous solutions
    # extrapolated in time (1 linear, 2 quadratic); 0 = the previous pressure
    warm_extrapolation: int = 0


def optimal_omega(nx: int, ny: Optional[int] = None) -> float:
    """simulation.ts:218-221"""
    ny = nx if ny is None else ny
    rho = 0.5 * (math.cos(math.pi / nx) + math.cos(math.pi / ny))
    return 2 / (1 + math.sqrt(1 - rho * rho))


def dye_keep_for(decay, dt: float):
    """simulation.ts:243-247"""
    k = (lambda d: math.exp(-d * dt) if d > 0 else 1.0)
    if isinstance(decay, (int, float)):
        return (k(decay),) * 3
    return tuple(k(d) for d in decay)


@dataclass
class Fields:
    """simulation.ts Fields (grid.ts:84-97), float64 host mirrors."""
    p: np.ndarray
    u: np.ndarray
    v: np.ndarray
    label: np.ndarray
    dye: list


@dataclass
class Stir:
    """stir.ts:21-34 -- world units."""
    x: float
    y: float
    r: float
    dx: float
    dy: float


class Profiler:
    """Per-phase device milliseconds of the last step (profiler.ts:50-113
    phases collapse to the device's: advect, forces, divergence, solve,
    project, dye, total), smoothed by the reference's EMA alpha 0.12."""

    ALPHA = 0.12

    def __init__(self):
        self.last: dict = {}
        self.ema: dict = {}

    def record(self, engine):
        t = engine.timings()
        if not any(t.values()):
            return
        self.last = t
        for k, v in t.items():
            self.ema[k] = v if k not in self.ema else self.ema[k] + self.ALPHA * (v - self.ema[k])


def _divergence(g: Grid, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """divergence.ts:17-25 -- the host-side residual diagnostic the reference
    computes from what came back from the device (simulation.ts:495-497)."""
    U = u.reshape(g.ny, g.nx + 1)
    V = v.reshape(g.ny + 1, g.nx)
    return ((1 / g.h) * (U[:, 1:] - U[:, :-1] + V[1:, :] - V[:-1, :])).ravel()


class Simulation:
    """One advect -> project loop over a MAC grid, stepped on the GPU."""

    def __init__(self, nx: int, ny: Optional[int] = None, params: Optional[dict] = None,
                 dye_scale: int = 1, timing: bool = False, readback: bool = True):
        ny = nx if ny is None else ny
        prm = SimulationParams(**(params or {}))
        if prm.omega == 0.0:
            prm = replace(prm, omega=optimal_omega(nx, ny))
        self.params = prm
        self.g = Grid(nx, ny, 1.0 / ny)
        s = max(1, int(round(dye_scale)))
        self.dyeG = self.g if s == 1 else Grid(nx * s, ny * s, 1.0 / (ny * s))
        nd = self.dyeG.nx * self.dyeG.ny
        self.f = Fields(p=np.zeros(nx * ny), u=np.zeros((nx + 1) * ny), v=np.zeros(nx * (ny + 1)),
                        label=np.zeros(nx * ny, np.uint8), dye=[np.zeros(nd) for _ in range(3)])
        self.div = np.zeros(nx * ny)
        self.time = 0.0
        self.dt = 0.0
        self.iters = 0
        self.stir: Optional[Stir] = None
        self.perf = Profiler()
        self.stepper = HipStepper(nx, ny, s, self.g.h, solver=prm.solver, readback=readback)
        if timing:
            self.stepper.engine.set_timing(True)
        if prm.warm_extrapolation:
            self.stepper.engine.set_warm_extrapolation(prm.warm_extrapolation)
        self._steps = 0
        self._dye_patch: Optional[DyePatch] = None

    def reset(self, scene: Optional[SceneSpec] = None) -> None:
        """simulation.ts:388-416: zero, labels, seed, dye, commitLabels, the
        dye source applied once at t = 0, then invalidate the device copy."""
        scene = scene or SceneSpec()
        g, f = self.g, self.f
        f.u.fill(0)
        f.v.fill(0)
        f.p.fill(0)
        f.label.fill(FLUID)
        for c in f.dye:
            c.fill(0)
        if scene.labels:
            scene.labels(g, f.label)
        if scene.seed:
            scene.seed(g, f.u, f.v)
        if scene.dye:
            scene.dye(self.dyeG, f.dye)
        commit_labels(g, f)
        self._dye_patch = scene.dye_source(self.dyeG, g) if scene.dye_source else None
        if self._dye_patch is not None:
            apply_dye_patch(self.dyeG, f.dye, self._dye_patch)
        self.stepper.set_dye_patch(self._dye_patch)
        self.stepper.set_buoyancy(self.params.buoyancy)
        self.stepper.set_confinement(self.params.confinement)
        self.time = 0.0
        self.dt = 0.0
        self.stepper.invalidate()

    def commit_label_edits(self) -> None:
        """simulation.ts:436-439"""
        commit_labels(self.g, self.f)
        self.stepper.set_labels(self.f.label)

    def max_face_speed(self) -> float:
        """simulation.ts:451-459 (device reduction once the fields are resident)."""
        if self.stepper._dirty:
            return float(max(np.abs(self.f.u).max(initial=0), np.abs(self.f.v).max(initial=0)))
        return self.stepper.max_face_speed()

    def step(self) -> None:
        """simulation.ts:467-497 -- the fused (stepper) branch."""
        g, p = self.g, self.params
        stir, self.stir = self.stir, None
        u_max = self.max_face_speed()
        dt = min(p.dt_max, p.cfl_target * g.h / u_max) if u_max > 1e-9 else p.dt_max
        scale = p.rho * g.h * g.h / dt
        grad_scale = dt / (p.rho * g.h)
        self.stepper.set_stir(stir)
        keep = dye_keep_for(p.dye_decay, dt)
        self.iters = self.stepper.step(self.f, p.scheme, dt, scale, grad_scale, keep, self.perf,
                                       iterations=p.pressure_iters, omega=p.omega, tol=p.tol)
        if self.stepper.readback and p.residual_every > 0 and self._steps % p.residual_every == 0:
            self.div = _divergence(g, self.f.u, self.f.v)
        self.dt = dt
        self.time += dt
        self._steps += 1

    def read_dye(self) -> list:
        self.stepper.read_dye(self.f.dye)
        return self.f.dye

    @property
    def cfl(self) -> float:
        return self.max_face_speed() * self.dt / self.g.h


def commit_labels(g: Grid, f: Fields) -> None:
    """boundaries.ts:45-64 on the host fields (scene setup, before upload)."""
    L = f.label.reshape(g.ny, g.nx)
    nonfluid = f.label != FLUID
    f.p[nonfluid] = 0
    solid = f.label == SOLID
    for c in f.dye:
        c[:solid.size][solid] = 0
    S = L == SOLID
    U = f.u.reshape(g.ny, g.nx + 1)
    su = np.zeros((g.ny, g.nx + 1), bool)
    su[:, 1:] |= S
    su[:, :-1] |= S
    U[su] = 0
    V = f.v.reshape(g.ny + 1, g.nx)
    sv = np.zeros((g.ny + 1, g.nx), bool)
    sv[1:, :] |= S
    sv[:-1, :] |= S
    V[sv] = 0


def apply_dye_patch(dg: Grid, dye: list, patch: DyePatch) -> None:
    """dye.ts:70-87 on the host fields (the t = 0 application of reset)."""
    cells = patch.nx * patch.ny
    cov = patch.data[3 * cells:].reshape(patch.ny, patch.nx).astype(np.float64)
    j0, j1 = max(patch.j0, 0), min(patch.j0 + patch.ny, dg.ny)
    i0, i1 = max(patch.i0, 0), min(patch.i0 + patch.nx, dg.nx)
    if j0 >= j1 or i0 >= i1:
        return
    cs = cov[j0 - patch.j0:j1 - patch.j0, i0 - patch.i0:i1 - patch.i0]
    m = cs != 0
    for c in range(3):
        val = patch.data[c * cells:(c + 1) * cells].reshape(patch.ny, patch.nx).astype(np.float64)
        vs = val[j0 - patch.j0:j1 - patch.j0, i0 - patch.i0:i1 - patch.i0]
        D = dye[c].reshape(dg.ny, dg.nx)[j0:j1, i0:i1]
        D[m] += cs[m] * (vs[m] - D[m])


def scene_fields(g: Grid, scene: SceneSpec) -> Fields:
    """The host fields Simulation.reset builds (simulation.ts:388-416) for a
    scene without a dye source: labels, velocity seed, dye seed, then
    commitLabels.  For callers that own no stepper (batches)."""
    n = g.nx * g.ny
    f = Fields(p=np.zeros(n), u=np.zeros((g.nx + 1) * g.ny), v=np.zeros(g.nx * (g.ny + 1)),
               label=np.full(n, FLUID, np.uint8), dye=[np.zeros(n) for _ in range(3)])
    if scene.labels:
        scene.labels(g, f.label)
    if scene.seed:
        scene.seed(g, f.u, f.v)
    if scene.dye:
        scene.dye(g, f.dye)
    commit_labels(g, f)
    return f
