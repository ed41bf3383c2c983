"""Many independent scenes advanced together on one GPU (libflowhip's
fh_batch_* entry points, include/flowhip.h).

Each scene steps as Simulation.step does with solver 'mg-fcg' (dt from its
own CFL, simulation.ts:467-587), bit for bit as a single-grid engine; a batch
step launches each phase once for all scenes over a (scene, block) work list.
The multigrid tile shape of each level index can be set or autotuned.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import hip
from .scenes import Grid, SceneSpec
from .simulation import Fields, scene_fields


def tile_names() -> list[str]:
    L = hip.lib()
    return [L.fh_batch_tile_name(t).decode() for t in range(L.fh_batch_num_tiles())]


class HipBatch:
    def __init__(self, grids: Sequence[Grid]):
        L = hip.lib()
        n = len(grids)
        self.grids = list(grids)
        nx = (C.c_int * n)(*[g.nx for g in grids])
        ny = (C.c_int * n)(*[g.ny for g in grids])
        h = (C.c_double * n)(*[g.h for g in grids])
        self._b = L.fh_batch_create(n, nx, ny, h)
        if not self._b:
            raise hip.HipError(hip.FH_ERR_NO_DEVICE, (L.fh_batch_last_error() or b"").decode())

    def close(self) -> None:
        if getattr(self, "_b", None):
            hip.lib().fh_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()

    def __len__(self) -> int:
        return len(self.grids)

    @property
    def cells(self) -> int:
        return int(hip.lib().fh_batch_cells(self._b))

    def upload(self, s: int, f: Fields) -> None:
        dye = np.concatenate(f.dye) if f.dye is not None else None
        hip.check_batch(hip.lib().fh_batch_upload(
            self._b, s, hip.dptr(f.u), hip.dptr(f.v), hip.dptr(f.p),
            f.label.ctypes.data_as(hip._u8p), hip.dptr(dye)))

    def reset(self, s: int, scene: SceneSpec) -> Fields:
        f = scene_fields(self.grids[s], scene)
        self.upload(s, f)
        return f

    def set_forces(self, s: int, buoyancy=0.0, confinement=0.0, decay=(0.0, 0.0, 0.0)) -> None:
        d = np.asarray(decay, np.float64)
        hip.check_batch(hip.lib().fh_batch_set_forces(self._b, s, buoyancy, confinement,
                                                      hip.dptr(d)))

    def set_warm_extrapolation(self, order: int) -> None:
        """Start every scene's solve from its last solutions extrapolated in
        time (1 linear, 2 quadratic; 0 = the previous pressure)."""
        hip.check_batch(hip.lib().fh_batch_set_warm_extrapolation(self._b, int(order)))

    def step(self, scheme="macCormack", cfl=0.5, dt_max=1.0, rho=1.0, tol=1e-6,
             max_cycles=60) -> None:
        hip.check_batch(hip.lib().fh_batch_step(self._b, hip.SCHEMES[scheme], cfl, dt_max, rho,
                                                tol, max_cycles))

    def results(self):
        """(dt, V-cycles, relative residual) per scene of the last step."""
        n = len(self)
        dt, res = np.zeros(n), np.zeros(n)
        it = np.zeros(n, np.int32)
        hip.check_batch(hip.lib().fh_batch_results(self._b, hip.dptr(dt),
                                                   it.ctypes.data_as(hip._ip), hip.dptr(res)))
        return dt, it, res

    def last_step_ms(self) -> float:
        ms = C.c_float()
        hip.check_batch(hip.lib().fh_batch_last_step_ms(self._b, C.byref(ms)))
        return float(ms.value)

    def read(self, s: int):
        """(u, v, p, dye[3]) of scene s, FP64 host copies."""
        g = self.grids[s]
        u = np.zeros((g.nx + 1) * g.ny)
        v = np.zeros(g.nx * (g.ny + 1))
        p = np.zeros(g.nx * g.ny)
        dye = np.zeros(3 * g.nx * g.ny)
        hip.check_batch(hip.lib().fh_batch_read(self._b, s, hip.dptr(u), hip.dptr(v),
                                                hip.dptr(p), hip.dptr(dye)))
        return u, v, p, np.split(dye, 3)

    def set_tile(self, shape: int | str, level: int = -1) -> None:
        if isinstance(shape, str):
            shape = tile_names().index(shape)
        hip.check_batch(hip.lib().fh_batch_set_tile(self._b, level, shape))

    def tiles(self) -> list[str]:
        names = tile_names()
        shapes = (C.c_int * 16)()
        n = C.c_int()
        hip.check_batch(hip.lib().fh_batch_get_tiles(self._b, shapes, C.byref(n)))
        return [names[shapes[l]] for l in range(n.value)]

    def tune(self, reps: int = 5) -> dict:
        """Autotune the tile shape of every tiled level; returns the chosen
        shapes and every candidate's ms per pass."""
        names = tile_names()
        nl = len(self.tiles())
        ms = np.zeros(max(1, nl) * len(names), np.float32)
        hip.check_batch(hip.lib().fh_batch_tune(self._b, reps, ms.ctypes.data_as(hip._fp)))
        return {"tiles": self.tiles(),
                "ms": [{names[t]: float(ms[l * len(names) + t]) for t in range(len(names))}
                       for l in range(nl)]}
