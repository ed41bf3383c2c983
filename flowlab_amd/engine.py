"""The reference's three engine seams, backed by libflowhip on the GPU.

  HipStepper         implements GpuStep        (src/core/simulation.ts:35-84)
  HipPressureSolver  implements PressureSolver (src/core/pressureSolver.ts:21-41)
  HipAdvector        implements Advector       (src/core/advector.ts:28-63)

Each owns one fh_engine (one grid on one HIP stream).  Arguments mean what
they mean in the reference; errors are raised with the reference's messages
(a re-entrant call -> "<Class>.<method>() is not re-entrant", a lost device
-> "GPU device lost").  Host fields are numpy float64 arrays in the
reference's layouts, labels uint8.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip
from .hip import _sized, check, dptr, fptr, u8ptr


class HipEngine:
    """One device-resident MAC grid (nx x ny, cell size h, dye refined by
    dye_scale).  The low-level handle every seam below wraps."""

    def __init__(self, nx: int, ny: int, dye_scale: int = 1, h: float | None = None):
        self.lib = hip.lib()
        self.nx, self.ny, self.ds = nx, ny, dye_scale
        self.h = 1.0 / ny if h is None else h
        ptr = self.lib.fh_create(nx, ny, dye_scale, self.h)
        if not ptr:
            raise hip.HipError(hip.FH_ERR_NO_DEVICE, hip.last_error())
        self._e = C.c_void_p(ptr)
        self.dnx, self.dny = nx * dye_scale, ny * dye_scale

    def close(self):
        if getattr(self, "_e", None):
            self.lib.fh_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # sizes
    @property
    def n_u(self):
        return (self.nx + 1) * self.ny

    @property
    def n_v(self):
        return self.nx * (self.ny + 1)

    @property
    def n_c(self):
        return self.nx * self.ny

    @property
    def n_d(self):
        return self.dnx * self.dny

    # ---- state ----------------------------------------------------------
    def upload(self, u=None, v=None, p=None, label=None, dye=None):
        check(self.lib.fh_upload(self._e, dptr(u, self.n_u, "u"), dptr(v, self.n_v, "v"),
                                 dptr(p, self.n_c, "p"), u8ptr(label, self.n_c),
                                 dptr(dye, 3 * self.n_d, "dye (3 planes)")))

    def set_labels(self, label: np.ndarray):
        if label is None:
            raise TypeError("set_labels: label is required")
        check(self.lib.fh_set_labels(self._e, u8ptr(label, self.n_c)))

    def set_dye_patch(self, patch):
        if patch is None:
            check(self.lib.fh_set_dye_patch(self._e, 0, 0, 0, 0, None))
        else:
            data = np.ascontiguousarray(patch.data, np.float32)
            check(self.lib.fh_set_dye_patch(self._e, patch.i0, patch.j0, patch.nx, patch.ny,
                                            fptr(data, 4 * patch.nx * patch.ny,
                                                 "dye patch (4 planes)")))

    def set_buoyancy(self, w: float):
        check(self.lib.fh_set_buoyancy(self._e, float(w)))

    def set_confinement(self, eps: float):
        check(self.lib.fh_set_confinement(self._e, float(eps)))

    def set_stir(self, stir):
        if stir is None:
            check(self.lib.fh_set_stir(self._e, 0, 0, 0, 0, 0, 0))
        else:
            x, y, r, dx, dy = stir
            check(self.lib.fh_set_stir(self._e, 1, x, y, r, dx, dy))

    def read_velocity(self, u=None, v=None):
        u = np.empty(self.n_u) if u is None else u
        v = np.empty(self.n_v) if v is None else v
        check(self.lib.fh_read_velocity(self._e, dptr(u, self.n_u, "u"), dptr(v, self.n_v, "v")))
        return u, v

    def read_dye(self, dye=None):
        dye = np.empty(3 * self.n_d) if dye is None else dye
        check(self.lib.fh_read_dye(self._e, dptr(dye, 3 * self.n_d, "dye (3 planes)")))
        return dye

    def read_pressure(self):
        p = np.empty(self.n_c)
        check(self.lib.fh_read_pressure(self._e, dptr(p, self.n_c, "p")))
        return p

    def read_divergence(self):
        d = np.empty(self.n_c)
        check(self.lib.fh_read_divergence(self._e, dptr(d, self.n_c, "div")))
        return d

    def max_face_speed(self) -> float:
        out = C.c_double()
        check(self.lib.fh_max_face_speed(self._e, C.byref(out)))
        return out.value

    # ---- the fused step and its phases -----------------------------------
    def step(self, scheme, dt, scale, grad_scale, keep=(1.0, 1.0, 1.0), solver="mg-dc",
             iterations=100, omega=1.0, tol=1e-6) -> int:
        k = (C.c_double * 3)(*keep)
        it = C.c_int()
        check(self.lib.fh_step(self._e, hip.SCHEMES[scheme], dt, scale, grad_scale,
                               C.cast(k, hip._dp), hip.SOLVERS[solver], iterations, omega, tol,
                               C.byref(it)))
        return it.value

    def phase_advect_velocity(self, scheme, dt):
        check(self.lib.fh_phase_advect_velocity(self._e, hip.SCHEMES[scheme], dt))

    def phase_forces(self, dt):
        check(self.lib.fh_phase_forces(self._e, dt))

    def phase_divergence(self):
        check(self.lib.fh_phase_divergence(self._e))

    def phase_solve(self, solver, scale, iterations, omega=1.0, tol=0.0) -> int:
        it = C.c_int()
        check(self.lib.fh_phase_solve(self._e, hip.SOLVERS[solver], scale, iterations, omega, tol,
                                      C.byref(it)))
        return it.value

    def phase_project(self, grad_scale):
        check(self.lib.fh_phase_project(self._e, grad_scale))

    def phase_dye(self, scheme, dt, keep=(1.0, 1.0, 1.0)):
        k = (C.c_double * 3)(*keep)
        check(self.lib.fh_phase_dye(self._e, hip.SCHEMES[scheme], dt, C.cast(k, hip._dp)))

    def solve_history(self, max_entries=512) -> np.ndarray:
        out = np.zeros(max_entries)
        n = self.lib.fh_solve_history(self._e, dptr(out), max_entries)
        return out[:n]

    def set_timing(self, on: bool):
        check(self.lib.fh_set_timing(self._e, 1 if on else 0))
        self._timing = bool(on)

    def timings(self) -> dict:
        ms = np.zeros(len(hip.PHASES))
        check(self.lib.fh_get_timings(self._e, dptr(ms)))
        return dict(zip(hip.PHASES, ms.tolist()))

    def synchronize(self):
        check(self.lib.fh_synchronize(self._e))

    def set_capture(self, unet):
        """Append every defect-correction solve to unet's training set."""
        check(self.lib.fh_set_capture(self._e, None if unet is None else unet._u))

    def set_warm_start(self, unet):
        """Learned warm start for defect-correction solves (None = p_prev)."""
        check(self.lib.fh_set_warm_start(self._e, None if unet is None else unet._u))

    def set_warm_extrapolation(self, order: int):
        """Start each step's solve from the previous solutions extrapolated
        in time (order 1 linear, 2 quadratic; 0 = the previous pressure)."""
        check(self.lib.fh_set_warm_extrapolation(self._e, int(order)))

    def set_fused_update(self, on: bool):
        """Defect-correction cycles run the finest prolongation + smoothing
        and the FP64 update as one kernel (default) or as two passes."""
        check(self.lib.fh_set_fused_update(self._e, 1 if on else 0))

    def last_solve_ms(self) -> float:
        ms = C.c_double()
        check(self.lib.fh_last_solve_ms(self._e, C.byref(ms)))
        return ms.value

    def trace_kernel(self, name: str | None):
        check(self.lib.fh_trace_kernel(self._e, 0 if name is None else hip.KERNELS[name]))

    def trace_stats(self):
        n, ms = C.c_int(), C.c_double()
        check(self.lib.fh_trace_stats(self._e, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def tile_names(self) -> list:
        return [self.lib.fh_tile_shape_name(t).decode()
                for t in range(self.lib.fh_num_tile_shapes())]

    def set_tile(self, shape, level: int = -1):
        if isinstance(shape, str):
            shape = self.tile_names().index(shape)
        check(self.lib.fh_set_tile(self._e, level, shape))

    def tiles(self) -> list:
        names = self.tile_names()
        shapes = (C.c_int * 16)()
        n = C.c_int()
        check(self.lib.fh_get_tiles(self._e, shapes, C.byref(n)))
        return [names[shapes[l]] for l in range(n.value)]

    def set_coarse_group(self, group: int):
        check(self.lib.fh_set_coarse_group(self._e, group))

    def set_advect_staging(self, on: bool):
        """LDS-staged advection kernels (default) or plain global gathers."""
        check(self.lib.fh_set_advect_staging(self._e, int(bool(on))))

    def set_dye_one_pass(self, on: bool):
        """MacCormack dye advection in one kernel (default) or as its two
        passes."""
        check(self.lib.fh_set_dye_one_pass(self._e, int(bool(on))))

    def set_faces_one_pass(self, on: bool):
        """MacCormack velocity advection in one kernel (default) or as its
        two passes."""
        check(self.lib.fh_set_faces_one_pass(self._e, int(bool(on))))

    def coarse_group_threads(self) -> int:
        return int(self.lib.fh_coarse_group_threads(self._e))

    def tune_tiles(self, reps: int = 10) -> dict:
        """Autotune the multigrid tile shape per level on this grid."""
        names = self.tile_names()
        n = len(self.tiles())
        ms = np.zeros(max(n, 1) * len(names), np.float32)
        check(self.lib.fh_tune_tiles(self._e, reps, ms.ctypes.data_as(C.POINTER(C.c_float))))
        return {"tiles": self.tiles(), "coarse_threads": self.coarse_group_threads(),
                "ms": [{names[t]: float(ms[l * len(names) + t]) for t in range(len(names))}
                       for l in range(n)]}

    def trace_overhead(self, reps: int = 200) -> float:
        """Device ms of an empty event bracket on this engine's stream."""
        ms = C.c_double()
        check(self.lib.fh_trace_overhead(self._e, reps, C.byref(ms)))
        return ms.value


class HipStepper:
    """GpuStep (simulation.ts:35-84) on libflowhip: the whole advect ->
    project -> dye stretch as one device pass with u, v, p and dye resident.

    step() returns the solver's iteration count; on return f.u and f.v hold
    the step's velocity (read back, as the fused WebGPU step does,
    stepGpu.ts:546-604) and f.p is NOT updated; the dye stays resident and
    is read with read_dye (simulation.ts:29-33, 75-83)."""

    name = "hip-fused"

    def __init__(self, nx, ny, dye_scale=1, h=None, solver="mg-dc", readback=True):
        self.engine = HipEngine(nx, ny, dye_scale, h)
        self.solver = solver
        self.readback = readback
        self._dirty = True

    def upload(self, f):
        """The re-upload invalidate() asks for (stepGpu.ts:419-433): every host
        field, with the labels' multigrid hierarchy rebuilt.  step() does it
        when needed; a caller that times steps on resident fields does it
        first."""
        self.engine.upload(f.u, f.v, f.p, f.label, np.concatenate(f.dye))
        self._dirty = False

    def step(self, f, scheme, dt, scale, grad_scale, dye_keep, perf=None, iterations=100,
             omega=1.0, tol=1e-6) -> int:
        e = self.engine
        if self._dirty:
            self.upload(f)
        iters = e.step(scheme, dt, scale, grad_scale, dye_keep, self.solver, iterations, omega,
                       tol)
        if self.readback:
            e.read_velocity(f.u, f.v)
        if perf is not None:
            perf.record(e)
        return iters

    def invalidate(self):
        self._dirty = True

    def set_dye_patch(self, patch):
        self.engine.set_dye_patch(patch)

    def set_buoyancy(self, w):
        self.engine.set_buoyancy(w)

    def set_confinement(self, eps):
        self.engine.set_confinement(eps)

    def set_stir(self, stir):
        self.engine.set_stir(None if stir is None else
                             (stir.x, stir.y, stir.r, stir.dx, stir.dy) if hasattr(stir, "x")
                             else stir)

    def set_labels(self, label):
        self.engine.set_labels(label)

    def read_dye(self, dye):
        buf = self.engine.read_dye()
        n = self.engine.n_d
        for c in range(3):
            dye[c][:] = buf[c * n:(c + 1) * n]

    def max_face_speed(self) -> float:
        return self.engine.max_face_speed()


class HipPressureSolver:
    """PressureSolver (pressureSolver.ts:21-41): solve() fills p (warm start
    in, answer out) and returns sweeps or V-cycles used."""

    def __init__(self, nx, ny, method="mg-dc", h=None):
        self.engine = HipEngine(nx, ny, 1, h)
        self.method = method
        self.name = {"rbsor": "hip-rbsor", "mg": "hip-mg", "mg-dc": "hip-mg-dc",
                     "mg-fcg": "hip-mg-fcg"}[method]

    def solve(self, g, p, div, label, scale, iterations, omega, tol) -> int:
        it = C.c_int()
        n = self.engine.n_c
        if p is None or div is None or label is None:
            raise TypeError("HipPressureSolver.solve: p, div and label are required")
        check(self.engine.lib.fh_solve_pressure(self.engine._e, hip.SOLVERS[self.method],
                                                dptr(p, n, "p"), dptr(div, n, "div"),
                                                u8ptr(label, n), scale, int(iterations), omega,
                                                tol, C.byref(it)))
        return it.value


class HipAdvector:
    """Advector (advector.ts:28-63).  velocity() runs before dye() within a
    step and labels do not change between them (advector.ts:21-24)."""

    name = "hip"

    def __init__(self, nx, ny, dye_scale=1, h=None):
        self.engine = HipEngine(nx, ny, dye_scale, h)

    def velocity(self, g, scheme, u_in, v_in, u_out, v_out, label, dt):
        E = self.engine
        if any(a is None for a in (u_in, v_in, u_out, v_out, label)):
            raise TypeError("HipAdvector.velocity: every field is required")
        check(E.lib.fh_advect_velocity(E._e, hip.SCHEMES[scheme], dptr(u_in, E.n_u, "uIn"),
                                       dptr(v_in, E.n_v, "vIn"), dptr(u_out, E.n_u, "uOut"),
                                       dptr(v_out, E.n_v, "vOut"), u8ptr(label, E.n_c), dt))

    def dye(self, g, scheme, u, v, q_in, q_out, label, dt, dg=None):
        """Advects len(q_in) independent channels (advector.ts:104-121), three
        per device pass; dg, when given, must be this engine's dye grid."""
        E = self.engine
        n = E.n_d
        if dg is not None and (dg.nx, dg.ny) != (E.dnx, E.dny):
            raise ValueError(f"HipAdvector.dye: dye grid {dg.nx}x{dg.ny} is not this "
                             f"engine's {E.dnx}x{E.dny} (dye scale {E.ds})")
        if len(q_out) < len(q_in):
            raise ValueError("HipAdvector.dye: fewer output channels than input channels")
        if u is None or v is None or label is None:
            raise TypeError("HipAdvector.dye: u, v and label are required")
        for c0 in range(0, len(q_in), 3):
            group = q_in[c0:c0 + 3]
            qi = np.zeros(3 * n)
            for c, q in enumerate(group):
                qi[c * n:(c + 1) * n] = _sized(np.asarray(q, np.float64).ravel(), n,
                                               f"dye channel {c0 + c}")
            qo = np.empty(3 * n)
            check(E.lib.fh_advect_dye(E._e, hip.SCHEMES[scheme], dptr(u, E.n_u, "u"),
                                      dptr(v, E.n_v, "v"), dptr(qi, 3 * n), dptr(qo, 3 * n),
                                      u8ptr(label, E.n_c), dt))
            for c in range(len(group)):
                q_out[c0 + c][:] = qo[c * n:(c + 1) * n]
