"""ctypes front end of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / timed CPU baseline, never by the
product path (flowlab_amd/ does not import it).

Each wrapper takes and returns numpy arrays in the reference's layouts
(src/core/grid.ts:84-141): p/label/div nx*ny, u (nx+1)*ny, v nx*(ny+1),
dye 3 planes of dnx*dny, all row-major with x fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

FLUID, AIR, SOLID = 0, 1, 2

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_fp = C.POINTER(C.c_float)


def build() -> None:
    """Compile liboracle.so from oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(_LIB_PATH):
        build()
    return C.CDLL(_LIB_PATH)


_lib = _load()


class StepParams(C.Structure):
    _fields_ = [
        ("scheme", C.c_int),
        ("solver", C.c_int),
        ("pressure_iters", C.c_int),
        ("residual_every", C.c_int),
        ("cfl_target", C.c_double),
        ("dt_max", C.c_double),
        ("tol", C.c_double),
        ("omega", C.c_double),
        ("rho", C.c_double),
        ("keep_decay", C.c_double * 3),
        ("buoyancy", C.c_double),
        ("confinement", C.c_double),
        ("stir_active", C.c_int),
        ("stir", C.c_double * 5),
        ("patch_i0", C.c_int),
        ("patch_j0", C.c_int),
        ("patch_nx", C.c_int),
        ("patch_ny", C.c_int),
        ("patch_data", _fp),
        ("dt_used", C.c_double),
        ("iters", C.c_int),
    ]


class _Sim(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dnx", C.c_int), ("dny", C.c_int),
        ("h", C.c_double), ("dh", C.c_double),
        ("u", _dp), ("v", _dp), ("p", _dp), ("dye", _dp),
        ("label", _u8p),
        ("uNext", _dp), ("vNext", _dp), ("uHat", _dp), ("vHat", _dp),
        ("dyeNext", _dp), ("qHat", _dp), ("div", _dp), ("work", _dp),
        ("steps", C.c_long),
    ]


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_i, _d = C.c_int, C.c_double
_sig("fo_sample_u", _d, _i, _i, _d, _dp, _d, _d)
_sig("fo_sample_v", _d, _i, _i, _d, _dp, _d, _d)
_sig("fo_sample_p", _d, _i, _i, _d, _dp, _d, _d)
_sig("fo_advect_velocity", None, _i, _i, _d, _dp, _dp, _dp, _dp, _u8p, _d)
_sig("fo_advect_velocity_maccormack", None, _i, _i, _d, _dp, _dp, _dp, _dp, _dp, _dp, _u8p, _d)
_sig("fo_advect_scalar", None, _i, _i, _d, _i, _i, _d, _dp, _dp, _dp, _dp, _u8p, _d)
_sig("fo_advect_scalar_maccormack", None, _i, _i, _d, _i, _i, _d, _dp, _dp, _dp, _dp, _dp, _u8p, _d)
_sig("fo_divergence", None, _i, _i, _d, _dp, _dp, _dp)
_sig("fo_gs_sweep", None, _i, _i, _dp, _dp, _u8p, _d, _d, _i, _i)
_sig("fo_rms_remaining_divergence", _d, _i, _i, _dp, _dp, _u8p, _d)
_sig("fo_fluid_div_rms", _d, _i, _dp, _u8p)
_sig("fo_solve_pressure", _i, _i, _i, _dp, _dp, _u8p, _d, _i, _d, _d)
_sig("fo_rb_sweep", None, _i, _i, _dp, _dp, _u8p, _d, _d, _i)
_sig("fo_solve_pressure_rb", _i, _i, _i, _dp, _dp, _u8p, _d, _i, _d, _d)
_sig("fo_mg_level_count", _i, _i, _i, _i)
_sig("fo_mg_level_size", None, _i, _i, _i, _i, C.POINTER(_i), C.POINTER(_i))
_sig("fo_coarsen_labels", None, _i, _i, _u8p, _i, _i, _u8p)
_sig("fo_mg_residual", None, _i, _i, _dp, _dp, _u8p, _dp)
_sig("fo_mg_restrict", None, _i, _i, _dp, _i, _i, _dp, _dp)
_sig("fo_mg_prolong_add", None, _i, _i, _dp, _u8p, _i, _i, _dp)
_sig("fo_mg_solve", _i, _i, _i, _dp, _dp, _u8p, _d, _i, _i, _d, _dp)
_sig("fo_mg_dc_cycle", None, _i, _i, _dp, _dp, _u8p)
_sig("fo_coarse_direct_solve", None, _i, _i, _u8p, _dp, _dp)
_sig("fo_subtract_gradient", None, _i, _i, _dp, _dp, _dp, _u8p, _d)
_sig("fo_apply_outflow", None, _i, _i, _dp, _dp, _u8p)
_sig("fo_commit_labels", None, _i, _i, _dp, _dp, _dp, _u8p, _dp, _i)
_sig("fo_apply_buoyancy", None, _i, _i, _d, _i, _i, _d, _dp, _dp, _u8p, _d, _d)
_sig("fo_apply_dye_patch", None, _i, _i, _dp, _i, _i, _i, _i, _i, _fp)
_sig("fo_apply_stir", None, _i, _i, _d, _dp, _dp, _u8p, _d, _d, _d, _d, _d)
_sig("fo_compute_vorticity", None, _i, _i, _d, _dp, _dp, _dp)
_sig("fo_vorticity_confinement", None, _i, _i, _d, _dp, _dp, _u8p, _dp, _d, _d)
_sig("fo_optimal_omega", _d, _i, _i)
_sig("fo_sim_step", None, C.POINTER(_Sim), C.POINTER(StepParams))
_sig("fo_max_face_speed", _d, _i, _i, _dp, _dp)
_sig("fo_lcg_fill", None, _i, _i, _dp)


def _f64(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _p(a: np.ndarray):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(_u8p)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    raise TypeError(a.dtype)


def _lbl(label) -> np.ndarray:
    return np.ascontiguousarray(label, dtype=np.uint8)


# ---- grid sampling (grid.ts:197-250) ---------------------------------------

def sample_u(nx, ny, h, u, x, y):
    u = _f64(u)
    return _lib.fo_sample_u(nx, ny, h, _p(u), x, y)


def sample_v(nx, ny, h, v, x, y):
    v = _f64(v)
    return _lib.fo_sample_v(nx, ny, h, _p(v), x, y)


def sample_p(nx, ny, h, q, x, y):
    q = _f64(q)
    return _lib.fo_sample_p(nx, ny, h, _p(q), x, y)


# ---- advection (advect.ts) ---------------------------------------------------

def advect_velocity(nx, ny, h, u, v, label, dt, maccormack=False):
    u, v, label = _f64(u), _f64(v), _lbl(label)
    uo = np.full_like(u, np.nan)
    vo = np.full_like(v, np.nan)
    if maccormack:
        uh = np.full_like(u, np.nan)
        vh = np.full_like(v, np.nan)
        _lib.fo_advect_velocity_maccormack(nx, ny, h, _p(u), _p(v), _p(uh), _p(vh), _p(uo),
                                           _p(vo), _p(label), dt)
    else:
        _lib.fo_advect_velocity(nx, ny, h, _p(u), _p(v), _p(uo), _p(vo), _p(label), dt)
    return uo, vo


def advect_scalar(nx, ny, h, u, v, q, label, dt, maccormack=False, dscale=1):
    u, v, q, label = _f64(u), _f64(v), _f64(q), _lbl(label)
    dnx, dny, dh = nx * dscale, ny * dscale, h / dscale
    out = np.full_like(q, np.nan)
    if maccormack:
        hat = np.full_like(q, np.nan)
        _lib.fo_advect_scalar_maccormack(nx, ny, h, dnx, dny, dh, _p(u), _p(v), _p(q), _p(hat),
                                         _p(out), _p(label), dt)
    else:
        _lib.fo_advect_scalar(nx, ny, h, dnx, dny, dh, _p(u), _p(v), _p(q), _p(out), _p(label), dt)
    return out


# ---- projection pieces -------------------------------------------------------

def divergence(nx, ny, h, u, v):
    u, v = _f64(u), _f64(v)
    out = np.empty(nx * ny)
    _lib.fo_divergence(nx, ny, h, _p(u), _p(v), _p(out))
    return out


def gs_sweep(nx, ny, p, div, label, scale, omega=1.0, reverse_x=False, reverse_y=False):
    p = _f64(p).copy()
    div, label = _f64(div), _lbl(label)
    _lib.fo_gs_sweep(nx, ny, _p(p), _p(div), _p(label), scale, omega, int(reverse_x),
                     int(reverse_y))
    return p


def rb_sweep(nx, ny, p, div, label, scale, omega, color):
    p = _f64(p).copy()
    div, label = _f64(div), _lbl(label)
    _lib.fo_rb_sweep(nx, ny, _p(p), _p(div), _p(label), scale, omega, color)
    return p


def rms_remaining_divergence(nx, ny, p, div, label, scale):
    p, div, label = _f64(p), _f64(div), _lbl(label)
    return _lib.fo_rms_remaining_divergence(nx, ny, _p(p), _p(div), _p(label), scale)


def fluid_div_rms(div, label):
    div, label = _f64(div), _lbl(label)
    return _lib.fo_fluid_div_rms(div.size, _p(div), _p(label))


def relative_residual(nx, ny, p, div, label, scale):
    """rmsRemainingDivergence / fluidDivRms: the solve's relative residual
    (pressure.ts:102-161)."""
    d = fluid_div_rms(div, label)
    return rms_remaining_divergence(nx, ny, p, div, label, scale) / d if d > 0 else 0.0


def solve_pressure(nx, ny, p, div, label, scale, iterations, omega=1.0, tol=1e-3):
    p = _f64(p).copy()
    div, label = _f64(div), _lbl(label)
    it = _lib.fo_solve_pressure(nx, ny, _p(p), _p(div), _p(label), scale, iterations, omega, tol)
    return p, it


def solve_pressure_rb(nx, ny, p, div, label, scale, iterations, omega=1.0, tol=1e-3):
    p = _f64(p).copy()
    div, label = _f64(div), _lbl(label)
    it = _lib.fo_solve_pressure_rb(nx, ny, _p(p), _p(div), _p(label), scale, iterations, omega,
                                   tol)
    return p, it


def mg_level_sizes(nx, ny, mode):
    n = _lib.fo_mg_level_count(nx, ny, mode)
    out = []
    for l in range(n):
        a, b = C.c_int(), C.c_int()
        _lib.fo_mg_level_size(nx, ny, mode, l, C.byref(a), C.byref(b))
        out.append((a.value, b.value))
    return out


def coarsen_labels(fnx, fny, fine, cnx, cny):
    fine = _lbl(fine)
    out = np.empty(cnx * cny, np.uint8)
    _lib.fo_coarsen_labels(fnx, fny, _p(fine), cnx, cny, _p(out))
    return out


def mg_residual(nx, ny, x, b, label):
    x, b, label = _f64(x), _f64(b), _lbl(label)
    r = np.empty(nx * ny)
    _lib.fo_mg_residual(nx, ny, _p(x), _p(b), _p(label), _p(r))
    return r


def mg_restrict(fnx, fny, r, cnx, cny):
    r = _f64(r)
    b = np.empty(cnx * cny)
    e = np.empty(cnx * cny)
    _lib.fo_mg_restrict(fnx, fny, _p(r), cnx, cny, _p(b), _p(e))
    return b


def mg_prolong_add(fnx, fny, x, label, cnx, cny, e):
    x = _f64(x).copy()
    label, e = _lbl(label), _f64(e)
    _lib.fo_mg_prolong_add(fnx, fny, _p(x), _p(label), cnx, cny, _p(e))
    return x


def mg_solve(nx, ny, p, div, label, scale, mode, cycles, tol):
    """Returns (p, cycles used, relative-residual history)."""
    p = _f64(p).copy()
    div, label = _f64(div), _lbl(label)
    hist = np.zeros(cycles + 1)
    it = _lib.fo_mg_solve(nx, ny, _p(p), _p(div), _p(label), scale, mode, cycles, tol, _p(hist))
    return p, it, hist[: it + 1]


def mg_dc_cycle(nx, ny, p, b, label):
    p = _f64(p).copy()
    b, label = _f64(b), _lbl(label)
    _lib.fo_mg_dc_cycle(nx, ny, _p(p), _p(b), _p(label))
    return p


def mg_apply_a(nx, ny, x, label):
    """A x of the defect-correction system (count x - sum of open neighbours
    on fluid cells with a non-zero count, 0 elsewhere): the negated residual
    of x against a zero right-hand side (oracle.c fo_mg_residual)."""
    return -mg_residual(nx, ny, x, np.zeros(nx * ny), label)


def mg_fcg_solve(nx, ny, p, div, label, scale, cycles, tol, precond=None, dc_first=True):
    """Flexible conjugate gradients with one stored direction (Y. Notay,
    "Flexible conjugate gradients", SIAM J. Sci. Comput. 22 (2000)), FP64,
    on the defect-correction system A p = b, b = -scale div, preconditioned
    by one defect-correction V-cycle from zero (precond(r) overrides it);
    the Krylov wrapper docs/WEBGPU.md:462-465 names for multigrid whose
    contraction is limited at Air/Solid boundaries.  Textbook form, every
    product computed explicitly:
        z = B r;  beta = -<z, A d'> / <d', A d'>;  d = z + beta d';
        alpha = <d, r> / <d, A d>;  p += alpha d;  r -= alpha A d.
    dc_first: the first cycle is a plain defect-correction cycle (p += B r,
    r recomputed from p) and CG starts from its result; False: CG from the
    start.  The zero-divergence guard of pressureMultigrid.ts:294-300
    applies.  Returns (p, iterations, relative-residual history as in
    mg_solve)."""
    p = _f64(p).copy()
    div, label = _f64(div), _lbl(label)
    if fluid_div_rms(div, label) == 0:
        return np.zeros(nx * ny), 0, np.zeros(0)
    B = precond or (lambda r: mg_dc_cycle(nx, ny, np.zeros(nx * ny), r, label))
    b = -scale * div
    r = mg_residual(nx, ny, p, b, label)
    hist = [relative_residual(nx, ny, p, div, label, scale)]
    start = 0
    if dc_first and cycles > 0:
        p = p + B(r)
        r = mg_residual(nx, ny, p, b, label)
        hist.append(relative_residual(nx, ny, p, div, label, scale))
        if tol > 0 and hist[-1] <= tol:
            return p, 1, np.array(hist)
        start = 1
    d = q = None
    for it in range(start, cycles):
        z = B(r)
        if d is None:
            d = z
        else:
            d = z - (z @ q) / (d @ q) * d
        q = mg_apply_a(nx, ny, d, label)
        alpha = (d @ r) / (d @ q)
        p = p + alpha * d
        r = r - alpha * q
        hist.append(relative_residual(nx, ny, p, div, label, scale))
        if tol > 0 and hist[-1] <= tol:
            return p, it + 1, np.array(hist)
    return p, cycles, np.array(hist)


def coarse_direct_solve(nx, ny, label, b):
    label, b = _lbl(label), _f64(b)
    e = np.empty(nx * ny)
    _lib.fo_coarse_direct_solve(nx, ny, _p(label), _p(b), _p(e))
    return e


def subtract_gradient(nx, ny, p, u, v, label, grad_scale):
    u, v = _f64(u).copy(), _f64(v).copy()
    p, label = _f64(p), _lbl(label)
    _lib.fo_subtract_gradient(nx, ny, _p(p), _p(u), _p(v), _p(label), grad_scale)
    return u, v


def apply_outflow(nx, ny, u, v, label):
    u, v = _f64(u).copy(), _f64(v).copy()
    label = _lbl(label)
    _lib.fo_apply_outflow(nx, ny, _p(u), _p(v), _p(label))
    return u, v


def commit_labels(nx, ny, p, u, v, label, dye=None):
    p, u, v = _f64(p).copy(), _f64(u).copy(), _f64(v).copy()
    label = _lbl(label)
    dye_arr = None if dye is None else _f64(dye).copy()
    plane = 0 if dye is None else dye_arr.size // 3
    _lib.fo_commit_labels(nx, ny, _p(p), _p(u), _p(v), _p(label),
                          None if dye_arr is None else _p(dye_arr), plane)
    return p, u, v, dye_arr


def apply_buoyancy(nx, ny, h, v, heat, label, w, dt, dscale=1):
    v = _f64(v).copy()
    heat, label = _f64(heat), _lbl(label)
    _lib.fo_apply_buoyancy(nx, ny, h, nx * dscale, ny * dscale, h / dscale, _p(v), _p(heat),
                           _p(label), w, dt)
    return v


def apply_dye_patch(dnx, dny, dye, i0, j0, pnx, pny, data):
    dye = _f64(dye).copy()
    data = np.ascontiguousarray(data, dtype=np.float32)
    _lib.fo_apply_dye_patch(dnx, dny, _p(dye), dye.size // 3, i0, j0, pnx, pny, _p(data))
    return dye


def apply_stir(nx, ny, h, u, v, label, x, y, r, dx, dy):
    u, v = _f64(u).copy(), _f64(v).copy()
    label = _lbl(label)
    _lib.fo_apply_stir(nx, ny, h, _p(u), _p(v), _p(label), x, y, r, dx, dy)
    return u, v


def compute_vorticity(nx, ny, h, u, v):
    u, v = _f64(u), _f64(v)
    out = np.empty(nx * ny)
    _lib.fo_compute_vorticity(nx, ny, h, _p(u), _p(v), _p(out))
    return out


def vorticity_confinement(nx, ny, h, u, v, label, eps, dt):
    """Returns (u, v) with the confinement force of (u, v) added."""
    u0, v0 = _f64(u), _f64(v)
    w = compute_vorticity(nx, ny, h, u0, v0)
    u1, v1 = u0.copy(), v0.copy()
    label = _lbl(label)
    _lib.fo_vorticity_confinement(nx, ny, h, _p(u1), _p(v1), _p(label), _p(w), eps, dt)
    return u1, v1


def optimal_omega(nx, ny=None):
    return _lib.fo_optimal_omega(nx, nx if ny is None else ny)


def max_face_speed(nx, ny, u, v):
    u, v = _f64(u), _f64(v)
    return _lib.fo_max_face_speed(nx, ny, _p(u), _p(v))


def lcg_fill(n, seed=12345):
    out = np.empty(n)
    _lib.fo_lcg_fill(n, seed, _p(out))
    return out


# ---- whole step (simulation.ts:467-587) ------------------------------------

SOLVERS = {"sor": 0, "rbsor": 1, "mg": 2, "mg-dc": 3}


class OracleSim:
    """FP64 CPU twin of Simulation's CPU path.  Holds its own buffers; the C
    step swaps the ping-pong pointers, so the current fields are re-read
    from the struct after every step."""

    def __init__(self, nx, ny, dye_scale=1, h=None):
        self.nx, self.ny = nx, ny
        self.h = 1.0 / ny if h is None else h
        self.dnx, self.dny = nx * dye_scale, ny * dye_scale
        self.dh = self.h / dye_scale
        nu, nv, npc, nd = (nx + 1) * ny, nx * (ny + 1), nx * ny, self.dnx * self.dny
        self._bufs = {
            "u": np.zeros(nu), "v": np.zeros(nv), "p": np.zeros(npc), "dye": np.zeros(3 * nd),
            "uNext": np.zeros(nu), "vNext": np.zeros(nv), "uHat": np.zeros(nu),
            "vHat": np.zeros(nv), "dyeNext": np.zeros(3 * nd), "qHat": np.zeros(nd),
            "div": np.zeros(npc), "work": np.zeros(npc),
        }
        self.label = np.zeros(npc, np.uint8)
        self._by_addr = {a.ctypes.data: a for a in self._bufs.values()}
        s = _Sim()
        s.nx, s.ny, s.dnx, s.dny, s.h, s.dh = nx, ny, self.dnx, self.dny, self.h, self.dh
        for k, a in self._bufs.items():
            setattr(s, k, a.ctypes.data_as(_dp))
        s.label = self.label.ctypes.data_as(_u8p)
        s.steps = 0
        self._s = s
        self.time = 0.0

    def _cur(self, name):
        addr = C.cast(getattr(self._s, name), C.c_void_p).value
        return self._by_addr[addr]

    @property
    def u(self):
        return self._cur("u")

    @property
    def v(self):
        return self._cur("v")

    @property
    def p(self):
        return self._cur("p")

    @property
    def dye(self):
        return self._cur("dye")

    @property
    def div(self):
        return self._cur("div")

    def step(self, prm: StepParams, patch: np.ndarray | None = None):
        if patch is not None:
            self._patch_keep = np.ascontiguousarray(patch, np.float32)
            prm.patch_data = self._patch_keep.ctypes.data_as(_fp)
        _lib.fo_sim_step(C.byref(self._s), C.byref(prm))
        self.time += prm.dt_used
        return prm.iters


def step_params(scheme="macCormack", solver="rbsor", pressure_iters=100, residual_every=1,
                cfl_target=2.0, dt_max=1 / 30, tol=5e-3, omega=1.0, rho=1.0,
                dye_decay=(0.0, 0.0, 0.0), buoyancy=0.0, confinement=0.0, stir=None,
                patch=None) -> StepParams:
    """Parameters of one step, defaults from simulation.ts:223-234."""
    p = StepParams()
    p.scheme = 1 if scheme == "macCormack" else 0
    p.solver = SOLVERS[solver]
    p.pressure_iters = pressure_iters
    p.residual_every = residual_every
    p.cfl_target, p.dt_max, p.tol, p.omega, p.rho = cfl_target, dt_max, tol, omega, rho
    for c in range(3):
        p.keep_decay[c] = dye_decay[c]
    p.buoyancy = buoyancy
    p.confinement = confinement
    p.stir_active = 0 if stir is None else 1
    if stir is not None:
        for k in range(5):
            p.stir[k] = stir[k]
    if patch is not None:
        i0, j0, pnx, pny = patch
        p.patch_i0, p.patch_j0, p.patch_nx, p.patch_ny = i0, j0, pnx, pny
    return p


def _reset_impl(sim: "OracleSim", g, scene):
    """simulation.ts:388-416 on the oracle's buffers: zero, labels, seed, dye,
    commitLabels, then the dye source applied once at t = 0."""
    from flowlab_amd.scenes import Grid
    sim.u[:] = 0
    sim.v[:] = 0
    sim.p[:] = 0
    sim.dye[:] = 0
    sim.label[:] = FLUID
    dg = Grid(sim.dnx, sim.dny, sim.dh)
    if scene.labels:
        scene.labels(g, sim.label)
    if scene.seed:
        scene.seed(g, sim.u, sim.v)
    nd = sim.dnx * sim.dny
    if scene.dye:
        scene.dye(dg, [sim.dye[c * nd:(c + 1) * nd] for c in range(3)])
    _lib.fo_commit_labels(sim.nx, sim.ny, _p(sim.p), _p(sim.u), _p(sim.v), _p(sim.label),
                          _p(sim.dye), nd)
    sim.patch = scene.dye_source(dg, g) if scene.dye_source else None
    if sim.patch is not None:
        pt = sim.patch
        data = np.ascontiguousarray(pt.data, np.float32)
        _lib.fo_apply_dye_patch(sim.dnx, sim.dny, _p(sim.dye), nd, pt.i0, pt.j0, pt.nx, pt.ny,
                                _p(data))
    sim.time = 0.0


OracleSim.reset = _reset_impl


# ---- learned warm start (oracle/unet.c, build-defined) ----------------------

_sig("fo_unet_num_params", _i)
_sig("fo_unet_forward", None, _i, _i, _i, _dp, _dp, _dp, _i)
_sig("fo_unet_backward", None, _i, _i, _i, _dp, _dp, _dp, _dp, _i)

UNET_LAYERS = [(8, 16, 1), (16, 16, 1), (16, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1),
               (64, 64, 1), (64, 64, 1), (128, 64, 1), (64, 64, 1), (96, 32, 1), (32, 32, 1),
               (48, 16, 1), (16, 16, 1), (16, 16, 0)]


def unet_num_params():
    return _lib.fo_unet_num_params()


def unet_forward(params, x, emulate_bf16=False):
    """x: [N][H][W][8] -> [N][H][W]"""
    params, x = _f64(params), _f64(x)
    N, H, W, _ = x.shape
    out = np.empty((N, H, W))
    _lib.fo_unet_forward(N, H, W, _p(params), _p(x), _p(out), int(emulate_bf16))
    return out


def unet_backward(params, x, dout, emulate_bf16=False):
    """Gradient of sum(dout * net(x)) wrt the parameters."""
    params, x, dout = _f64(params), _f64(x), _f64(dout)
    N, H, W, _ = x.shape
    g = np.empty(params.size)
    _lib.fo_unet_backward(N, H, W, _p(params), _p(x), _p(dout), _p(g), int(emulate_bf16))
    return g


def unet_init(seed=0, head_scale=0.1):
    """He-normal weights, zero biases (head: only output channel 0 live)."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (cin, cout, _) in enumerate(UNET_LAYERS):
        std = np.sqrt(2.0 / (9 * cin))
        w = rng.standard_normal((cout, 3, 3, cin)) * std
        b = rng.standard_normal(cout) * 0.05
        if li == len(UNET_LAYERS) - 1:
            w[1:] = 0
            b[1:] = 0
            w *= head_scale
        out += [w.ravel(), b]
    return np.concatenate(out)
