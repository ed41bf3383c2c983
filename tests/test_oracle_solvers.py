"""Multigrid and whole-step pins for the FP64 oracle.

Mode 0 is the reference CpuMultigridSolver (pressureMultigrid.ts:219-320),
pinned by pressureMultigrid.test.ts's contraction and SOR-agreement tests.
Mode 1 (defect correction with a dense FP64 coarsest solve) is build-defined;
it is pinned by the same contraction/agreement bars, by an exact check of the
coarse solve against numpy's pseudo-inverse, and by convergence to 1e-10.
"""
import numpy as np
import pytest

from flowlab_amd import scenes as S
from oracle import pyoracle as O

SCALE = 0.02


def rough(n, seed=12345):
    return O.lcg_fill(n, seed)


def boundary_fixture(nx, ny):  # pressureMultigrid.test.ts:201-209
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[ny >> 2:ny >> 1, nx >> 2:(nx * 3) >> 3] = O.SOLID
    return lab.ravel(), rough(nx * ny)


def contraction(nx, ny, lab, div, cycles, mode):  # pressureMultigrid.test.ts:212-233
    p = np.zeros(nx * ny)
    prev = O.fluid_div_rms(div, lab)
    out = []
    for _ in range(cycles):
        p, _, _ = O.mg_solve(nx, ny, p, div, lab, SCALE, mode, 1, 0)
        res = O.rms_remaining_divergence(nx, ny, p, div, lab, SCALE)
        out.append(res / prev)
        prev = res
    return out


def test_level_sizes():
    assert O.mg_level_sizes(2048, 1024, 1)[-1] == (16, 8)
    assert len(O.mg_level_sizes(2048, 1024, 1)) == 8
    assert O.mg_level_sizes(2048, 1024, 0)[-1] == (4, 2)  # pressureMultigrid.ts:71-78
    assert O.mg_level_sizes(33, 17, 0)[1] == (17, 9)      # ceil keeps the ragged edge


def test_coarsen_priority():  # pressureMultigrid.ts:92-110: Air > Fluid > Solid
    fine = np.array([[2, 2, 0, 2, 1],
                     [2, 2, 2, 2, 2]], np.uint8).ravel()
    c = O.coarsen_labels(5, 2, fine, 3, 1)
    assert list(c) == [O.SOLID, O.FLUID, O.AIR]


@pytest.mark.parametrize("mode", [0, 1])
def test_interior_contraction(mode):  # pressureMultigrid.test.ts:235-256
    for nx, ny in ((32, 32), (64, 48), (128, 96)):
        div = rough(nx * ny, 999)
        div -= div.mean()
        f = contraction(nx, ny, np.zeros(nx * ny, np.uint8), div, 5, mode)
        assert max(f) < 0.15, (nx, ny, f)


@pytest.mark.parametrize("mode", [0, 1])
def test_boundary_contraction(mode):  # pressureMultigrid.test.ts:258-275
    for nx, ny in ((32, 32), (64, 48), (128, 96)):
        lab, div = boundary_fixture(nx, ny)
        f = contraction(nx, ny, lab, div, 6, mode)
        assert max(f) < 0.7, (nx, ny, f)
        assert np.prod(f) < 1e-2


@pytest.mark.parametrize("mode", [0, 1])
def test_mg_matches_sor(mode):  # pressureMultigrid.test.ts:277-299
    nx, ny = 64, 48
    lab, div = boundary_fixture(nx, ny)
    exp, _ = O.solve_pressure(nx, ny, np.zeros(nx * ny), div, lab, SCALE, 20000, 1.9, 1e-12)
    p, _, _ = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, mode, 40, 1e-12)
    pmax = np.abs(exp).max()
    assert pmax > 1e-3
    assert np.abs(exp - p).max() < 1e-8 * pmax


def _dense(nx, ny, lab):
    """The coarse operator as a dense matrix over fluid cells with count>0."""
    L = lab.reshape(ny, nx)

    def so(i, j):
        return i < 0 or j < 0 or i >= nx or j >= ny or L[j, i] == O.SOLID

    cells = []
    for j in range(ny):
        for i in range(nx):
            if L[j, i] != O.FLUID:
                continue
            cnt = sum(not so(i + di, j + dj) for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)))
            if cnt > 0:
                cells.append((i, j))
    idx = {c: a for a, c in enumerate(cells)}
    A = np.zeros((len(cells), len(cells)))
    for a, (i, j) in enumerate(cells):
        for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            if not so(i + di, j + dj):
                A[a, a] += 1
                if (i + di, j + dj) in idx:
                    A[a, idx[(i + di, j + dj)]] -= 1
    return A, [i + j * nx for i, j in cells]


@pytest.mark.parametrize("case", ["anchored", "neumann", "two_components"])
def test_coarse_direct_is_pseudo_inverse(case):
    nx, ny = 16, 8
    lab = np.zeros((ny, nx), np.uint8)
    if case == "anchored":
        lab[:, nx - 1] = O.AIR
        lab[2:5, 4:7] = O.SOLID
    elif case == "two_components":
        lab[:, 7] = O.SOLID          # a wall splits the box in two closed halves
        lab[3, 12] = O.AIR           # the right half is anchored, the left is not
    lab = lab.ravel()
    b = rough(nx * ny, 4242)
    e = O.coarse_direct_solve(nx, ny, lab, b)
    A, cells = _dense(nx, ny, lab)
    want = np.linalg.pinv(A) @ b[cells]
    assert np.abs(e[cells] - want).max() < 1e-10 * max(1.0, np.abs(want).max())
    others = np.setdiff1d(np.arange(nx * ny), cells)
    assert np.all(e[others] == 0)


def test_dc_reaches_1e10():
    nx, ny = 96, 64
    lab, div = boundary_fixture(nx, ny)
    p, it, hist = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, 1, 80, 1e-10)
    assert hist[-1] <= 1e-10 and it < 80
    assert np.all(np.diff(np.log(hist)) < 0)  # monotone


def test_dc_cycle_is_one_solve_cycle():
    nx, ny = 64, 48
    lab, div = boundary_fixture(nx, ny)
    p1, _, _ = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, SCALE, 1, 1, 0)
    p2 = O.mg_dc_cycle(nx, ny, np.zeros(nx * ny), -SCALE * div, lab)
    assert np.array_equal(p1, p2)


# ---- whole step (simulation.test.ts, buoyancy.test.ts, emitters.test.ts) ----

def make(nx, ny, scene=None):
    sim = O.OracleSim(nx, ny)
    g = S.Grid(nx, ny, 1.0 / ny)
    sim.reset(g, scene or S.SceneSpec())
    return sim, g


def test_long_run_finite_and_divergence_free():  # simulation.test.ts:8-20
    sim, g = make(32, 32, S.SceneSpec(seed=S.add_vortex_pair))
    prm = O.step_params(scheme="semiLagrangian", solver="sor", pressure_iters=40,
                        omega=O.optimal_omega(32), tol=5e-3)
    for _ in range(200):
        sim.step(prm)
    assert np.isfinite(sim.u).all() and np.isfinite(sim.v).all()
    assert np.abs(sim.div).max() < 1e-2
    assert sim.time > 0


def test_outlet_never_inflows():  # simulation.test.ts:22-39
    def seed(g, u, v):
        u.fill(-1)
    sim, g = make(32, 24, S.SceneSpec(labels=S.open_right(), seed=seed))
    prm = O.step_params(scheme="semiLagrangian", solver="sor", pressure_iters=60,
                        omega=O.optimal_omega(32, 24))
    for _ in range(5):
        sim.step(prm)
        assert np.all(sim.u.reshape(24, 33)[:, 31] >= 0)


def test_cfl_and_dt_cap():  # simulation.test.ts:41-51
    sim, g = make(32, 32, S.SceneSpec(seed=S.add_vortex_pair))
    prm = O.step_params(scheme="semiLagrangian", solver="sor", cfl_target=1.0,
                        omega=O.optimal_omega(32))
    for _ in range(50):
        sim.step(prm)
        assert prm.dt_used <= 1 / 30 + 1e-12
    # the CFL bound holds for the field dt was chosen on
    u0 = O.max_face_speed(32, 32, sim.u, sim.v)
    sim.step(prm)
    assert u0 * prm.dt_used / g.h <= 1.0 + 1e-9


def test_reproducible():  # simulation.test.ts:53-67
    a, _ = make(24, 24, S.SceneSpec(seed=S.add_vortex_cluster))
    b, _ = make(24, 24, S.SceneSpec(seed=S.add_vortex_cluster))
    prm = O.step_params(scheme="semiLagrangian", solver="sor", pressure_iters=20,
                        omega=O.optimal_omega(24))
    for _ in range(25):
        a.step(prm)
        b.step(prm)
    assert np.array_equal(a.u, b.u) and a.time == b.time


def test_dye_decay_exponential_in_time():  # simulation.test.ts:69-82
    def dye(dg, planes):
        planes[0].fill(1)
    sim, g = make(16, 16, S.SceneSpec(dye=dye))
    prm = O.step_params(scheme="semiLagrangian", solver="sor", dye_decay=(2, 2, 2),
                        omega=O.optimal_omega(16))
    for _ in range(30):
        sim.step(prm)
    assert abs(sim.dye[0] - np.exp(-2 * sim.time)) < 1e-12


def test_optimal_omega():  # simulation.test.ts:84-91
    for n in (16, 64, 256):
        assert abs(O.optimal_omega(n, n) - 2 / (1 + np.sin(np.pi / n))) < 1e-12
    assert O.optimal_omega(512, 128) > O.optimal_omega(128, 128)


def _hot(w, blob=None, N=48):
    def dye(dg, planes):
        if blob is None:
            planes[0].fill(1)
        else:
            j, i = np.mgrid[0:dg.ny, 0:dg.nx]
            planes[0][:] = blob(i, j).ravel()
    sim, g = make(N, N, S.SceneSpec(dye=dye))
    prm = O.step_params(scheme="semiLagrangian", solver="sor", buoyancy=w, tol=1e-8,
                        omega=O.optimal_omega(N))
    return sim, prm


def test_uniform_heat_moves_nothing():  # buoyancy.test.ts:33-43
    sim, prm = _hot(2)
    for _ in range(20):
        sim.step(prm)
    assert O.max_face_speed(48, 48, sim.u, sim.v) < 1e-3


def test_hot_blob_rises():  # buoyancy.test.ts:45-63
    N = 48

    def blob(i, j):
        dx = (i + 0.5) / N - 0.5
        dy = (j + 0.5) / N - 0.3
        return (dx * dx + dy * dy < 0.01).astype(float)
    sim, prm = _hot(4, blob)
    for _ in range(30):
        sim.step(prm)
    V = sim.v.reshape(N + 1, N)
    assert V[round(0.3 * N), N >> 1] > 0.05
    assert V[round(0.3 * N), round(0.12 * N)] < 0


def test_uniform_through_flow_is_steady():  # emitters.test.ts:160-183
    def seed(g, u, v):
        u.fill(1)
    sim, g = make(40, 20, S.SceneSpec(labels=S.open_right(), seed=seed))
    u0 = sim.u.copy()
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=20, tol=1e-6)
    for _ in range(20):
        sim.step(prm)
    assert np.array_equal(sim.u, u0)


def test_vortex_dipole_self_propels():  # advect.test.ts:209-258
    N, dt = 48, 0.05
    h = 1 / N
    g = S.Grid(N, N, h)
    u = np.zeros((N + 1) * N)
    v = np.zeros(N * (N + 1))
    S.add_vortex_pair(g, u, v)
    lab = np.zeros(N * N, np.uint8)
    p = np.zeros(N * N)

    def cy(u, v):
        U = u.reshape(N, N + 1)
        V = v.reshape(N + 1, N)
        cu = 0.5 * (U[:, :-1] + U[:, 1:])
        cv = 0.5 * (V[:-1, :] + V[1:, :])
        e = cu * cu + cv * cv
        y = ((np.arange(N) + 0.5) * h)[:, None]
        return (e * y).sum() / e.sum()

    y0 = cy(u, v)
    for _ in range(150):
        u, v = O.advect_velocity(N, N, h, u, v, lab, dt)
        div = O.divergence(N, N, h, u, v)
        p, _ = O.solve_pressure(N, N, p, div, lab, h * h / dt, 40, 1.8)
        u, v = O.subtract_gradient(N, N, p, u, v, lab, dt / h)
    assert cy(u, v) - y0 > 0.04


# ---- vorticity confinement (build-defined; no reference counterpart) -------

def _field(nx, ny, h, fu, fv):
    xu = (np.arange(nx + 1) * h)[None, :]
    yu = ((np.arange(ny) + 0.5) * h)[:, None]
    xv = ((np.arange(nx) + 0.5) * h)[None, :]
    yv = (np.arange(ny + 1) * h)[:, None]
    return fu(xu, yu).ravel().astype(float), fv(xv, yv).ravel().astype(float)


def test_confinement_eps_zero_and_uniform():
    nx, ny, h = 32, 24, 1 / 24
    lab = np.zeros(nx * ny, np.uint8)
    u, v = _field(nx, ny, h, lambda x, y: np.sin(3 * x) * np.cos(2 * y) + 0 * x,
                  lambda x, y: np.cos(x + y) + 0 * x)
    u1, v1 = O.vorticity_confinement(nx, ny, h, u, v, lab, 0.0, 0.1)
    assert np.array_equal(u1, u) and np.array_equal(v1, v)
    u, v = np.full_like(u, 0.7), np.full_like(v, -0.3)
    u1, v1 = O.vorticity_confinement(nx, ny, h, u, v, lab, 0.5, 0.1)
    assert np.array_equal(u1, u) and np.array_equal(v1, v)


def test_confinement_solid_body_rotation_is_zero():
    nx, ny, h = 40, 40, 1 / 40
    lab = np.zeros(nx * ny, np.uint8)
    u, v = _field(nx, ny, h, lambda x, y: -(y - 0.5) + 0 * x, lambda x, y: (x - 0.5) + 0 * y)
    u1, v1 = O.vorticity_confinement(nx, ny, h, u, v, lab, 0.3, 0.1)
    assert np.abs(u1 - u).max() < 1e-12 and np.abs(v1 - v).max() < 1e-12


def test_confinement_mirror_symmetry():
    nx, ny, h = 36, 30, 1 / 30
    lab = np.zeros(nx * ny, np.uint8)
    Lx = nx * h
    fu = lambda x, y: np.exp(-((x - 0.4 * Lx) ** 2 + (y - 0.5) ** 2) / 0.02) * (y - 0.5)
    fv = lambda x, y: -np.exp(-((x - 0.4 * Lx) ** 2 + (y - 0.5) ** 2) / 0.02) * (x - 0.4 * Lx)
    u, v = _field(nx, ny, h, fu, fv)
    # mirror x -> Lx - x: u'(x) = -u(Lx-x), v'(x) = v(Lx-x)
    U = u.reshape(ny, nx + 1)
    V = v.reshape(ny + 1, nx)
    um = (-U[:, ::-1]).ravel().copy()
    vm = V[:, ::-1].ravel().copy()
    u1, v1 = O.vorticity_confinement(nx, ny, h, u, v, lab, 0.4, 0.2)
    u2, v2 = O.vorticity_confinement(nx, ny, h, um, vm, lab, 0.4, 0.2)
    fu1 = (u1 - u).reshape(ny, nx + 1)
    fv1 = (v1 - v).reshape(ny + 1, nx)
    fu2 = (u2 - um).reshape(ny, nx + 1)
    fv2 = (v2 - vm).reshape(ny + 1, nx)
    assert np.abs(fu1).max() > 1e-4
    assert np.abs(fu2 + fu1[:, ::-1]).max() < 1e-12
    assert np.abs(fv2 - fv1[:, ::-1]).max() < 1e-12


def test_confinement_matches_continuous_formula():
    # A Gaussian vortex omega(r) = A exp(-r^2/s^2): |grad|omega|| points
    # inward, so f = eps*h*(N x omega) is tangential with magnitude
    # eps*h*omega, directed with the rotation.  Check the discrete face force
    # against the continuous one away from the core and the walls.
    nx = ny = 128
    h = 1 / ny
    s2 = 0.01
    A = 2.0
    # streamfunction psi = -(A s2/4) exp(-r^2/s2); omega = -lap psi
    # = -A exp(-r^2/s2) (1 - r^2/s2)  (negative, clockwise core)
    psi = lambda x, y: -(A * s2 / 4) * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / s2)
    g = S.Grid(nx, ny, h)
    xc = (np.arange(nx + 1) * h)[None, :]
    yc = (np.arange(ny + 1) * h)[:, None]
    u = np.zeros((nx + 1) * ny)
    v = np.zeros(nx * (ny + 1))
    S.add_curl_of_stream(g, u, v, psi(xc, yc), 1.0)
    lab = np.zeros(nx * ny, np.uint8)
    eps, dt = 0.5, 1.0
    u1, _ = O.vorticity_confinement(nx, ny, h, u, v, lab, eps, dt)
    fu = ((u1 - u) / dt).reshape(ny, nx + 1)
    # continuous: omega at (x,y); N = -rhat (|omega| decreasing outward)
    # f_x = eps*h*N_y*omega = eps*h*(-(y-.5)/r)*omega
    x = (np.arange(nx + 1) * h)[None, :]
    y = ((np.arange(ny) + 0.5) * h)[:, None]
    r2 = (x - 0.5) ** 2 + (y - 0.5) ** 2
    om = -A * np.exp(-r2 / s2) * (1 - r2 / s2)
    r = np.sqrt(r2)
    # |omega| decreases outward inside the core (r < s): use the ring
    # 0.3 s < r < 0.8 s, clear of the centre where N is undefined
    ring = (r > 0.3 * np.sqrt(s2)) & (r < 0.8 * np.sqrt(s2))
    fx = eps * h * (-(y - 0.5) / np.maximum(r, 1e-12)) * om
    err = np.abs(fu - fx)[ring].max() / np.abs(fx[ring]).max()
    assert err < 0.05
