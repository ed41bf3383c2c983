"""Pins the FP64 oracle against the closed forms and invariants the
reference's own CPU tests assert (src/core/*.test.ts).  Each test names the
reference test it restates; the oracle is the checker every GPU parity test
relies on, so a plausible slip (a dropped term, a wrong stride, a flipped
sign) must fail one of these."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O

H = 0.5  # grid.test.ts:5 -- not 1.0, so a missing /h cannot hide


def zeros(nx, ny):
    return np.zeros((nx + 1) * ny), np.zeros(nx * (ny + 1)), np.zeros(nx * ny, np.uint8)


def iu(nx, i, j):
    return i + j * (nx + 1)


def iv(nx, i, j):
    return i + j * nx


# ---- grid.test.ts ----------------------------------------------------------

def _ramp(nx, ny):
    return np.array([i + 10 * j for j in range(ny) for i in range(nx)], dtype=np.float64)


def test_sample_p_exact_at_centres():  # grid.test.ts:57-71
    q = _ramp(4, 4)
    for j in range(4):
        for i in range(4):
            assert abs(O.sample_p(4, 4, H, q, (i + 0.5) * H, (j + 0.5) * H) - (i + 10 * j)) < 1e-12


def test_sample_p_midpoints():  # grid.test.ts:73-82
    q = _ramp(4, 4)
    assert abs(O.sample_p(4, 4, H, q, 2 * H, 1.5 * H) - 11.5) < 1e-12
    assert abs(O.sample_p(4, 4, H, q, 1.5 * H, 2 * H) - 16) < 1e-12


def test_sample_p_clamps():  # grid.test.ts:84-93
    q = _ramp(4, 4)
    assert O.sample_p(4, 4, H, q, -5 * H, -5 * H) == 0
    assert O.sample_p(4, 4, H, q, 9 * H, 9 * H) == 33


def test_sample_u_v_offsets():
    # u[i,j] sits at (i h, (j+.5) h), v[i,j] at ((i+.5) h, j h) (grid.ts:193-225)
    nx, ny = 5, 4
    u = np.arange((nx + 1) * ny, dtype=np.float64)
    v = np.arange(nx * (ny + 1), dtype=np.float64) * 3
    for j in range(ny):
        for i in range(nx + 1):
            assert O.sample_u(nx, ny, H, u, i * H, (j + 0.5) * H) == u[iu(nx, i, j)]
    for j in range(ny + 1):
        for i in range(nx):
            assert O.sample_v(nx, ny, H, v, (i + 0.5) * H, j * H) == v[iv(nx, i, j)]


# ---- advect.test.ts --------------------------------------------------------

def test_every_face_written():  # advect.test.ts:25-46
    u, v, lab = zeros(4, 4)
    u[:] = 0.3
    v[:] = -0.2
    uo, vo = O.advect_velocity(4, 4, H, u, v, lab, 0.25 * H)
    assert np.isfinite(uo).all() and np.isfinite(vo).all()


@pytest.mark.parametrize("mac", [False, True])
def test_uniform_velocity_advects_to_itself(mac):  # advect.test.ts:48-78, 432-457
    nx = ny = 6 if mac else 4
    u, v, lab = zeros(nx, ny)
    u[:] = 0.75
    v[:] = -0.25
    uo, vo = O.advect_velocity(nx, ny, H, u, v, lab, 0.3 * H, maccormack=mac)
    U = uo.reshape(ny, nx + 1)
    V = vo.reshape(ny + 1, nx)
    assert np.abs(U[:, 1:nx] - 0.75).max() < 1e-12
    assert np.abs(V[1:ny, :] + 0.25).max() < 1e-12


def test_rk2_ramp_u():  # advect.test.ts:88-111
    nx, ny = 6, 4
    u, v, lab = zeros(nx, ny)
    u[:] = np.tile(np.arange(nx + 1, dtype=np.float64), ny)
    dt = 0.5 * H
    r = dt / H
    factor = 1 - r + 0.5 * r * r
    uo, _ = O.advect_velocity(nx, ny, H, u, v, lab, dt)
    U = uo.reshape(ny, nx + 1)
    for i in range(1, nx):
        assert np.abs(U[:, i] - i * factor).max() < 1e-12
        assert np.abs(U[:, i] - i * (1 - r)).min() > 1e-9  # not forward Euler


def test_rk2_ramp_v():  # advect.test.ts:113-137
    nx, ny = 4, 6
    u, v, lab = zeros(nx, ny)
    v[:] = np.repeat(np.arange(ny + 1, dtype=np.float64), nx)
    dt = 0.5 * H
    r = dt / H
    factor = 1 - r + 0.5 * r * r
    _, vo = O.advect_velocity(nx, ny, H, u, v, lab, dt)
    V = vo.reshape(ny + 1, nx)
    for j in range(1, ny):
        assert np.abs(V[j, :] - j * factor).max() < 1e-12


def test_walls_stay_zero():  # advect.test.ts:139-168
    nx = ny = 6
    u, v, lab = zeros(nx, ny)
    U = u.reshape(ny, nx + 1)
    V = v.reshape(ny + 1, nx)
    U[:] = np.sin(np.pi * np.arange(nx + 1) / nx)[None, :]
    V[:] = np.sin(np.pi * np.arange(ny + 1) / ny)[:, None]
    uo, vo = O.advect_velocity(nx, ny, H, u, v, lab, 0.4 * H)
    Uo = uo.reshape(ny, nx + 1)
    Vo = vo.reshape(ny + 1, nx)
    assert np.abs(Uo[:, 0]).max() < 1e-12 and np.abs(Uo[:, nx]).max() < 1e-12
    assert np.abs(Vo[0, :]).max() < 1e-12 and np.abs(Vo[ny, :]).max() < 1e-12


def _wavy(nx, ny):
    u, v, lab = zeros(nx, ny)
    U = u.reshape(ny, nx + 1)
    V = v.reshape(ny + 1, nx)
    for j in range(ny):
        for i in range(nx + 1):
            U[j, i] = math.sin(i * 1.7) * math.cos(j * 2.3)
    for j in range(ny + 1):
        for i in range(nx):
            V[j, i] = math.cos(i * 1.1) * math.sin(j * 0.9)
    return u, v, lab


@pytest.mark.parametrize("mac", [False, True])
def test_huge_dt_bounded(mac):  # advect.test.ts:170-200, 459-484
    u, v, lab = _wavy(8, 8)
    uo, vo = O.advect_velocity(8, 8, H, u, v, lab, 50 * H, maccormack=mac)
    assert np.isfinite(uo).all() and np.abs(uo).max() <= np.abs(u).max() + 1e-12
    assert np.isfinite(vo).all() and np.abs(vo).max() <= np.abs(v).max() + 1e-12


def test_stripe_translates_one_cell_x():  # advect.test.ts:267-287
    nx = ny = 8
    u, v, lab = zeros(nx, ny)
    u[:] = 0.75
    q = np.zeros(nx * ny)
    q.reshape(ny, nx)[:, 3] = 1
    out = O.advect_scalar(nx, ny, H, u, v, q, lab, H / 0.75)
    want = np.zeros((ny, nx))
    want[:, 4] = 1
    assert np.abs(out.reshape(ny, nx) - want).max() < 1e-12


def test_stripe_translates_one_cell_y():  # advect.test.ts:289-312
    nx = ny = 8
    u, v, lab = zeros(nx, ny)
    v[:] = 0.5
    q = np.zeros(nx * ny)
    q.reshape(ny, nx)[3, :] = 1
    out = O.advect_scalar(nx, ny, H, u, v, q, lab, H / 0.5)
    want = np.zeros((ny, nx))
    want[4, :] = 1
    assert np.abs(out.reshape(ny, nx) - want).max() < 1e-12


def _vortex_pair(nx, ny, h):
    from flowlab_amd.scenes import Grid, add_vortex_pair
    u, v, lab = zeros(nx, ny)
    add_vortex_pair(Grid(nx, ny, h), u, v)
    return u, v, lab


def test_scalar_no_new_extrema_and_uniform():  # advect.test.ts:314-350
    u, v, lab = _vortex_pair(16, 16, H)
    q = np.zeros(256)
    q.reshape(16, 16)[:, :8] = 1
    out = O.advect_scalar(16, 16, H, u, v, q, lab, 50 * H)
    assert np.isfinite(out).all() and out.min() >= -1e-12 and out.max() <= 1 + 1e-12
    q[:] = 0.7
    out = O.advect_scalar(16, 16, H, u, v, q, lab, 0.3 * H)
    assert np.abs(out - 0.7).max() < 1e-12


def test_scalar_rk2_ramps():  # advect.test.ts:366-428
    dt = 0.5 * H
    r = dt / H
    factor = 1 - r + 0.5 * r * r
    nx, ny = 8, 4
    u, v, lab = zeros(nx, ny)
    u[:] = np.tile(np.arange(nx + 1, dtype=np.float64), ny)
    q = np.tile(np.arange(nx) + 0.5, ny)
    out = O.advect_scalar(nx, ny, H, u, v, q, lab, dt).reshape(ny, nx)
    for i in range(1, nx):
        assert np.abs(out[:, i] - (i + 0.5) * factor).max() < 1e-12
        assert np.abs(out[:, i] - (i + 0.5) * (1 - r)).min() > 1e-9
    nx, ny = 4, 8
    u, v, lab = zeros(nx, ny)
    v[:] = np.repeat(np.arange(ny + 1, dtype=np.float64), nx)
    q = np.repeat(np.arange(ny) + 0.5, nx)
    out = O.advect_scalar(nx, ny, H, u, v, q, lab, dt).reshape(ny, nx)
    for j in range(1, ny):
        assert np.abs(out[j, :] - (j + 0.5) * factor).max() < 1e-12


def test_maccormack_keeps_peak():  # advect.test.ts:492-525
    nx, ny = 64, 16
    u, v, lab = zeros(nx, ny)
    u[:] = 1.0
    sigma = 4 * H
    xs = (np.arange(nx) + 0.5) * H - 8 * H
    ys = (np.arange(ny) + 0.5) * H - 8 * H
    q0 = np.exp(-(xs[None, :] ** 2 + ys[:, None] ** 2) / (2 * sigma * sigma)).ravel()

    def run(mac):
        q = q0.copy()
        for _ in range(40):
            q = O.advect_scalar(nx, ny, H, u, v, q, lab, 0.5 * H, maccormack=mac)
        return q.max()

    sl, mc = run(False), run(True)
    assert mc > sl + 0.05
    assert mc <= 1 + 1e-12


def test_dye_grid_refinement_uniform_translation():
    # advect.ts:241-270 with dg = g refined x2: a constant carrier moving one
    # DYE cell per step shifts the dye by exactly one dye cell.
    nx, ny, s = 8, 8, 2
    u, v, lab = zeros(nx, ny)
    u[:] = 0.5
    q = np.zeros(nx * s * ny * s)
    q.reshape(ny * s, nx * s)[:, 5] = 1
    out = O.advect_scalar(nx, ny, H, u, v, q, lab, (H / s) / 0.5, dscale=s)
    want = np.zeros((ny * s, nx * s))
    want[:, 6] = 1
    assert np.abs(out.reshape(ny * s, nx * s) - want).max() < 1e-12


# ---- divergence.test.ts ----------------------------------------------------

def test_divergence_closed_forms():  # divergence.test.ts:8-55
    u, v, _ = zeros(4, 4)
    u[:] = 1
    assert np.all(O.divergence(4, 4, H, u, v) == 0)
    u[:] = np.tile(np.arange(5) * H, 4)
    assert np.abs(O.divergence(4, 4, H, u, v) - 1).max() < 1e-12
    u[:] = 0
    v[:] = np.repeat(np.arange(5) * H, 4)
    assert np.abs(O.divergence(4, 4, H, u, v) - 1).max() < 1e-12


# ---- pressure.test.ts ------------------------------------------------------

SCALE, DIV = 7.0, 3.0


def _sweep_once(nb_label):  # pressure.test.ts:24-34
    lab = np.zeros(16, np.uint8)
    lab[1 + 1 * 4] = nb_label
    div = np.zeros(16)
    div[2 + 1 * 4] = DIV
    return O.gs_sweep(4, 4, np.zeros(16), div, lab, SCALE, 1.0)


def test_gs_hand_values():  # pressure.test.ts:36-63
    assert abs(_sweep_once(O.FLUID)[6] - (-SCALE * DIV) / 4) < 1e-15
    assert abs(_sweep_once(O.SOLID)[6] - (-SCALE * DIV) / 3) < 1e-15
    assert abs(_sweep_once(O.AIR)[6] - (-SCALE * DIV) / 4) < 1e-15


def test_gs_air_untouched_and_corner():  # pressure.test.ts:65-91
    lab = np.zeros(16, np.uint8)
    lab[5] = O.AIR
    p = O.gs_sweep(4, 4, np.zeros(16), np.full(16, DIV), lab, SCALE, 1.0)
    assert p[5] == 0
    div = np.zeros(16)
    div[0] = DIV
    p = O.gs_sweep(4, 4, np.zeros(16), div, np.zeros(16, np.uint8), SCALE, 1.0)
    assert abs(p[0] - (-SCALE * DIV) / 2) < 1e-15


def test_rb_hand_values():
    # the red-black sweep solves the same row (pressureRedBlack.ts:32-72):
    # the probe (2,1) is colour 1, its neighbours colour 0 and still 0
    for nb, cnt in ((O.FLUID, 4), (O.SOLID, 3), (O.AIR, 4)):
        lab = np.zeros(16, np.uint8)
        lab[5] = nb
        div = np.zeros(16)
        div[6] = DIV
        p = O.rb_sweep(4, 4, np.zeros(16), div, lab, SCALE, 1.0, 1)
        assert abs(p[6] - (-SCALE * DIV) / cnt) < 1e-15


def lcg_fixture(nx=40, ny=32):
    """redBlack.gputest.ts:38-52: Air column at i = nx-1, Solid block
    i in [10,16), j in [12,20), LCG(12345) divergence."""
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[12:20, 10:16] = O.SOLID
    return lab.ravel(), O.lcg_fill(nx * ny, 12345)


def test_lcg_matches_js_double_semantics():
    # seed = (seed*1103515245 + 12345) & 0x7fffffff in doubles, then ToInt32
    seed, vals = 12345, []
    for _ in range(50):
        d = float(seed) * 1103515245.0 + 12345.0
        seed = int(d) % (1 << 32) & 0x7FFFFFFF
        vals.append(seed / 0x7FFFFFFF * 2 - 1)
    assert np.array_equal(O.lcg_fill(50, 12345), np.array(vals))
    # the rounding branch actually fires: exact integer arithmetic differs
    s, exact = 12345, []
    for _ in range(50):
        s = (s * 1103515245 + 12345) & 0x7FFFFFFF
        exact.append(s / 0x7FFFFFFF * 2 - 1)
    assert not np.array_equal(np.array(vals), np.array(exact))


def test_rb_equals_lexicographic():  # pressure.test.ts:106-140
    nx, ny = 40, 32
    lab, div = lcg_fixture(nx, ny)
    lex, _ = O.solve_pressure(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 4000, 1.0, 0)
    rb, _ = O.solve_pressure_rb(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 4000, 1.0, 0)
    pmax = np.abs(lex).max()
    assert pmax > 1e-3
    assert np.abs(lex - rb).max() < 1e-4 * pmax


def test_rb_solves_air_ring():  # pressure.test.ts:151-167
    n = 32
    lab = np.zeros((n, n), np.uint8)
    lab[0, :] = lab[-1, :] = lab[:, 0] = lab[:, -1] = O.AIR
    lab = lab.ravel()
    div = np.ones(n * n)
    p, _ = O.solve_pressure_rb(n, n, np.zeros(n * n), div, lab, 0.05, 2000, 1.0, 0)
    assert O.rms_remaining_divergence(n, n, p, div, lab, 0.05) < 1e-6


def test_solver_guards_zero_divergence():  # pressure.ts:192-199
    lab = np.zeros(16, np.uint8)
    p0 = np.full(16, 3.0)
    for solve in (O.solve_pressure, O.solve_pressure_rb):
        p, it = solve(4, 4, p0, np.zeros(16), lab, 1.0, 10)
        assert it == 0 and np.all(p == 0)


def test_tol_checks_every_12():  # pressure.ts:163-212
    nx, ny = 40, 32
    lab, div = lcg_fixture(nx, ny)
    _, it = O.solve_pressure(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 5000, 1.9, 1e-3)
    assert it % 12 == 0 and 0 < it < 5000
    _, it = O.solve_pressure_rb(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 5000, 1.9, 1e-3)
    assert it % 12 == 0 and 0 < it < 5000


# ---- subtractGradient / boundaries.test.ts ---------------------------------

def test_project_removes_pure_gradient():
    # A field that IS the discrete gradient of phi (subtractGradient.ts:40-61
    # bounds, testFields.ts addGradient) projects to ~0 once solved.
    nx, ny, h = 24, 20, 1 / 20
    phi = np.cos(2 * np.pi * (np.arange(nx) + 0.5) * h)[None, :] * \
        np.cos(2 * np.pi * (np.arange(ny) + 0.5) * h)[:, None]
    u, v, lab = zeros(nx, ny)
    U = u.reshape(ny, nx + 1)
    V = v.reshape(ny + 1, nx)
    U[:, 1:nx] = (phi[:, 1:] - phi[:, :-1]) / h
    V[1:ny, :] = (phi[1:, :] - phi[:-1, :]) / h
    div = O.divergence(nx, ny, h, u, v)
    dt = 0.1
    p, _, _ = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, h * h / dt, 1, 60, 1e-12)
    u2, v2 = O.subtract_gradient(nx, ny, p, u, v, lab, dt / h)
    assert np.abs(u2).max() < 1e-8 and np.abs(v2).max() < 1e-8


def test_commit_labels():  # boundaries.test.ts:8-61
    g = 6
    u, v, lab = zeros(g, g)
    u[:] = 1
    v[:] = 1
    lab[2 + 2 * g] = O.SOLID
    _, u2, v2, _ = O.commit_labels(g, g, np.zeros(36), u, v, lab)
    assert u2[iu(g, 2, 2)] == 0 and u2[iu(g, 3, 2)] == 0
    assert v2[iv(g, 2, 2)] == 0 and v2[iv(g, 2, 3)] == 0
    assert u2[iu(g, 1, 2)] == 1
    u[:] = 2
    lab[:] = 0
    lab[3 + 3 * g] = O.SOLID
    _, u2, _, _ = O.commit_labels(g, g, np.zeros(36), u, v, lab)
    U = u2.reshape(g, g + 1)
    assert np.all(U[:, 0] == 2) and np.all(U[:, g] == 2)
    p = np.full(36, 5.0)
    dye = np.ones(3 * 36)
    lab[:] = 0
    lab[5 + 1 * g] = O.AIR
    lab[2 + 2 * g] = O.SOLID
    p2, _, _, d2 = O.commit_labels(g, g, p, u, v, lab, dye)
    assert p2[5 + g] == 0
    assert all(d2[c * 36 + 2 + 2 * g] == 0 for c in range(3))
    assert p2[1 + g] == 5 and d2[1 + g] == 1


def test_outflow():  # boundaries.test.ts:63-124
    g = 6
    u, v, lab = zeros(g, g)
    U = u.reshape(g, g + 1)
    U[:, g - 1] = np.arange(g) + 1
    lab.reshape(g, g)[:3, g - 1] = O.AIR
    u2, _ = O.apply_outflow(g, g, u, v, lab)
    U2 = u2.reshape(g, g + 1)
    assert np.all(U2[:3, g] == np.arange(3) + 1) and np.all(U2[3:, g] == 0)
    # closed box: no-op
    u[:] = 3
    v[:] = 4
    lab[:] = 0
    u2, v2 = O.apply_outflow(g, g, u, v, lab)
    assert np.array_equal(u2, u) and np.array_equal(v2, v)
    # backflow refused
    u[:] = 0
    lab.reshape(g, g)[:, g - 1] = O.AIR
    U[0, g - 1] = 2
    U[1, g - 1] = -2
    u2, _ = O.apply_outflow(g, g, u, v, lab)
    U2 = u2.reshape(g, g + 1)
    assert U2[0, g - 1] == 2 and U2[0, g] == 2 and U2[1, g - 1] == 0 and U2[1, g] == 0
    # low edge clamps the other way
    u[:] = 0
    lab[:] = 0
    lab.reshape(g, g)[:, 0] = O.AIR
    U[0, 1] = -2
    U[1, 1] = 2
    u2, _ = O.apply_outflow(g, g, u, v, lab)
    U2 = u2.reshape(g, g + 1)
    assert U2[0, 1] == -2 and U2[0, 0] == -2 and U2[1, 1] == 0 and U2[1, 0] == 0


def test_outflow_vertical_edges():
    # boundaries.ts:131-142: the v loops mirror the u loops on the low/high rows
    g = 5
    u, v, lab = zeros(g, g)
    V = v.reshape(g + 1, g)
    L = lab.reshape(g, g)
    L[0, :] = O.AIR
    L[g - 1, :] = O.AIR
    V[1, :] = [-1, 2, -3, 4, 0]
    V[g - 1, :] = [1, -2, 3, -4, 0]
    _, v2 = O.apply_outflow(g, g, u, v, lab)
    V2 = v2.reshape(g + 1, g)
    assert np.array_equal(V2[1], [-1, 0, -3, 0, 0]) and np.array_equal(V2[0], V2[1])
    assert np.array_equal(V2[g - 1], [1, 0, 3, 0, 0]) and np.array_equal(V2[g], V2[g - 1])


# ---- buoyancy.test.ts ------------------------------------------------------

def test_buoyancy_faces():  # buoyancy.test.ts:82-109
    N = 48
    heat = np.ones(N * N)
    lab = np.zeros((N, N), np.uint8)
    lab[10:14, 10:14] = O.SOLID
    v = O.apply_buoyancy(N, N, 1 / N, np.zeros(N * (N + 1)), heat, lab.ravel(), 1, 0.1)
    V = v.reshape(N + 1, N)
    assert np.all(V[10:15, 10:14] == 0)
    assert np.all(V[0] == 0) and np.all(V[N] == 0)
    assert abs(V[20, 30] - 0.1) < 1e-12


def test_buoyancy_per_sim_time():  # buoyancy.test.ts:111-122
    N = 48
    heat = np.ones(N * N)
    lab = np.zeros(N * N, np.uint8)
    one = O.apply_buoyancy(N, N, 1 / N, np.zeros(N * (N + 1)), heat, lab, 3, 0.4)
    many = np.zeros(N * (N + 1))
    for _ in range(4):
        many = O.apply_buoyancy(N, N, 1 / N, many, heat, lab, 3, 0.1)
    assert abs(one[iv(N, 20, 20)] - many[iv(N, 20, 20)]) < 1e-12


def test_buoyancy_dye_grid_finer():  # buoyancy.test.ts:124-142
    heat = np.zeros(256)
    heat.reshape(16, 16)[8:, :] = 1
    v = O.apply_buoyancy(8, 8, 1 / 8, np.zeros(72), heat, np.zeros(64, np.uint8), 1, 1, dscale=2)
    assert abs(v[iv(8, 3, 4)] - 0.5) < 1e-12
    assert abs(v[iv(8, 3, 6)] - 1) < 1e-12
    assert v[iv(8, 3, 2)] == 0


# ---- dye.ts ----------------------------------------------------------------

def test_dye_patch_lerp_and_clip():
    dnx = dny = 6
    dye = np.full(3 * 36, 0.5)
    pnx, pny = 3, 2
    data = np.zeros(4 * pnx * pny, np.float32)
    data[0:6] = 1.0        # channel 0 value
    data[6:12] = 0.25      # channel 1 value
    data[18:24] = [1, 0.5, 0, 1, 1, 1]  # coverage
    out = O.apply_dye_patch(dnx, dny, dye, 4, 4, pnx, pny, data)  # clipped at the edge
    D0 = out[:36].reshape(6, 6)
    assert D0[4, 4] == 1.0 and D0[4, 5] == 0.75 and D0[5, 4] == 1.0 and D0[0, 0] == 0.5
    D1 = out[36:72].reshape(6, 6)
    assert D1[4, 4] == 0.25


# ---- stir.test.ts ----------------------------------------------------------

def test_stir_local_and_signed():  # stir.test.ts:8-22
    nx, ny = 96, 48
    h = 1 / ny
    u, v, lab = zeros(nx, ny)
    s = (0.9, 0.5, 0.06, 0.8, -0.4)
    u2, v2 = O.apply_stir(nx, ny, h, u, v, lab, *s)
    i = round(s[0] / h)
    j = round(s[1] / h - 0.5)
    assert u2[iu(nx, i, j)] > 0.5 * s[3]
    assert v2[iv(nx, i, j)] < 0.5 * s[4]
    far = round((s[0] - 4 * s[2]) / h)
    assert abs(u2[iu(nx, far, j)]) < 1e-3


def test_stir_skips_boundary_and_solids():  # stir.test.ts:24-49
    nx, ny = 96, 48
    h = 1 / ny
    u, v, lab = zeros(nx, ny)
    u[:] = 1
    u2, _ = O.apply_stir(nx, ny, h, u, v, lab, 0.0, 0.5, 0.06, 0.8, -0.4)
    assert np.all(u2.reshape(ny, nx + 1)[:, 0] == 1)
    L = lab.reshape(ny, nx)
    L[20:28, 40:48] = O.SOLID
    u2, v2 = O.apply_stir(nx, ny, h, np.zeros_like(u), v, lab, 0.92, 0.5, 0.06, 0.8, -0.4)
    U = u2.reshape(ny, nx + 1)
    assert np.all(U[20:28, 40:49] == 0)
