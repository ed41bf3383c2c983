"""GPU parity: every libflowhip kernel/phase diffed against the FP64 oracle
from identical state.

Inputs are rounded to FP32 first (the device's storage precision) and the
oracle runs on those same values, so each diff measures the kernel's own
arithmetic.  The bar is a per-kernel relative L2 <= 1e-5 in
FP32, plus the reference gputests' max|diff| < 1e-5*max|p| and residuals
within 1% for the pressure solvers (src/gpu/redBlack.gputest.ts:155-164,
multigrid.gputest.ts:231-238).
"""
import numpy as np
import pytest

from flowlab_amd import scenes as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def E():
    from flowlab_amd.engine import HipEngine
    return HipEngine


def lcg_fixture(nx=40, ny=32):
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[12:20, 10:16] = O.SOLID
    return lab.ravel(), O.lcg_fill(nx * ny, 12345)


def test_rbsor_matches_oracle(E):
    # redBlack.gputest.ts: 40x32 fixture, 200 sweeps, omega 1.85, scale 0.02
    nx, ny = 40, 32
    lab, div = lcg_fixture(nx, ny)
    e = E(nx, ny)
    from flowlab_amd.engine import HipPressureSolver
    sol = HipPressureSolver(nx, ny, "rbsor")
    p = np.zeros(nx * ny)
    it = sol.solve(None, p, div, lab, 0.02, 200, 1.85, 0)
    assert it == 200
    exp, _ = O.solve_pressure_rb(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 200, 1.85, 0)
    pmax = np.abs(exp).max()
    assert pmax > 1e-3
    assert np.abs(p - exp).max() < 1e-5 * pmax
    assert rel_l2(p, exp) < REL_L2
    r_cpu = O.rms_remaining_divergence(nx, ny, exp, div, lab, 0.02)
    r_gpu = O.rms_remaining_divergence(nx, ny, p, div, lab, 0.02)
    assert abs(r_gpu - r_cpu) < 0.01 * r_cpu
    del e


def test_mg_reference_mode_matches_oracle():
    # multigrid.gputest.ts: V(2,2) to <=4 cells, 8 coarse pairs, 3 cycles
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = 40, 32
    lab, div = lcg_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, "mg")
    p = np.zeros(nx * ny)
    it = sol.solve(None, p, div, lab, 0.02, 3, 1.0, 0)
    assert it == 3
    exp, _, _ = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 0, 3, 0)
    pmax = np.abs(exp).max()
    assert np.abs(p - exp).max() < 1e-5 * pmax
    r_cpu = O.rms_remaining_divergence(nx, ny, exp, div, lab, 0.02)
    r_gpu = O.rms_remaining_divergence(nx, ny, p, div, lab, 0.02)
    assert abs(r_gpu - r_cpu) < 0.01 * r_cpu


def boundary_fixture(nx, ny):
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[ny >> 2:ny >> 1, nx >> 2:(nx * 3) >> 3] = O.SOLID
    return lab.ravel(), O.lcg_fill(nx * ny, 12345)


@pytest.mark.parametrize("shape", [(64, 48), (96, 40), (37, 29)])
def test_mg_dc_one_cycle_matches_oracle(shape):
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = shape
    lab, div = boundary_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, "mg-dc")
    p = np.zeros(nx * ny)
    it = sol.solve(None, p, div, lab, 0.02, 1, 1.0, 0)
    assert it == 1
    exp = O.mg_dc_cycle(nx, ny, np.zeros(nx * ny), -0.02 * div, lab)
    assert rel_l2(p, exp) < REL_L2


@pytest.mark.parametrize("closed", [False, True])
def test_mg_dc_reaches_1e6(closed):
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = 256, 192
    if closed:
        lab = np.zeros(nx * ny, np.uint8)
        lab.reshape(ny, nx)[60:100, 80:120] = O.SOLID
        div = O.lcg_fill(nx * ny, 777)
        div[lab == O.SOLID] = 0
        div[lab == O.FLUID] -= div[lab == O.FLUID].mean()  # compatible RHS
    else:
        lab, div = boundary_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, "mg-dc")
    p = np.zeros(nx * ny)
    it = sol.solve(None, p, div, lab, 0.02, 60, 1.0, 1e-6)
    rel = O.relative_residual(nx, ny, p, div, lab, 0.02)
    assert rel <= 1e-6 and it < 60
    _, it_cpu, _ = O.mg_solve(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 1, 60, 1e-6)
    assert abs(it - it_cpu) <= 1
    hist = sol.engine.solve_history()
    assert len(hist) == it + 1 and hist[-1] <= 1e-6


@pytest.mark.parametrize("solver", ["mg-dc", "mg-fcg"])
@pytest.mark.parametrize("shape", [(256, 192), (197, 131)])
def test_fused_update_gives_the_same_bits(solver, shape):
    """Defect-correction cycles whose finest prolongation + post-smoothing
    and FP64 update run as one kernel give the same pressure bits and cycle
    counts as the two passes, on every tile shape of the finest level (odd
    sizes: partial tiles at two edges), and reach 1e-6 by the oracle."""
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = shape
    lab, div = boundary_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, solver)
    e = sol.engine
    p0 = np.zeros(nx * ny)
    sol.solve(None, p0.copy(), div, lab, 0.02, 1, 1.0, 0)  # labels, hierarchy
    assert len(e.tiles()) >= 1, "the finest level must be tiled for the fused pass"
    ref = {}
    for fused in (False, True):
        e.set_fused_update(fused)
        for t in range(len(e.tile_names())):
            e.set_tile(t, level=0)
            p = p0.copy()
            it = sol.solve(None, p, div, lab, 0.02, 60, 1.0, 1e-6)
            if not fused:
                ref[t] = (p, it)
                continue
            rp, rit = ref[t]
            assert it == rit, (t, it, rit)
            assert np.array_equal(p, rp), t
    assert O.relative_residual(nx, ny, p, div, lab, 0.02) <= 1e-6
    e.set_fused_update(True)


def _velocity_state(nx, ny, seed=12345):
    g = S.Grid(nx, ny, 1.0 / ny)
    u = np.zeros((nx + 1) * ny)
    v = np.zeros(nx * (ny + 1))
    S.add_vortex_cluster(g, u, v, amp=1.0, seed=seed)
    u += 0.3
    lab = np.zeros((ny, nx), np.uint8)
    lab[:, nx - 1] = O.AIR
    lab[ny // 3:ny // 2, nx // 4:nx // 3] = O.SOLID
    lab = lab.ravel()
    _, u, v, _ = O.commit_labels(nx, ny, np.zeros(nx * ny), u, v, lab)
    return g, f32(u), f32(v), lab


@pytest.mark.parametrize("scheme", ["semiLagrangian", "macCormack"])
def test_advect_velocity_matches_oracle(scheme):
    from flowlab_amd.engine import HipAdvector
    nx, ny = 96, 64
    g, u, v, lab = _velocity_state(nx, ny)
    dt = 1.5 * g.h / max(np.abs(u).max(), np.abs(v).max())
    adv = HipAdvector(nx, ny)
    uo, vo = np.empty_like(u), np.empty_like(v)
    adv.velocity(g, scheme, u, v, uo, vo, lab, dt)
    eu, ev = O.advect_velocity(nx, ny, g.h, u, v, lab, dt, maccormack=scheme == "macCormack")
    assert rel_l2(uo, eu) < REL_L2 and rel_l2(vo, ev) < REL_L2


@pytest.mark.parametrize("scheme,ds", [("semiLagrangian", 1), ("macCormack", 1),
                                       ("macCormack", 2)])
def test_advect_dye_matches_oracle(scheme, ds):
    from flowlab_amd.engine import HipAdvector
    nx, ny = 80, 48
    g, u, v, lab = _velocity_state(nx, ny, seed=99)
    dt = 1.2 * g.h / max(np.abs(u).max(), np.abs(v).max())
    dnx, dny = nx * ds, ny * ds
    q = []
    for c in range(3):
        x = (np.arange(dnx) + 0.5) / dnx
        y = (np.arange(dny) + 0.5) / dny
        q.append(f32((np.sin((c + 2) * 7 * x)[None, :] * np.cos(5 * y + c)[:, None] > 0.2)
                     .astype(float).ravel()))
    adv = HipAdvector(nx, ny, ds)
    out = [np.empty(dnx * dny) for _ in range(3)]
    adv.dye(g, scheme, u, v, q, out, lab, dt)
    for c in range(3):
        exp = O.advect_scalar(nx, ny, g.h, u, v, q[c], lab, dt,
                              maccormack=scheme == "macCormack", dscale=ds)
        assert rel_l2(out[c], exp) < REL_L2, c


def _resident(nx, ny, ds=1, seed=7):
    from flowlab_amd.engine import HipEngine
    g, u, v, lab = _velocity_state(nx, ny, seed)
    dnx, dny = nx * ds, ny * ds
    dye = np.zeros(3 * dnx * dny)
    x = (np.arange(dnx) + 0.5) / dnx
    y = (np.arange(dny) + 0.5) / dny
    dye[:dnx * dny] = (np.exp(-((x[None, :] - 0.4) ** 2 + (y[:, None] - 0.4) ** 2) / 0.02)).ravel()
    dye = f32(dye)
    p = f32(O.lcg_fill(nx * ny, 31) * 1e-3)
    p[lab != O.FLUID] = 0
    e = HipEngine(nx, ny, ds)
    e.upload(u, v, p, lab, dye)
    return e, g, u, v, p, lab, dye


def test_forces_match_oracle():
    nx, ny = 96, 64
    e, g, u, v, p, lab, dye = _resident(nx, ny)
    dt, w, eps = 0.01, 2.0, 0.3
    stir = (0.6, 0.5, 0.08, 0.7, -0.4)
    e.set_buoyancy(w)
    e.set_confinement(eps)
    e.set_stir(stir)
    e.phase_forces(dt)
    gu, gv = e.read_velocity()
    # oracle, in the CPU step's order: confinement from the advected field,
    # then stir, buoyancy, confinement added
    wc = O.compute_vorticity(nx, ny, g.h, u, v)
    ou, ov = O.apply_stir(nx, ny, g.h, u, v, lab, *stir)
    ov = O.apply_buoyancy(nx, ny, g.h, ov, dye[:nx * ny], lab, w, dt)
    import ctypes as C
    ou = ou.copy()
    ov = ov.copy()
    O._lib.fo_vorticity_confinement(nx, ny, g.h, O._p(ou), O._p(ov), O._p(lab), O._p(wc), eps, dt)
    du_g, du_o = gu - u, ou - u
    dv_g, dv_o = gv - v, ov - v
    assert np.abs(du_o).max() > 1e-4 and np.abs(dv_o).max() > 1e-4
    assert rel_l2(du_g, du_o) < 1e-4 and rel_l2(dv_g, dv_o) < 1e-4
    assert rel_l2(gu, ou) < REL_L2 and rel_l2(gv, ov) < REL_L2


@pytest.mark.parametrize("ds", [1, 2])
def test_buoyancy_alone_matches_oracle(ds):
    """Buoyancy without stir or confinement (its own kernel): u untouched bit
    for bit, v against the oracle's buoyancy on the (refined) dye grid."""
    nx, ny = 96, 64
    e, g, u, v, p, lab, dye = _resident(nx, ny, ds)
    dt, w = 0.01, 2.0
    e.set_buoyancy(w)
    e.phase_forces(dt)
    gu, gv = e.read_velocity()
    assert np.array_equal(gu, u)
    ov = O.apply_buoyancy(nx, ny, g.h, v, dye[:nx * ny * ds * ds], lab, w, dt, dscale=ds)
    dv_g, dv_o = gv - v, ov - v
    assert np.abs(dv_o).max() > 1e-4
    assert rel_l2(dv_g, dv_o) < 1e-5
    assert rel_l2(gv, ov) < REL_L2


@pytest.mark.parametrize("solver", ["mg-fcg", "mg-dc"])
def test_fused_step_equals_its_phases(solver):
    """fh_step applies buoyancy alone inside the MacCormack face pass and
    forms the divergence inside the solve's initial defect pass; the
    phase-by-phase path runs the separate buoyancy and divergence kernels.
    Two steps each way from the same state agree bit for bit (and the
    divergence both leave behind), so the fusions change no value."""
    nx, ny = 96, 64
    dt, w = 0.01, 2.0
    scale, gs = 1.0 / (64 * 64) / dt, dt * 64
    keep = (0.98, 1.0, 0.95)
    engines = []
    for _ in range(2):
        e, g, u, v, p, lab, dye = _resident(nx, ny)
        e.set_buoyancy(w)
        engines.append(e)
    fused, phased = engines
    for _ in range(2):
        it_f = fused.step("macCormack", dt, scale, gs, keep, solver, 60, 1.0, 1e-6)
        phased.phase_advect_velocity("macCormack", dt)
        phased.phase_forces(dt)
        phased.phase_divergence()
        it_p = phased.phase_solve(solver, scale, 60, 1.0, 1e-6)
        phased.phase_project(gs)
        phased.phase_dye("macCormack", dt, keep)
        assert it_f == it_p
        for a, b in zip(fused.read_velocity(), phased.read_velocity()):
            assert np.array_equal(a, b)
        assert np.array_equal(fused.read_pressure(), phased.read_pressure())
        assert np.array_equal(fused.read_divergence(), phased.read_divergence())
        assert np.array_equal(fused.read_dye(), phased.read_dye())


def test_divergence_subtract_outflow_match_oracle():
    nx, ny = 96, 64
    e, g, u, v, p, lab, dye = _resident(nx, ny)
    e.phase_divergence()
    d = e.read_divergence()
    od = O.divergence(nx, ny, g.h, u, v)
    assert rel_l2(d, od) < 1e-12
    gs = 0.37
    e.phase_project(gs)
    gu, gv = e.read_velocity()
    ou, ov = O.subtract_gradient(nx, ny, p, u, v, lab, gs)
    ou, ov = O.apply_outflow(nx, ny, ou, ov, lab)
    assert rel_l2(gu, ou) < REL_L2 and rel_l2(gv, ov) < REL_L2
    assert np.all(gu.reshape(ny, nx + 1)[:, nx - 1] >= 0)


@pytest.mark.parametrize("nx,ny,edges", [(96, 64, "lrbt"), (67, 45, "lt"), (3, 3, "lrbt"),
                                          (5, 4, "rb"), (2, 5, "lr"), (40, 2, "bt")])
def test_projection_outflow_edges_and_max_speed(nx, ny, edges):
    """The projection pass (gradient subtraction, outflow at Air edge cells,
    and the largest face speed of the result, taken in one pass where nx,
    ny >= 3) against the oracle's subtractGradient then applyOutflow, with
    Air on the given edges (l, r, b, t) and a solid block; the reported
    face speed is the maximum of the velocity read back, exactly."""
    from flowlab_amd.engine import HipEngine
    g, u, v, lab = _velocity_state(nx, ny, 11)
    L = lab.reshape(ny, nx).copy()
    L[:] = O.FLUID
    if "l" in edges:
        L[:, 0] = O.AIR
    if "r" in edges:
        L[:, nx - 1] = O.AIR
    if "b" in edges:
        L[0, :] = O.AIR
    if "t" in edges:
        L[ny - 1, :] = O.AIR
    if nx >= 8 and ny >= 8:
        L[ny // 3:ny // 2, nx // 3:nx // 2] = O.SOLID
    lab = L.ravel()
    p = f32(O.lcg_fill(nx * ny, 5) * 1e-2)
    p[lab != O.FLUID] = 0
    e = HipEngine(nx, ny)
    e.upload(u, v, p, lab, np.zeros(3 * nx * ny))
    gs = 0.41
    e.phase_project(gs)
    gu, gv = e.read_velocity()
    ou, ov = O.subtract_gradient(nx, ny, p, u, v, lab, gs)
    ou, ov = O.apply_outflow(nx, ny, ou, ov, lab)
    assert rel_l2(gu, ou) < REL_L2 and rel_l2(gv, ov) < REL_L2
    assert np.abs(gu - ou).max() <= 1e-6 * max(np.abs(ou).max(), 1.0)
    assert np.abs(gv - ov).max() <= 1e-6 * max(np.abs(ov).max(), 1.0)
    assert e.max_face_speed() == max(np.abs(gu).max(), np.abs(gv).max())


@pytest.mark.parametrize("scheme", ["semiLagrangian", "macCormack"])
def test_dye_phase_with_decay_and_patch(scheme):
    nx, ny = 64, 48
    e, g, u, v, p, lab, dye = _resident(nx, ny)
    patch = S.heat_disk_source(0.5, 0.2, 0.1)(g, g)
    e.set_dye_patch(patch)
    keep = (0.9, 0.8, 1.0)
    dt = 0.7 * g.h
    e.phase_dye(scheme, dt, keep)
    got = e.read_dye()
    n = nx * ny
    exp = np.zeros(3 * n)
    for c in range(3):
        q = O.advect_scalar(nx, ny, g.h, u, v, dye[c * n:(c + 1) * n], lab, dt,
                            maccormack=scheme == "macCormack")
        exp[c * n:(c + 1) * n] = q * keep[c]
    exp = O.apply_dye_patch(nx, ny, exp, patch.i0, patch.j0, patch.nx, patch.ny, patch.data)
    assert rel_l2(got, exp) < REL_L2


def _oracle_step_params(sim_params, **kw):
    return O.step_params(**kw)


def test_plume_128_whole_step_matches_oracle():
    """Thermal plume 128^2: MacCormack, buoyancy 1, confinement 0.1, RB-SOR
    omega 1.9 x 200, dt = 0.5 h.  Whole fused steps vs the oracle's step."""
    from flowlab_amd.simulation import Simulation
    g, scene = S.thermal_plume_128()
    params = dict(scheme="macCormack", cfl_target=1e30, dt_max=0.5 * g.h, solver="rbsor",
                  pressure_iters=200, omega=1.9, tol=0.0, buoyancy=1.0, confinement=0.1,
                  residual_every=0)
    sim = Simulation(128, 128, params)
    sim.reset(scene)
    osim = O.OracleSim(128, 128)
    osim.reset(g, scene)
    prm = O.step_params(scheme="macCormack", solver="rbsor", pressure_iters=200, omega=1.9,
                        tol=0.0, cfl_target=1e30, dt_max=0.5 * g.h, buoyancy=1.0,
                        confinement=0.1, residual_every=0,
                        patch=(osim.patch.i0, osim.patch.j0, osim.patch.nx, osim.patch.ny))
    errs = []
    for n in range(10):
        sim.step()
        osim.step(prm, osim.patch.data)
        errs.append((rel_l2(sim.f.u, osim.u), rel_l2(sim.f.v, osim.v)))
    dye = sim.read_dye()
    assert max(errs[0]) < REL_L2, errs[0]
    assert max(errs[-1]) < 1e-4, errs
    assert rel_l2(np.concatenate(dye), osim.dye) < 1e-4
    assert sim.f.v.max() > 0.01  # the plume rises


def _reseed(osim, sim):
    """The oracle takes the device's current fields: velocity, dye and the
    pressure the next solve starts from."""
    osim.u[:] = sim.f.u
    osim.v[:] = sim.f.v
    osim.dye[:] = np.concatenate(sim.read_dye())
    osim.p[:] = sim.stepper.engine.read_pressure()


def test_plume_128_every_step_from_device_state():
    """All 200 steps of the 128^2 plume (RB-SOR omega 1.9 x 200, buoyancy 1,
    confinement 0.1, MacCormack): before each device step the oracle takes
    the device's fields and steps once in FP64, and the device's step must
    match it at the per-kernel bar.  Re-seeding every step keeps the FP32
    rounding of earlier steps from compounding through the flow, so the bar
    stays 1e-5 over the whole run."""
    from flowlab_amd.simulation import Simulation
    g, scene = S.thermal_plume_128()
    params = dict(scheme="macCormack", cfl_target=1e30, dt_max=0.5 * g.h, solver="rbsor",
                  pressure_iters=200, omega=1.9, tol=0.0, buoyancy=1.0, confinement=0.1,
                  residual_every=0)
    sim = Simulation(128, 128, params)
    sim.reset(scene)
    osim = O.OracleSim(128, 128)
    osim.reset(g, scene)
    prm = O.step_params(scheme="macCormack", solver="rbsor", pressure_iters=200, omega=1.9,
                        tol=0.0, cfl_target=1e30, dt_max=0.5 * g.h, buoyancy=1.0,
                        confinement=0.1, residual_every=0,
                        patch=(osim.patch.i0, osim.patch.j0, osim.patch.nx, osim.patch.ny))
    worst = (0.0, 0.0, 0.0)
    for n in range(200):
        _reseed(osim, sim)
        sim.step()
        osim.step(prm, osim.patch.data)
        e = (rel_l2(sim.f.u, osim.u), rel_l2(sim.f.v, osim.v),
             rel_l2(np.concatenate(sim.read_dye()), osim.dye))
        worst = tuple(max(a, b) for a, b in zip(worst, e))
        if n % 50 == 0:
            print(f"step {n}: u {e[0]:.2e} v {e[1]:.2e} dye {e[2]:.2e}", flush=True)
    print("worst", worst)
    assert max(worst) < REL_L2, worst
    assert sim.f.v.max() > 0.05  # the plume has risen


@pytest.mark.parametrize("solver", ["mg-dc", "mg-fcg"])
def test_karman_every_step_from_device_state(solver):
    """Karman street 512x256, 20 steps at CFL 0.5 with every solve to 1e-6:
    each device step against one oracle step (defect-correction multigrid)
    from the device's fields.  The device's dt comes from its own max face
    speed and the oracle's from the same values, so both take the same dt;
    the two solves differ by at most their 1e-6 residuals."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 512, 256
    g, scene = S.karman_street(nx, ny)
    sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver=solver,
                                  pressure_iters=60, tol=1e-6, residual_every=0))
    sim.reset(scene)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    worst = (0.0, 0.0)
    for n in range(20):
        _reseed(osim, sim)
        prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=1e-6,
                            cfl_target=0.5, residual_every=0)
        sim.step()
        osim.step(prm)
        assert sim.dt == prm.dt_used, (n, sim.dt, prm.dt_used)
        e = (rel_l2(sim.f.u, osim.u), rel_l2(sim.f.v, osim.v))
        worst = tuple(max(a, b) for a, b in zip(worst, e))
    print("worst", worst)
    assert worst[0] < REL_L2 and worst[1] < REL_L2, worst


def test_karman_whole_step_mg_dc():
    """Karman street at 512x256 with MG-DC to 1e-6: per-step parity with the
    oracle and every solve at 1e-6 by the oracle's own FP64 residual."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 512, 256
    g, scene = S.karman_street(nx, ny)
    params = dict(scheme="macCormack", cfl_target=0.5, solver="mg-dc", pressure_iters=60,
                  tol=1e-6, residual_every=0)
    sim = Simulation(nx, ny, params)
    sim.reset(scene)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=1e-6,
                        cfl_target=0.5, residual_every=0)
    for n in range(3):
        sim.step()
        osim.step(prm)
        p = sim.stepper.engine.read_pressure()
        d = sim.stepper.engine.read_divergence()
        scale = g.h * g.h / sim.dt
        assert O.relative_residual(nx, ny, p, d, sim.f.label, scale) <= 1e-6
        assert abs(sim.iters - prm.iters) <= 1
        assert rel_l2(sim.f.u, osim.u) < 1e-5 and rel_l2(sim.f.v, osim.v) < 1e-4


def test_uniform_through_flow_steady_on_gpu():
    # emitters.test.ts:160-183: uniform through-flow is an exact steady state
    from flowlab_amd.simulation import Simulation

    def seed(g, u, v):
        u.fill(1)
    sim = Simulation(64, 32, dict(scheme="macCormack", solver="mg-dc", tol=1e-6,
                                  pressure_iters=30))
    sim.reset(S.SceneSpec(labels=S.open_right(), seed=seed))
    u0 = sim.f.u.copy()
    for _ in range(10):
        sim.step()
    assert np.abs(sim.f.u - u0).max() < 1e-6


def test_full_size_karman_projection_property():
    """The Karman street at 2048x1024: after a fused step the solve met 1e-6 by
    the oracle's FP64 residual, and the projected field is divergence-free to
    that level -- a size-independent check where the full oracle step would
    take too long."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 2048, 1024
    g, scene = S.karman_street(nx, ny)
    sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver="mg-dc",
                                  pressure_iters=60, tol=1e-6, residual_every=1))
    sim.reset(scene)
    for _ in range(2):
        sim.step()
        e = sim.stepper.engine
        p, d = e.read_pressure(), e.read_divergence()
        scale = g.h * g.h / sim.dt
        assert O.relative_residual(nx, ny, p, d, sim.f.label, scale) <= 1e-6
    fl = sim.f.label == O.FLUID
    interior = fl.copy().reshape(ny, nx)
    interior[:, nx - 2:] = False  # the outlet clamp leaves one layer (boundaries.ts:100-106)
    div_after = sim.div.reshape(ny, nx)[interior]
    div_before = d.reshape(ny, nx)[interior]
    assert np.sqrt(np.mean(div_after ** 2)) < 1e-4 * max(np.sqrt(np.mean(div_before ** 2)), 1)


@pytest.mark.parametrize("solver", ["mg-dc", "mg-fcg"])
@pytest.mark.parametrize("iterations,tol", [(60, 1e-6), (1, 0.0), (5, 0.0), (0, 1e-6)])
def test_mg_dc_zero_divergence_guard(iterations, tol, solver):
    """pressureMultigrid.ts:294-300: zero divergence gives p = 0 and 0 cycles,
    whatever the warm start, and the next solve on the engine runs normally."""
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = 96, 64
    lab, div = boundary_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, solver)
    p = np.linspace(-1.0, 1.0, nx * ny)  # a leftover warm start
    it = sol.solve(None, p, np.zeros(nx * ny), lab, 0.02, iterations, 1.0, tol)
    assert it == 0 and not p.any()
    assert len(sol.engine.solve_history()) == 0
    it = sol.solve(None, p, div, lab, 0.02, 60, 1.0, 1e-6)
    assert 0 < it < 60
    assert O.relative_residual(nx, ny, p, div, lab, 0.02) <= 1e-6


@pytest.mark.parametrize("solver", ["mg-dc", "mg-fcg"])
def test_mg_dc_stops_at_the_first_cycle_within_tolerance(solver):
    """Cycle count is minimal however the host schedules the cycles: every
    history entry but the last is above the tolerance, repeated warm-started
    solves included."""
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = 256, 192
    lab, div = boundary_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, solver)
    p = np.zeros(nx * ny)
    for k in range(4):
        d = div * (1.0 + 0.05 * k)
        it = sol.solve(None, p, d, lab, 0.02, 60, 1.0, 1e-6)
        hist = sol.engine.solve_history()
        assert len(hist) == it + 1
        assert hist[-1] <= 1e-6 and all(h > 1e-6 for h in hist[:-1])
        assert O.relative_residual(nx, ny, p, d, lab, 0.02) <= 1e-6


def test_step_max_face_speed_tracks_the_velocity():
    """simulation.ts:451-459: the largest face speed of the CURRENT velocity.
    A step reduces its final velocity for the next step's CFL; any later
    write of the velocity (here an upload) must not see that value."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 256, 128
    g, scene = S.karman_street(nx, ny)
    sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver="mg-dc",
                                  pressure_iters=60, tol=1e-6, residual_every=0), readback=True)
    sim.reset(scene)
    e = sim.stepper.engine
    for _ in range(3):
        sim.step()
        exp = max(np.abs(sim.f.u).max(), np.abs(sim.f.v).max())  # read back, f32 values
        assert e.max_face_speed() == exp
    u2, v2 = 3.0 * sim.f.u, 3.0 * sim.f.v
    e.upload(u2, v2, None, None, None)
    assert e.max_face_speed() == max(np.abs(f32(u2)).max(), np.abs(f32(v2)).max())


@pytest.mark.parametrize("closed", [False, True])
def test_mg_fcg_reaches_1e6(closed):
    """Flexible CG around the V-cycle: the oracle's FP64 residual of the
    device's p is within the tolerance, the cycle count is the oracle
    restatement's within one, and the residual the device reports (updated
    recursively) is the true one."""
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = 256, 192
    if closed:
        lab = np.zeros(nx * ny, np.uint8)
        lab.reshape(ny, nx)[60:100, 80:120] = O.SOLID
        div = O.lcg_fill(nx * ny, 777)
        div[lab == O.SOLID] = 0
        div[lab == O.FLUID] -= div[lab == O.FLUID].mean()  # compatible RHS
    else:
        lab, div = boundary_fixture(nx, ny)
    sol = HipPressureSolver(nx, ny, "mg-fcg")
    p = np.zeros(nx * ny)
    it = sol.solve(None, p, div, lab, 0.02, 60, 1.0, 1e-6)
    rel = O.relative_residual(nx, ny, p, div, lab, 0.02)
    assert rel <= 1e-6 and it < 60
    _, it_cpu, _ = O.mg_fcg_solve(nx, ny, np.zeros(nx * ny), div, lab, 0.02, 60, 1e-6)
    assert abs(it - it_cpu) <= 1
    hist = sol.engine.solve_history()
    assert len(hist) == it + 1 and hist[-1] <= 1e-6
    assert hist[-1] == pytest.approx(rel, rel=1e-3)
    dc = HipPressureSolver(nx, ny, "mg-dc")
    it_dc = dc.solve(None, np.zeros(nx * ny), div, lab, 0.02, 60, 1.0, 1e-6)
    assert it < it_dc


def test_mg_fcg_steps_agree_with_mg_dc():
    """Whole Karman steps with either solver: both pressures meet 1e-6, so
    the velocities agree to the level that tolerance allows."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 256, 128
    out = {}
    for solver in ("mg-dc", "mg-fcg"):
        g, scene = S.karman_street(nx, ny)
        sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver=solver,
                                      pressure_iters=60, tol=1e-6, residual_every=0))
        sim.reset(scene)
        for _ in range(3):
            sim.step()
            assert sim.stepper.engine.solve_history()[-1] <= 1e-6
        out[solver] = (sim.f.u.copy(), sim.f.v.copy())
    assert rel_l2(out["mg-fcg"][0], out["mg-dc"][0]) < 1e-4
    assert rel_l2(out["mg-fcg"][1], out["mg-dc"][1]) < 1e-4


def test_mg_fcg_first_step_is_the_mg_dc_cycle():
    """The first flexible-CG step takes alpha = 1: the same bits as one
    defect-correction cycle from the same warm start."""
    from flowlab_amd.engine import HipPressureSolver
    nx, ny = 160, 96
    lab, div = boundary_fixture(nx, ny)
    p0 = O.lcg_fill(nx * ny, 5) * (lab == O.FLUID)
    out = []
    for solver in ("mg-dc", "mg-fcg"):
        p = p0.copy()
        assert HipPressureSolver(nx, ny, solver).solve(None, p, div, lab, 0.02, 1, 1.0, 0) == 1
        out.append(p)
    assert np.array_equal(out[0], out[1])
