"""LDS-staged advection kernels against the plain global-gather kernels.

The staged kernels read each bilinear stencil from a workgroup's LDS window
when it lies inside it and from global memory otherwise, so their outputs
must equal the plain kernels' bit for bit at any CFL number: below the halo
(only LDS reads), and well beyond it (the global fallback).  The grids are
not multiples of the 64x16 tile, so partial tiles at the edges are covered;
dye scale 2 runs the plain dye kernels either way."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(nx, ny, ds, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, (nx + 1) * ny)
    v = rng.uniform(-1, 1, nx * (ny + 1))
    lbl = rng.choice(np.array([0, 1, 2], np.uint8), nx * ny, p=[0.8, 0.1, 0.1])
    q = [rng.uniform(0, 1, nx * ds * ny * ds) for _ in range(3)]
    return u, v, lbl, q


def _advect(adv, staged, scheme, u, v, lbl, q, dt):
    adv.engine.set_advect_staging(staged)
    uo, vo = np.empty_like(u), np.empty_like(v)
    adv.velocity(None, scheme, u, v, uo, vo, lbl, dt)
    qo = [np.empty_like(c) for c in q]
    adv.dye(None, scheme, u, v, q, qo, lbl, dt)
    return uo, vo, qo


@pytest.mark.parametrize("scheme", ["semiLagrangian", "macCormack"])
@pytest.mark.parametrize("cfl", [0.4, 1.2, 3.7])
@pytest.mark.parametrize("nx,ny,ds", [(300, 170, 1), (64, 16, 1), (130, 70, 2)])
def test_staged_advection_matches_global_gathers(scheme, cfl, nx, ny, ds):
    from flowlab_amd.engine import HipAdvector
    adv = HipAdvector(nx, ny, dye_scale=ds)
    u, v, lbl, q = _fields(nx, ny, ds, seed=nx * 7 + ny)
    dt = cfl * adv.engine.h  # speeds are at most 1, so c = dt/h = cfl
    a = _advect(adv, False, scheme, u, v, lbl, q, dt)
    b = _advect(adv, True, scheme, u, v, lbl, q, dt)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for ca, cb in zip(a[2], b[2]):
        assert np.array_equal(ca, cb)
    # the step moved something (not a vacuous comparison)
    assert not np.array_equal(a[0], u)


@pytest.mark.parametrize("nx,ny", [(300, 170), (64, 16), (517, 263)])
@pytest.mark.parametrize("fast_band", [False, True])
def test_one_pass_maccormack_matches_two_passes(nx, ny, fast_band):
    """MacCormack velocity and dye, each in one kernel
    (fh_set_faces_one_pass, fh_set_dye_one_pass), against their two
    LDS-staged passes and against the plain global gathers: the same bits.
    With a band of fast flow, the tiles whose windows hold it exceed the
    staged reach, so the one-pass kernel marks its launch and the two-pass
    kernels redo every tile after it; the other tiles pass."""
    from flowlab_amd.engine import HipAdvector
    adv = HipAdvector(nx, ny)
    u, v, lbl, q = _fields(nx, ny, 1, seed=nx + 3 * ny)
    if fast_band:
        u.reshape(ny, nx + 1)[ny // 3:ny // 3 + 3] *= 4.0
    dt = 0.9 * adv.engine.h
    outs = []
    for staged, one in ((False, False), (True, False), (True, True)):
        adv.engine.set_advect_staging(staged)
        adv.engine.set_dye_one_pass(one)
        adv.engine.set_faces_one_pass(one)
        outs.append(_advect(adv, staged, "macCormack", u, v, lbl, q, dt))
    for k in (1, 2):
        for x, y in zip(outs[0][:2] + tuple(outs[0][2]), outs[k][:2] + tuple(outs[k][2])):
            assert np.array_equal(x, y), (k, np.flatnonzero(x != y)[:5])
    assert not np.array_equal(outs[0][0], u) and not np.array_equal(outs[0][2][0], q[0])


def _steps(scene_name, one, steps):
    from flowlab_amd import scenes as S
    from flowlab_amd.simulation import Simulation
    if scene_name == "karman":
        nx, ny = 330, 170
        g, scene = S.karman_street(nx, ny)
        prm = dict(scheme="macCormack", cfl_target=0.5, solver="mg-dc", pressure_iters=60,
                   tol=1e-6, residual_every=0)
    else:
        nx = ny = 330
        g, scene = S.large_plume(nx, cell_scale=nx / 8192)
        prm = dict(scheme="macCormack", cfl_target=0.5, dt_max=4.0 * g.h, solver="mg-dc",
                   pressure_iters=60, tol=1e-6, residual_every=0, buoyancy=1.0)
    sim = Simulation(nx, ny, prm)
    sim.reset(scene)
    if scene_name == "karman":
        rng = np.random.default_rng(17)
        for c in range(3):
            sim.f.dye[c][:] = rng.uniform(0, 1, nx * ny)
        sim.stepper.invalidate()
    dye0 = np.concatenate([d.copy() for d in sim.f.dye])
    sim.stepper.engine.set_dye_one_pass(one)
    sim.stepper.engine.set_faces_one_pass(one)
    dts = []
    for _ in range(steps):
        sim.step()
        dts.append(sim.dt)
    e = sim.stepper.engine
    return (sim.f.u.copy(), sim.f.v.copy(), e.read_pressure(), np.concatenate(sim.read_dye()),
            dts, dye0)


@pytest.mark.parametrize("scene", ["karman", "plume"])
def test_one_pass_steps_match_two_passes(scene):
    """Twelve whole steps with the MacCormack velocity and dye in one pass
    each or in two: u, v, p, dye and every dt the same bits.  The Karman
    street at 330x170 with a random dye field runs at CFL 0.5, so every dt
    comes from the max face speed that the dye pass reports; the plume at
    330x330 applies buoyancy in the velocity pass."""
    a = _steps(scene, False, 12)
    b = _steps(scene, True, 12)
    assert a[4] == b[4]
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[3], a[5])
    if scene == "karman":
        assert len(set(a[4])) > 3, a[4]  # the CFL bound moved dt
