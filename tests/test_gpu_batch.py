"""Batched scenes on the GPU (fh_batch_*): scenes of uneven size stepped
together, each phase one launch over a (scene, block) work list, with dt from
each scene's CFL.

Each scene's blocks take the indices and partials slots of that scene's own
single-grid launch, so a scene must equal the same scene stepped alone bit for
bit; both are held to the FP64 oracle at the per-kernel bar (relative L2 <=
1e-5 in FP32) with every pressure solve at 1e-6 by the oracle's own residual.
"""
import numpy as np
import pytest

from flowlab_amd import scenes as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-6


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_scenes():
    """Uneven sizes, closed boxes with obstacles and one open outlet."""
    return [S.decaying_turbulence(64, 64, 11), S.decaying_turbulence(128, 64, 12),
            S.karman_street(96, 48), S.decaying_turbulence(64, 128, 13),
            S.decaying_turbulence(160, 96, 14), S.karman_street(128, 64)]


def make_batch(scenes):
    from flowlab_amd.batch import HipBatch
    b = HipBatch([g for g, _ in scenes])
    fields = [b.reset(s, sc) for s, (_, sc) in enumerate(scenes)]
    return b, fields


def test_batch_equals_single_engine_bitwise():
    from flowlab_amd.simulation import Simulation
    scenes = small_scenes()
    b, _ = make_batch(scenes)
    sims = []
    for g, sc in scenes:
        sim = Simulation(g.nx, g.ny, dict(scheme="macCormack", cfl_target=0.5, dt_max=1.0,
                                         solver="mg-fcg", pressure_iters=60, tol=TOL,
                                         residual_every=0))
        sim.reset(sc)
        sims.append(sim)
    for n in range(3):
        print(f"step {n}", flush=True)
        b.step(cfl=0.5, dt_max=1.0, tol=TOL, max_cycles=60)
        dt, iters, res = b.results()
        for s, sim in enumerate(sims):
            sim.step()
            u, v, p, dye = b.read(s)
            assert dt[s] == sim.dt
            assert iters[s] == sim.iters, (s, iters[s], sim.iters)
            assert np.array_equal(u, sim.f.u) and np.array_equal(v, sim.f.v), s
            assert np.array_equal(p, sim.stepper.engine.read_pressure()), s
            assert res[s] <= TOL


@pytest.mark.parametrize("order", [1, 2])
def test_batch_extrapolated_start_equals_single_engine_bitwise(order):
    """Every scene's solve from its last solutions extrapolated in time:
    each scene still equals the same scene stepped alone with the same
    setting, bit for bit, over 7 steps, through a re-upload of two scenes
    after step 4 (their histories restart, the others' do not)."""
    from flowlab_amd.simulation import Simulation
    scenes = small_scenes()
    b, fields = make_batch(scenes)
    b.set_warm_extrapolation(order)
    sims = []
    for g, sc in scenes:
        sim = Simulation(g.nx, g.ny, dict(scheme="macCormack", cfl_target=0.5, dt_max=1.0,
                                         solver="mg-fcg", pressure_iters=60, tol=TOL,
                                         residual_every=0, warm_extrapolation=order))
        sim.reset(sc)
        sims.append(sim)
    total = 0
    for n in range(7):
        if n == 4:
            for s in (1, 4):
                u, v, p, dye = b.read(s)
                f = fields[s]
                f.u[:], f.v[:], f.p[:] = u, v, p
                for c in range(3):
                    f.dye[c][:] = dye[c]
                b.upload(s, f)
                sim = sims[s]
                sim.f.u[:], sim.f.v[:], sim.f.p[:] = u, v, p
                sim.stepper.engine.upload(u, v, p, sim.f.label, np.concatenate(dye))
        b.step(cfl=0.5, dt_max=1.0, tol=TOL, max_cycles=60)
        dt, iters, res = b.results()
        for s, sim in enumerate(sims):
            sim.step()
            u, v, p, _ = b.read(s)
            assert dt[s] == sim.dt, (n, s)
            assert iters[s] == sim.iters, (n, s, iters[s], sim.iters)
            assert np.array_equal(u, sim.f.u) and np.array_equal(v, sim.f.v), (n, s)
            assert np.array_equal(p, sim.stepper.engine.read_pressure()), (n, s)
            assert res[s] <= TOL
            total += int(iters[s])
    print("V-cycles over all scenes and steps:", total)


def test_batch_matches_oracle():
    scenes = small_scenes()[:4]
    b, _ = make_batch(scenes)
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=TOL,
                        cfl_target=0.5, dt_max=1.0, residual_every=0)
    osims = []
    for g, sc in scenes:
        o = O.OracleSim(g.nx, g.ny)
        o.reset(g, sc)
        osims.append(o)
    for n in range(2):
        b.step(cfl=0.5, dt_max=1.0, tol=TOL, max_cycles=60)
        dt, iters, res = b.results()
        for s, (g, _) in enumerate(scenes):
            osims[s].step(prm)
            u, v, p, _ = b.read(s)
            # dt = cfl h / max|u|: the device's FP32 velocity after a solve to
            # 1e-6 against the oracle's FP64 one, so the velocity's bar
            assert abs(dt[s] - prm.dt_used) <= 1e-5 * prm.dt_used
            # flexible CG against the oracle's defect correction: no more cycles
            assert iters[s] <= prm.iters + 1, (s, iters[s], prm.iters)
            assert rel_l2(u, osims[s].u) < 1e-5, s
            assert rel_l2(v, osims[s].v) < 1e-4, s
            assert res[s] <= TOL


def test_repeated_batches_give_the_same_bits():
    """Thread interleaving changes nothing: two batches of the same scenes
    agree field for field, step after step."""
    scenes = small_scenes()
    a, _ = make_batch(scenes)
    c, _ = make_batch(scenes)
    for n in range(3):
        a.step(tol=TOL)
        c.step(tol=TOL)
        ra, rc = a.results(), c.results()
        assert np.array_equal(ra[1], rc[1]) and np.array_equal(ra[0], rc[0])
        for s in range(len(scenes)):
            for x, y in zip(a.read(s)[:3], c.read(s)[:3]):
                assert np.array_equal(x, y), (n, s)


def test_tile_shapes_and_autotuner_do_not_change_results():
    from flowlab_amd.batch import tile_names
    scenes = small_scenes()[:2]
    ref = None
    for shape in tile_names():
        b, _ = make_batch(scenes)
        b.set_tile(shape)
        assert set(b.tiles()) == {shape}
        for _ in range(2):
            b.step(tol=TOL)
        out = [b.read(s)[:3] for s in range(len(scenes))]
        if ref is None:
            ref = out
        else:
            for x, y in zip(ref, out):
                for a_, b_ in zip(x, y):
                    assert np.array_equal(a_, b_), shape
    b, _ = make_batch(scenes)
    rep = b.tune(reps=2)
    assert len(rep["tiles"]) >= 1 and all(t in tile_names() for t in rep["tiles"])
    assert all(v > 0 for lvl in rep["ms"] for v in lvl.values())
    for _ in range(2):
        b.step(tol=TOL)
    for x, y in zip(ref, [b.read(s)[:3] for s in range(len(scenes))]):
        for a_, b_ in zip(x, y):
            assert np.array_equal(a_, b_)


def test_batch_rejects_bad_arguments():
    from flowlab_amd import hip
    from flowlab_amd.batch import HipBatch
    b = HipBatch([S.Grid(32, 32, 1 / 32)])
    with pytest.raises(hip.HipError):
        b.step(cfl=-1.0)
    with pytest.raises(hip.HipError):
        b.set_tile(99)


def heated(scene, cx=0.5, cy=0.3, sigma=0.15):
    """The scene plus a Gaussian heat blob (dye channel 0), smoke in channel 1
    and a uniform channel 2, so buoyancy and decay act."""
    g0, sc = scene

    def dye(g, d):
        x = (np.arange(g.nx) + 0.5) * g.h
        y = (np.arange(g.ny) + 0.5) * g.h
        r2 = (x[None, :] - cx * g.nx * g.h) ** 2 + (y[:, None] - cy) ** 2
        blob = np.exp(-r2 / (sigma * sigma)).ravel()
        d[0][:] = blob
        d[1][:] = 0.5 * blob
        d[2][:] = 0.25
    return g0, S.SceneSpec(labels=sc.labels, seed=sc.seed, dye=dye)


# (buoyancy, confinement, per-channel decay) per scene: buoyancy alone (fused
# into the MacCormack face pass), buoyancy with confinement (the vorticity and
# force passes), confinement with decay, decay alone, buoyancy on an open outlet
FORCES = [(2.0, 0.0, (0.0, 0.0, 0.0)), (2.0, 0.3, (0.0, 0.0, 0.0)),
          (0.0, 0.3, (0.5, 1.0, 0.0)), (0.0, 0.0, (0.2, 0.2, 0.2)),
          (1.0, 0.0, (0.0, 0.3, 0.0))]


@pytest.mark.parametrize("scheme", ["macCormack", "semiLagrangian"])
def test_batch_forces_equal_single_engine_bitwise(scheme):
    """Buoyancy, vorticity confinement and dye decay set per scene step each
    scene exactly as a single-grid engine with the same parameters does."""
    from flowlab_amd.simulation import Simulation
    scenes = [heated(S.decaying_turbulence(64, 64, 21)),
              heated(S.decaying_turbulence(128, 64, 22)),
              heated(S.decaying_turbulence(64, 96, 23)),
              heated(S.decaying_turbulence(96, 64, 24)),
              heated(S.karman_street(128, 64), cx=0.25, cy=0.5)]
    b, _ = make_batch(scenes)
    sims = []
    for s, ((g, sc), (buoy, conf, decay)) in enumerate(zip(scenes, FORCES)):
        b.set_forces(s, buoyancy=buoy, confinement=conf, decay=decay)
        sim = Simulation(g.nx, g.ny, dict(scheme=scheme, cfl_target=0.5, dt_max=1.0,
                                         solver="mg-fcg", pressure_iters=60, tol=TOL,
                                         residual_every=0, buoyancy=buoy, confinement=conf,
                                         dye_decay=list(decay)))
        sim.reset(sc)
        sims.append(sim)
    for n in range(3):
        print(f"{scheme} step {n}", flush=True)
        b.step(scheme=scheme, cfl=0.5, dt_max=1.0, tol=TOL, max_cycles=60)
        dt, iters, res = b.results()
        for s, sim in enumerate(sims):
            sim.step()
            u, v, p, dye = b.read(s)
            sdye = sim.read_dye()
            assert dt[s] == sim.dt, s
            assert iters[s] == sim.iters, (s, iters[s], sim.iters)
            assert np.array_equal(u, sim.f.u) and np.array_equal(v, sim.f.v), s
            assert np.array_equal(p, sim.stepper.engine.read_pressure()), s
            for c in range(3):
                assert np.array_equal(dye[c], sdye[c]), (s, c)
            assert res[s] <= TOL
    # the forces acted: a heated scene's velocity differs from the unforced one
    plain, _ = make_batch(scenes[:1])
    for n in range(3):
        plain.step(scheme=scheme, cfl=0.5, dt_max=1.0, tol=TOL, max_cycles=60)
    assert not np.array_equal(plain.read(0)[1], b.read(0)[1])
