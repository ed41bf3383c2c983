"""Tile shapes of the LDS-tiled multigrid smoothers and their autotuner.

Temporal blocking computes exactly the global red-black sweep inside every
tile, so the tile shape must not change a single bit of a step; the tuner
only picks among shapes by their measured time."""
import numpy as np
import pytest

from flowlab_amd import scenes as S

pytestmark = pytest.mark.gpu


def _names():
    from flowlab_amd import hip
    L = hip.lib()
    return [L.fh_tile_shape_name(t).decode() for t in range(L.fh_num_tile_shapes())]


def _run(shape=None, tune=False, steps=2):
    from flowlab_amd.simulation import Simulation
    nx, ny = 512, 256
    g, scene = S.karman_street(nx, ny)
    sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver="mg-dc",
                                  pressure_iters=60, tol=1e-6, residual_every=0))
    sim.reset(scene)
    e = sim.stepper.engine
    if shape is not None:
        e.set_tile(shape)
    sim.step()
    rep = e.tune_tiles(reps=2) if tune else None
    for _ in range(steps - 1):
        sim.step()
    return sim.f.u.copy(), sim.f.v.copy(), e.read_pressure(), rep, e


def test_every_tile_shape_gives_the_same_bits():
    ref = None
    names = None
    for shape in range(len(_names())):
        u, v, p, _, e = _run(shape)
        names = e.tile_names()
        assert set(e.tiles()) == {names[shape]}
        if ref is None:
            ref = (u, v, p)
        else:
            assert np.array_equal(u, ref[0]) and np.array_equal(v, ref[1]), names[shape]
            assert np.array_equal(p, ref[2]), names[shape]


def test_tuner_picks_valid_shapes_and_changes_nothing():
    u0, v0, p0, _, _ = _run(0)
    u, v, p, rep, e = _run(None, tune=True)
    names = e.tile_names()
    assert len(rep["tiles"]) >= 1 and all(t in names for t in rep["tiles"])
    assert all(ms > 0 for lvl in rep["ms"] for ms in lvl.values())
    assert np.array_equal(u, u0) and np.array_equal(v, v0) and np.array_equal(p, p0)


def test_coarse_workgroup_sizes_give_the_same_bits():
    from flowlab_amd.simulation import Simulation
    out = []
    for group in range(3):
        nx, ny = 256, 128
        g, scene = S.karman_street(nx, ny)
        sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver="mg-dc",
                                      pressure_iters=60, tol=1e-6, residual_every=0))
        sim.reset(scene)
        sim.stepper.engine.set_coarse_group(group)
        for _ in range(2):
            sim.step()
        out.append((sim.f.u.copy(), sim.stepper.engine.read_pressure()))
    for u, p in out[1:]:
        assert np.array_equal(u, out[0][0]) and np.array_equal(p, out[0][1])
