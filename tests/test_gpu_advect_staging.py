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
