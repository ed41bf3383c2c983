"""GPU: the warm start extrapolated in time (fh_set_warm_extrapolation).

Each solve of fh_step may start from the last solutions extrapolated to the
step's time by Lagrange interpolation over the steps' own dt values, instead
of from the previous pressure (the reference's start, stepGpu.ts:19-24).
The solve still runs to its tolerance, so only the start and the cycle
count change.  Checked here, step by step from the device's own state:

* the start the device solved from: its initial relative residual equals the
  oracle's residual of the host's extrapolation of the device's earlier
  pressures (1e-9 relative), through a history restart mid-run;
* the step: u and v against one oracle step (FP64 defect correction to 1e-6
  from that same start) at the per-kernel bar, with the solve at 1e-6 by the
  oracle's own residual.
"""
import numpy as np
import pytest

from flowlab_amd import scenes as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def lagrange(order, tau, a, b):
    """Weights of p_n, p_{n-1}, p_{n-2} (nodes 0, -a, -a-b) at time tau."""
    if order == 0:
        return (1.0, 0.0, 0.0)
    if order == 1:
        return (1.0 + tau / a, -tau / a, 0.0)
    return ((tau + a) * (tau + a + b) / (a * (a + b)), -tau * (tau + a + b) / (a * b),
            tau * (tau + a) / (b * (a + b)))


def test_lagrange_weights_are_exact_on_polynomials():
    rng = np.random.default_rng(3)
    for order in (1, 2):
        for _ in range(20):
            tau, a, b = rng.uniform(0.2, 2.0, 3)
            c = rng.normal(size=order + 1)
            f = lambda t: sum(c[k] * t ** k for k in range(order + 1))
            w = lagrange(order, tau, a, b)
            got = w[0] * f(0.0) + w[1] * f(-a) + w[2] * f(-a - b)
            assert abs(got - f(tau)) < 1e-12 * (1 + abs(f(tau)))


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("solver", ["mg-dc", "mg-fcg"])
def test_extrapolated_start_every_step_from_device_state(order, solver):
    """Karman street 256x128 at CFL 0.5 (dt changes every step), 14 steps,
    the history restarted by a re-upload after step 7."""
    from flowlab_amd.simulation import Simulation
    nx, ny = 256, 128
    g, scene = S.karman_street(nx, ny)
    sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver=solver,
                                  pressure_iters=60, tol=1e-6, residual_every=0,
                                  warm_extrapolation=order))
    sim.reset(scene)
    e = sim.stepper.engine
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    hist = []  # (p, dt) of the solves since the last restart, newest last
    prev_dt = None
    worst, dts = (0.0, 0.0, 0.0), set()
    for n in range(14):
        if n == 8:
            # re-upload the device's own state: the history restarts
            sim.f.p[:] = e.read_pressure()
            sim.stepper.invalidate()
            hist = []
        osim.u[:] = sim.f.u
        osim.v[:] = sim.f.v
        osim.dye[:] = np.concatenate(sim.read_dye())
        p_now = e.read_pressure()
        sim.step()
        dt = sim.dt
        dts.add(dt)
        # the device's start: the current p (a step's solution unless the
        # history just restarted) extrapolated over the solutions before it
        if hist:
            k = min(len(hist) - 1, order)
            a = hist[-1][1]
            b = hist[-2][1] if len(hist) >= 2 else 1.0
            w = lagrange(k, dt, a, b)
            start = w[0] * hist[-1][0]
            if k >= 1:
                start = start + w[1] * hist[-2][0]
            if k >= 2:
                start = start + w[2] * hist[-3][0]
        else:
            start = p_now
        osim.p[:] = start
        prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=1e-6,
                            cfl_target=0.5, residual_every=0)
        osim.step(prm)
        assert prm.dt_used == dt, (n, prm.dt_used, dt)
        div = e.read_divergence()
        scale = g.h * g.h / dt
        r_dev = float(e.solve_history()[0])
        r_host = O.relative_residual(nx, ny, start, div, sim.f.label, scale)
        p = e.read_pressure()
        err = (abs(r_dev - r_host) / r_host, rel_l2(sim.f.u, osim.u), rel_l2(sim.f.v, osim.v))
        worst = tuple(max(x, y) for x, y in zip(worst, err))
        assert O.relative_residual(nx, ny, p, div, sim.f.label, scale) <= 1e-6
        hist.append((p, dt))
        prev_dt = dt
    print(f"order {order} {solver}: worst start residual {worst[0]:.2e}, "
          f"u {worst[1]:.2e}, v {worst[2]:.2e} over {len(dts)} distinct dt")
    assert prev_dt is not None and len(dts) > 3
    assert worst[0] < 1e-9, worst
    assert worst[1] < REL_L2 and worst[2] < REL_L2, worst


def test_extrapolated_start_cuts_cycles_on_a_plume():
    """The 8192^2 plume's scene at 512^2 (radii scaled), 40 steps from rest:
    the linear extrapolation needs fewer V-cycles in all than the previous
    pressure, every solve at 1e-6."""
    from flowlab_amd.simulation import Simulation
    n = 512
    g, scene = S.large_plume(n, cell_scale=n / 8192)
    total = {}
    for order in (0, 1):
        sim = Simulation(n, n, dict(scheme="macCormack", cfl_target=0.5, dt_max=0.5 * g.h,
                                    solver="mg-fcg", pressure_iters=60, tol=1e-6,
                                    residual_every=0, buoyancy=1.0, warm_extrapolation=order),
                         readback=False)
        sim.reset(scene)
        its = []
        for _ in range(40):
            sim.step()
            its.append(sim.iters)
            assert sim.stepper.engine.solve_history()[-1] <= 1e-6
        total[order] = sum(its)
    print("V-cycles over 40 steps:", total)
    assert total[1] < total[0], total
