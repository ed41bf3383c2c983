"""The N-API boundary: the addon a TypeScript host binds.

CPU: the addon loads in this image's Node and exports every entry point the
JS seams use.  GPU: a scene stepped through HipStepper from Node matches the
FP64 oracle with the fused WebGPU step's semantics (stepGpu.ts:393-604): the
velocity readback runs one step behind, invalidate() drops an in-flight
readback, setters made while one is in flight apply to the next step, the
profiler gets the readback/upload/solve marks and the device timings are
reported; an overlapping step() rejects with the reference's re-entrancy
message (stepGpu.ts:402); arrays of the wrong type or length are rejected
before any library call; Advector.dye takes any number of channels."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "flowlab_amd", "node")
DRIVER = os.path.join(ROOT, "tests", "node", "run_steps.js")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(
    os.path.join(NODE_DIR, "flowhip.node")), reason="node or the built addon is absent")

EXPORTS = {"deviceCount", "create", "upload", "setLabels", "setDyePatch", "setBuoyancy",
           "setConfinement", "setStir", "setTiming", "getTimings", "maxFaceSpeed", "step",
           "solvePressure", "readVelocity", "readDye", "readPressure", "advectVelocity",
           "advectDye", "setWarmExtrapolation"}


def test_addon_loads_and_exports():
    out = subprocess.run([NODE, "-e", "const m=require('./flowlab_amd/node/flowhip.node');"
                          "console.log(JSON.stringify(Object.keys(m)));"
                          "const h=require('./flowlab_amd/node/hipEngine.js');"
                          "console.log(JSON.stringify(Object.keys(h)))"],
                         cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert set(json.loads(lines[0])) == EXPORTS
    assert {"HipStepper", "HipPressureSolver", "HipAdvector"} <= set(json.loads(lines[1]))


def test_addon_rejects_non_engine_handles():
    """No GPU needed: a call with something that is not an engine handle
    throws TypeError before reaching the library."""
    out = subprocess.run([NODE, "-e", "const m=require('./flowlab_amd/node/flowhip.node');"
                          "try { m.setBuoyancy({}, 1); } catch (e) { console.log(e.name); }"],
                         cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "TypeError"


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("f32", [False, True])
def test_node_stepper_matches_oracle(tmp_path, f32):
    from flowlab_amd import scenes as S
    from oracle import pyoracle as O
    nx, ny, steps = 128, 64, 3
    g, scene = S.karman_street(nx, ny)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    dt = 0.5 * g.h
    u0 = osim.u.copy()
    for name, arr in (("u", osim.u), ("v", osim.v), ("p", osim.p), ("dye", osim.dye)):
        np.asarray(arr, "<f8").tofile(tmp_path / f"{name}.f64")
    osim.label.tofile(tmp_path / "label.u8")
    case = dict(nx=nx, ny=ny, h=g.h, dyeScale=1, scheme="macCormack", solver="mg-dc",
                iterations=60, omega=1.0, tol=1e-6, steps=steps, dt=dt, buoyancy=0.0,
                confinement=0.0, keep=[1, 1, 1], patch=None, f32=f32)
    (tmp_path / "case.json").write_text(json.dumps(case))
    r = subprocess.run([NODE, DRIVER, str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads((tmp_path / "result.json").read_text())
    err = res["errors"]
    assert err["reentrant"] == "Error: HipStepper.step() is not re-entrant"
    assert err["shortArray"].startswith("TypeError: readVelocity u: expected")
    assert err["wrongType"].startswith("TypeError: flowhip.readDye: expected a Float64Array")
    assert err["notEngine"].startswith("TypeError")
    assert err["uploadType"].startswith("TypeError: upload u: expected a Float64Array")
    assert err["dyeGrid"].startswith("Error: HipAdvector.dye: dye grid 3x3")
    assert res["dyeChannels"] == {"one": True, "four": True}
    # perf marks of every step, device timings of the drained steps
    assert res["marks"] == [["readback", "upload", "solve"]] * steps
    assert res["timings"]["device"] > 0 and res["timings"]["upload"] >= 0
    ms = res["deviceMs"]
    assert len(ms) == 7 and ms[6] > 0 and abs(sum(ms[:6]) - ms[6]) < 0.05 * ms[6] + 1e-3

    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=1e-6,
                        cfl_target=1e30, dt_max=dt, residual_every=0)
    us, vs = [], []
    for _ in range(steps):
        osim.step(prm)
        us.append(osim.u.copy())
        vs.append(osim.v.copy())
    tol_u = 1e-5
    # one step stale: after step k returned, f.u held step k-1's velocity
    stale0 = np.fromfile(tmp_path / "stale_u_0.f64", "<f8")
    assert np.array_equal(stale0, np.asarray(u0, np.float32 if f32 else np.float64))
    for k in range(1, steps):
        sk = np.fromfile(tmp_path / f"stale_u_{k}.f64", "<f8")
        assert rel(sk, us[k - 1]) < tol_u, k
    u = np.fromfile(tmp_path / "out_u.f64", "<f8")
    v = np.fromfile(tmp_path / "out_v.f64", "<f8")
    assert rel(u, us[-1]) < tol_u
    assert rel(v, vs[-1]) < 1e-4
    assert all(abs(a - prm.iters) <= 1 for a in res["iters"][-1:])
    # invalidate() with a readback in flight: the stale values were dropped
    # (f.u still the host's after the next step returned) and that step ran
    # from the host's fields, the same bits as the first step from them
    assert res["inflightBeforeInvalidate"] is True
    assert res["uAfterInvalidateIsHost"] is True
    inv_u = np.fromfile(tmp_path / "inv_u.f64", "<f8")
    stale1 = np.fromfile(tmp_path / "stale_u_1.f64", "<f8")
    assert np.array_equal(inv_u, stale1)
    assert rel(inv_u, us[0]) < tol_u


@pytest.mark.gpu
def test_node_stepper_extrapolated_start_matches_python_engine(tmp_path):
    """HipStepper({warmExtrapolation: 1}) from Node and the Python engine with
    the same setting step the same scene through the same library: the same
    V-cycles and the same velocity bits after 6 steps, and fewer V-cycles in
    all than the previous-pressure start."""
    from flowlab_amd import scenes as S
    from flowlab_amd.engine import HipEngine
    from oracle import pyoracle as O
    nx, ny, steps = 128, 64, 6
    g, scene = S.karman_street(nx, ny)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    dt = 0.5 * g.h
    for name, arr in (("u", osim.u), ("v", osim.v), ("p", osim.p), ("dye", osim.dye)):
        np.asarray(arr, "<f8").tofile(tmp_path / f"{name}.f64")
    osim.label.tofile(tmp_path / "label.u8")
    case = dict(nx=nx, ny=ny, h=g.h, dyeScale=1, scheme="macCormack", solver="mg-fcg",
                iterations=60, omega=1.0, tol=1e-6, steps=steps, dt=dt, buoyancy=0.0,
                confinement=0.0, keep=[1, 1, 1], patch=None, f32=False, warm=1)
    (tmp_path / "case.json").write_text(json.dumps(case))
    r = subprocess.run([NODE, DRIVER, str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads((tmp_path / "result.json").read_text())
    its = {}
    for order in (1, 0):
        e = HipEngine(nx, ny, 1, g.h)
        e.upload(osim.u, osim.v, osim.p, osim.label, osim.dye)
        e.set_warm_extrapolation(order)
        its[order] = [e.step("macCormack", dt, g.h * g.h / dt, dt / g.h, (1.0, 1.0, 1.0),
                             "mg-fcg", 60, 1.0, 1e-6) for _ in range(steps)]
        if order == 1:
            u, v = e.read_velocity()
        e.close()
    assert res["iters"] == its[1], (res["iters"], its)
    assert np.array_equal(np.fromfile(tmp_path / "out_u.f64", "<f8"), u)
    assert np.array_equal(np.fromfile(tmp_path / "out_v.f64", "<f8"), v)
    assert sum(its[1]) < sum(its[0]), its

