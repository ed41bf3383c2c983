"""The N-API boundary: the addon a TypeScript host binds.

CPU: the addon loads in this image's Node and exports every entry point the
JS seams use.  GPU: a scene stepped through HipStepper from Node matches the
FP64 oracle, and an overlapping step() rejects with the reference's
re-entrancy message (src/gpu/stepGpu.ts:402)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "flowlab_amd", "node")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(
    os.path.join(NODE_DIR, "flowhip.node")), reason="node or the built addon is absent")

EXPORTS = {"deviceCount", "create", "upload", "setLabels", "setDyePatch", "setBuoyancy",
           "setConfinement", "setStir", "maxFaceSpeed", "step", "solvePressure", "readVelocity",
           "readDye", "readPressure", "advectVelocity", "advectDye"}


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


@pytest.mark.gpu
def test_node_stepper_matches_oracle(tmp_path):
    from flowlab_amd import scenes as S
    from oracle import pyoracle as O
    nx, ny = 128, 64
    g, scene = S.karman_street(nx, ny)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, scene)
    dt = 0.5 * g.h
    for name, arr in (("u", osim.u), ("v", osim.v), ("p", osim.p), ("dye", osim.dye)):
        np.asarray(arr, "<f8").tofile(tmp_path / f"{name}.f64")
    osim.label.tofile(tmp_path / "label.u8")
    case = dict(nx=nx, ny=ny, h=g.h, dyeScale=1, scheme="macCormack", solver="mg-dc",
                iterations=60, omega=1.0, tol=1e-6, steps=3, dt=dt, buoyancy=0.0,
                confinement=0.0, keep=[1, 1, 1], patch=None)
    (tmp_path / "case.json").write_text(json.dumps(case))
    r = subprocess.run([NODE, os.path.join(NODE_DIR, "run_steps.js"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads((tmp_path / "result.json").read_text())
    assert res["reentrant"] == "HipStepper.step() is not re-entrant"
    prm = O.step_params(scheme="macCormack", solver="mg-dc", pressure_iters=60, tol=1e-6,
                        cfl_target=1e30, dt_max=dt, residual_every=0)
    for _ in range(3):
        osim.step(prm)
    u = np.fromfile(tmp_path / "out_u.f64", "<f8")
    v = np.fromfile(tmp_path / "out_v.f64", "<f8")
    assert np.linalg.norm(u - osim.u) / np.linalg.norm(osim.u) < 1e-5
    assert np.linalg.norm(v - osim.v) / max(np.linalg.norm(osim.v), 1e-30) < 1e-4
    assert all(abs(a - prm.iters) <= 1 for a in res["iters"][-1:])
