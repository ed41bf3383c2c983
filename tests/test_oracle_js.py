"""The plain-JavaScript FP64 CPU step (oracle/js/cpuStep.js, the CPU baseline
run on the GPU host's Node) against the C oracle: whole steps of scenes with
solids, an Air outlet, buoyancy, vorticity confinement and a dye inlet, with
each pressure solver, agree bit for bit.  Both restate the reference's CPU
path (src/core/simulation.ts:467-587) operation for operation, so any dropped
term, wrong index or reordered update shows as a difference."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from flowlab_amd import scenes as S
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "oracle", "js", "cpuStep.js")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is absent")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def scene(nx, ny):
    g = S.Grid(nx, ny, 1.0 / ny)

    def labels(gg, label):
        L = label.reshape(gg.ny, gg.nx)
        L[:, gg.nx - 1] = O.AIR
        L[gg.ny // 3:gg.ny // 2, gg.nx // 4:gg.nx // 3] = O.SOLID

    def seed(gg, u, v):
        S.add_vortex_cluster(gg, u, v, amp=0.8, seed=4321)

    return g, S.SceneSpec(labels=labels, seed=seed,
                          dye_source=S.heat_disk_source(0.35, 0.2, 0.08))


@pytest.mark.parametrize("solver,scheme,iters,tol", [
    ("mg-dc", "macCormack", 40, 1e-8),
    ("mg", "macCormack", 6, 0.0),
    ("rbsor", "semiLagrangian", 60, 1e-3),
    ("sor", "macCormack", 40, 0.0),
])
def test_js_step_matches_c_oracle(tmp_path, solver, scheme, iters, tol):
    nx, ny, steps = 48, 32, 3
    g, sc = scene(nx, ny)
    osim = O.OracleSim(nx, ny)
    osim.reset(g, sc)
    pt = osim.patch
    for name, arr in (("u", osim.u), ("v", osim.v), ("p", osim.p), ("dye", osim.dye)):
        np.asarray(arr, "<f8").tofile(tmp_path / f"{name}.f64")
    osim.label.tofile(tmp_path / "label.u8")
    np.asarray(pt.data, "<f4").tofile(tmp_path / "patch.f32")
    omega = 1.7
    case = dict(nx=nx, ny=ny, h=g.h, dyeScale=1, scheme=scheme, solver=solver,
                pressureIters=iters, omega=omega, tol=tol, cflTarget=1.5, dtMax=1 / 30,
                buoyancy=1.3, confinement=0.2, decay=[0, 0, 0],
                patch=dict(i0=pt.i0, j0=pt.j0, nx=pt.nx, ny=pt.ny))
    (tmp_path / "case.json").write_text(json.dumps(case))
    r = subprocess.run([NODE, SCRIPT, str(tmp_path), "0", str(steps)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    prm = O.step_params(scheme=scheme, solver=solver, pressure_iters=iters, omega=omega, tol=tol,
                        cfl_target=1.5, dt_max=1 / 30, buoyancy=1.3, confinement=0.2,
                        residual_every=0, patch=(pt.i0, pt.j0, pt.nx, pt.ny))
    its = []
    for _ in range(steps):
        osim.step(prm, pt.data)
        its.append(prm.iters)
    assert res["iters"] == its
    for name, ref in (("u", osim.u), ("v", osim.v), ("p", osim.p), ("dye", osim.dye)):
        got = np.fromfile(tmp_path / f"out_{name}.f64", "<f8")
        assert np.abs(ref).max() > 1e-6, name
        # the same operations in the same order: equal bit for bit
        assert np.array_equal(got, ref), (name, rel(got, ref))
