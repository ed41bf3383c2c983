"""Field dumps (raw arrays + JSON header) and run logs: format round trips on
the CPU; a dump of a stepped GPU simulation against the engine's own reads."""
import json
import os

import numpy as np
import pytest

from flowlab_amd import fielddump as FD


def _arrays(nx, ny, ds=1, seed=0):
    rng = np.random.default_rng(seed)
    sh = FD.shapes(nx, ny, ds)
    return {
        "u": rng.standard_normal(int(np.prod(sh["u"]))),
        "v": rng.standard_normal(int(np.prod(sh["v"]))),
        "p": rng.standard_normal(int(np.prod(sh["p"]))),
        "label": rng.integers(0, 3, int(np.prod(sh["label"]))).astype(np.uint8),
        "dye": rng.random(int(np.prod(sh["dye"]))),
    }


@pytest.mark.parametrize("ds", [1, 2])
def test_round_trip_and_layout(tmp_path, ds):
    nx, ny = 7, 5
    a = _arrays(nx, ny, ds)
    types = {"u": "f32", "v": "f32", "p": "f64", "label": "u8", "dye": "f32"}
    prefix = str(tmp_path / "run" / "frame0")
    h = FD.write(prefix, {"nx": nx, "ny": ny, "h": 1 / ny, "dye_scale": ds, "time": 0.5},
                 a, types)
    assert h["format"] == "flowlab-field-dump" and h["version"] == 1
    # raw files: exactly count * itemsize bytes, little-endian values
    for name, d in h["fields"].items():
        path = tmp_path / "run" / d["file"]
        item = {"f32": 4, "f64": 8, "u8": 1}[d["dtype"]]
        assert os.path.getsize(path) == d["count"] * item
    raw_u = np.frombuffer((tmp_path / "run" / "frame0.u.f32").read_bytes(), "<f4")
    assert np.array_equal(raw_u, a["u"].astype(np.float32))
    header, out = FD.load(prefix)
    assert header["time"] == 0.5 and header["dye_scale"] == ds
    assert out["u"].shape == (ny, nx + 1) and out["v"].shape == (ny + 1, nx)
    assert out["dye"].shape == (3, ny * ds, nx * ds)
    assert np.array_equal(out["p"].ravel(), a["p"])  # f64 is exact
    assert np.array_equal(out["label"].ravel(), a["label"])
    assert np.array_equal(out["v"].ravel(), a["v"].astype(np.float32))
    # u(i, j) sits at row j, column i (grid.ts idxU = i + j (nx+1))
    i, j = 3, 2
    assert out["u"][j, i] == np.float32(a["u"][i + j * (nx + 1)])


def test_wrong_size_and_foreign_file_are_rejected(tmp_path):
    a = _arrays(4, 3)
    with pytest.raises(ValueError):
        FD.write(str(tmp_path / "x"), {"nx": 5, "ny": 3}, {"u": a["u"]}, {"u": "f32"})
    (tmp_path / "y.json").write_text(json.dumps({"format": "other"}))
    with pytest.raises(ValueError):
        FD.load(str(tmp_path / "y"))


@pytest.mark.gpu
def test_dump_of_a_stepped_simulation(tmp_path):
    from flowlab_amd import scenes as S
    from flowlab_amd.simulation import Simulation
    nx, ny = 128, 64
    g, scene = S.karman_street(nx, ny)
    sim = Simulation(nx, ny, dict(scheme="macCormack", cfl_target=0.5, solver="mg-fcg",
                                  pressure_iters=60, tol=1e-6, residual_every=0), readback=False)
    sim.reset(scene)
    log_path = str(tmp_path / "log.jsonl")
    with FD.RunLog(log_path) as log:
        for _ in range(3):
            sim.step()
            log.record(sim)
    prefix = str(tmp_path / "karman")
    FD.dump(prefix, sim)
    header, f = FD.load(prefix)
    e = sim.stepper.engine
    u, v = e.read_velocity()
    assert np.array_equal(f["u"].ravel(), u.astype(np.float32))  # device values, bit for bit
    assert np.array_equal(f["v"].ravel(), v.astype(np.float32))
    assert np.array_equal(f["p"].ravel(), e.read_pressure())
    assert header["steps"] == 3 and header["solver"]["name"] == "mg-fcg"
    hist = header["solver"]["residual_history"]
    assert len(hist) == header["solver"]["iterations"] + 1 and hist[-1] <= 1e-6
    recs = FD.read_log(log_path)
    assert [r["step"] for r in recs] == [1, 2, 3]
    assert all(r["residual_history"][-1] <= 1e-6 for r in recs)
    assert recs[-1]["time"] == pytest.approx(header["time"])
