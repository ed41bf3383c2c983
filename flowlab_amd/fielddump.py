"""Field dumps and run logs, for offline comparison with other runs.

A dump is `<prefix>.json` plus one raw little-endian file per field,
`<prefix>.<name>.<f32|f64|u8>`, in the reference's host layouts
(src/core/grid.ts:84-141): p and label nx*ny with stride nx; u (nx+1)*ny with
stride nx+1; v nx*(ny+1) with stride nx; dye 3 planes of dnx*dny; x fastest.
Each file is one typed array, so any reader -- a run of the reference itself
under Node >= 22.6 included -- loads a field with a single view of the bytes,
e.g. `new Float32Array(fs.readFileSync(prefix + '.u.f32').buffer)`.  The
header records the grid, the step state, the solver's iteration count and
residual history, and for every field its file, element type and shape.

A run log is a JSON-lines file with one record per step: step index, time,
dt, pressure iterations, the relative-residual history of the solve
(`pressure.ts:102-161`) and, when timing is on, the per-phase device ms.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, is_dataclass

import numpy as np

FORMAT = "flowlab-field-dump"
VERSION = 1
_DTYPES = {"f32": np.dtype("<f4"), "f64": np.dtype("<f8"), "u8": np.dtype("u1")}


def shapes(nx: int, ny: int, dye_scale: int = 1) -> dict:
    """(rows, cols) of each field in its host layout."""
    dnx, dny = nx * dye_scale, ny * dye_scale
    return {"u": (ny, nx + 1), "v": (ny + 1, nx), "p": (ny, nx), "label": (ny, nx),
            "dye": (3, dny, dnx)}


def write(prefix: str, meta: dict, arrays: dict, dtypes: dict) -> dict:
    """Writes `arrays` (name -> flat numpy array) as raw files of the given
    element types (name -> 'f32' | 'f64' | 'u8') and the JSON header; meta
    must hold nx, ny and dye_scale.  Returns the header."""
    nx, ny, ds = int(meta["nx"]), int(meta["ny"]), int(meta.get("dye_scale", 1))
    shp = shapes(nx, ny, ds)
    os.makedirs(os.path.dirname(os.path.abspath(prefix)), exist_ok=True)
    fields = {}
    for name, a in arrays.items():
        if name not in shp:
            raise ValueError(f"unknown field {name!r}")
        t = dtypes[name]
        want = int(np.prod(shp[name]))
        a = np.asarray(a).ravel()
        if a.size != want:
            raise ValueError(f"{name}: {a.size} values, layout needs {want}")
        fn = f"{os.path.basename(prefix)}.{name}.{t}"
        a.astype(_DTYPES[t]).tofile(os.path.join(os.path.dirname(os.path.abspath(prefix)), fn))
        fields[name] = {"file": fn, "dtype": t, "shape": list(shp[name]), "count": want,
                        "order": "row-major, x fastest, little-endian"}
    header = {"format": FORMAT, "version": VERSION, **meta, "fields": fields}
    with open(prefix + ".json", "w") as f:
        json.dump(header, f, indent=1, default=_json_default)
    return header


def load(prefix: str):
    """(header, {name: array shaped as in `shapes`})."""
    with open(prefix + ".json") as f:
        header = json.load(f)
    if header.get("format") != FORMAT:
        raise ValueError(f"{prefix}.json is not a {FORMAT}")
    base = os.path.dirname(os.path.abspath(prefix))
    out = {}
    for name, d in header["fields"].items():
        a = np.fromfile(os.path.join(base, d["file"]), dtype=_DTYPES[d["dtype"]])
        if a.size != d["count"]:
            raise ValueError(f"{d['file']}: {a.size} values, header says {d['count']}")
        out[name] = a.reshape(d["shape"])
    return header, out


def dump(prefix: str, sim, fields=("u", "v", "p", "label", "dye"), precision: str = "f32") -> dict:
    """Dumps a `Simulation`'s current fields from the device: velocity and dye
    as `precision` ('f32' is what the device holds, bit for bit), the FP64
    pressure of the last solve as f64, labels as u8."""
    e = sim.stepper.engine
    e.synchronize()
    arrays, dtypes = {}, {}
    if "u" in fields or "v" in fields:
        u, v = e.read_velocity()
        if "u" in fields:
            arrays["u"], dtypes["u"] = u, precision
        if "v" in fields:
            arrays["v"], dtypes["v"] = v, precision
    if "p" in fields:
        arrays["p"], dtypes["p"] = e.read_pressure(), "f64"
    if "label" in fields:
        arrays["label"], dtypes["label"] = sim.f.label, "u8"
    if "dye" in fields:
        arrays["dye"], dtypes["dye"] = e.read_dye(), precision
    meta = {
        "nx": sim.g.nx, "ny": sim.g.ny, "h": sim.g.h, "dye_scale": e.ds,
        "time": sim.time, "steps": sim._steps, "dt": sim.dt,
        "params": sim.params,
        "solver": {"name": sim.params.solver, "iterations": sim.iters,
                   "residual_history": [float(x) for x in e.solve_history()]},
    }
    return write(prefix, meta, arrays, dtypes)


class RunLog:
    """Appends one JSON line per step: `log.record(sim)` after `sim.step()`."""

    def __init__(self, path: str):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        self._f = open(path, "a")

    def record(self, sim, **extra) -> dict:
        e = sim.stepper.engine
        rec = {"step": sim._steps, "time": sim.time, "dt": sim.dt, "iterations": sim.iters,
               "residual_history": [float(x) for x in e.solve_history()]}
        if getattr(e, "_timing", False):
            rec["phase_ms"] = e.timings()
        rec.update(extra)
        self._f.write(json.dumps(rec, default=_json_default) + "\n")
        self._f.flush()
        return rec

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_log(path: str) -> list:
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def _json_default(o):
    if is_dataclass(o):
        return asdict(o)
    if isinstance(o, (np.floating, np.integer)):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    if isinstance(o, tuple):
        return list(o)
    raise TypeError(f"not JSON-serialisable: {type(o).__name__}")
