'use strict';
/**
 * Times the fused step through the N-API HipStepper, as a TypeScript host
 * drives it (the HIP counterpart of src/bench/step.ts): host upload after
 * invalidate(), the one-step-stale velocity readback into the host fields
 * every step, and the profiler marks, all inside the timed loop.
 *
 *   node bench_step.js <dir> <warmup> <steps>
 *
 * <dir> holds a scene as written by the Python side (case.json with nx, ny,
 * h, scheme, solver, tol, cfl, dtMax, warm (the extrapolation order of the
 * solve's start, 0 = the previous pressure); u.f64 v.f64 p.f64 label.u8
 * dye.f64).
 * dt follows Simulation.step (simulation.ts:474-479) from the host's
 * one-step-stale velocity, as the reference's fused path computes it.
 * Prints one JSON line.
 */
const fs = require('fs');
const os = require('os');
const path = require('path');
const { performance } = require('perf_hooks');
const { HipStepper, PHASES } = require('./hipEngine.js');

function readF64(p) {
  const b = fs.readFileSync(p);
  return new Float64Array(b.buffer, b.byteOffset, b.length / 8).slice();
}

function maxFaceSpeed(f) {  // simulation.ts:451-459, on the host fields
  let m = 0;
  for (let k = 0; k < f.u.length; k++) m = Math.max(m, Math.abs(f.u[k]));
  for (let k = 0; k < f.v.length; k++) m = Math.max(m, Math.abs(f.v[k]));
  return m;
}

class Phases {  // profiler.ts's mark/begin surface, summing each phase
  constructor() { this.sum = {}; this.t = 0; }
  begin() { this.t = performance.now(); }
  mark(label) {
    const now = performance.now();
    this.sum[label] = (this.sum[label] || 0) + (now - this.t);
    this.t = now;
  }
}

async function main() {
  const [dir, warmupArg, stepsArg] = process.argv.slice(2);
  const warmup = Number(warmupArg), steps = Number(stepsArg);
  const c = JSON.parse(fs.readFileSync(path.join(dir, 'case.json'), 'utf8'));
  const g = { nx: c.nx, ny: c.ny, h: c.h };
  const dye = readF64(path.join(dir, 'dye.f64'));
  const n = dye.length / 3;
  const f = {
    u: readF64(path.join(dir, 'u.f64')), v: readF64(path.join(dir, 'v.f64')),
    p: readF64(path.join(dir, 'p.f64')),
    label: new Uint8Array(fs.readFileSync(path.join(dir, 'label.u8'))),
    dye: [0, 1, 2].map((k) => dye.slice(k * n, (k + 1) * n)),
  };
  const st = new HipStepper(g, 1, { solver: c.solver, iterations: 100, tol: c.tol,
                                   warmExtrapolation: c.warm || 0 });
  const keep = [1, 1, 1];
  const perf = new Phases();
  const one = async () => {
    perf.begin();
    const uMax = maxFaceSpeed(f);
    const dt = uMax > 1e-9 ? Math.min(c.dtMax, (c.cfl * g.h) / uMax) : c.dtMax;
    perf.mark('cfl');
    return st.step(f, c.scheme, dt, (g.h * g.h) / dt, dt / g.h, keep, perf);
  };
  for (let s = 0; s < warmup; s++) await one();
  await st.drainReadback(f);
  perf.sum = {};
  const iters = [];
  const t0 = performance.now();
  for (let s = 0; s < steps; s++) iters.push(await one());
  await st.drainReadback(f);
  const ms = (performance.now() - t0) / steps;
  const host = {};
  for (const [k, v] of Object.entries(perf.sum)) host[k] = v / steps;
  const dev = {};
  if (st.deviceMs) PHASES.forEach((p, k) => { dev[p] = st.deviceMs[k]; });
  console.log(JSON.stringify({
    ms_per_step: ms, cell_steps_per_s: (c.nx * c.ny) / (ms * 1e-3), steps, warmup,
    vcycles_per_solve: iters.reduce((a, b) => a + b, 0) / steps,
    host_phase_ms_per_step: host, last_step_device_ms: dev, timings: st.timings,
    node: process.version, cpus: os.cpus().length,
  }));
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
