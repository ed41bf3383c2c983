'use strict';
/**
 * Drives the N-API seams as a TypeScript host would, for the Python tests.
 *
 *   node run_steps.js <dir>
 *
 * <dir>/case.json: {nx, ny, h, dyeScale, scheme, solver, iterations, omega,
 *   tol, steps, dt, buoyancy, confinement, keep:[3], patch:{i0,j0,nx,ny}|null,
 *   f32: step Float32Array fields instead of Float64Array}
 * and raw little-endian arrays u.f64 v.f64 p.f64 label.u8 dye.f64 (3 planes)
 * patch.f32.  dt, scale and gradScale follow Simulation.step
 * (simulation.ts:477-479) with the fixed dt given.
 *
 * Writes, for the Python side to diff against the oracle:
 *   stale_u_<k>.f64   f.u right after step k returned (the one-step-stale
 *                     readback: step k-1's velocity, the initial one for k=0)
 *   out_u.f64 out_v.f64 out_dye.f64  after the last step and a drain
 *   inv_u.f64 inv_v.f64  one step from the initial host fields, taken after
 *                     invalidate() while a readback was in flight
 *   result.json       iterations, marks, timings, and every error message
 */
const fs = require('fs');
const path = require('path');

const ROOT = path.join(__dirname, '..', '..');
const { HipStepper, HipAdvector, addon } = require(path.join(ROOT, 'flowlab_amd', 'node',
  'hipEngine.js'));

function readF64(p) {
  const b = fs.readFileSync(p);
  return new Float64Array(b.buffer, b.byteOffset, b.length / 8).slice();
}

function write(dir, name, a) {
  const d = a instanceof Float64Array ? a : Float64Array.from(a);
  fs.writeFileSync(path.join(dir, name), Buffer.from(d.buffer, d.byteOffset, d.byteLength));
}

/** profiler.ts's Profiler surface: records the phase labels of each step. */
class MarkLog {
  constructor() { this.steps = []; }
  begin() { this.steps.push([]); }
  mark(label) { this.steps[this.steps.length - 1].push(label); }
}

async function errorOf(fn) {
  try {
    await fn();
  } catch (e) {
    return `${e.name}: ${e.message}`;
  }
  return null;
}

async function main() {
  const dir = process.argv[2];
  const c = JSON.parse(fs.readFileSync(path.join(dir, 'case.json'), 'utf8'));
  const g = { nx: c.nx, ny: c.ny, h: c.h };
  const Arr = c.f32 ? Float32Array : Float64Array;
  const dye0 = readF64(path.join(dir, 'dye.f64'));
  const n = dye0.length / 3;
  const init = {
    u: readF64(path.join(dir, 'u.f64')),
    v: readF64(path.join(dir, 'v.f64')),
    p: readF64(path.join(dir, 'p.f64')),
    label: new Uint8Array(fs.readFileSync(path.join(dir, 'label.u8'))),
  };
  const fields = () => ({
    u: Arr.from(init.u), v: Arr.from(init.v), p: Arr.from(init.p), label: init.label.slice(),
    dye: [0, 1, 2].map((k) => Arr.from(dye0.subarray(k * n, (k + 1) * n))),
  });
  const f = fields();
  const st = new HipStepper(g, c.dyeScale, {
    solver: c.solver, iterations: c.iterations, omega: c.omega, tol: c.tol,
    warmExtrapolation: c.warm || 0,
  });
  let patch = null;
  if (c.patch) {
    const b = fs.readFileSync(path.join(dir, 'patch.f32'));
    patch = { ...c.patch, data: new Float32Array(b.buffer, b.byteOffset, b.length / 4).slice() };
    st.setDyePatch(patch);
  }
  st.setBuoyancy(c.buoyancy);
  st.setConfinement(c.confinement);
  const res = { iters: [], marks: null, errors: {} };
  const perf = new MarkLog();
  const scale = (g.h * g.h) / c.dt;
  const gradScale = c.dt / g.h;
  for (let s = 0; s < c.steps; s++) {
    perf.begin();
    // the setters a frame makes before its step, while the last step's
    // readback is still in flight: held and applied by this step
    st.setBuoyancy(c.buoyancy);
    st.setStir(null);
    const pending = st.step(f, c.scheme, c.dt, scale, gradScale, c.keep, perf);
    if (s === 0) {
      // a second step() while the first is in flight must reject
      res.errors.reentrant = await errorOf(() => st.step(f, c.scheme, c.dt, scale, gradScale,
        c.keep, perf));
    }
    res.iters.push(await pending);
    write(dir, `stale_u_${s}.f64`, f.u);
  }
  await st.drainReadback(f);
  res.marks = perf.steps;
  res.timings = st.timings;
  res.deviceMs = st.deviceMs ? Array.from(st.deviceMs) : null;
  await st.readDye(f.dye);
  const out = new Float64Array(3 * n);
  for (let k = 0; k < 3; k++) out.set(f.dye[k], k * n);
  write(dir, 'out_u.f64', f.u);
  write(dir, 'out_v.f64', f.v);
  write(dir, 'out_dye.f64', out);

  // invalidate() with a readback in flight: the host's fields win
  const st2 = new HipStepper(g, c.dyeScale, {
    solver: c.solver, iterations: c.iterations, omega: c.omega, tol: c.tol,
  });
  if (patch) st2.setDyePatch(patch);
  st2.setBuoyancy(c.buoyancy);
  st2.setConfinement(c.confinement);
  const f2 = fields();
  await st2.step(f2, c.scheme, c.dt, scale, gradScale, c.keep, null);
  await st2.step(f2, c.scheme, c.dt, scale, gradScale, c.keep, null);
  res.inflightBeforeInvalidate = st2.velMap !== null;
  const fresh = fields();  // the host resets to the initial state
  f2.u.set(fresh.u);
  f2.v.set(fresh.v);
  f2.p.set(fresh.p);
  for (let k = 0; k < 3; k++) f2.dye[k].set(fresh.dye[k]);
  st2.invalidate();
  await st2.step(f2, c.scheme, c.dt, scale, gradScale, c.keep, null);
  res.uAfterInvalidateIsHost = f2.u.every((x, k) => x === fresh.u[k]);
  await st2.drainReadback(f2);
  write(dir, 'inv_u.f64', f2.u);
  write(dir, 'inv_v.f64', f2.v);

  // argument checks at the boundary
  const e = addon.create(g.nx, g.ny, c.dyeScale, g.h);
  res.errors.shortArray = await errorOf(() => addon.readVelocity(e, new Float64Array(3),
    new Float64Array((g.nx) * (g.ny + 1))));
  res.errors.wrongType = await errorOf(() => addon.readDye(e, new Float32Array(3 * n)));
  res.errors.notEngine = await errorOf(() => addon.readDye({}, new Float64Array(3 * n)));
  res.errors.uploadType = await errorOf(() => addon.upload(e, new Float32Array(init.u.length),
    null, null, null, null));

  // Advector.dye over qIn.length channels (1 and 4), against the 3-channel pass
  const adv = new HipAdvector(g, c.dyeScale);
  const q = [0, 1, 2].map((k) => dye0.slice(k * n, (k + 1) * n));
  const q3 = q.map(() => new Float64Array(n));
  await adv.dye(g, c.scheme, init.u, init.v, q, q3, init.label, c.dt);
  const q1 = [new Float64Array(n)];
  await adv.dye(g, c.scheme, init.u, init.v, [q[1]], q1, init.label, c.dt);
  const q4 = [0, 1, 2, 3].map(() => new Float64Array(n));
  await adv.dye(g, c.scheme, init.u, init.v, [q[0], q[1], q[2], q[1]], q4, init.label, c.dt);
  const eq = (a, b) => a.every((x, k) => x === b[k]);
  res.dyeChannels = { one: eq(q1[0], q3[1]), four: eq(q4[0], q3[0]) && eq(q4[3], q3[1]) &&
    eq(q4[2], q3[2]) };
  res.errors.dyeGrid = await errorOf(() => adv.dye(g, c.scheme, init.u, init.v, q, q3,
    init.label, c.dt, { nx: 3, ny: 3 }));
  fs.writeFileSync(path.join(dir, 'result.json'), JSON.stringify(res));
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
