'use strict';
/**
 * Steps a scene through the N-API HipStepper, as a TypeScript host would.
 *
 *   node run_steps.js <dir>
 *
 * <dir>/case.json: {nx, ny, h, dyeScale, scheme, solver, iterations, omega,
 *   tol, steps, dt, buoyancy, confinement, keep:[3], patch:{i0,j0,nx,ny}|null}
 * and raw little-endian arrays u.f64 v.f64 p.f64 label.u8 dye.f64 (3 planes)
 * patch.f32.  Writes out_u.f64 out_v.f64 out_dye.f64 and result.json.
 * dt, scale and gradScale follow Simulation.step (simulation.ts:477-479)
 * with the fixed dt given.
 */
const fs = require('fs');
const path = require('path');
const { HipStepper } = require('./hipEngine.js');

function readF64(p) {
  const b = fs.readFileSync(p);
  return new Float64Array(b.buffer, b.byteOffset, b.length / 8).slice();
}

async function main() {
  const dir = process.argv[2];
  const c = JSON.parse(fs.readFileSync(path.join(dir, 'case.json'), 'utf8'));
  const g = { nx: c.nx, ny: c.ny, h: c.h };
  const dye = readF64(path.join(dir, 'dye.f64'));
  const n = dye.length / 3;
  const f = {
    u: readF64(path.join(dir, 'u.f64')),
    v: readF64(path.join(dir, 'v.f64')),
    p: readF64(path.join(dir, 'p.f64')),
    label: new Uint8Array(fs.readFileSync(path.join(dir, 'label.u8'))),
    dye: [dye.slice(0, n), dye.slice(n, 2 * n), dye.slice(2 * n)],
  };
  const st = new HipStepper(g, c.dyeScale, {
    solver: c.solver, iterations: c.iterations, omega: c.omega, tol: c.tol,
  });
  if (c.patch) {
    const b = fs.readFileSync(path.join(dir, 'patch.f32'));
    st.setDyePatch({ ...c.patch, data: new Float32Array(b.buffer, b.byteOffset, b.length / 4) });
  }
  st.setBuoyancy(c.buoyancy);
  st.setConfinement(c.confinement);
  const iters = [];
  let reentrant = null;
  for (let s = 0; s < c.steps; s++) {
    const scale = (g.h * g.h) / c.dt;
    const gradScale = c.dt / g.h;
    const pending = st.step(f, c.scheme, c.dt, scale, gradScale, c.keep);
    if (s === 0) {
      // a second step() while the first is in flight must reject
      try {
        await st.step(f, c.scheme, c.dt, scale, gradScale, c.keep);
      } catch (e) {
        reentrant = e.message;
      }
    }
    iters.push(await pending);
  }
  await st.readDye(f.dye);
  const out = new Float64Array(3 * n);
  for (let k = 0; k < 3; k++) out.set(f.dye[k], k * n);
  fs.writeFileSync(path.join(dir, 'out_u.f64'), Buffer.from(f.u.buffer));
  fs.writeFileSync(path.join(dir, 'out_v.f64'), Buffer.from(f.v.buffer));
  fs.writeFileSync(path.join(dir, 'out_dye.f64'), Buffer.from(out.buffer));
  fs.writeFileSync(path.join(dir, 'result.json'), JSON.stringify({ iters, reentrant }));
}

main().catch((e) => {
  console.error(e.stack || e.message);
  process.exit(1);
});
