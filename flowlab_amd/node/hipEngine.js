'use strict';
/**
 * HIP engine seams for flowlab's Simulation, over the flowhip N-API addon.
 *
 *   HipStepper         implements GpuStep        (src/core/simulation.ts:35-84)
 *   HipPressureSolver  implements PressureSolver (src/core/pressureSolver.ts:21-41)
 *   HipAdvector        implements Advector       (src/core/advector.ts:28-63)
 *
 * Plain JavaScript (erasable TypeScript-compatible shapes) so it runs on the
 * Node the GPU host has.  Fields are the reference's Float64Array /
 * Uint8Array layouts; the device computes in FP32 with an FP64 pressure.
 */
const path = require('path');

const addon = require(path.join(__dirname, 'flowhip.node'));

const SCHEME = { semiLagrangian: 0, macCormack: 1 };
const SOLVER = { rbsor: 1, mg: 2, 'mg-dc': 3, 'mg-fcg': 4 };

function ensure(handle, cls) {
  if (!handle) throw new Error(`${cls}: no GPU (libflowhip could not create an engine)`);
  return handle;
}

class HipStepper {
  /**
   * @param {{nx:number, ny:number, h:number}} g  the velocity grid
   * @param {number} dyeScale whole-number dye refinement
   * @param {object} opts {solver: 'mg-dc'|'mg-fcg'|'mg'|'rbsor', iterations, omega, tol}
   */
  constructor(g, dyeScale = 1, opts = {}) {
    this.name = 'hip-fused';
    this.e = ensure(addon.create(g.nx, g.ny, dyeScale, g.h), 'HipStepper');
    this.solver = SOLVER[opts.solver || 'mg-dc'];
    this.iterations = opts.iterations || 100;
    this.omega = opts.omega || 1.0;
    this.tol = opts.tol === undefined ? 1e-6 : opts.tol;
    this.dirty = true;
    this.busy = false;
  }

  /** GpuStep.step: resolves with the solver's iteration count; f.u / f.v
   *  hold the step's velocity, the dye stays resident (readDye). */
  async step(f, scheme, dt, scale, gradScale, dyeKeep) {
    if (this.busy) throw new Error('HipStepper.step() is not re-entrant');
    this.busy = true;
    try {
      if (this.dirty) {
        const n = f.dye[0].length;
        const dye = new Float64Array(3 * n);
        for (let c = 0; c < 3; c++) dye.set(f.dye[c], c * n);
        addon.upload(this.e, f.u, f.v, f.p, f.label, dye);
        this.dirty = false;
      }
      const iters = await addon.step(this.e, SCHEME[scheme], dt, scale, gradScale,
        [dyeKeep[0], dyeKeep[1], dyeKeep[2]], this.solver, this.iterations, this.omega, this.tol);
      await addon.readVelocity(this.e, f.u, f.v);
      return iters;
    } finally {
      this.busy = false;
    }
  }

  invalidate() { this.dirty = true; }

  setDyePatch(patch) {
    if (patch) addon.setDyePatch(this.e, patch.i0, patch.j0, patch.nx, patch.ny, patch.data);
    else addon.setDyePatch(this.e, 0, 0, 0, 0, null);
  }

  setBuoyancy(w) { addon.setBuoyancy(this.e, w); }

  setConfinement(eps) { addon.setConfinement(this.e, eps); }

  setStir(stir) {
    if (stir) addon.setStir(this.e, 1, stir.x, stir.y, stir.r, stir.dx, stir.dy);
    else addon.setStir(this.e, 0, 0, 0, 0, 0, 0);
  }

  setLabels(label) { addon.setLabels(this.e, label); }

  async readDye(dye) {
    const n = dye[0].length;
    const buf = new Float64Array(3 * n);
    await addon.readDye(this.e, buf);
    for (let c = 0; c < 3; c++) dye[c].set(buf.subarray(c * n, (c + 1) * n));
  }

  maxFaceSpeed() { return addon.maxFaceSpeed(this.e); }
}

class HipPressureSolver {
  constructor(g, method = 'mg-dc') {
    this.name = `hip-${method}`;
    this.method = SOLVER[method];
    this.e = ensure(addon.create(g.nx, g.ny, 1, g.h), 'HipPressureSolver');
  }

  /** PressureSolver.solve: p is the warm start in and the answer out. */
  solve(g, p, div, label, scale, iterations, omega, tol) {
    return addon.solvePressure(this.e, this.method, p, div, label, scale, iterations, omega, tol);
  }
}

class HipAdvector {
  constructor(g, dyeScale = 1) {
    this.name = 'hip';
    this.e = ensure(addon.create(g.nx, g.ny, dyeScale, g.h), 'HipAdvector');
  }

  velocity(g, scheme, uIn, vIn, uOut, vOut, label, dt) {
    return addon.advectVelocity(this.e, SCHEME[scheme], uIn, vIn, uOut, vOut, label, dt);
  }

  async dye(g, scheme, u, v, qIn, qOut, label, dt) {
    const n = qIn[0].length;
    const a = new Float64Array(3 * n);
    const b = new Float64Array(3 * n);
    for (let c = 0; c < 3; c++) a.set(qIn[c], c * n);
    await addon.advectDye(this.e, SCHEME[scheme], u, v, a, b, label, dt);
    for (let c = 0; c < qOut.length; c++) qOut[c].set(b.subarray(c * n, (c + 1) * n));
  }
}

module.exports = { HipStepper, HipPressureSolver, HipAdvector, addon, SCHEME, SOLVER };
