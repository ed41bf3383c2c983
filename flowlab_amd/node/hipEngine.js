'use strict';
/**
 * HIP engine seams for flowlab's Simulation, over the flowhip N-API addon.
 *
 *   HipStepper         implements GpuStep        (src/core/simulation.ts:35-84)
 *   HipPressureSolver  implements PressureSolver (src/core/pressureSolver.ts:21-41)
 *   HipAdvector        implements Advector       (src/core/advector.ts:28-63)
 *
 * Plain JavaScript (erasable TypeScript-compatible shapes) so it runs on the
 * Node the GPU host has.  Fields are the reference's FieldArray layouts
 * (Float64Array or Float32Array, Uint8Array labels); the device computes in
 * FP32 with an FP64 pressure.
 */
const path = require('path');
const { performance } = require('perf_hooks');

const addon = require(path.join(__dirname, 'flowhip.node'));

const SCHEME = { semiLagrangian: 0, macCormack: 1 };
const SOLVER = { rbsor: 1, mg: 2, 'mg-dc': 3, 'mg-fcg': 4 };
/** Device phases of getTimings(), in order (include/flowhip.h FH_PHASE_*). */
const PHASES = ['advect', 'forces', 'divergence', 'solve', 'project', 'dye', 'total'];

/** profiler.ts:39-41: the EMA every displayed duration goes through. */
const ALPHA = 0.12;
function smooth(avg, sample) {
  return Number.isFinite(sample) ? avg + ALPHA * (sample - avg) : avg;
}

function ensure(handle, cls) {
  if (!handle) throw new Error(`${cls}: no GPU (libflowhip could not create an engine)`);
  return handle;
}

/** The addon takes Float64Array fields; a Float32Array FieldArray is widened. */
function f64(a) {
  return a instanceof Float64Array ? a : Float64Array.from(a);
}

/** Resolves when p settles, whatever the outcome (p may be null). */
function settled(p) {
  return p ? p.then(() => undefined, () => undefined) : Promise.resolve();
}

class HipStepper {
  /**
   * @param {{nx:number, ny:number, h:number}} g  the velocity grid
   * @param {number} dyeScale whole-number dye refinement
   * @param {object} opts {solver: 'mg-dc'|'mg-fcg'|'mg'|'rbsor', iterations, omega, tol,
   *   deviceTiming: per-phase device timestamps of every step (default true),
   *   warmExtrapolation: 0 (default: each solve starts from the previous
   *   pressure, as stepGpu.ts does), 1 or 2 (from the last solutions
   *   extrapolated linearly or quadratically in time)}
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
    this.lost = false;
    // the step's velocity is read into these while the host goes on
    // (stepGpu.ts:546-551) and copied into f.u / f.v at the top of the next
    // step (drainReadback); velMap is that readback's promise
    this.velU = new Float64Array((g.nx + 1) * g.ny);
    this.velV = new Float64Array(g.nx * (g.ny + 1));
    this.velMap = null;
    // setters that arrive while a readback is in flight are held here and
    // applied at the top of the next step (the engine takes one call at a time)
    this.pending = new Map();
    /** GpuTimings (pressureGpu.ts:56-66): host upload, host wait for the
     *  readback, device ms of the step; smoothed like the profiler's phases. */
    this.timings = { upload: 0, wait: 0, device: NaN };
    /** Per-phase device ms of the last drained step (PHASES order). */
    this.deviceMs = null;
    this.deviceTiming = opts.deviceTiming !== false;
    if (this.deviceTiming) addon.setTiming(this.e, 1);
    if (opts.warmExtrapolation) addon.setWarmExtrapolation(this.e, opts.warmExtrapolation);
  }

  /** GpuStep.step: resolves with the solver's iteration count.  Like the
   *  fused WebGPU step (stepGpu.ts:393-559) the velocity readback is started
   *  and NOT awaited: f.u / f.v receive it at the top of the next step, so
   *  on return they hold the previous step's velocity; the dye stays
   *  resident (readDye) and f.p is not updated (simulation.ts:29-33). */
  async step(f, scheme, dt, scale, gradScale, dyeKeep, perf) {
    if (this.busy) throw new Error('HipStepper.step() is not re-entrant');
    if (this.lost) throw new Error('GPU device lost');
    this.busy = true;
    try {
      const tDrain = performance.now();
      await this.drainReadback(f);
      const t0 = performance.now();
      this.timings.wait = smooth(this.timings.wait, t0 - tDrain);
      if (perf) perf.mark('readback');

      this.applyPending();
      if (this.dirty) {
        const n = f.dye[0].length;
        const dye = new Float64Array(3 * n);
        for (let c = 0; c < 3; c++) dye.set(f.dye[c], c * n);
        addon.upload(this.e, f64(f.u), f64(f.v), f64(f.p), f.label, dye);
        this.dirty = false;
      }
      const t1 = performance.now();
      if (perf) perf.mark('upload');
      this.timings.upload = smooth(this.timings.upload, t1 - t0);

      const iters = await addon.step(this.e, SCHEME[scheme], dt, scale, gradScale,
        [dyeKeep[0], dyeKeep[1], dyeKeep[2]], this.solver, this.iterations, this.omega,
        this.tol);
      // started, not awaited: drained at the top of the next step
      this.velMap = addon.readVelocity(this.e, this.velU, this.velV);
      this.velMap.catch(() => undefined);  // the drain reports it
      if (perf) perf.mark('solve');
      return iters;
    } catch (err) {
      if (err && err.message === 'GPU device lost') this.lost = true;
      throw err;
    } finally {
      this.busy = false;
    }
  }

  /**
   * The last step's velocity (and its device timings) into the host fields
   * (stepGpu.ts:578-604).  After invalidate() the host holds the truth and
   * the in-flight values are dropped, not copied: writing them into f.u /
   * f.v would undo what the host just did.
   */
  async drainReadback(f) {
    if (!this.velMap) return;
    const m = this.velMap;
    this.velMap = null;
    try {
      await m;
    } catch (err) {
      if (err && err.message === 'GPU device lost') this.lost = true;
      throw err;
    }
    if (!this.dirty && f) {
      f.u.set(this.velU);
      f.v.set(this.velV);
    }
    if (this.deviceTiming) {
      const ms = addon.getTimings(this.e);  // that step has completed
      if (ms[6] > 0) {
        this.deviceMs = ms;
        this.timings.device = Number.isFinite(this.timings.device)
          ? smooth(this.timings.device, ms[6]) : ms[6];
      }
    }
  }

  /** Host fields changed under the engine (reset, CPU frames): the next step
   *  re-uploads u, v, p, label and dye, and drops an in-flight readback. */
  invalidate() { this.dirty = true; }

  /** A setter that must not reach the engine while a readback runs is held
   *  until the next step; otherwise it applies now. */
  defer(key, fn) {
    if (this.velMap || this.busy) this.pending.set(key, fn);
    else fn();
  }

  applyPending() {
    for (const fn of this.pending.values()) fn();
    this.pending.clear();
  }

  setDyePatch(patch) {
    this.defer('patch', () => {
      if (patch) addon.setDyePatch(this.e, patch.i0, patch.j0, patch.nx, patch.ny, patch.data);
      else addon.setDyePatch(this.e, 0, 0, 0, 0, null);
    });
  }

  setBuoyancy(w) { this.defer('buoyancy', () => addon.setBuoyancy(this.e, w)); }

  setConfinement(eps) { this.defer('confinement', () => addon.setConfinement(this.e, eps)); }

  setStir(stir) {
    this.defer('stir', () => {
      if (stir) addon.setStir(this.e, 1, stir.x, stir.y, stir.r, stir.dx, stir.dy);
      else addon.setStir(this.e, 0, 0, 0, 0, 0, 0);
    });
  }

  /** GpuStep.setLabels: labels alone, then commitLabels on the device fields. */
  setLabels(label) {
    const copy = Uint8Array.from(label);
    this.defer('labels', () => addon.setLabels(this.e, copy));
  }

  /** GpuStep.readDye: the resident dye into the host arrays. */
  async readDye(dye) {
    await settled(this.velMap);
    this.applyPending();
    const n = dye[0].length;
    const buf = new Float64Array(3 * n);
    await addon.readDye(this.e, buf);
    for (let c = 0; c < 3; c++) dye[c].set(buf.subarray(c * n, (c + 1) * n));
  }

  /** Largest face speed of the device's current velocity. */
  async maxFaceSpeed() {
    await settled(this.velMap);
    return addon.maxFaceSpeed(this.e);
  }
}

class HipPressureSolver {
  constructor(g, method = 'mg-dc') {
    this.name = `hip-${method}`;
    this.method = SOLVER[method];
    this.e = ensure(addon.create(g.nx, g.ny, 1, g.h), 'HipPressureSolver');
    this.busy = false;
  }

  /** PressureSolver.solve: p is the warm start in and the answer out. */
  async solve(g, p, div, label, scale, iterations, omega, tol) {
    if (this.busy) throw new Error('HipPressureSolver.solve() is not re-entrant');
    this.busy = true;
    try {
      const pd = f64(p);
      const it = await addon.solvePressure(this.e, this.method, pd, f64(div), label, scale,
        iterations, omega, tol);
      if (pd !== p) p.set(pd);
      return it;
    } finally {
      this.busy = false;
    }
  }
}

class HipAdvector {
  constructor(g, dyeScale = 1) {
    this.name = 'hip';
    this.ds = dyeScale;
    this.nx = g.nx;
    this.ny = g.ny;
    this.e = ensure(addon.create(g.nx, g.ny, dyeScale, g.h), 'HipAdvector');
  }

  async velocity(g, scheme, uIn, vIn, uOut, vOut, label, dt) {
    const uo = f64(uOut), vo = f64(vOut);
    await addon.advectVelocity(this.e, SCHEME[scheme], f64(uIn), f64(vIn), uo, vo, label, dt);
    if (uo !== uOut) uOut.set(uo);
    if (vo !== vOut) vOut.set(vo);
  }

  /** Advector.dye: qIn.length independent channels (advector.ts:104-121),
   *  three per device pass; dg, when given, must be this engine's dye grid. */
  async dye(g, scheme, u, v, qIn, qOut, label, dt, dg) {
    const dnx = this.nx * this.ds, dny = this.ny * this.ds;
    if (dg && (dg.nx !== dnx || dg.ny !== dny))
      throw new Error(`HipAdvector.dye: dye grid ${dg.nx}x${dg.ny} is not this engine's ` +
        `${dnx}x${dny} (dye scale ${this.ds})`);
    if (qOut.length < qIn.length) throw new Error('HipAdvector.dye: fewer outputs than inputs');
    const n = dnx * dny;
    const ud = f64(u), vd = f64(v);
    for (let c0 = 0; c0 < qIn.length; c0 += 3) {
      const k = Math.min(3, qIn.length - c0);
      const a = new Float64Array(3 * n);
      const b = new Float64Array(3 * n);
      for (let c = 0; c < k; c++) a.set(qIn[c0 + c], c * n);
      await addon.advectDye(this.e, SCHEME[scheme], ud, vd, a, b, label, dt);
      for (let c = 0; c < k; c++) qOut[c0 + c].set(b.subarray(c * n, (c + 1) * n));
    }
  }
}

module.exports = { HipStepper, HipPressureSolver, HipAdvector, addon, SCHEME, SOLVER, PHASES, smooth };
