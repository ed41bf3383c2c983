'use strict';
/**
 * FP64 CPU step in plain JavaScript: the reference's CPU path
 * (src/core/simulation.ts:467-587, the engine with no stepper) restated for
 * the Node on the GPU host, one thread, Float64Array fields.  TEST AND
 * BASELINE INFRASTRUCTURE ONLY: run by tests/ (pinned against the C oracle,
 * oracle/oracle.c, operation for operation) and by bench.py's CPU baseline;
 * the product path never loads it.
 *
 * Each function cites the reference it follows.  Vorticity confinement and
 * the defect-correction multigrid (solver 'mg-dc': coarsen to 16, FP64
 * direct coarsest solve) are build-defined, as in the C oracle.
 *
 *   node cpuStep.js <dir> <warmup> <steps>
 * reads a scene from <dir> (case.json: nx, ny, h, dyeScale, scheme, solver,
 * pressureIters, omega, tol, cflTarget, dtMax, buoyancy, confinement,
 * decay[3], patch{i0,j0,nx,ny}|null; u.f64 v.f64 p.f64 label.u8 dye.f64
 * patch.f32), steps it warmup + steps times, writes out_u.f64 out_v.f64
 * out_p.f64 out_dye.f64 and prints one JSON line with the timing.
 */
const FLUID = 0, AIR = 1, SOLID = 2;

/* grid.ts:159-161 -- the bounds test runs before the lookup */
function solidOrOutside(nx, ny, l, i, j) {
  return i < 0 || j < 0 || i >= nx || j >= ny || l[i + j * nx] === SOLID;
}

/* boundaries.ts:23-25 */
function isSolidCell(nx, ny, l, i, j) {
  return i >= 0 && j >= 0 && i < nx && j < ny && l[i + j * nx] === SOLID;
}

const dmin = (a, b) => (a < b ? a : b);
const dmax = (a, b) => (a > b ? a : b);

/* grid.ts:173-177 (returns i0 in AX.i, f in AX.f) */
const AX = { i: 0, f: 0 }, AY = { i: 0, f: 0 };
function clampedAxis(pos, count, out) {
  const c = dmin(dmax(pos, 0.0), count - 1);
  let i = Math.floor(c);
  if (i > count - 2) i = count - 2;
  out.i = i;
  out.f = c - i;
}

/* grid.ts:180-191 */
function bilerp(v00, v10, v01, v11, fx, fy) {
  const lo = v00 + fx * (v10 - v00);
  const hi = v01 + fx * (v11 - v01);
  return lo + fy * (hi - lo);
}

/* grid.ts:197-208 */
function sampleU(nx, ny, h, u, x, y) {
  clampedAxis(x / h, nx + 1, AX);
  clampedAxis(y / h - 0.5, ny, AY);
  const s = nx + 1, k = AX.i + AY.i * s;
  return bilerp(u[k], u[k + 1], u[k + s], u[k + s + 1], AX.f, AY.f);
}

/* grid.ts:214-225 */
function sampleV(nx, ny, h, v, x, y) {
  clampedAxis(x / h - 0.5, nx, AX);
  clampedAxis(y / h, ny + 1, AY);
  const k = AX.i + AY.i * nx;
  return bilerp(v[k], v[k + 1], v[k + nx], v[k + nx + 1], AX.f, AY.f);
}

/* grid.ts:239-250 */
function sampleP(nx, ny, h, q, x, y) {
  clampedAxis(x / h - 0.5, nx, AX);
  clampedAxis(y / h - 0.5, ny, AY);
  const k = AX.i + AY.i * nx;
  return bilerp(q[k], q[k + 1], q[k + nx], q[k + nx + 1], AX.f, AY.f);
}

/* grid.ts:268-270 */
function clampToRange(q, a, b, c, d) {
  const lo = dmin(dmin(a, b), dmin(c, d));
  const hi = dmax(dmax(a, b), dmax(c, d));
  return dmin(dmax(q, lo), hi);
}

/* grid.ts:273-315 */
function clampToStencilU(nx, ny, h, u, x, y, q) {
  clampedAxis(x / h, nx + 1, AX);
  clampedAxis(y / h - 0.5, ny, AY);
  const s = nx + 1, k = AX.i + AY.i * s;
  return clampToRange(q, u[k], u[k + 1], u[k + s], u[k + s + 1]);
}
function clampToStencilV(nx, ny, h, v, x, y, q) {
  clampedAxis(x / h - 0.5, nx, AX);
  clampedAxis(y / h, ny + 1, AY);
  const k = AX.i + AY.i * nx;
  return clampToRange(q, v[k], v[k + 1], v[k + nx], v[k + nx + 1]);
}
function clampToStencilP(nx, ny, h, q, x, y, value) {
  clampedAxis(x / h - 0.5, nx, AX);
  clampedAxis(y / h - 0.5, ny, AY);
  const k = AX.i + AY.i * nx;
  return clampToRange(value, q[k], q[k + 1], q[k + nx], q[k + nx + 1]);
}

/* advect.ts:41-55: RK2 midpoint backtrace into BT */
const BT = { x: 0, y: 0 };
function backtrace(nx, ny, h, uCar, vCar, x, y, u0, v0, dt) {
  const midX = x - 0.5 * dt * u0;
  const midY = y - 0.5 * dt * v0;
  const us = sampleU(nx, ny, h, uCar, midX, midY);
  const vs = sampleV(nx, ny, h, vCar, midX, midY);
  BT.x = x - dt * us;
  BT.y = y - dt * vs;
}

/* advect.ts:72-112 */
function advectFaces(nx, ny, h, uCar, vCar, uSrc, vSrc, uOut, vOut, label, dt) {
  for (let j = 0; j < ny; j++) {
    const y = (j + 0.5) * h;
    for (let i = 0; i <= nx; i++) {
      const k = i + j * (nx + 1);
      if (solidOrOutside(nx, ny, label, i - 1, j) || solidOrOutside(nx, ny, label, i, j)) {
        uOut[k] = uSrc[k];
        continue;
      }
      const x = i * h;
      backtrace(nx, ny, h, uCar, vCar, x, y, uCar[k], sampleV(nx, ny, h, vCar, x, y), dt);
      uOut[k] = sampleU(nx, ny, h, uSrc, BT.x, BT.y);
    }
  }
  for (let j = 0; j <= ny; j++) {
    const y = j * h;
    for (let i = 0; i < nx; i++) {
      const k = i + j * nx;
      if (solidOrOutside(nx, ny, label, i, j - 1) || solidOrOutside(nx, ny, label, i, j)) {
        vOut[k] = vSrc[k];
        continue;
      }
      const x = (i + 0.5) * h;
      backtrace(nx, ny, h, uCar, vCar, x, y, sampleU(nx, ny, h, uCar, x, y), vCar[k], dt);
      vOut[k] = sampleV(nx, ny, h, vSrc, BT.x, BT.y);
    }
  }
}

/* advect.ts:166-211 */
function advectVelocityMacCormack(nx, ny, h, uIn, vIn, uHat, vHat, uOut, vOut, label, dt) {
  advectFaces(nx, ny, h, uIn, vIn, uIn, vIn, uHat, vHat, label, dt);
  advectFaces(nx, ny, h, uIn, vIn, uHat, vHat, uOut, vOut, label, -dt);
  for (let j = 0; j < ny; j++) {
    const y = (j + 0.5) * h;
    for (let i = 0; i <= nx; i++) {
      const k = i + j * (nx + 1);
      if (solidOrOutside(nx, ny, label, i - 1, j) || solidOrOutside(nx, ny, label, i, j)) {
        uOut[k] = uIn[k];
        continue;
      }
      const x = i * h;
      const corrected = uHat[k] + 0.5 * (uIn[k] - uOut[k]);
      backtrace(nx, ny, h, uIn, vIn, x, y, uIn[k], sampleV(nx, ny, h, vIn, x, y), dt);
      uOut[k] = clampToStencilU(nx, ny, h, uIn, BT.x, BT.y, corrected);
    }
  }
  for (let j = 0; j <= ny; j++) {
    const y = j * h;
    for (let i = 0; i < nx; i++) {
      const k = i + j * nx;
      if (solidOrOutside(nx, ny, label, i, j - 1) || solidOrOutside(nx, ny, label, i, j)) {
        vOut[k] = vIn[k];
        continue;
      }
      const x = (i + 0.5) * h;
      const corrected = vHat[k] + 0.5 * (vIn[k] - vOut[k]);
      backtrace(nx, ny, h, uIn, vIn, x, y, sampleU(nx, ny, h, uIn, x, y), vIn[k], dt);
      vOut[k] = clampToStencilV(nx, ny, h, vIn, BT.x, BT.y, corrected);
    }
  }
}

/* advect.ts:241-270 */
function advectScalar(nx, ny, h, dnx, dny, dh, u, v, qIn, qOut, label, dt) {
  const ratio = dh / h;
  for (let j = 0; j < dny; j++) {
    const y = (j + 0.5) * dh;
    const lj = Math.floor((j + 0.5) * ratio);
    for (let i = 0; i < dnx; i++) {
      const k = i + j * dnx;
      if (solidOrOutside(nx, ny, label, Math.floor((i + 0.5) * ratio), lj)) {
        qOut[k] = qIn[k];
        continue;
      }
      const x = (i + 0.5) * dh;
      const u0 = sampleU(nx, ny, h, u, x, y);
      backtrace(nx, ny, h, u, v, x, y, u0, sampleV(nx, ny, h, v, x, y), dt);
      qOut[k] = sampleP(dnx, dny, dh, qIn, BT.x, BT.y);
    }
  }
}

/* advect.ts:280-312 */
function advectScalarMacCormack(nx, ny, h, dnx, dny, dh, u, v, qIn, qHat, qOut, label, dt) {
  advectScalar(nx, ny, h, dnx, dny, dh, u, v, qIn, qHat, label, dt);
  advectScalar(nx, ny, h, dnx, dny, dh, u, v, qHat, qOut, label, -dt);
  const ratio = dh / h;
  for (let j = 0; j < dny; j++) {
    const y = (j + 0.5) * dh;
    const lj = Math.floor((j + 0.5) * ratio);
    for (let i = 0; i < dnx; i++) {
      const k = i + j * dnx;
      if (solidOrOutside(nx, ny, label, Math.floor((i + 0.5) * ratio), lj)) {
        qOut[k] = qIn[k];
        continue;
      }
      const x = (i + 0.5) * dh;
      const corrected = qHat[k] + 0.5 * (qIn[k] - qOut[k]);
      const u0 = sampleU(nx, ny, h, u, x, y);
      backtrace(nx, ny, h, u, v, x, y, u0, sampleV(nx, ny, h, v, x, y), dt);
      qOut[k] = clampToStencilP(dnx, dny, dh, qIn, BT.x, BT.y, corrected);
    }
  }
}

/* divergence.ts:17-25 */
function divergence(nx, ny, h, u, v, out) {
  const coef = 1.0 / h;
  for (let j = 0; j < ny; j++)
    for (let i = 0; i < nx; i++)
      out[i + j * nx] = coef * (u[i + 1 + j * (nx + 1)] - u[i + j * (nx + 1)] +
        v[i + (j + 1) * nx] - v[i + j * nx]);
}

/* pressure.ts:40-91.  Neighbour order +x, +y, -x, -y. */
function gsSweep(nx, ny, p, div, label, scale, omega, reverseX, reverseY) {
  const iStep = reverseX ? -1 : 1, iFrom = reverseX ? nx - 1 : 0, iTo = reverseX ? -1 : nx;
  const jStep = reverseY ? -1 : 1, jFrom = reverseY ? ny - 1 : 0, jTo = reverseY ? -1 : ny;
  for (let j = jFrom; j !== jTo; j += jStep) {
    for (let i = iFrom; i !== iTo; i += iStep) {
      const k = i + j * nx;
      if (label[k] !== FLUID) continue;
      let count = 0, sum = 0;
      if (!solidOrOutside(nx, ny, label, i + 1, j)) { count++; sum += p[k + 1]; }
      if (!solidOrOutside(nx, ny, label, i, j + 1)) { count++; sum += p[k + nx]; }
      if (!solidOrOutside(nx, ny, label, i - 1, j)) { count++; sum += p[k - 1]; }
      if (!solidOrOutside(nx, ny, label, i, j - 1)) { count++; sum += p[k - nx]; }
      if (count === 0) continue;
      const pGS = (sum - scale * div[k]) / count;
      p[k] = (1 - omega) * p[k] + omega * pGS;
    }
  }
}

/* pressure.ts:102-142 */
function rmsRemainingDivergence(nx, ny, p, div, label, scale) {
  let sumSq = 0, n = 0;
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      const k = i + j * nx;
      if (label[k] !== FLUID) continue;
      let count = 0, sum = 0;
      if (!solidOrOutside(nx, ny, label, i + 1, j)) { count++; sum += p[k + 1]; }
      if (!solidOrOutside(nx, ny, label, i, j + 1)) { count++; sum += p[k + nx]; }
      if (!solidOrOutside(nx, ny, label, i - 1, j)) { count++; sum += p[k - 1]; }
      if (!solidOrOutside(nx, ny, label, i, j - 1)) { count++; sum += p[k - nx]; }
      if (count === 0) continue;
      const r = div[k] + (count * p[k] - sum) / scale;
      sumSq += r * r;
      n++;
    }
  }
  return n > 0 ? Math.sqrt(sumSq / n) : 0;
}

/* pressure.ts:152-161 */
function fluidDivRms(div, label) {
  let sumSq = 0, m = 0;
  for (let k = 0; k < div.length; k++) {
    if (label[k] !== FLUID) continue;
    sumSq += div[k] * div[k];
    m++;
  }
  return m > 0 ? Math.sqrt(sumSq / m) : 0;
}

/* pressure.ts:182-212 */
function solvePressure(nx, ny, p, div, label, scale, iterations, omega, tol) {
  const divRms = fluidDivRms(div, label);
  if (divRms === 0) { p.fill(0); return 0; }
  const target = tol > 0 ? tol * divRms : -1;
  for (let k = 0; k < iterations; k++) {
    gsSweep(nx, ny, p, div, label, scale, omega, k % 2 === 1, k % 4 >= 2);
    if (target >= 0 && k % 12 === 11 &&
        rmsRemainingDivergence(nx, ny, p, div, label, scale) <= target) return k + 1;
  }
  return iterations;
}

/* pressureRedBlack.ts:32-72.  Neighbour order +x, -x, +y, -y. */
function rbSweep(nx, ny, p, div, label, scale, omega, color) {
  for (let j = 0; j < ny; j++) {
    for (let i = (j & 1) ^ color; i < nx; i += 2) {
      const k = i + j * nx;
      if (label[k] !== FLUID) continue;
      let count = 0, sum = 0;
      if (!solidOrOutside(nx, ny, label, i + 1, j)) { count++; sum += p[k + 1]; }
      if (!solidOrOutside(nx, ny, label, i - 1, j)) { count++; sum += p[k - 1]; }
      if (!solidOrOutside(nx, ny, label, i, j + 1)) { count++; sum += p[k + nx]; }
      if (!solidOrOutside(nx, ny, label, i, j - 1)) { count++; sum += p[k - nx]; }
      if (count === 0) continue;
      const pGS = (sum - scale * div[k]) / count;
      p[k] = (1 - omega) * p[k] + omega * pGS;
    }
  }
}

/* pressureRedBlack.ts:88-113 */
function solvePressureRedBlack(nx, ny, p, div, label, scale, iterations, omega, tol) {
  const divRms = fluidDivRms(div, label);
  if (divRms === 0) { p.fill(0); return 0; }
  const target = tol > 0 ? tol * divRms : -1;
  for (let k = 0; k < iterations; k++) {
    rbSweep(nx, ny, p, div, label, scale, omega, 0);
    rbSweep(nx, ny, p, div, label, scale, omega, 1);
    if (target >= 0 && k % 12 === 11 &&
        rmsRemainingDivergence(nx, ny, p, div, label, scale) <= target) return k + 1;
  }
  return iterations;
}

/* ---- multigrid (pressureMultigrid.ts) ------------------------------------ */

const MG_PRE = 2, MG_POST = 2, MG_COARSE_SWEEPS = 8;  /* pressureMultigrid.ts:47-63 */

/* pressureMultigrid.ts:65-78 (mode 0: coarsest 4); mode 1 stops at 16 */
function levelSizes(nx, ny, mode) {
  const c = mode === 0 ? 4 : 16;
  const out = [[nx, ny]];
  while (nx > c || ny > c) {
    nx = (nx + 1) >> 1;
    ny = (ny + 1) >> 1;
    out.push([nx, ny]);
  }
  return out;
}

/* pressureMultigrid.ts:92-110: Air > Fluid > Solid */
function coarsenLabels(fnx, fny, fine, cnx, cny, coarse) {
  for (let J = 0; J < cny; J++) {
    for (let I = 0; I < cnx; I++) {
      let air = false, fluid = false;
      for (let dj = 0; dj < 2; dj++) {
        for (let di = 0; di < 2; di++) {
          const i = 2 * I + di, j = 2 * J + dj;
          if (i >= fnx || j >= fny) continue;
          const l = fine[i + j * fnx];
          air = air || l === AIR;
          fluid = fluid || l === FLUID;
        }
      }
      coarse[I + J * cnx] = air ? AIR : fluid ? FLUID : SOLID;
    }
  }
}

/* pressureMultigrid.ts:117-151.  Neighbour order +x, -x, +y, -y. */
function computeResidual(nx, ny, x, b, label, r) {
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      const k = i + j * nx;
      r[k] = 0;
      if (label[k] !== FLUID) continue;
      let count = 0, sum = 0;
      if (!solidOrOutside(nx, ny, label, i + 1, j)) { count++; sum += x[k + 1]; }
      if (!solidOrOutside(nx, ny, label, i - 1, j)) { count++; sum += x[k - 1]; }
      if (!solidOrOutside(nx, ny, label, i, j + 1)) { count++; sum += x[k + nx]; }
      if (!solidOrOutside(nx, ny, label, i, j - 1)) { count++; sum += x[k - nx]; }
      if (count === 0) continue;
      r[k] = b[k] - (count * x[k] - sum);
    }
  }
}

/* pressureMultigrid.ts:156-178 */
function restrictResidual(fnx, fny, r, cnx, cny, bC, eC) {
  for (let J = 0; J < cny; J++) {
    for (let I = 0; I < cnx; I++) {
      let sum = 0;
      for (let dj = 0; dj < 2; dj++)
        for (let di = 0; di < 2; di++) {
          const i = 2 * I + di, j = 2 * J + dj;
          if (i < fnx && j < fny) sum += r[i + j * fnx];
        }
      bC[I + J * cnx] = sum;
      eC[I + J * cnx] = 0;
    }
  }
}

/* pressureMultigrid.ts:192-207 */
function prolongAdd(fnx, fny, x, label, cnx, cny, e) {
  for (let j = 0; j < fny; j++) {
    const J = j >> 1;
    const J2 = (j & 1) ? (J + 1 < cny - 1 ? J + 1 : cny - 1) : (J - 1 > 0 ? J - 1 : 0);
    for (let i = 0; i < fnx; i++) {
      const k = i + j * fnx;
      if (label[k] !== FLUID) continue;
      const I = i >> 1;
      const I2 = (i & 1) ? (I + 1 < cnx - 1 ? I + 1 : cnx - 1) : (I - 1 > 0 ? I - 1 : 0);
      x[k] += 0.5625 * e[I + J * cnx] + 0.1875 * (e[I2 + J * cnx] + e[I + J2 * cnx]) +
        0.0625 * e[I2 + J2 * cnx];
    }
  }
}

/* Build-defined (mode 1): FP64 dense solve of the coarsest level; pure-Neumann
 * fluid components regularised by (1/n) 1 1^T, their mean removed after. */
function coarseFactor(nx, ny, label) {
  const N = nx * ny;
  const count = (i, j) => (!solidOrOutside(nx, ny, label, i + 1, j)) +
    (!solidOrOutside(nx, ny, label, i - 1, j)) + (!solidOrOutside(nx, ny, label, i, j + 1)) +
    (!solidOrOutside(nx, ny, label, i, j - 1));
  const map = new Int32Array(N).fill(-1);
  const cell = [];
  for (let k = 0; k < N; k++) {
    const i = k % nx, j = (k / nx) | 0;
    if (label[k] === FLUID && count(i, j) > 0) { map[k] = cell.length; cell.push(k); }
  }
  const n = cell.length;
  const comp = new Int32Array(n).fill(-1);
  const neumann = [], size = [];
  const di = [1, -1, 0, 0], dj = [0, 0, 1, -1];
  for (let a = 0; a < n; a++) {
    if (comp[a] >= 0) continue;
    const c = neumann.length;
    let anchored = false, sz = 0;
    const stack = [a];
    comp[a] = c;
    while (stack.length) {
      const b = stack.pop();
      sz++;
      const k = cell[b], i = k % nx, j = (k / nx) | 0;
      for (let t = 0; t < 4; t++) {
        const ii = i + di[t], jj = j + dj[t];
        if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
        const kk = ii + jj * nx;
        if (label[kk] === AIR) anchored = true;
        const m = map[kk];
        if (m >= 0 && comp[m] < 0) { comp[m] = c; stack.push(m); }
      }
    }
    neumann.push(!anchored);
    size.push(sz);
  }
  const A = new Float64Array(n * n);
  for (let a = 0; a < n; a++) {
    const k = cell[a], i = k % nx, j = (k / nx) | 0;
    A[a * n + a] = count(i, j);
    for (let t = 0; t < 4; t++) {
      const ii = i + di[t], jj = j + dj[t];
      if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
      const m = map[ii + jj * nx];
      if (m >= 0) A[a * n + m] -= 1.0;
    }
  }
  for (let a = 0; a < n; a++)
    for (let b = 0; b < n; b++)
      if (comp[a] === comp[b] && neumann[comp[a]]) A[a * n + b] += 1.0 / size[comp[a]];
  for (let c = 0; c < n; c++) {  // Cholesky, lower triangle in place
    let s = A[c * n + c];
    for (let t = 0; t < c; t++) s -= A[c * n + t] * A[c * n + t];
    const d = Math.sqrt(s);
    A[c * n + c] = d;
    for (let r = c + 1; r < n; r++) {
      let q = A[r * n + c];
      for (let t = 0; t < c; t++) q -= A[r * n + t] * A[c * n + t];
      A[r * n + c] = q / d;
    }
  }
  return { n, cell, comp, neumann, size, L: A };
}

function coarseApply(nx, ny, cf, b, e) {
  e.fill(0);
  const { n, cell, comp, neumann, size, L } = cf;
  if (n === 0) return;
  const y = new Float64Array(n);
  for (let r = 0; r < n; r++) {
    let s = b[cell[r]];
    for (let t = 0; t < r; t++) s -= L[r * n + t] * y[t];
    y[r] = s / L[r * n + r];
  }
  for (let r = n - 1; r >= 0; r--) {
    let s = y[r];
    for (let t = r + 1; t < n; t++) s -= L[t * n + r] * y[t];
    y[r] = s / L[r * n + r];
  }
  const mean = new Float64Array(neumann.length);
  for (let a = 0; a < n; a++) if (neumann[comp[a]]) mean[comp[a]] += y[a];
  for (let a = 0; a < n; a++) {
    const c = comp[a];
    let v = y[a];
    if (neumann[c]) v -= mean[c] / size[c];
    e[cell[a]] = v;
  }
}

/** CpuMultigridSolver (pressureMultigrid.ts:219-320) in mode 0; the build's
 *  defect correction with a direct coarsest solve in mode 1. */
class Multigrid {
  constructor(nx, ny, mode) {
    this.mode = mode;
    this.levels = levelSizes(nx, ny, mode).map(([lx, ly]) => ({
      nx: lx, ny: ly, label: new Uint8Array(lx * ly), x: new Float64Array(lx * ly),
      b: new Float64Array(lx * ly), r: new Float64Array(lx * ly),
    }));
    this.labelKey = null;
  }

  /* coarsenLabels chain (pressureMultigrid.ts:302-311), redone on a change */
  setLabels(label) {
    const L0 = this.levels[0];
    let same = this.labelKey !== null;
    if (same) for (let k = 0; k < label.length && same; k++) same = L0.label[k] === label[k];
    if (same) return;
    L0.label.set(label);
    for (let l = 1; l < this.levels.length; l++) {
      const F = this.levels[l - 1], C = this.levels[l];
      coarsenLabels(F.nx, F.ny, F.label, C.nx, C.ny, C.label);
    }
    if (this.mode === 1) {
      const C = this.levels[this.levels.length - 1];
      this.cf = coarseFactor(C.nx, C.ny, C.label);
    }
    this.labelKey = true;
  }

  /* pressureMultigrid.ts:253-275 */
  vcycle(l, x) {
    const L = this.levels[l];
    if (l === this.levels.length - 1) {
      if (this.mode === 0) {
        for (let k = 0; k < MG_COARSE_SWEEPS; k++) {
          rbSweep(L.nx, L.ny, x, L.b, L.label, -1, 1.0, 0);
          rbSweep(L.nx, L.ny, x, L.b, L.label, -1, 1.0, 1);
        }
      } else {
        coarseApply(L.nx, L.ny, this.cf, L.b, x);
      }
      return;
    }
    for (let k = 0; k < MG_PRE; k++) {
      rbSweep(L.nx, L.ny, x, L.b, L.label, -1, 1.0, 0);
      rbSweep(L.nx, L.ny, x, L.b, L.label, -1, 1.0, 1);
    }
    computeResidual(L.nx, L.ny, x, L.b, L.label, L.r);
    const N = this.levels[l + 1];
    restrictResidual(L.nx, L.ny, L.r, N.nx, N.ny, N.b, N.x);
    this.vcycle(l + 1, N.x);
    prolongAdd(L.nx, L.ny, x, L.label, N.nx, N.ny, N.x);
    for (let k = 0; k < MG_POST; k++) {
      rbSweep(L.nx, L.ny, x, L.b, L.label, -1, 1.0, 0);
      rbSweep(L.nx, L.ny, x, L.b, L.label, -1, 1.0, 1);
    }
  }

  /* pressureMultigrid.ts:284-319 (mode 0) / defect correction (mode 1) */
  solve(nx, ny, p, div, label, scale, cycles, tol) {
    const divRms = fluidDivRms(div, label);
    if (divRms === 0) { p.fill(0); return 0; }
    this.setLabels(label);
    const L0 = this.levels[0];
    const b = new Float64Array(nx * ny);
    for (let k = 0; k < b.length; k++) b[k] = -scale * div[k];
    if (this.mode === 0) L0.b.set(b);
    const target = tol > 0 ? tol * divRms : -1;
    for (let c = 0; c < cycles; c++) {
      if (this.mode === 0) {
        this.vcycle(0, p);
      } else {
        const rr = new Float64Array(nx * ny);
        computeResidual(nx, ny, p, b, L0.label, rr);
        L0.b.set(rr);
        L0.x.fill(0);
        this.vcycle(0, L0.x);
        for (let k = 0; k < p.length; k++) p[k] += L0.x[k];
      }
      if (target >= 0 && rmsRemainingDivergence(nx, ny, p, div, label, scale) <= target)
        return c + 1;
    }
    return cycles;
  }
}

/* subtractGradient.ts:40-61 */
function subtractGradient(nx, ny, p, u, v, label, gradScale) {
  for (let j = 0; j < ny; j++)
    for (let i = 1; i < nx; i++) {
      if (solidOrOutside(nx, ny, label, i - 1, j) || solidOrOutside(nx, ny, label, i, j)) continue;
      u[i + j * (nx + 1)] -= gradScale * (p[i + j * nx] - p[i - 1 + j * nx]);
    }
  for (let j = 1; j < ny; j++)
    for (let i = 0; i < nx; i++) {
      if (solidOrOutside(nx, ny, label, i, j - 1) || solidOrOutside(nx, ny, label, i, j)) continue;
      v[i + j * nx] -= gradScale * (p[i + j * nx] - p[i + (j - 1) * nx]);
    }
}

/* boundaries.ts:115-143 */
function applyOutflow(nx, ny, u, v, label) {
  const su = nx + 1;
  for (let j = 0; j < ny; j++) {
    if (label[j * nx] === AIR) {
      const out = dmin(u[1 + j * su], 0.0);
      u[1 + j * su] = out;
      u[j * su] = out;
    }
    if (label[nx - 1 + j * nx] === AIR) {
      const out = dmax(u[nx - 1 + j * su], 0.0);
      u[nx - 1 + j * su] = out;
      u[nx + j * su] = out;
    }
  }
  for (let i = 0; i < nx; i++) {
    if (label[i] === AIR) {
      const out = dmin(v[i + nx], 0.0);
      v[i + nx] = out;
      v[i] = out;
    }
    if (label[i + (ny - 1) * nx] === AIR) {
      const out = dmax(v[i + (ny - 1) * nx], 0.0);
      v[i + (ny - 1) * nx] = out;
      v[i + ny * nx] = out;
    }
  }
}

/* buoyancy.ts:70-88 */
function applyBuoyancy(nx, ny, h, dnx, dny, dh, v, heat, label, w, dt) {
  if (w === 0) return;
  const gain = w * dt;
  for (let j = 1; j < ny; j++) {
    const y = j * h;
    for (let i = 0; i < nx; i++) {
      if (isSolidCell(nx, ny, label, i, j - 1) || isSolidCell(nx, ny, label, i, j)) continue;
      v[i + j * nx] += gain * sampleP(dnx, dny, dh, heat, (i + 0.5) * h, y);
    }
  }
}

/* dye.ts:70-87 */
function applyDyePatch(dnx, dny, dye, patch) {
  const cells = patch.nx * patch.ny, d = patch.data;
  for (let j = 0; j < patch.ny; j++) {
    const dj = patch.j0 + j;
    if (dj < 0 || dj >= dny) continue;
    for (let i = 0; i < patch.nx; i++) {
      const di2 = patch.i0 + i;
      if (di2 < 0 || di2 >= dnx) continue;
      const s = i + j * patch.nx;
      const cov = d[3 * cells + s];
      if (cov === 0) continue;
      const k = di2 + dj * dnx;
      for (let c = 0; c < dye.length; c++) dye[c][k] += cov * (d[c * cells + s] - dye[c][k]);
    }
  }
}

/* vorticity.ts:9-42 */
const iclamp = (a, lo, hi) => (a < lo ? lo : a > hi ? hi : a);
function curlAtCorner(nx, h, u, v, i, j) {
  return (v[i + j * nx] - v[i - 1 + j * nx]) / h - (u[i + j * (nx + 1)] - u[i + (j - 1) * (nx + 1)]) / h;
}
function computeVorticity(nx, ny, h, u, v, out) {
  const hiI = nx - 1, hiJ = ny - 1;
  for (let j = 0; j < ny; j++) {
    const j0 = iclamp(j, 1, hiJ), j1 = iclamp(j + 1, 1, hiJ);
    for (let i = 0; i < nx; i++) {
      const i0 = iclamp(i, 1, hiI), i1 = iclamp(i + 1, 1, hiI);
      out[i + j * nx] = 0.25 * (curlAtCorner(nx, h, u, v, i0, j0) + curlAtCorner(nx, h, u, v, i1, j0) +
        curlAtCorner(nx, h, u, v, i0, j1) + curlAtCorner(nx, h, u, v, i1, j1));
    }
  }
}

/* Build-defined vorticity confinement (the C oracle's conf_force). */
const CF = { x: 0, y: 0 };
function confForce(nx, ny, h, w, eps, i, j) {
  const ip = i + 1 < nx ? i + 1 : nx - 1, im = i - 1 > 0 ? i - 1 : 0;
  const jp = j + 1 < ny ? j + 1 : ny - 1, jm = j - 1 > 0 ? j - 1 : 0;
  const a = Math.abs(w[i + j * nx]);
  const gx = (Math.abs(w[ip + j * nx]) - Math.abs(w[im + j * nx])) / ((ip - im) * h);
  const gy = (Math.abs(w[i + jp * nx]) - Math.abs(w[i + jm * nx])) / ((jp - jm) * h);
  const len = Math.sqrt(gx * gx + gy * gy);
  const inv = 1.0 / (len + 1e-4 * a / h + 1e-30);
  const om = w[i + j * nx];
  CF.x = eps * h * ((gy * inv) * om);
  CF.y = eps * h * (-(gx * inv) * om);
}
function vorticityConfinement(nx, ny, h, u, v, label, w, eps, dt) {
  if (eps === 0) return;
  for (let j = 0; j < ny; j++)
    for (let i = 1; i < nx; i++) {
      if (isSolidCell(nx, ny, label, i - 1, j) || isSolidCell(nx, ny, label, i, j)) continue;
      confForce(nx, ny, h, w, eps, i - 1, j);
      const fa = CF.x;
      confForce(nx, ny, h, w, eps, i, j);
      u[i + j * (nx + 1)] += dt * (0.5 * (fa + CF.x));
    }
  for (let j = 1; j < ny; j++)
    for (let i = 0; i < nx; i++) {
      if (isSolidCell(nx, ny, label, i, j - 1) || isSolidCell(nx, ny, label, i, j)) continue;
      confForce(nx, ny, h, w, eps, i, j - 1);
      const fa = CF.y;
      confForce(nx, ny, h, w, eps, i, j);
      v[i + j * nx] += dt * (0.5 * (fa + CF.y));
    }
}

/* simulation.ts:451-459 */
function maxFaceSpeed(u, v) {
  let m = 0;
  for (let k = 0; k < u.length; k++) m = dmax(m, Math.abs(u[k]));
  for (let k = 0; k < v.length; k++) m = dmax(m, Math.abs(v[k]));
  return m;
}

const SOLVERS = { sor: 0, rbsor: 1, mg: 2, 'mg-dc': 3 };

/** The simulation state of the CPU path (simulation.ts:259-386). */
function createSim(nx, ny, dyeScale = 1, h = 1 / ny) {
  const dnx = nx * dyeScale, dny = ny * dyeScale, nd = dnx * dny;
  const nu = (nx + 1) * ny, nv = nx * (ny + 1), nc = nx * ny;
  return {
    nx, ny, h, dnx, dny, dh: h / dyeScale, steps: 0,
    u: new Float64Array(nu), v: new Float64Array(nv), p: new Float64Array(nc),
    label: new Uint8Array(nc), dye: [0, 1, 2].map(() => new Float64Array(nd)),
    uNext: new Float64Array(nu), vNext: new Float64Array(nv), uHat: new Float64Array(nu),
    vHat: new Float64Array(nv), dyeNext: [0, 1, 2].map(() => new Float64Array(nd)),
    qHat: new Float64Array(nd), div: new Float64Array(nc), work: new Float64Array(nc),
    mg: null,
  };
}

/** Simulation.step, CPU path (simulation.ts:467-587), the C oracle's order. */
function step(s, prm) {
  const { nx, ny, h, dnx, dny, dh } = s;
  const uMax = maxFaceSpeed(s.u, s.v);
  const dt = uMax > 1e-9 ? dmin(prm.dtMax, prm.cflTarget * h / uMax) : prm.dtMax;
  const scale = prm.rho * h * h / dt;
  const gradScale = dt / (prm.rho * h);
  if (prm.scheme === 'macCormack')
    advectVelocityMacCormack(nx, ny, h, s.u, s.v, s.uHat, s.vHat, s.uNext, s.vNext, s.label, dt);
  else advectFaces(nx, ny, h, s.u, s.v, s.u, s.v, s.uNext, s.vNext, s.label, dt);
  [s.u, s.uNext] = [s.uNext, s.u];
  [s.v, s.vNext] = [s.vNext, s.v];
  if (prm.confinement !== 0) computeVorticity(nx, ny, h, s.u, s.v, s.work);
  applyBuoyancy(nx, ny, h, dnx, dny, dh, s.v, s.dye[0], s.label, prm.buoyancy, dt);
  if (prm.confinement !== 0)
    vorticityConfinement(nx, ny, h, s.u, s.v, s.label, s.work, prm.confinement, dt);
  divergence(nx, ny, h, s.u, s.v, s.div);
  let iters;
  const solver = SOLVERS[prm.solver];
  if (solver === 0) {
    iters = solvePressure(nx, ny, s.p, s.div, s.label, scale, prm.pressureIters, prm.omega, prm.tol);
  } else if (solver === 1) {
    iters = solvePressureRedBlack(nx, ny, s.p, s.div, s.label, scale, prm.pressureIters, prm.omega,
      prm.tol);
  } else {
    const mode = solver === 2 ? 0 : 1;
    if (!s.mg || s.mg.mode !== mode) s.mg = new Multigrid(nx, ny, mode);
    iters = s.mg.solve(nx, ny, s.p, s.div, s.label, scale, prm.pressureIters, prm.tol);
  }
  subtractGradient(nx, ny, s.p, s.u, s.v, s.label, gradScale);
  applyOutflow(nx, ny, s.u, s.v, s.label);
  if (prm.residualEvery > 0 && s.steps % prm.residualEvery === 0)
    divergence(nx, ny, h, s.u, s.v, s.div);
  for (let c = 0; c < 3; c++) {
    if (prm.scheme === 'macCormack')
      advectScalarMacCormack(nx, ny, h, dnx, dny, dh, s.u, s.v, s.dye[c], s.qHat, s.dyeNext[c],
        s.label, dt);
    else advectScalar(nx, ny, h, dnx, dny, dh, s.u, s.v, s.dye[c], s.dyeNext[c], s.label, dt);
  }
  [s.dye, s.dyeNext] = [s.dyeNext, s.dye];
  for (let c = 0; c < 3; c++) {  /* simulation.ts:243-247, 569-576 */
    const d = prm.decay[c];
    const keep = d > 0 ? Math.exp(-d * dt) : 1.0;
    if (keep === 1.0) continue;
    const q = s.dye[c];
    for (let k = 0; k < q.length; k++) q[k] *= keep;
  }
  if (prm.patch) applyDyePatch(dnx, dny, s.dye, prm.patch);
  s.steps++;
  return { dt, iters };
}

module.exports = { createSim, step, sampleU, sampleV, sampleP, divergence, Multigrid,
  rmsRemainingDivergence, fluidDivRms };

if (require.main === module) {
  const fs = require('fs');
  const os = require('os');
  const path = require('path');
  const { performance } = require('perf_hooks');
  const [dir, warmupArg, stepsArg] = process.argv.slice(2);
  const c = JSON.parse(fs.readFileSync(path.join(dir, 'case.json'), 'utf8'));
  const rd = (name) => {
    const b = fs.readFileSync(path.join(dir, name));
    return new Float64Array(b.buffer, b.byteOffset, b.length / 8).slice();
  };
  const s = createSim(c.nx, c.ny, c.dyeScale || 1, c.h);
  s.u.set(rd('u.f64'));
  s.v.set(rd('v.f64'));
  s.p.set(rd('p.f64'));
  s.label.set(fs.readFileSync(path.join(dir, 'label.u8')));
  const dye = rd('dye.f64');
  const nd = s.dnx * s.dny;
  for (let k = 0; k < 3; k++) s.dye[k].set(dye.subarray(k * nd, (k + 1) * nd));
  let patch = null;
  if (c.patch) {
    const b = fs.readFileSync(path.join(dir, 'patch.f32'));
    patch = Object.assign({}, c.patch,
      { data: new Float32Array(b.buffer, b.byteOffset, b.length / 4).slice() });
  }
  const prm = {
    scheme: c.scheme, solver: c.solver, pressureIters: c.pressureIters, omega: c.omega || 1,
    tol: c.tol, cflTarget: c.cflTarget, dtMax: c.dtMax, rho: 1, buoyancy: c.buoyancy || 0,
    confinement: c.confinement || 0, decay: c.decay || [0, 0, 0], residualEvery: 0, patch,
  };
  const warmup = Number(warmupArg || 0), steps = Number(stepsArg || 1);
  const iters = [];
  for (let k = 0; k < warmup; k++) step(s, prm);
  const t0 = performance.now();
  for (let k = 0; k < steps; k++) iters.push(step(s, prm).iters);
  const secs = (performance.now() - t0) / 1e3;
  const wr = (name, a) => fs.writeFileSync(path.join(dir, name), Buffer.from(a.buffer, a.byteOffset,
    a.byteLength));
  wr('out_u.f64', s.u);
  wr('out_v.f64', s.v);
  wr('out_p.f64', s.p);
  const out = new Float64Array(3 * nd);
  for (let k = 0; k < 3; k++) out.set(s.dye[k], k * nd);
  wr('out_dye.f64', out);
  const cpu = os.cpus()[0];
  console.log(JSON.stringify({
    seconds: secs, steps, warmup, iters, cell_steps_per_s: (c.nx * c.ny * steps) / secs,
    node: process.version, cpus: os.cpus().length, cpu_model: cpu ? cpu.model : null,
  }));
}
