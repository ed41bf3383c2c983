/*
 * FP64 CPU oracle for flowlab's projection-method step.  TEST INFRASTRUCTURE
 * ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, never by the product path.  See oracle.h for the contract.
 *
 * Built with -ffp-contract=off so no multiply-add is fused: the reference is
 * JavaScript, whose arithmetic is IEEE double with no contraction.
 *
 * Parity pins (tests/test_oracle_*.py): every function below is checked
 * against the closed forms and invariants the reference's own tests assert,
 * plus cross-solver agreement; see each function's comment.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int solid_or_outside(int nx, int ny, const uint8_t* l, int i, int j) {
  /* grid.ts:159-161 -- the bounds test runs before the lookup */
  return i < 0 || j < 0 || i >= nx || j >= ny || l[i + j * nx] == FO_SOLID;
}

static inline int is_solid_cell(int nx, int ny, const uint8_t* l, int i, int j) {
  /* boundaries.ts:23-25 */
  return i >= 0 && j >= 0 && i < nx && j < ny && l[i + j * nx] == FO_SOLID;
}

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* grid.ts:173-177 */
static inline void clamped_axis(double pos, int count, int* i0, double* f) {
  double c = dmin(dmax(pos, 0.0), (double)(count - 1));
  int i = (int)floor(c);
  if (i > count - 2) i = count - 2;
  *i0 = i;
  *f = c - (double)i;
}

/* grid.ts:180-191 */
static inline double bilerp(double v00, double v10, double v01, double v11, double fx, double fy) {
  double lo = v00 + fx * (v10 - v00);
  double hi = v01 + fx * (v11 - v01);
  return lo + fy * (hi - lo);
}

/* grid.ts:197-208 */
double fo_sample_u(int nx, int ny, double h, const double* u, double x, double y) {
  int i0, j0;
  double fx, fy;
  clamped_axis(x / h, nx + 1, &i0, &fx);
  clamped_axis(y / h - 0.5, ny, &j0, &fy);
  int s = nx + 1;
  return bilerp(u[i0 + j0 * s], u[i0 + 1 + j0 * s], u[i0 + (j0 + 1) * s],
                u[i0 + 1 + (j0 + 1) * s], fx, fy);
}

/* grid.ts:214-225 */
double fo_sample_v(int nx, int ny, double h, const double* v, double x, double y) {
  int i0, j0;
  double fx, fy;
  clamped_axis(x / h - 0.5, nx, &i0, &fx);
  clamped_axis(y / h, ny + 1, &j0, &fy);
  return bilerp(v[i0 + j0 * nx], v[i0 + 1 + j0 * nx], v[i0 + (j0 + 1) * nx],
                v[i0 + 1 + (j0 + 1) * nx], fx, fy);
}

/* grid.ts:239-250 */
double fo_sample_p(int nx, int ny, double h, const double* q, double x, double y) {
  int i0, j0;
  double fx, fy;
  clamped_axis(x / h - 0.5, nx, &i0, &fx);
  clamped_axis(y / h - 0.5, ny, &j0, &fy);
  return bilerp(q[i0 + j0 * nx], q[i0 + 1 + j0 * nx], q[i0 + (j0 + 1) * nx],
                q[i0 + 1 + (j0 + 1) * nx], fx, fy);
}

/* grid.ts:268-270: min(max(q, min(a,b,c,d)), max(a,b,c,d)) */
static inline double clamp_to_range(double q, double a, double b, double c, double d) {
  double lo = dmin(dmin(a, b), dmin(c, d));
  double hi = dmax(dmax(a, b), dmax(c, d));
  return dmin(dmax(q, lo), hi);
}

/* grid.ts:273-283 */
static double clamp_to_stencil_u(int nx, int ny, double h, const double* u, double x, double y,
                                 double q) {
  int i0, j0;
  double fx, fy;
  clamped_axis(x / h, nx + 1, &i0, &fx);
  clamped_axis(y / h - 0.5, ny, &j0, &fy);
  int s = nx + 1;
  return clamp_to_range(q, u[i0 + j0 * s], u[i0 + 1 + j0 * s], u[i0 + (j0 + 1) * s],
                        u[i0 + 1 + (j0 + 1) * s]);
}

/* grid.ts:286-296 */
static double clamp_to_stencil_v(int nx, int ny, double h, const double* v, double x, double y,
                                 double q) {
  int i0, j0;
  double fx, fy;
  clamped_axis(x / h - 0.5, nx, &i0, &fx);
  clamped_axis(y / h, ny + 1, &j0, &fy);
  return clamp_to_range(q, v[i0 + j0 * nx], v[i0 + 1 + j0 * nx], v[i0 + (j0 + 1) * nx],
                        v[i0 + 1 + (j0 + 1) * nx]);
}

/* grid.ts:299-315 */
static double clamp_to_stencil_p(int nx, int ny, double h, const double* q, double x, double y,
                                 double value) {
  int i0, j0;
  double fx, fy;
  clamped_axis(x / h - 0.5, nx, &i0, &fx);
  clamped_axis(y / h - 0.5, ny, &j0, &fy);
  return clamp_to_range(value, q[i0 + j0 * nx], q[i0 + 1 + j0 * nx], q[i0 + (j0 + 1) * nx],
                        q[i0 + 1 + (j0 + 1) * nx]);
}

/* advect.ts:41-55: RK2 midpoint backtrace */
static inline void backtrace(int nx, int ny, double h, const double* uCar, const double* vCar,
                             double x, double y, double u0, double v0, double dt, double* bx,
                             double* by) {
  double midX = x - 0.5 * dt * u0;
  double midY = y - 0.5 * dt * v0;
  *bx = x - dt * fo_sample_u(nx, ny, h, uCar, midX, midY);
  *by = y - dt * fo_sample_v(nx, ny, h, vCar, midX, midY);
}

/* advect.ts:72-112 */
static void advect_faces(int nx, int ny, double h, const double* uCar, const double* vCar,
                         const double* uSrc, const double* vSrc, double* uOut, double* vOut,
                         const uint8_t* label, double dt) {
  double bx, by;
  for (int j = 0; j < ny; j++) {
    double y = (j + 0.5) * h;
    for (int i = 0; i <= nx; i++) {
      int k = i + j * (nx + 1);
      if (solid_or_outside(nx, ny, label, i - 1, j) || solid_or_outside(nx, ny, label, i, j)) {
        uOut[k] = uSrc[k];
        continue;
      }
      double x = i * h;
      backtrace(nx, ny, h, uCar, vCar, x, y, uCar[k], fo_sample_v(nx, ny, h, vCar, x, y), dt,
                &bx, &by);
      uOut[k] = fo_sample_u(nx, ny, h, uSrc, bx, by);
    }
  }
  for (int j = 0; j <= ny; j++) {
    double y = j * h;
    for (int i = 0; i < nx; i++) {
      int k = i + j * nx;
      if (solid_or_outside(nx, ny, label, i, j - 1) || solid_or_outside(nx, ny, label, i, j)) {
        vOut[k] = vSrc[k];
        continue;
      }
      double x = (i + 0.5) * h;
      backtrace(nx, ny, h, uCar, vCar, x, y, fo_sample_u(nx, ny, h, uCar, x, y), vCar[k], dt,
                &bx, &by);
      vOut[k] = fo_sample_v(nx, ny, h, vSrc, bx, by);
    }
  }
}

/* advect.ts:126-136 */
void fo_advect_velocity(int nx, int ny, double h, const double* uIn, const double* vIn,
                        double* uOut, double* vOut, const uint8_t* label, double dt) {
  advect_faces(nx, ny, h, uIn, vIn, uIn, vIn, uOut, vOut, label, dt);
}

/* advect.ts:166-211 */
void fo_advect_velocity_maccormack(int nx, int ny, double h, const double* uIn,
                                   const double* vIn, double* uHat, double* vHat,
                                   double* uOut, double* vOut, const uint8_t* label,
                                   double dt) {
  advect_faces(nx, ny, h, uIn, vIn, uIn, vIn, uHat, vHat, label, dt);
  advect_faces(nx, ny, h, uIn, vIn, uHat, vHat, uOut, vOut, label, -dt);
  double bx, by;
  for (int j = 0; j < ny; j++) {
    double y = (j + 0.5) * h;
    for (int i = 0; i <= nx; i++) {
      int k = i + j * (nx + 1);
      if (solid_or_outside(nx, ny, label, i - 1, j) || solid_or_outside(nx, ny, label, i, j)) {
        uOut[k] = uIn[k];
        continue;
      }
      double x = i * h;
      double corrected = uHat[k] + 0.5 * (uIn[k] - uOut[k]);
      backtrace(nx, ny, h, uIn, vIn, x, y, uIn[k], fo_sample_v(nx, ny, h, vIn, x, y), dt, &bx,
                &by);
      uOut[k] = clamp_to_stencil_u(nx, ny, h, uIn, bx, by, corrected);
    }
  }
  for (int j = 0; j <= ny; j++) {
    double y = j * h;
    for (int i = 0; i < nx; i++) {
      int k = i + j * nx;
      if (solid_or_outside(nx, ny, label, i, j - 1) || solid_or_outside(nx, ny, label, i, j)) {
        vOut[k] = vIn[k];
        continue;
      }
      double x = (i + 0.5) * h;
      double corrected = vHat[k] + 0.5 * (vIn[k] - vOut[k]);
      backtrace(nx, ny, h, uIn, vIn, x, y, fo_sample_u(nx, ny, h, uIn, x, y), vIn[k], dt, &bx,
                &by);
      vOut[k] = clamp_to_stencil_v(nx, ny, h, vIn, bx, by, corrected);
    }
  }
}

/* advect.ts:241-270 */
void fo_advect_scalar(int nx, int ny, double h, int dnx, int dny, double dh, const double* u,
                      const double* v, const double* qIn, double* qOut, const uint8_t* label,
                      double dt) {
  double ratio = dh / h;
  double bx, by;
  for (int j = 0; j < dny; j++) {
    double y = (j + 0.5) * dh;
    int lj = (int)floor((j + 0.5) * ratio);
    for (int i = 0; i < dnx; i++) {
      int k = i + j * dnx;
      if (solid_or_outside(nx, ny, label, (int)floor((i + 0.5) * ratio), lj)) {
        qOut[k] = qIn[k];
        continue;
      }
      double x = (i + 0.5) * dh;
      backtrace(nx, ny, h, u, v, x, y, fo_sample_u(nx, ny, h, u, x, y),
                fo_sample_v(nx, ny, h, v, x, y), dt, &bx, &by);
      qOut[k] = fo_sample_p(dnx, dny, dh, qIn, bx, by);
    }
  }
}

/* advect.ts:280-312 */
void fo_advect_scalar_maccormack(int nx, int ny, double h, int dnx, int dny, double dh,
                                 const double* u, const double* v, const double* qIn,
                                 double* qHat, double* qOut, const uint8_t* label, double dt) {
  fo_advect_scalar(nx, ny, h, dnx, dny, dh, u, v, qIn, qHat, label, dt);
  fo_advect_scalar(nx, ny, h, dnx, dny, dh, u, v, qHat, qOut, label, -dt);
  double ratio = dh / h;
  double bx, by;
  for (int j = 0; j < dny; j++) {
    double y = (j + 0.5) * dh;
    int lj = (int)floor((j + 0.5) * ratio);
    for (int i = 0; i < dnx; i++) {
      int k = i + j * dnx;
      if (solid_or_outside(nx, ny, label, (int)floor((i + 0.5) * ratio), lj)) {
        qOut[k] = qIn[k];
        continue;
      }
      double x = (i + 0.5) * dh;
      double corrected = qHat[k] + 0.5 * (qIn[k] - qOut[k]);
      backtrace(nx, ny, h, u, v, x, y, fo_sample_u(nx, ny, h, u, x, y),
                fo_sample_v(nx, ny, h, v, x, y), dt, &bx, &by);
      qOut[k] = clamp_to_stencil_p(dnx, dny, dh, qIn, bx, by, corrected);
    }
  }
}

/* divergence.ts:17-25 */
void fo_divergence(int nx, int ny, double h, const double* u, const double* v, double* out) {
  double coef = 1.0 / h;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      out[i + j * nx] = coef * (u[i + 1 + j * (nx + 1)] - u[i + j * (nx + 1)] +
                                v[i + (j + 1) * nx] - v[i + j * nx]);
}

/* pressure.ts:40-91.  Neighbour order +x, +y, -x, -y. */
void fo_gs_sweep(int nx, int ny, double* p, const double* div, const uint8_t* label,
                 double scale, double omega, int reverseX, int reverseY) {
  int iStep = reverseX ? -1 : 1, iFrom = reverseX ? nx - 1 : 0, iTo = reverseX ? -1 : nx;
  int jStep = reverseY ? -1 : 1, jFrom = reverseY ? ny - 1 : 0, jTo = reverseY ? -1 : ny;
  for (int j = jFrom; j != jTo; j += jStep) {
    for (int i = iFrom; i != iTo; i += iStep) {
      int k = i + j * nx;
      if (label[k] != FO_FLUID) continue;
      int count = 0;
      double sum = 0;
      if (!solid_or_outside(nx, ny, label, i + 1, j)) { count++; sum += p[k + 1]; }
      if (!solid_or_outside(nx, ny, label, i, j + 1)) { count++; sum += p[k + nx]; }
      if (!solid_or_outside(nx, ny, label, i - 1, j)) { count++; sum += p[k - 1]; }
      if (!solid_or_outside(nx, ny, label, i, j - 1)) { count++; sum += p[k - nx]; }
      if (count == 0) continue;
      double pGS = (sum - scale * div[k]) / count;
      p[k] = (1 - omega) * p[k] + omega * pGS;
    }
  }
}

/* pressure.ts:102-142 */
double fo_rms_remaining_divergence(int nx, int ny, const double* p, const double* div,
                                   const uint8_t* label, double scale) {
  double sumSq = 0;
  long n = 0;
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int k = i + j * nx;
      if (label[k] != FO_FLUID) continue;
      int count = 0;
      double sum = 0;
      if (!solid_or_outside(nx, ny, label, i + 1, j)) { count++; sum += p[k + 1]; }
      if (!solid_or_outside(nx, ny, label, i, j + 1)) { count++; sum += p[k + nx]; }
      if (!solid_or_outside(nx, ny, label, i - 1, j)) { count++; sum += p[k - 1]; }
      if (!solid_or_outside(nx, ny, label, i, j - 1)) { count++; sum += p[k - nx]; }
      if (count == 0) continue;
      double r = div[k] + (count * p[k] - sum) / scale;
      sumSq += r * r;
      n++;
    }
  }
  return n > 0 ? sqrt(sumSq / n) : 0;
}

/* pressure.ts:152-161 */
double fo_fluid_div_rms(int n, const double* div, const uint8_t* label) {
  double sumSq = 0;
  long m = 0;
  for (int k = 0; k < n; k++) {
    if (label[k] != FO_FLUID) continue;
    sumSq += div[k] * div[k];
    m++;
  }
  return m > 0 ? sqrt(sumSq / m) : 0;
}

/* pressure.ts:182-212 */
int fo_solve_pressure(int nx, int ny, double* p, const double* div, const uint8_t* label,
                      double scale, int iterations, double omega, double tol) {
  double divRms = fo_fluid_div_rms(nx * ny, div, label);
  if (divRms == 0) {
    memset(p, 0, sizeof(double) * nx * ny);
    return 0;
  }
  double target = tol > 0 ? tol * divRms : -1;
  for (int k = 0; k < iterations; k++) {
    fo_gs_sweep(nx, ny, p, div, label, scale, omega, k % 2 == 1, k % 4 >= 2);
    if (target >= 0 && k % 12 == 11) {
      if (fo_rms_remaining_divergence(nx, ny, p, div, label, scale) <= target) return k + 1;
    }
  }
  return iterations;
}

/* pressureRedBlack.ts:32-72.  Neighbour order +x, -x, +y, -y. */
void fo_rb_sweep(int nx, int ny, double* p, const double* div, const uint8_t* label,
                 double scale, double omega, int color) {
  for (int j = 0; j < ny; j++) {
    for (int i = (j & 1) ^ color; i < nx; i += 2) {
      int k = i + j * nx;
      if (label[k] != FO_FLUID) continue;
      int count = 0;
      double sum = 0;
      if (!solid_or_outside(nx, ny, label, i + 1, j)) { count++; sum += p[k + 1]; }
      if (!solid_or_outside(nx, ny, label, i - 1, j)) { count++; sum += p[k - 1]; }
      if (!solid_or_outside(nx, ny, label, i, j + 1)) { count++; sum += p[k + nx]; }
      if (!solid_or_outside(nx, ny, label, i, j - 1)) { count++; sum += p[k - nx]; }
      if (count == 0) continue;
      double pGS = (sum - scale * div[k]) / count;
      p[k] = (1 - omega) * p[k] + omega * pGS;
    }
  }
}

/* pressureRedBlack.ts:88-113 */
int fo_solve_pressure_rb(int nx, int ny, double* p, const double* div, const uint8_t* label,
                         double scale, int iterations, double omega, double tol) {
  double divRms = fo_fluid_div_rms(nx * ny, div, label);
  if (divRms == 0) {
    memset(p, 0, sizeof(double) * nx * ny);
    return 0;
  }
  double target = tol > 0 ? tol * divRms : -1;
  for (int k = 0; k < iterations; k++) {
    fo_rb_sweep(nx, ny, p, div, label, scale, omega, 0);
    fo_rb_sweep(nx, ny, p, div, label, scale, omega, 1);
    if (target >= 0 && k % 12 == 11) {
      if (fo_rms_remaining_divergence(nx, ny, p, div, label, scale) <= target) return k + 1;
    }
  }
  return iterations;
}

/* ---- multigrid ---------------------------------------------------------- */

/* pressureMultigrid.ts:65-78 (mode 0: COARSEST = 4).  Mode 1 stops at 16. */
static int coarsest_side(int mode) { return mode == 0 ? 4 : 16; }

int fo_mg_level_count(int nx, int ny, int mode) {
  int c = coarsest_side(mode), n = 1;
  while (nx > c || ny > c) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
    n++;
  }
  return n;
}

void fo_mg_level_size(int nx, int ny, int mode, int level, int* lnx, int* lny) {
  (void)mode;
  for (int l = 0; l < level; l++) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
  }
  *lnx = nx;
  *lny = ny;
}

/* pressureMultigrid.ts:92-110: priority Air > Fluid > Solid */
void fo_coarsen_labels(int fnx, int fny, const uint8_t* fine, int cnx, int cny,
                       uint8_t* coarse) {
  for (int J = 0; J < cny; J++) {
    for (int I = 0; I < cnx; I++) {
      int air = 0, fluid = 0;
      for (int dj = 0; dj < 2; dj++) {
        for (int di = 0; di < 2; di++) {
          int i = 2 * I + di, j = 2 * J + dj;
          if (i >= fnx || j >= fny) continue;
          uint8_t l = fine[i + j * fnx];
          air = air || l == FO_AIR;
          fluid = fluid || l == FO_FLUID;
        }
      }
      coarse[I + J * cnx] = air ? FO_AIR : fluid ? FO_FLUID : FO_SOLID;
    }
  }
}

/* pressureMultigrid.ts:117-151.  Neighbour order +x, -x, +y, -y. */
void fo_mg_residual(int nx, int ny, const double* x, const double* b, const uint8_t* label,
                    double* r) {
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int k = i + j * nx;
      r[k] = 0;
      if (label[k] != FO_FLUID) continue;
      int count = 0;
      double sum = 0;
      if (!solid_or_outside(nx, ny, label, i + 1, j)) { count++; sum += x[k + 1]; }
      if (!solid_or_outside(nx, ny, label, i - 1, j)) { count++; sum += x[k - 1]; }
      if (!solid_or_outside(nx, ny, label, i, j + 1)) { count++; sum += x[k + nx]; }
      if (!solid_or_outside(nx, ny, label, i, j - 1)) { count++; sum += x[k - nx]; }
      if (count == 0) continue;
      r[k] = b[k] - (count * x[k] - sum);
    }
  }
}

/* pressureMultigrid.ts:156-178 */
void fo_mg_restrict(int fnx, int fny, const double* r, int cnx, int cny, double* bC,
                    double* eC) {
  for (int J = 0; J < cny; J++) {
    for (int I = 0; I < cnx; I++) {
      double sum = 0;
      for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
          int i = 2 * I + di, j = 2 * J + dj;
          if (i < fnx && j < fny) sum += r[i + j * fnx];
        }
      bC[I + J * cnx] = sum;
      eC[I + J * cnx] = 0;
    }
  }
}

/* pressureMultigrid.ts:192-207 */
void fo_mg_prolong_add(int fnx, int fny, double* x, const uint8_t* label, int cnx, int cny,
                       const double* e) {
  for (int j = 0; j < fny; j++) {
    int J = j >> 1;
    int J2 = (j & 1) ? (J + 1 < cny - 1 ? J + 1 : cny - 1) : (J - 1 > 0 ? J - 1 : 0);
    for (int i = 0; i < fnx; i++) {
      int k = i + j * fnx;
      if (label[k] != FO_FLUID) continue;
      int I = i >> 1;
      int I2 = (i & 1) ? (I + 1 < cnx - 1 ? I + 1 : cnx - 1) : (I - 1 > 0 ? I - 1 : 0);
      x[k] += 0.5625 * e[I + J * cnx] + 0.1875 * (e[I2 + J * cnx] + e[I + J2 * cnx]) +
              0.0625 * e[I2 + J2 * cnx];
    }
  }
}

/* BUILD-DEFINED: dense FP64 solve of the coarsest-level system
 *   count*x - sum(non-solid nbr x) = b   on fluid cells with count > 0,
 * Air neighbours contributing x = 0.  A fluid component that touches no Air
 * is pure Neumann and singular; it is regularised by (1/n) 1 1^T and the
 * component mean is removed afterwards, which yields the minimum-norm
 * least-squares solution A^+ b on that component. */
typedef struct {
  int n;        /* unknowns */
  int* cell;    /* unknown -> cell index */
  int* comp;    /* unknown -> component id */
  int ncomp;
  int* comp_neumann;
  int* comp_size;
  double* L;    /* Cholesky factor, n*n */
} coarse_fact;

static int coarse_count(int nx, int ny, const uint8_t* label, int i, int j) {
  int c = 0;
  c += !solid_or_outside(nx, ny, label, i + 1, j);
  c += !solid_or_outside(nx, ny, label, i - 1, j);
  c += !solid_or_outside(nx, ny, label, i, j + 1);
  c += !solid_or_outside(nx, ny, label, i, j - 1);
  return c;
}

static void coarse_factor(int nx, int ny, const uint8_t* label, coarse_fact* cf) {
  int N = nx * ny;
  int* map = (int*)malloc(sizeof(int) * N);
  cf->cell = (int*)malloc(sizeof(int) * N);
  cf->comp = (int*)malloc(sizeof(int) * N);
  cf->comp_neumann = (int*)malloc(sizeof(int) * N);
  cf->comp_size = (int*)malloc(sizeof(int) * N);
  int n = 0;
  for (int k = 0; k < N; k++) {
    map[k] = -1;
    int i = k % nx, j = k / nx;
    if (label[k] == FO_FLUID && coarse_count(nx, ny, label, i, j) > 0) {
      map[k] = n;
      cf->cell[n++] = k;
    }
  }
  cf->n = n;
  /* components over fluid-fluid adjacency; a component is Dirichlet-anchored
   * if any member has an Air neighbour */
  int* stack = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int a = 0; a < n; a++) cf->comp[a] = -1;
  cf->ncomp = 0;
  for (int a = 0; a < n; a++) {
    if (cf->comp[a] >= 0) continue;
    int c = cf->ncomp++;
    int sp = 0, size = 0, anchored = 0;
    stack[sp++] = a;
    cf->comp[a] = c;
    while (sp) {
      int b = stack[--sp];
      size++;
      int k = cf->cell[b], i = k % nx, j = k / nx;
      int di[4] = {1, -1, 0, 0}, dj[4] = {0, 0, 1, -1};
      for (int t = 0; t < 4; t++) {
        int ii = i + di[t], jj = j + dj[t];
        if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
        int kk = ii + jj * nx;
        if (label[kk] == FO_AIR) anchored = 1;
        int m = map[kk];
        if (m >= 0 && cf->comp[m] < 0) {
          cf->comp[m] = c;
          stack[sp++] = m;
        }
      }
    }
    cf->comp_neumann[c] = !anchored;
    cf->comp_size[c] = size;
  }
  free(stack);
  double* A = (double*)calloc((size_t)n * n, sizeof(double));
  for (int a = 0; a < n; a++) {
    int k = cf->cell[a], i = k % nx, j = k / nx;
    A[(size_t)a * n + a] = coarse_count(nx, ny, label, i, j);
    int di[4] = {1, -1, 0, 0}, dj[4] = {0, 0, 1, -1};
    for (int t = 0; t < 4; t++) {
      int ii = i + di[t], jj = j + dj[t];
      if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
      int m = map[ii + jj * nx];
      if (m >= 0) A[(size_t)a * n + m] -= 1.0;
    }
  }
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++)
      if (cf->comp[a] == cf->comp[b] && cf->comp_neumann[cf->comp[a]])
        A[(size_t)a * n + b] += 1.0 / cf->comp_size[cf->comp[a]];
  /* Cholesky, lower triangle in place */
  for (int c = 0; c < n; c++) {
    double s = A[(size_t)c * n + c];
    for (int t = 0; t < c; t++) s -= A[(size_t)c * n + t] * A[(size_t)c * n + t];
    double d = sqrt(s);
    A[(size_t)c * n + c] = d;
    for (int r = c + 1; r < n; r++) {
      double q = A[(size_t)r * n + c];
      for (int t = 0; t < c; t++) q -= A[(size_t)r * n + t] * A[(size_t)c * n + t];
      A[(size_t)r * n + c] = q / d;
    }
  }
  cf->L = A;
  free(map);
}

static void coarse_free(coarse_fact* cf) {
  free(cf->cell);
  free(cf->comp);
  free(cf->comp_neumann);
  free(cf->comp_size);
  free(cf->L);
}

static void coarse_apply(int nx, int ny, const coarse_fact* cf, const double* b, double* e) {
  int n = cf->n;
  const double* L = cf->L;
  memset(e, 0, sizeof(double) * nx * ny);
  if (n == 0) return;
  double* y = (double*)malloc(sizeof(double) * n);
  for (int r = 0; r < n; r++) {
    double s = b[cf->cell[r]];
    for (int t = 0; t < r; t++) s -= L[(size_t)r * n + t] * y[t];
    y[r] = s / L[(size_t)r * n + r];
  }
  for (int r = n - 1; r >= 0; r--) {
    double s = y[r];
    for (int t = r + 1; t < n; t++) s -= L[(size_t)t * n + r] * y[t];
    y[r] = s / L[(size_t)r * n + r];
  }
  double* mean = (double*)calloc(cf->ncomp, sizeof(double));
  for (int a = 0; a < n; a++)
    if (cf->comp_neumann[cf->comp[a]]) mean[cf->comp[a]] += y[a];
  for (int a = 0; a < n; a++) {
    int c = cf->comp[a];
    double v = y[a];
    if (cf->comp_neumann[c]) v -= mean[c] / cf->comp_size[c];
    e[cf->cell[a]] = v;
  }
  free(mean);
  free(y);
}

void fo_coarse_direct_solve(int nx, int ny, const uint8_t* label, const double* b, double* e) {
  coarse_fact cf;
  coarse_factor(nx, ny, label, &cf);
  coarse_apply(nx, ny, &cf, b, e);
  coarse_free(&cf);
}

typedef struct {
  int nx, ny;
  uint8_t* label;
  double *x, *b, *r;
} mg_level;

typedef struct {
  int mode, nlev;
  mg_level* lv;
  coarse_fact cf;
} mg_state;

static void mg_build(mg_state* s, int nx, int ny, const uint8_t* label, int mode) {
  s->mode = mode;
  s->nlev = fo_mg_level_count(nx, ny, mode);
  s->lv = (mg_level*)calloc(s->nlev, sizeof(mg_level));
  for (int l = 0; l < s->nlev; l++) {
    int lnx, lny;
    fo_mg_level_size(nx, ny, mode, l, &lnx, &lny);
    size_t n = (size_t)lnx * lny;
    mg_level* L = &s->lv[l];
    L->nx = lnx;
    L->ny = lny;
    L->label = (uint8_t*)malloc(n);
    L->x = (double*)calloc(n, sizeof(double));
    L->b = (double*)calloc(n, sizeof(double));
    L->r = (double*)calloc(n, sizeof(double));
    if (l == 0)
      memcpy(L->label, label, n);
    else
      fo_coarsen_labels(s->lv[l - 1].nx, s->lv[l - 1].ny, s->lv[l - 1].label, lnx, lny,
                        L->label);
  }
  if (mode == 1) {
    mg_level* C = &s->lv[s->nlev - 1];
    coarse_factor(C->nx, C->ny, C->label, &s->cf);
  }
}

static void mg_free(mg_state* s) {
  for (int l = 0; l < s->nlev; l++) {
    free(s->lv[l].label);
    free(s->lv[l].x);
    free(s->lv[l].b);
    free(s->lv[l].r);
  }
  free(s->lv);
  if (s->mode == 1) coarse_free(&s->cf);
}

/* pressureMultigrid.ts:47-63 */
enum { MG_PRE = 2, MG_POST = 2, MG_COARSE_SWEEPS = 8 };

/* pressureMultigrid.ts:253-275 */
static void mg_vcycle(mg_state* s, int l, double* x) {
  mg_level* L = &s->lv[l];
  if (l == s->nlev - 1) {
    if (s->mode == 0) {
      for (int k = 0; k < MG_COARSE_SWEEPS; k++) {
        fo_rb_sweep(L->nx, L->ny, x, L->b, L->label, -1, 1.0, 0);
        fo_rb_sweep(L->nx, L->ny, x, L->b, L->label, -1, 1.0, 1);
      }
    } else {
      coarse_apply(L->nx, L->ny, &s->cf, L->b, x);
    }
    return;
  }
  for (int k = 0; k < MG_PRE; k++) {
    fo_rb_sweep(L->nx, L->ny, x, L->b, L->label, -1, 1.0, 0);
    fo_rb_sweep(L->nx, L->ny, x, L->b, L->label, -1, 1.0, 1);
  }
  fo_mg_residual(L->nx, L->ny, x, L->b, L->label, L->r);
  mg_level* N = &s->lv[l + 1];
  fo_mg_restrict(L->nx, L->ny, L->r, N->nx, N->ny, N->b, N->x);
  mg_vcycle(s, l + 1, N->x);
  fo_mg_prolong_add(L->nx, L->ny, x, L->label, N->nx, N->ny, N->x);
  for (int k = 0; k < MG_POST; k++) {
    fo_rb_sweep(L->nx, L->ny, x, L->b, L->label, -1, 1.0, 0);
    fo_rb_sweep(L->nx, L->ny, x, L->b, L->label, -1, 1.0, 1);
  }
}

/* Mode 1 outer step: r = b - A p;  e = V(r) from e = 0;  p += e. */
static void mg_dc_cycle(mg_state* s, double* p, const double* b) {
  mg_level* L = &s->lv[0];
  size_t n = (size_t)L->nx * L->ny;
  double* rr = (double*)malloc(sizeof(double) * n);
  fo_mg_residual(L->nx, L->ny, p, b, L->label, rr);
  memcpy(L->b, rr, sizeof(double) * n);
  memset(L->x, 0, sizeof(double) * n);
  mg_vcycle(s, 0, L->x);
  for (size_t k = 0; k < n; k++) p[k] += L->x[k];
  free(rr);
}

void fo_mg_dc_cycle(int nx, int ny, double* p, const double* b, const uint8_t* label) {
  mg_state s;
  mg_build(&s, nx, ny, label, 1);
  mg_dc_cycle(&s, p, b);
  mg_free(&s);
}

/* pressureMultigrid.ts:284-319 (mode 0) and the build's mode 1.  Returns the
 * V-cycles used; resid_hist (optional, cycles+1 entries) receives the
 * relative residual rmsRemainingDivergence/fluidDivRms before cycle 0 and
 * after every cycle. */
int fo_mg_solve(int nx, int ny, double* p, const double* div, const uint8_t* label, double scale,
                int mode, int cycles, double tol, double* resid_hist) {
  double divRms = fo_fluid_div_rms(nx * ny, div, label);
  if (divRms == 0) {
    memset(p, 0, sizeof(double) * nx * ny);
    return 0;
  }
  mg_state s;
  mg_build(&s, nx, ny, label, mode);
  size_t n = (size_t)nx * ny;
  double* b = (double*)malloc(sizeof(double) * n);
  for (size_t k = 0; k < n; k++) b[k] = -scale * div[k];
  if (mode == 0) memcpy(s.lv[0].b, b, sizeof(double) * n);
  if (resid_hist) resid_hist[0] = fo_rms_remaining_divergence(nx, ny, p, div, label, scale) / divRms;
  double target = tol > 0 ? tol * divRms : -1;
  int used = cycles;
  for (int c = 0; c < cycles; c++) {
    if (mode == 0)
      mg_vcycle(&s, 0, p);
    else
      mg_dc_cycle(&s, p, b);
    double rr = fo_rms_remaining_divergence(nx, ny, p, div, label, scale);
    if (resid_hist) resid_hist[c + 1] = rr / divRms;
    if (target >= 0 && rr <= target) {
      used = c + 1;
      break;
    }
  }
  free(b);
  mg_free(&s);
  return used;
}

/* ---- projection and boundaries ------------------------------------------ */

/* subtractGradient.ts:40-61 */
void fo_subtract_gradient(int nx, int ny, const double* p, double* u, double* v,
                          const uint8_t* label, double gradScale) {
  for (int j = 0; j < ny; j++)
    for (int i = 1; i < nx; i++) {
      if (solid_or_outside(nx, ny, label, i - 1, j) || solid_or_outside(nx, ny, label, i, j))
        continue;
      u[i + j * (nx + 1)] -= gradScale * (p[i + j * nx] - p[i - 1 + j * nx]);
    }
  for (int j = 1; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      if (solid_or_outside(nx, ny, label, i, j - 1) || solid_or_outside(nx, ny, label, i, j))
        continue;
      v[i + j * nx] -= gradScale * (p[i + j * nx] - p[i + (j - 1) * nx]);
    }
}

/* boundaries.ts:115-143 */
void fo_apply_outflow(int nx, int ny, double* u, double* v, const uint8_t* label) {
  int su = nx + 1;
  for (int j = 0; j < ny; j++) {
    if (label[0 + j * nx] == FO_AIR) {
      double out = dmin(u[1 + j * su], 0.0);
      u[1 + j * su] = out;
      u[0 + j * su] = out;
    }
    if (label[nx - 1 + j * nx] == FO_AIR) {
      double out = dmax(u[nx - 1 + j * su], 0.0);
      u[nx - 1 + j * su] = out;
      u[nx + j * su] = out;
    }
  }
  for (int i = 0; i < nx; i++) {
    if (label[i] == FO_AIR) {
      double out = dmin(v[i + 1 * nx], 0.0);
      v[i + 1 * nx] = out;
      v[i] = out;
    }
    if (label[i + (ny - 1) * nx] == FO_AIR) {
      double out = dmax(v[i + (ny - 1) * nx], 0.0);
      v[i + (ny - 1) * nx] = out;
      v[i + ny * nx] = out;
    }
  }
}

/* boundaries.ts:45-64.  Dye is indexed by the label index k, as the
 * reference does (exact when the dye grid is the velocity grid). */
void fo_commit_labels(int nx, int ny, double* p, double* u, double* v, const uint8_t* label,
                      double* dye, int dye_plane) {
  for (int k = 0; k < nx * ny; k++) {
    if (label[k] == FO_FLUID) continue;
    p[k] = 0;
    if (label[k] == FO_SOLID && dye)
      for (int c = 0; c < 3; c++) dye[(size_t)c * dye_plane + k] = 0;
  }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i <= nx; i++)
      if (is_solid_cell(nx, ny, label, i - 1, j) || is_solid_cell(nx, ny, label, i, j))
        u[i + j * (nx + 1)] = 0;
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i < nx; i++)
      if (is_solid_cell(nx, ny, label, i, j - 1) || is_solid_cell(nx, ny, label, i, j))
        v[i + j * nx] = 0;
}

/* buoyancy.ts:70-88 */
void fo_apply_buoyancy(int nx, int ny, double h, int dnx, int dny, double dh, double* v,
                       const double* heat, const uint8_t* label, double w, double dt) {
  if (w == 0) return;
  double gain = w * dt;
  for (int j = 1; j < ny; j++) {
    double y = j * h;
    for (int i = 0; i < nx; i++) {
      if (is_solid_cell(nx, ny, label, i, j - 1) || is_solid_cell(nx, ny, label, i, j)) continue;
      v[i + j * nx] += gain * fo_sample_p(dnx, dny, dh, heat, (i + 0.5) * h, y);
    }
  }
}

/* dye.ts:70-87.  data: 3 value planes then the coverage plane. */
void fo_apply_dye_patch(int dnx, int dny, double* dye, int dye_plane, int i0, int j0, int pnx,
                        int pny, const float* data) {
  int cells = pnx * pny;
  for (int j = 0; j < pny; j++) {
    int dj = j0 + j;
    if (dj < 0 || dj >= dny) continue;
    for (int i = 0; i < pnx; i++) {
      int di = i0 + i;
      if (di < 0 || di >= dnx) continue;
      int s = i + j * pnx;
      double cov = data[3 * cells + s];
      if (cov == 0) continue;
      int k = di + dj * dnx;
      for (int c = 0; c < 3; c++) {
        double* q = dye + (size_t)c * dye_plane;
        q[k] += cov * ((double)data[c * cells + s] - q[k]);
      }
    }
  }
}

static inline int iclamp(int a, int lo, int hi) { return a < lo ? lo : a > hi ? hi : a; }

/* stir.ts:49-83 */
void fo_apply_stir(int nx, int ny, double h, double* u, double* v, const uint8_t* label,
                   double sx, double sy, double sr, double sdx, double sdy) {
  double invR2 = 1 / (sr * sr);
  double reach = 3 * sr;
  int c0 = (int)floor((sx - reach) / h), c1 = (int)ceil((sx + reach) / h);
  int r0 = (int)floor((sy - reach) / h), r1 = (int)ceil((sy + reach) / h);
  for (int j = iclamp(r0, 0, ny - 1); j <= iclamp(r1, 0, ny - 1); j++) {
    double dy = (j + 0.5) * h - sy;
    for (int i = iclamp(c0, 1, nx - 1); i <= iclamp(c1, 1, nx - 1); i++) {
      if (is_solid_cell(nx, ny, label, i - 1, j) || is_solid_cell(nx, ny, label, i, j)) continue;
      double dx = i * h - sx;
      u[i + j * (nx + 1)] += sdx * exp(-(dx * dx + dy * dy) * invR2);
    }
  }
  for (int j = iclamp(r0, 1, ny - 1); j <= iclamp(r1, 1, ny - 1); j++) {
    double dy = j * h - sy;
    for (int i = iclamp(c0, 0, nx - 1); i <= iclamp(c1, 0, nx - 1); i++) {
      if (is_solid_cell(nx, ny, label, i, j - 1) || is_solid_cell(nx, ny, label, i, j)) continue;
      double dx = (i + 0.5) * h - sx;
      v[i + j * nx] += sdy * exp(-(dx * dx + dy * dy) * invR2);
    }
  }
}

/* vorticity.ts:9-16 */
static inline double curl_at_corner(int nx, double h, const double* u, const double* v, int i,
                                    int j) {
  return (v[i + j * nx] - v[i - 1 + j * nx]) / h - (u[i + j * (nx + 1)] - u[i + (j - 1) * (nx + 1)]) / h;
}

/* vorticity.ts:24-42 */
void fo_compute_vorticity(int nx, int ny, double h, const double* u, const double* v,
                          double* out) {
  int hiI = nx - 1, hiJ = ny - 1;
  for (int j = 0; j < ny; j++) {
    int j0 = iclamp(j, 1, hiJ), j1 = iclamp(j + 1, 1, hiJ);
    for (int i = 0; i < nx; i++) {
      int i0 = iclamp(i, 1, hiI), i1 = iclamp(i + 1, 1, hiI);
      out[i + j * nx] = 0.25 * (curl_at_corner(nx, h, u, v, i0, j0) +
                                curl_at_corner(nx, h, u, v, i1, j0) +
                                curl_at_corner(nx, h, u, v, i0, j1) +
                                curl_at_corner(nx, h, u, v, i1, j1));
    }
  }
}

/* BUILD-DEFINED vorticity confinement.  omega_c is the cell-centred
 * vorticity of computeVorticity.  At each cell centre:
 *   a = |omega_c|,  g = grad a by (clamped) central differences,
 *   N = g / (|g| + 1e-4 * a / h + 1e-30),
 *   f = eps * h * (N_y * omega_c, -N_x * omega_c).
 * Faces take dt times the mean of their two cells' force, on the same face
 * set buoyancy and stir write: u i in [1,nx-1], v j in [1,ny-1], skipping
 * faces that touch an in-domain Solid. */
static inline void conf_force(int nx, int ny, double h, const double* w, double eps, int i,
                              int j, double* fx, double* fy) {
  int ip = i + 1 < nx ? i + 1 : nx - 1, im = i - 1 > 0 ? i - 1 : 0;
  int jp = j + 1 < ny ? j + 1 : ny - 1, jm = j - 1 > 0 ? j - 1 : 0;
  double a = fabs(w[i + j * nx]);
  double gx = (fabs(w[ip + j * nx]) - fabs(w[im + j * nx])) / ((ip - im) * h);
  double gy = (fabs(w[i + jp * nx]) - fabs(w[i + jm * nx])) / ((jp - jm) * h);
  double len = sqrt(gx * gx + gy * gy);
  double inv = 1.0 / (len + 1e-4 * a / h + 1e-30);
  double nxv = gx * inv, nyv = gy * inv;
  double om = w[i + j * nx];
  *fx = eps * h * (nyv * om);
  *fy = eps * h * (-nxv * om);
}

void fo_vorticity_confinement(int nx, int ny, double h, double* u, double* v,
                              const uint8_t* label, const double* omega_c, double eps, double dt) {
  if (eps == 0) return;
  double fxa, fya, fxb, fyb;
  for (int j = 0; j < ny; j++)
    for (int i = 1; i < nx; i++) {
      if (is_solid_cell(nx, ny, label, i - 1, j) || is_solid_cell(nx, ny, label, i, j)) continue;
      conf_force(nx, ny, h, omega_c, eps, i - 1, j, &fxa, &fya);
      conf_force(nx, ny, h, omega_c, eps, i, j, &fxb, &fyb);
      u[i + j * (nx + 1)] += dt * (0.5 * (fxa + fxb));
    }
  for (int j = 1; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      if (is_solid_cell(nx, ny, label, i, j - 1) || is_solid_cell(nx, ny, label, i, j)) continue;
      conf_force(nx, ny, h, omega_c, eps, i, j - 1, &fxa, &fya);
      conf_force(nx, ny, h, omega_c, eps, i, j, &fxb, &fyb);
      v[i + j * nx] += dt * (0.5 * (fya + fyb));
    }
}

/* simulation.ts:218-221 */
double fo_optimal_omega(int nx, int ny) {
  double rho = 0.5 * (cos(M_PI / nx) + cos(M_PI / ny));
  return 2 / (1 + sqrt(1 - rho * rho));
}

/* simulation.ts:451-459 */
double fo_max_face_speed(int nx, int ny, const double* u, const double* v) {
  double m = 0;
  for (long k = 0; k < (long)(nx + 1) * ny; k++) m = dmax(m, fabs(u[k]));
  for (long k = 0; k < (long)nx * (ny + 1); k++) m = dmax(m, fabs(v[k]));
  return m;
}

static void swap_ptr(double** a, double** b) {
  double* t = *a;
  *a = *b;
  *b = t;
}

/* simulation.ts:467-587, CPU path (no stepper). */
void fo_sim_step(fo_sim* s, fo_step_params* prm) {
  int nx = s->nx, ny = s->ny, dnx = s->dnx, dny = s->dny;
  double h = s->h, dh = s->dh;
  double uMax = fo_max_face_speed(nx, ny, s->u, s->v);
  double dt = uMax > 1e-9 ? dmin(prm->dt_max, prm->cfl_target * h / uMax) : prm->dt_max;
  double scale = prm->rho * h * h / dt;
  double gradScale = dt / (prm->rho * h);

  if (prm->scheme == 1)
    fo_advect_velocity_maccormack(nx, ny, h, s->u, s->v, s->uHat, s->vHat, s->uNext, s->vNext,
                                  s->label, dt);
  else
    fo_advect_velocity(nx, ny, h, s->u, s->v, s->uNext, s->vNext, s->label, dt);
  swap_ptr(&s->u, &s->uNext);
  swap_ptr(&s->v, &s->vNext);

  /* confinement reads the advected velocity, before any body force */
  if (prm->confinement != 0) fo_compute_vorticity(nx, ny, h, s->u, s->v, s->work);
  if (prm->stir_active)
    fo_apply_stir(nx, ny, h, s->u, s->v, s->label, prm->stir[0], prm->stir[1], prm->stir[2],
                  prm->stir[3], prm->stir[4]);
  fo_apply_buoyancy(nx, ny, h, dnx, dny, dh, s->v, s->dye, s->label, prm->buoyancy, dt);
  if (prm->confinement != 0)
    fo_vorticity_confinement(nx, ny, h, s->u, s->v, s->label, s->work, prm->confinement, dt);

  fo_divergence(nx, ny, h, s->u, s->v, s->div);
  int iters = 0;
  switch (prm->solver) {
    case 0:
      iters = fo_solve_pressure(nx, ny, s->p, s->div, s->label, scale, prm->pressure_iters,
                                prm->omega, prm->tol);
      break;
    case 1:
      iters = fo_solve_pressure_rb(nx, ny, s->p, s->div, s->label, scale, prm->pressure_iters,
                                   prm->omega, prm->tol);
      break;
    case 2:
      iters = fo_mg_solve(nx, ny, s->p, s->div, s->label, scale, 0, prm->pressure_iters,
                          prm->tol, NULL);
      break;
    default:
      iters = fo_mg_solve(nx, ny, s->p, s->div, s->label, scale, 1, prm->pressure_iters,
                          prm->tol, NULL);
      break;
  }
  fo_subtract_gradient(nx, ny, s->p, s->u, s->v, s->label, gradScale);
  fo_apply_outflow(nx, ny, s->u, s->v, s->label);
  if (prm->residual_every > 0 && s->steps % prm->residual_every == 0)
    fo_divergence(nx, ny, h, s->u, s->v, s->div);

  size_t dn = (size_t)dnx * dny;
  for (int c = 0; c < 3; c++) {
    if (prm->scheme == 1)
      fo_advect_scalar_maccormack(nx, ny, h, dnx, dny, dh, s->u, s->v, s->dye + c * dn, s->qHat,
                                  s->dyeNext + c * dn, s->label, dt);
    else
      fo_advect_scalar(nx, ny, h, dnx, dny, dh, s->u, s->v, s->dye + c * dn,
                       s->dyeNext + c * dn, s->label, dt);
  }
  swap_ptr(&s->dye, &s->dyeNext);
  /* simulation.ts:243-247, 569-576 */
  for (int c = 0; c < 3; c++) {
    double d = prm->keep_decay[c];
    double keep = d > 0 ? exp(-d * dt) : 1.0;
    if (keep == 1.0) continue;
    double* q = s->dye + c * dn;
    for (size_t k = 0; k < dn; k++) q[k] *= keep;
  }
  if (prm->patch_data)
    fo_apply_dye_patch(dnx, dny, s->dye, (int)dn, prm->patch_i0, prm->patch_j0, prm->patch_nx,
                       prm->patch_ny, prm->patch_data);
  prm->dt_used = dt;
  prm->iters = iters;
  s->steps++;
}

void fo_lcg_fill(int n, int seed, double* out) {
  long long s = seed;
  for (int k = 0; k < n; k++) {
    double d = (double)s * 1103515245.0 + 12345.0;
    unsigned long long w = (unsigned long long)d; /* exact: d is an integer below 2^63 */
    s = (long long)(w & 0x7fffffffULL);
    out[k] = ((double)s / 2147483647.0) * 2 - 1;
  }
}
