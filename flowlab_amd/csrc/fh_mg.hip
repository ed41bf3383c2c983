// Pressure solvers: red-black SOR sweeps, the geometric-multigrid transfer
// operators and the FP64 defect-correction update with its residual norm
// (src/core/pressureRedBlack.ts, pressureMultigrid.ts, pressure.ts:102-161).
//
// Every level solves  count*x - sum(x over non-solid neighbours) = b  on
// fluid cells, with Air neighbours contributing x = 0 and off-grid/Solid
// neighbours absent (pressureMultigrid.ts:22-42).
#include <atomic>
#include <type_traits>

#include "fh_batchdesc.hpp"
#include "fh_device.hpp"
#include "fh_launch.hpp"
#include "fh_mgtile.hpp"

namespace fh {

constexpr int BX = 64, BY = 4;

static inline dim3 grid2(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }

// One SOR update of a fluid cell with cnt > 0 open neighbours summing to
// `sum` (pressureRedBlack.ts:52-66), shared by both red-black kernels.  The
// fused multiply-adds are written out: left to contraction, the compiler
// fused (1 - omega) old + omega pgs one way in one kernel and the other way
// in the other, and the two kernels' bits differed.
__device__ __forceinline__ float rb_value(float sum, int cnt, float bk, float old, float omega,
                                          float scale) {
  const float pgs = fmaf(-scale, bk, sum) / (float)cnt;
  return fmaf(omega, pgs, (1.f - omega) * old);
}

// pressureRedBlack.ts:32-72.  Thread t of row j owns the cell pair
// (2t, 2t+1) and updates the one of this colour, so every lane is busy and a
// wavefront covers 128 contiguous cells.  Neighbour order +x, -x, +y, -y.
__global__ __launch_bounds__(256) void k_rb_sweep(float* __restrict__ x,
                                                  const float* __restrict__ b,
                                                  const uint8_t* __restrict__ lbl, int nx, int ny,
                                                  int color, float omega, float scale, int first) {
  int t = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (j >= ny) return;
  int i = 2 * t + ((j + color) & 1);
  if (first) {
    int io = 2 * t + ((j + color + 1) & 1);
    if (io < nx && lbl[io + j * nx] == FLUID) x[io + j * nx] = 0.f;
  }
  if (i >= nx) return;
  int k = i + j * nx;
  if (lbl[k] != FLUID) return;
  int cnt = 0;
  float sum = 0.f;
  if (!solid_or_outside(lbl, nx, ny, i + 1, j)) {
    cnt++;
    if (!first) sum += x[k + 1];
  }
  if (!solid_or_outside(lbl, nx, ny, i - 1, j)) {
    cnt++;
    if (!first) sum += x[k - 1];
  }
  if (!solid_or_outside(lbl, nx, ny, i, j + 1)) {
    cnt++;
    if (!first) sum += x[k + nx];
  }
  if (!solid_or_outside(lbl, nx, ny, i, j - 1)) {
    cnt++;
    if (!first) sum += x[k - nx];
  }
  if (cnt == 0) return;
  x[k] = rb_value(sum, cnt, b[k], first ? 0.f : x[k], omega, scale);
}

// The whole red-black SOR solve of a grid small enough for one workgroup's
// LDS (x and b FP32 and the cell info byte, 9 B per cell: 128 x 128 fits the
// 160 KB): `iters` iterations of both colours in one launch, a barrier
// between half-sweeps, instead of two dependent launches per iteration.
// Each thread finds its cells of each colour once (LDS index and info byte
// packed in one register, at most kSorItems per colour); a half-sweep then
// reads every neighbour of its cells before writing any.  Each cell's update is
// k_rb_sweep's: rb_value, the same neighbours (the info byte's open bits are
// its solid_or_outside tests) summed in the same order, a closed neighbour
// adding +0 (which leaves the sum unchanged: the sum starts at +0 and so is
// never -0).  A colour reads only the other colour, so the bits are those of
// the launch-per-half-sweep solve.
constexpr int kSorThreads = 1024, kSorItems = 8;

__global__ __launch_bounds__(kSorThreads) void k_rb_sor_lds(float* __restrict__ x,
                                                            const float* __restrict__ b,
                                                            const uint8_t* __restrict__ lbl,
                                                            int nx, int ny, int iters,
                                                            float omega, float scale) {
  extern __shared__ unsigned char smem[];
  const int n = nx * ny, tid = threadIdx.x;
  float* sx = reinterpret_cast<float*>(smem);
  float* sb = sx + n;
  uint8_t* si = reinterpret_cast<uint8_t*>(sb + n);
  for (int k = tid; k < n; k += kSorThreads) {
    const int j = k / nx, i = k - j * nx;
    sx[k] = x[k];
    sb[k] = b[k];
    si[k] = cell_info(lbl, nx, ny, i, j);
  }
  __syncthreads();
  const int P = (nx + 1) / 2, items = P * ny;
  uint32_t pk[2][kSorItems];  // LDS index | info << 24; 0: no cell to update
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int r = 0; r < kSorItems; r++) {
      const int q = tid + r * kSorThreads;
      const int j = q / P, t = q - j * P;
      const int i = 2 * t + ((j + c) & 1);
      const int k = i + j * nx;
      const int f = (q < items && i < nx) ? si[k] : 0;
      pk[c][r] = ((f & 0x80) && (f & 7)) ? (uint32_t)k | ((uint32_t)f << 24) : 0u;
    }
  // A cell's own value changes only when its one owner (this thread, this
  // slot) updates it, so the owner keeps it and the cell's b in registers:
  // a half-sweep reads only the four neighbours from LDS.  Each neighbour
  // read is unconditional, from the cell itself when that side is closed
  // (always in the window), and the closed side's +0 is selected after.
  float xo[2][kSorItems], bo[2][kSorItems];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int r = 0; r < kSorItems; r++) {
      const int k = (int)(pk[c][r] & 0xffffffu);
      xo[c][r] = sx[k];
      bo[c][r] = sb[k];
    }
  auto half = [&](auto cc) {
    constexpr int c = decltype(cc)::value;
#pragma unroll
    for (int r = 0; r < kSorItems; r++) {
      const int k = (int)(pk[c][r] & 0xffffffu), f = (int)(pk[c][r] >> 24);
      const float ve = sx[(f & 0x08) ? k + 1 : k];
      const float vw = sx[(f & 0x10) ? k - 1 : k];
      const float vn = sx[(f & 0x20) ? k + nx : k];
      const float vs = sx[(f & 0x40) ? k - nx : k];
      const float sum = (((0.f + ((f & 0x08) ? ve : 0.f)) + ((f & 0x10) ? vw : 0.f)) +
                         ((f & 0x20) ? vn : 0.f)) +
                        ((f & 0x40) ? vs : 0.f);
      if (f) xo[c][r] = rb_value(sum, f & 7, bo[c][r], xo[c][r], omega, scale);
    }
#pragma unroll
    for (int r = 0; r < kSorItems; r++)
      if (pk[c][r]) sx[pk[c][r] & 0xffffffu] = xo[c][r];
    __syncthreads();
  };
  for (int it = 0; it < iters; it++) {
    half(std::integral_constant<int, 0>{});
    half(std::integral_constant<int, 1>{});
  }
  for (int k = tid; k < n; k += kSorThreads) x[k] = sx[k];
}

// pressureMultigrid.ts:117-151 at one cell; 0 off the fluid.
__device__ __forceinline__ float residual_at(const float* __restrict__ x,
                                             const float* __restrict__ b,
                                             const uint8_t* __restrict__ lbl, int nx, int ny,
                                             int i, int j) {
  int k = i + j * nx;
  if (lbl[k] != FLUID) return 0.f;
  int cnt = 0;
  float sum = 0.f;
  if (!solid_or_outside(lbl, nx, ny, i + 1, j)) { cnt++; sum += x[k + 1]; }
  if (!solid_or_outside(lbl, nx, ny, i - 1, j)) { cnt++; sum += x[k - 1]; }
  if (!solid_or_outside(lbl, nx, ny, i, j + 1)) { cnt++; sum += x[k + nx]; }
  if (!solid_or_outside(lbl, nx, ny, i, j - 1)) { cnt++; sum += x[k - nx]; }
  if (cnt == 0) return 0.f;
  return b[k] - ((float)cnt * x[k] - sum);
}

// pressureMultigrid.ts:117-178 fused: the residual is never stored, each
// coarse cell sums its (up to four) children's residuals directly.
__global__ __launch_bounds__(256) void k_residual_restrict(
    const float* __restrict__ x, const float* __restrict__ b, const uint8_t* __restrict__ lbl,
    int nx, int ny, float* __restrict__ bc, float* __restrict__ xc, int cnx, int cny) {
  int I = blockIdx.x * BX + threadIdx.x;
  int J = blockIdx.y * BY + threadIdx.y;
  if (I >= cnx || J >= cny) return;
  float s = 0.f;
#pragma unroll
  for (int dj = 0; dj < 2; dj++)
#pragma unroll
    for (int di = 0; di < 2; di++) {
      int i = 2 * I + di, j = 2 * J + dj;
      if (i < nx && j < ny) s += residual_at(x, b, lbl, nx, ny, i, j);
    }
  bc[I + J * cnx] = s;
  xc[I + J * cnx] = 0.f;
}

// pressureMultigrid.ts:192-207
__global__ __launch_bounds__(256) void k_prolong(float* __restrict__ x,
                                                 const uint8_t* __restrict__ lbl, int nx, int ny,
                                                 const float* __restrict__ e, int cnx, int cny) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  int k = i + j * nx;
  if (lbl[k] != FLUID) return;
  int J = j >> 1, I = i >> 1;
  int J2 = (j & 1) ? min(J + 1, cny - 1) : max(J - 1, 0);
  int I2 = (i & 1) ? min(I + 1, cnx - 1) : max(I - 1, 0);
  x[k] += 0.5625f * e[I + J * cnx] + 0.1875f * (e[I2 + J * cnx] + e[I + J2 * cnx]) +
          0.0625f * e[I2 + J2 * cnx];
}

// pressureMultigrid.ts:92-110: Air > Fluid > Solid
__global__ __launch_bounds__(256) void k_coarsen_labels(const uint8_t* __restrict__ f, int fnx,
                                                        int fny, uint8_t* __restrict__ c, int cnx,
                                                        int cny) {
  int I = blockIdx.x * BX + threadIdx.x;
  int J = blockIdx.y * BY + threadIdx.y;
  if (I >= cnx || J >= cny) return;
  bool air = false, fluid = false;
  for (int dj = 0; dj < 2; dj++)
    for (int di = 0; di < 2; di++) {
      int i = 2 * I + di, j = 2 * J + dj;
      if (i >= fnx || j >= fny) continue;
      uint8_t l = f[i + j * fnx];
      air = air || l == AIR;
      fluid = fluid || l == FLUID;
    }
  c[I + J * cnx] = air ? AIR : fluid ? FLUID : SOLID;
}

// Coarsest level: x[cells[a]] = sum_b M[a][b] b[cells[b]] in FP64, one
// workgroup; the right-hand side is staged in LDS.
__global__ __launch_bounds__(256) void k_coarse_solve(const float* __restrict__ bc,
                                                      float* __restrict__ xc,
                                                      const double* __restrict__ M,
                                                      const int* __restrict__ cells, int n) {
  __shared__ double rhs[1024];
  for (int a = threadIdx.x; a < n; a += blockDim.x) rhs[a] = (double)bc[cells[a]];
  __syncthreads();
  for (int a = threadIdx.x; a < n; a += blockDim.x) {
    const double* row = M + (size_t)a * n;
    double s = 0.0;
    for (int q = 0; q < n; q++) s += row[q] * rhs[q];
    xc[cells[a]] = (float)s;
  }
}

__global__ void k_rhs32(const double* __restrict__ div, float* __restrict__ out, long n,
                        double factor) {
  long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k < n) out[k] = (float)(factor * div[k]);
}

__global__ void k_f64_to_f32(const double* __restrict__ a, float* __restrict__ out, long n) {
  long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k < n) out[k] = (float)a[k];
}

__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ out, long n) {
  long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k < n) out[k] = (double)a[k];
}

// Warm start extrapolated in time on its own pass (the solvers whose
// initial pass is not dc_update_body's): out = p extrapolated over the
// history (order 0: a copy of p).  Two cells per thread as double2; every
// array holds an even number of doubles (the engine pads).
template <int X>
__global__ __launch_bounds__(256) void k_extrapolate(const double* __restrict__ p,
                                                    Extrapolation xp, double* __restrict__ out,
                                                    long n2) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= n2) return;
  double2 t = reinterpret_cast<const double2*>(p)[k];
  if constexpr (X >= 1) {
    const double2 a = reinterpret_cast<const double2*>(xp.h0)[k];
    double2 b = make_double2(0.0, 0.0);
    if constexpr (X >= 2) b = reinterpret_cast<const double2*>(xp.h1)[k];
    t.x = extrap_value<X>(xp, t.x, a.x, b.x);
    t.y = extrap_value<X>(xp, t.y, a.y, b.y);
  }
  reinterpret_cast<double2*>(out)[k] = t;
}

// Reduction grids are capped at kMaxPartials blocks (grid-stride over rows
// beyond that); the one-block finalize reads up to kMaxPartials/1024
// partials per thread, 8 in flight at a time.  At 8192^2 the cap leaves each
// thread 8 rows; caps of 8192 / 32768 left it 32 / 16, and the defect update
// there took 0.643 / 0.615 ms against 0.59 now.
constexpr int kMaxPartials = 65536;

// Grid of the reduction passes (defect update, CG dots and update, norms):
// 64-wide blocks along x, each thread walking rows with stride gy * BY.
// Four BY-row groups per thread: the per-block FP64 reduction is then paid
// once per four rows, and each wave has more loads to overlap (measured at
// 2048x1024: the CG update 21.3 -> 15.9 us, its dot-product pass 25.9 ->
// 14.8 us; two or eight groups were slower).  At least ~1024 blocks are kept
// where the grid has them, and at most kMaxPartials.
static inline dim3 reduce_grid(int nx, int ny) {
  const int gx = (nx + BX - 1) / BX;
  const int rows = (ny + BY - 1) / BY;
  int gy = (rows + 3) / 4;
  const int want = (1024 + gx - 1) / gx;
  if (gy < want) gy = rows < want ? rows : want;
  int cap = kMaxPartials / gx;
  if (cap < 1) cap = 1;
  return dim3(gx, gy < cap ? gy : cap);
}

// The defect-correction and CG passes walk kRedRows consecutive rows per
// thread and iteration, every stencil value of those rows loaded before any
// arithmetic or store (consecutive rows share their north/south loads, and a
// CU keeps enough bytes in flight to stream HBM).  The block count is capped
// at kMaxPartials (rows beyond it are walked with the grid's stride).  Grids
// below ~1 M cells have fewer blocks than CUs this way; one row per thread
// over more blocks did not make such a grid's solve faster on its own (it is
// bound by launch latency there), and in a batched launch the scenes
// together fill the chip, where one row per thread ran the 64-scene batch's
// CG update at 0.47 of the HBM peak.
constexpr int kRedRows = 4;

static inline dim3 reduce_grid_rows(int nx, int ny) {
  const int gx = (nx + BX - 1) / BX;
  const int trows = (ny + kRedRows - 1) / kRedRows;
  const int gy = (trows + BY - 1) / BY;
  int cap = kMaxPartials / gx;
  if (cap < 1) cap = 1;
  return dim3(gx, gy < cap ? gy : cap);
}

// Column i of a field over the rows j0-1 .. j0+R (clamped to the domain: c),
// and its east and west neighbours (clamped columns) over rows j0 .. j0+R-1.
// For row j0+r: centre c[r+1], north c[r+2], south c[r], east e[r], west
// w[r] -- at the domain edge the clamped index is the cell itself, as the
// one-row kernels' kN/kS/kE/kW are.
template <int R, class T>
struct Col5 {
  T c[R + 2], e[R], w[R];
};

template <int R, class T>
__device__ __forceinline__ void load_col5(Col5<R, T>& o, const T* __restrict__ a, int i, int j0,
                                          int nx, int ny) {
  const int iE = i + 1 < nx ? i + 1 : i, iW = i > 0 ? i - 1 : i;
#pragma unroll
  for (int q = 0; q < R + 2; q++) o.c[q] = a[i + min(max(j0 - 1 + q, 0), ny - 1) * nx];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int row = min(j0 + r, ny - 1) * nx;
    o.e[r] = a[iE + row];
    o.w[r] = a[iW + row];
  }
}

// One value of the extrapolated start: w0 t + w1 a (+ w2 b for X = 2), with
// its multiply-adds written out so every pass that forms it gives the same
// bits.
template <int X>
__device__ __forceinline__ double extrap_value(const Extrapolation& xp, double t, double a,
                                               double b) {
  double r = __fma_rn(xp.w1, a, xp.w0 * t);
  if constexpr (X >= 2) r = __fma_rn(xp.w2, b, r);
  return r;
}

// FP64 defect-correction update, one thread per cell:
//   pn = po + e  (e = 0 off the fluid),  r = b - (count*pn - sum pn),
//   b = -scale*div,  r32 = (float) r  (the next V-cycle's right-hand side).
// Reads po and e at the 5-point stencil; pn goes to a second buffer so no
// thread reads a neighbour that another thread already advanced.  The
// neighbour set comes from the level's info byte (open-neighbour mask and
// count), so every load of a cell is issued at once, unconditionally, from
// clamped in-domain indices, and closed neighbours add 0 (sum order +x, -x,
// +y, -y, as in the reference's residual).
//
// kDiv: the solve's initial pass (no e) forming the divergence itself from
// the faces u, v (div_value, as k_divergence does) and storing it to div_out,
// instead of reading a divergence a separate pass wrote.
// X >= 1 (kDiv only): the pass starts from po extrapolated in time over the
// history xp.h0 (and xp.h1 for X = 2), extrap_value at every stencil point,
// and stores that start to pn.
template <int R, bool kDiv = false, int X = 0>
__device__ __forceinline__ void dc_update_body(
    const double* __restrict__ po, double* __restrict__ pn, const float* __restrict__ e,
    const double* __restrict__ div, const uint8_t* __restrict__ info, float* __restrict__ r32,
    double* __restrict__ r64, int nx, int ny, double scale, double* partials,
    const int* __restrict__ done, int bx, int by, int gy, int bid,
    const float* __restrict__ u = nullptr, const float* __restrict__ v = nullptr,
    double inv_h = 0.0, double* __restrict__ div_out = nullptr, Extrapolation xp = {}) {
  if (done && *done) return;
  const int i = bx * BX + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j0 = (by * BY + threadIdx.y) * R; j0 < ny; j0 += gy * BY * R) {
    if (i >= nx) break;
    Col5<R, double> P;
    load_col5(P, po, i, j0, nx, ny);
    Col5<R, float> E;
    if (e) load_col5(E, e, i, j0, nx, ny);
    if constexpr (X >= 1) {
      Col5<R, double> H0, H1;
      load_col5(H0, xp.h0, i, j0, nx, ny);
      if constexpr (X >= 2) load_col5(H1, xp.h1, i, j0, nx, ny);
#pragma unroll
      for (int q = 0; q < R + 2; q++)
        P.c[q] = extrap_value<X>(xp, P.c[q], H0.c[q], X >= 2 ? H1.c[q] : 0.0);
#pragma unroll
      for (int r = 0; r < R; r++) {
        P.e[r] = extrap_value<X>(xp, P.e[r], H0.e[r], X >= 2 ? H1.e[r] : 0.0);
        P.w[r] = extrap_value<X>(xp, P.w[r], H0.w[r], X >= 2 ? H1.w[r] : 0.0);
      }
    }
    int fr[R];
    double dr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int jj = min(j0 + r, ny - 1), k = i + jj * nx;
      fr[r] = info[k];
      if constexpr (kDiv)
        dr[r] = div_value(u[i + jj * (nx + 1)], u[i + 1 + jj * (nx + 1)], v[k], v[k + nx], inv_h);
      else
        dr[r] = div[k];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = j0 + r;
      if (j >= ny) break;
      const int k = i + j * nx;
      const int f = fr[r];
      const double d = dr[r];
      if constexpr (kDiv) div_out[k] = d;
      double pc = P.c[r + 1], pE = P.e[r], pW = P.w[r], pN = P.c[r + 2], pS = P.c[r];
      if (e) {
        pc += (double)E.c[r + 1];
        pE += (double)E.e[r];
        pW += (double)E.w[r];
        pN += (double)E.c[r + 2];
        pS += (double)E.c[r];
      }
      if (pn) pn[k] = pc;
      const double rr = dc_cell(f, d, pc, pE, pW, pN, pS, scale, acc);
      r32[k] = (float)rr;
      if (r64) r64[k] = rr;
    }
  }
  block_sum_store_at<4>(acc, partials, bid);
}

template <int R>
__global__ __launch_bounds__(256) void k_dc_update(
    const double* __restrict__ po, double* __restrict__ pn, const float* __restrict__ e,
    const double* __restrict__ div, const uint8_t* __restrict__ info, float* __restrict__ r32,
    double* __restrict__ r64, int nx, int ny, double scale, double* partials,
    const int* __restrict__ done) {
  int bx, by;
  xcd_block(&bx, &by);
  dc_update_body<R>(po, pn, e, div, info, r32, r64, nx, ny, scale, partials, done, bx, by,
                    gridDim.y, blockIdx.x + blockIdx.y * gridDim.x);
}

// The solve's initial defect pass with the divergence formed in the same
// pass (dc_update_body<R, true>): the divergence kernel and this pass's read
// of div become one launch.
// X >= 1: from po extrapolated over xp's history, the start stored to pn
// (pn non-null with X = 0: po copied there).
template <int R, int X = 0>
__global__ __launch_bounds__(256) void k_div_defect(
    const float* __restrict__ u, const float* __restrict__ v, double inv_h, double* __restrict__ div,
    const double* __restrict__ po, const uint8_t* __restrict__ info, float* __restrict__ r32,
    double* __restrict__ r64, int nx, int ny, double scale, double* partials,
    double* __restrict__ pn, Extrapolation xp) {
  int bx, by;
  xcd_block(&bx, &by);
  dc_update_body<R, true, X>(po, pn, nullptr, nullptr, info, r32, r64, nx, ny, scale, partials,
                             nullptr, bx, by, gridDim.y, blockIdx.x + blockIdx.y * gridDim.x, u,
                             v, inv_h, div, xp);
}

// ---- flexible conjugate gradients (FCG(1)) around the V-cycle -------------
// The preconditioner B is one V(2,2) cycle from zero on r (FP32); it is not
// symmetric (pre- and post-smoothing sweep in the same colour order, and
// restriction is not the transpose of prolongation), so the direction is
// A-orthogonalised against the previous one explicitly (Notay's flexible
// CG with one stored direction) instead of by the Fletcher-Reeves ratio:
//   z = B r;  beta = -<z, A d'> / <d', A d'>;  d = z + beta d';
//   alpha = <d, r> / <d, A d>;  p += alpha d;  r -= alpha A d.
// A solve's first cycle is a plain defect-correction cycle (p += B r, the
// residual recomputed from p), so a solve that one V-cycle completes costs
// what defect correction costs; CG starts from its result.  <d, r> is formed
// explicitly rather than assumed 0 against d'.
// A is the 5-point operator of the fluid cells with a non-zero count
// (count x - sum of open neighbours; Air neighbours hold 0), symmetric, so
// every product above follows from four dot products taken in one pass
// after the V-cycle:
//   s = {<z, r>, <z, A z>, <z, A d'>, <d', r>},
//   <d, A d> = <z,Az> + 2 beta <z,Ad'> + beta^2 <d',Ad'>,  <d, r> = <z,r> + beta <d',r>,
// and A d is formed on the fly in the update pass as A z + beta A d'.  The
// vectors d and z are FP32, p and r FP64 with FP64 arithmetic throughout: p
// and r are updated with the same (exactly converted) d, so the recursive r
// stays the true residual b - A p to FP64 rounding.

// (A x)_k for one cell from its info byte and the five stencil values.
__device__ __forceinline__ double apply_a(int f, double c, double e, double w, double n,
                                          double s) {
  const int cnt = f & 7;
  if (!(f & 0x80) || cnt == 0) return 0.0;
  double sum = 0.0;
  sum += (f & 0x08) ? e : 0.0;
  sum += (f & 0x10) ? w : 0.0;
  sum += (f & 0x20) ? n : 0.0;
  sum += (f & 0x40) ? s : 0.0;
  return (double)cnt * c - sum;
}

// Pass 1: partials of s (first: d' is not read and counts as 0).
template <int R>
__device__ __forceinline__ void fcg_dots_body(const float* __restrict__ z,
                                              const float* __restrict__ dprev,
                                              const double* __restrict__ r,
                                              const uint8_t* __restrict__ info, int nx, int ny,
                                              int first, double* partials,
                                              const int* __restrict__ done, int bx, int by,
                                              int gy, int bid) {
  if (done && *done) return;
  const int i = bx * BX + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j0 = (by * BY + threadIdx.y) * R; j0 < ny; j0 += gy * BY * R) {
    if (i >= nx) break;
    Col5<R, float> Z, D;
    load_col5(Z, z, i, j0, nx, ny);
    if (!first) load_col5(D, dprev, i, j0, nx, ny);
    int fr[R];
    double rr[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int k = i + min(j0 + q, ny - 1) * nx;
      fr[q] = info[k];
      rr[q] = r[k];
    }
#pragma unroll
    for (int q = 0; q < R; q++) {
      if (j0 + q >= ny) break;
      const int f = fr[q];
      const double zc = Z.c[q + 1], rc = rr[q];
      const double az = apply_a(f, zc, Z.e[q], Z.w[q], Z.c[q + 2], Z.c[q]);
      acc[0] += zc * rc;
      acc[1] += zc * az;
      if (!first) {
        const double dc = D.c[q + 1];
        const double ad = apply_a(f, dc, D.e[q], D.w[q], D.c[q + 2], D.c[q]);
        acc[2] += zc * ad;
        acc[3] += dc * rc;
      }
    }
  }
  block_sum_store_at<4>(acc, partials, bid);
}

template <int R>
__global__ __launch_bounds__(256) void k_fcg_dots(const float* __restrict__ z,
                                                  const float* __restrict__ dprev,
                                                  const double* __restrict__ r,
                                                  const uint8_t* __restrict__ info, int nx,
                                                  int ny, int first, double* partials,
                                                  const int* __restrict__ done) {
  int bx, by;
  xcd_block(&bx, &by);
  fcg_dots_body<R>(z, dprev, r, info, nx, ny, first, partials, done, bx, by, gridDim.y,
                   blockIdx.x + blockIdx.y * gridDim.x);
}

// The step coefficients from pass 1's partials: state = {alpha, beta,
// <d, A d>} (the last carried to the next iteration).  A non-positive or
// non-finite <d, A d> (a direction the operator does not see) gives alpha 0.
__device__ __forceinline__ void fcg_coef_body(const double* __restrict__ partials, int nblocks,
                                              double* __restrict__ state, int first,
                                              const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double sh[4][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = tid; b < nblocks; b += 1024)
    for (int q = 0; q < 4; q++) acc[q] += partials[4 * b + q];
  for (int q = 0; q < 4; q++) {
    double t = wave_sum(acc[q]);
    if (lane == 0) sh[q][wid] = t;
  }
  __syncthreads();
  if (tid == 0) {
    double v[4];
    for (int q = 0; q < 4; q++) {
      double a = 0.0;
      for (int w = 0; w < 16; w++) a += sh[q][w];
      v[q] = a;
    }
    const double dad_prev = state[2];
    const double beta = (first || !(dad_prev > 0.0)) ? 0.0 : -v[2] / dad_prev;
    const double dad = v[1] + 2.0 * beta * v[2] + beta * beta * dad_prev;
    const double dr = v[0] + beta * v[3];
    const double alpha = (dad > 0.0 && isfinite(dad)) ? dr / dad : 0.0;
    state[0] = alpha;
    state[1] = beta;
    state[2] = dad;
  }
}

__global__ __launch_bounds__(1024) void k_fcg_coef(const double* __restrict__ partials,
                                                   int nblocks, double* __restrict__ state,
                                                   int first, const int* __restrict__ done) {
  fcg_coef_body(partials, nblocks, state, first, done);
}

// Pass 2: d = z + beta d';  p += alpha d;  r -= alpha (A z + beta A d');
// r32 = (float) r;  partials {sum r^2, n} over the fluid cells with a
// non-zero count (the divergence terms of the norm are the initial pass's).
template <int R>
__device__ __forceinline__ void fcg_update_body(
    const float* __restrict__ z, const float* __restrict__ dprev, float* __restrict__ dnext,
    double* __restrict__ p, double* __restrict__ r, const uint8_t* __restrict__ info,
    float* __restrict__ r32, const double* __restrict__ state, int nx, int ny, int first,
    double* partials, const int* __restrict__ done, int bx, int by, int gy, int bid) {
  if (done && *done) return;
  const int i = bx * BX + threadIdx.x;
  const double alpha = state[0], beta = first ? 0.0 : state[1];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j0 = (by * BY + threadIdx.y) * R; j0 < ny; j0 += gy * BY * R) {
    if (i >= nx) break;
    Col5<R, float> Z, D;
    load_col5(Z, z, i, j0, nx, ny);
    if (!first) load_col5(D, dprev, i, j0, nx, ny);
    int fr[R];
    double pr[R], rr[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int k = i + min(j0 + q, ny - 1) * nx;
      fr[q] = info[k];
      pr[q] = p[k];
      rr[q] = r[k];
    }
#pragma unroll
    for (int q = 0; q < R; q++) {
      if (j0 + q >= ny) break;
      const int k = i + (j0 + q) * nx;
      const int f = fr[q];
      double dc = Z.c[q + 1], dE = Z.e[q], dW = Z.w[q], dN = Z.c[q + 2], dS = Z.c[q];
      if (!first) {
        dc += beta * (double)D.c[q + 1];
        dE += beta * (double)D.e[q];
        dW += beta * (double)D.w[q];
        dN += beta * (double)D.c[q + 2];
        dS += beta * (double)D.c[q];
      }
      // the stored direction is FP32: p, r and A d all use its rounded values
      const float d32 = (float)dc;
      dnext[k] = d32;
      const double ad = apply_a(f, (double)d32, (double)(float)dE, (double)(float)dW,
                                (double)(float)dN, (double)(float)dS);
      p[k] = pr[q] + alpha * (double)d32;
      const double rn = rr[q] - alpha * ad;
      r[k] = rn;
      r32[k] = (float)rn;
      if ((f & 0x80) && (f & 7)) {
        acc[0] += rn * rn;
        acc[1] += 1.0;
      }
    }
  }
  block_sum_store_at<4>(acc, partials, bid);
}

template <int R>
__global__ __launch_bounds__(256) void k_fcg_update(
    const float* __restrict__ z, const float* __restrict__ dprev, float* __restrict__ dnext,
    double* __restrict__ p, double* __restrict__ r, const uint8_t* __restrict__ info,
    float* __restrict__ r32, const double* __restrict__ state, int nx, int ny, int first,
    double* partials, const int* __restrict__ done) {
  int bx, by;
  xcd_block(&bx, &by);
  fcg_update_body<R>(z, dprev, dnext, p, r, info, r32, state, nx, ny, first, partials, done, bx,
                     by, gridDim.y, blockIdx.x + blockIdx.y * gridDim.x);
}

// pressure.ts:102-161 norms of a pressure held in FP32 (p32) or FP64 (p64):
// sums = {sum rr^2, n, sum div^2, n_fluid}, rr = div + (count p - sum)/scale.
// Neighbour order +x, +y, -x, -y as in rmsRemainingDivergence.
__global__ __launch_bounds__(256) void k_norms(const float* __restrict__ p32,
                                               const double* __restrict__ p64,
                                               const double* __restrict__ div,
                                               const uint8_t* __restrict__ lbl, int nx, int ny,
                                               double scale, double* partials) {
  int bx, by;
  xcd_block(&bx, &by);
  const int i = bx * BX + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = by * BY + threadIdx.y; j < ny; j += gridDim.y * BY) {
    if (i >= nx) break;
    int k = i + j * nx;
    if (lbl[k] != FLUID) continue;
    double d = div[k];
    acc[2] += d * d;
    acc[3] += 1.0;
    auto P = [&](int q) { return p32 ? (double)p32[q] : p64[q]; };
    int cnt = 0;
    double sum = 0.0;
    if (!solid_or_outside(lbl, nx, ny, i + 1, j)) { cnt++; sum += P(k + 1); }
    if (!solid_or_outside(lbl, nx, ny, i, j + 1)) { cnt++; sum += P(k + nx); }
    if (!solid_or_outside(lbl, nx, ny, i - 1, j)) { cnt++; sum += P(k - 1); }
    if (!solid_or_outside(lbl, nx, ny, i, j - 1)) { cnt++; sum += P(k - nx); }
    if (cnt > 0) {
      double rr = d + ((double)cnt * P(k) - sum) / scale;
      acc[0] += rr * rr;
      acc[1] += 1.0;
    }
  }
  block_sum_store<4>(acc, partials);
}

// Reductions with more than kFoldAbove block partials are folded first:
// block b of kFoldBlocks sums the partials of one fixed contiguous range of
// producing blocks in a fixed order, written after the originals (the
// partials buffers hold 4 kFoldBlocks doubles more), so the one-block kernels
// below read kFoldBlocks partials instead of up to 65536 (2 MB through one
// CU: 45 us at 8192^2) and stay deterministic.  Skips like the kernel that
// consumes it (done set, outside the initial pass).
__global__ __launch_bounds__(256) void k_fold_partials(const double* __restrict__ in, int n,
                                                       double* __restrict__ out,
                                                       const int* __restrict__ done, int init) {
  if (!init && done && *done) return;
  const int chunk = (n + kFoldBlocks - 1) / kFoldBlocks;
  const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  const double2* p2 = reinterpret_cast<const double2*>(in);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = b0 + (int)threadIdx.x; b < b1; b += 256) {
    const double2 lo = p2[2 * b], hi = p2[2 * b + 1];
    acc[0] += lo.x;
    acc[1] += lo.y;
    acc[2] += hi.x;
    acc[3] += hi.y;
  }
  block_sum_store<4>(acc, out);
}

// The partials a one-block reduction reads: folded first when there are many.
static inline const double* folded(hipStream_t s, const double* partials, int* n,
                                   const int* done, int init) {
  if (*n <= kFoldAbove) return partials;
  double* out = const_cast<double*>(partials) + 4 * (size_t)*n;
  k_fold_partials<<<kFoldBlocks, 256, 0, s>>>(partials, *n, out, done, init);
  *n = kFoldBlocks;
  return out;
}

// One block sums the 4-slot partials in block order (deterministic) into
// out[0..3].  With tol > 0 it also writes out[4] = the relative residual
// sqrt(out0/out1) / (|scale| sqrt(out2/out3)) (the defect-correction norm)
// and sets *done (and out[5] = 1) once it is <= tol.  host_out, if given,
// is a mapped pinned host buffer that receives the same 6 values, so the host
// reads them after an event without a copy on the stream.  nblocks <=
// kMaxPartials.
__device__ __forceinline__ void finalize_sum4_body(const double* __restrict__ partials,
                                                   int nblocks, double* __restrict__ out,
                                                   double* __restrict__ host_out,
                                                   int* __restrict__ done, double scale,
                                                   double tol, int init,
                                                   const double* __restrict__ fixed23) {
  if (!init && done && *done) return;
  __shared__ double sh[4][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int K = 8;
  const double2* p2 = reinterpret_cast<const double2*>(partials);
  const int chunks = (nblocks + 1023) / 1024;  // block-uniform
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < chunks; c += K) {
    double2 lo[K], hi[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      int b = tid + (c + k) * 1024;
      lo[k] = b < nblocks ? p2[2 * b] : make_double2(0.0, 0.0);
      hi[k] = b < nblocks ? p2[2 * b + 1] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      acc[0] += lo[k].x;
      acc[1] += lo[k].y;
      acc[2] += hi[k].x;
      acc[3] += hi[k].y;
    }
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    double t = wave_sum(acc[s]);
    if (lane == 0) sh[s][wid] = t;
  }
  __syncthreads();
  if (tid == 0) {
    double v[4];
    for (int s = 0; s < 4; s++) {
      double a = 0.0;
      for (int w = 0; w < 16; w++) a += sh[s][w];
      v[s] = a;
    }
    if (fixed23) {  // the divergence terms, from the pass before the first cycle
      v[2] = fixed23[2];
      v[3] = fixed23[3];
    }
    for (int s = 0; s < 4; s++) out[s] = v[s];
    double rel = 0.0;
    if (v[1] > 0 && v[2] > 0) rel = sqrt(v[0] / v[1]) / (fabs(scale) * sqrt(v[2] / v[3]));
    out[4] = rel;
    // init (the pass before the first cycle): *done is (re)set from the
    // zero-divergence guard alone; otherwise it is set once rel <= tol.
    bool stop = init ? v[2] == 0.0 : (tol > 0 && rel <= tol);
    out[5] = stop ? 1.0 : 0.0;
    if (host_out) {
      for (int s = 0; s < 4; s++) host_out[s] = v[s];
      host_out[4] = rel;
      host_out[5] = out[5];
    }
    if (done && (stop || init)) *done = stop ? 1 : 0;
  }
}

__global__ __launch_bounds__(1024) void k_finalize_sum4(const double* __restrict__ partials,
                                                        int nblocks, double* __restrict__ out,
                                                        double* __restrict__ host_out,
                                                        int* __restrict__ done, double scale,
                                                        double tol, int init,
                                                        const double* __restrict__ fixed23) {
  finalize_sum4_body(partials, nblocks, out, host_out, done, scale, tol, init, fixed23);
}

void launch_finalize_sum4(hipStream_t s, const double* partials, int nblocks, double* out) {
  partials = folded(s, partials, &nblocks, nullptr, 1);
  k_finalize_sum4<<<1, 1024, 0, s>>>(partials, nblocks, out, nullptr, nullptr, 1.0, 0.0, 0,
                                     nullptr);
}

__global__ void k_set_int(int* p, int v) { *p = v; }

void launch_set_int(hipStream_t s, int* p, int v) { k_set_int<<<1, 1, 0, s>>>(p, v); }

int device_max_lds() {
  // the workgroup LDS limit of the current device, read once per device;
  // engines on other host threads may race here, so the cache is atomic
  // (they can only store the same value)
  constexpr int kMaxDevices = 64;
  static std::atomic<int> limit[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = -1;
  int v = dev >= 0 ? limit[dev].load(std::memory_order_relaxed) : 0;
  if (v <= 0) {
    if (dev < 0 || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                       hipSuccess)
      v = 65536;
    if (dev >= 0) limit[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

bool rb_sor_lds_fits(int nx, int ny) {
  return (size_t)nx * ny * 9 <= (size_t)device_max_lds() &&
         (long)((nx + 1) / 2) * ny <= (long)kSorThreads * kSorItems;
}

void launch_rb_sor_lds(hipStream_t s, float* x, const float* b, const uint8_t* lbl, int nx, int ny,
                       int iters, float omega, float scale) {
  k_rb_sor_lds<<<1, kSorThreads, (size_t)nx * ny * 9, s>>>(x, b, lbl, nx, ny, iters, omega,
                                                           scale);
}

void launch_rb_sweep(hipStream_t s, float* x, const float* b, const uint8_t* lbl, int nx, int ny,
                     int color, float omega, float scale, int first) {
  k_rb_sweep<<<grid2((nx + 1) / 2, ny), dim3(BX, BY), 0, s>>>(x, b, lbl, nx, ny, color, omega,
                                                              scale, first);
}

void launch_residual_restrict(hipStream_t s, const float* x, const float* b, const uint8_t* lbl,
                              int nx, int ny, float* bc, float* xc, int cnx, int cny) {
  k_residual_restrict<<<grid2(cnx, cny), dim3(BX, BY), 0, s>>>(x, b, lbl, nx, ny, bc, xc, cnx,
                                                               cny);
}

void launch_prolong(hipStream_t s, float* x, const uint8_t* lbl, int nx, int ny, const float* e,
                    int cnx, int cny) {
  k_prolong<<<grid2(nx, ny), dim3(BX, BY), 0, s>>>(x, lbl, nx, ny, e, cnx, cny);
}

void launch_coarsen_labels(hipStream_t s, const uint8_t* f, int fnx, int fny, uint8_t* c, int cnx,
                           int cny) {
  k_coarsen_labels<<<grid2(cnx, cny), dim3(BX, BY), 0, s>>>(f, fnx, fny, c, cnx, cny);
}

void launch_coarse_solve(hipStream_t s, const float* b, float* x, const double* M,
                         const int* cells, int n) {
  k_coarse_solve<<<1, 256, 0, s>>>(b, x, M, cells, n);
}

void launch_rhs32(hipStream_t s, const double* div, float* out, int n, double factor) {
  k_rhs32<<<(n + 255) / 256, 256, 0, s>>>(div, out, n, factor);
}

void launch_f64_to_f32(hipStream_t s, const double* a, float* out, int n) {
  k_f64_to_f32<<<(n + 255) / 256, 256, 0, s>>>(a, out, n);
}

void launch_f32_to_f64(hipStream_t s, const float* a, double* out, int n) {
  k_f32_to_f64<<<(n + 255) / 256, 256, 0, s>>>(a, out, n);
}

void launch_extrapolate(hipStream_t s, const double* p, const Extrapolation& xp, int order,
                        double* out, long n) {
  const long n2 = (n + 1) / 2;
  const unsigned g = (unsigned)((n2 + 255) / 256);
  if (order <= 0)
    k_extrapolate<0><<<g, 256, 0, s>>>(p, xp, out, n2);
  else if (order == 1)
    k_extrapolate<1><<<g, 256, 0, s>>>(p, xp, out, n2);
  else
    k_extrapolate<2><<<g, 256, 0, s>>>(p, xp, out, n2);
}

void launch_dc_update(hipStream_t s, const double* po, double* pn, const float* e,
                      const double* div, const uint8_t* info, float* r32, int nx, int ny,
                      double scale, double* partials, double* sums, double* host_sums, int* done,
                      double tol) {
  launch_dc_correct(s, po, pn, e, div, info, r32, nx, ny, scale, partials, done);
  launch_dc_norm(s, partials, nx, ny, sums, host_sums, done, scale, tol);
}

void launch_dc_correct(hipStream_t s, const double* po, double* pn, const float* e,
                       const double* div, const uint8_t* info, float* r32, int nx, int ny,
                       double scale, double* partials, const int* done, double* r64) {
  k_dc_update<kRedRows><<<reduce_grid_rows(nx, ny), dim3(BX, BY), 0, s>>>(
      po, pn, e, div, info, r32, r64, nx, ny, scale, partials, done);
}

void launch_div_defect(hipStream_t s, const float* u, const float* v, double inv_h, double* div,
                       const double* po, const uint8_t* info, float* r32, double* r64, int nx,
                       int ny, double scale, double* partials, double* pn,
                       const Extrapolation* xp, int order) {
  const dim3 g = reduce_grid_rows(nx, ny), b(BX, BY);
  const Extrapolation x = xp ? *xp : Extrapolation{};
  if (!pn || order <= 0)
    k_div_defect<kRedRows, 0><<<g, b, 0, s>>>(u, v, inv_h, div, po, info, r32, r64, nx, ny,
                                               scale, partials, pn, x);
  else if (order == 1)
    k_div_defect<kRedRows, 1><<<g, b, 0, s>>>(u, v, inv_h, div, po, info, r32, r64, nx, ny,
                                               scale, partials, pn, x);
  else
    k_div_defect<kRedRows, 2><<<g, b, 0, s>>>(u, v, inv_h, div, po, info, r32, r64, nx, ny,
                                               scale, partials, pn, x);
}

void launch_dc_norm(hipStream_t s, const double* partials, int nx, int ny, double* sums,
                    double* host_sums, int* done, double scale, double tol) {
  const dim3 g = reduce_grid_rows(nx, ny);
  int n = g.x * g.y;
  partials = folded(s, partials, &n, done, 0);
  k_finalize_sum4<<<1, 1024, 0, s>>>(partials, n, sums, host_sums, done, scale, tol, 0, nullptr);
}

void launch_dc_norm_n(hipStream_t s, const double* partials, int n, double* sums,
                      double* host_sums, int* done, double scale, double tol) {
  partials = folded(s, partials, &n, done, 0);
  k_finalize_sum4<<<1, 1024, 0, s>>>(partials, n, sums, host_sums, done, scale, tol, 0, nullptr);
}

void launch_dc_norm_init(hipStream_t s, const double* partials, int nx, int ny, double* sums,
                         double* host_sums, int* done, double scale) {
  const dim3 g = reduce_grid_rows(nx, ny);
  int n = g.x * g.y;
  partials = folded(s, partials, &n, done, 1);
  k_finalize_sum4<<<1, 1024, 0, s>>>(partials, n, sums, host_sums, done, scale, 0.0, 1, nullptr);
}

void launch_fcg_dots(hipStream_t s, const float* z, const float* dprev, const double* r,
                     const uint8_t* info, int nx, int ny, int first, double* partials,
                     double* state, const int* done) {
  const dim3 g = reduce_grid_rows(nx, ny);
  k_fcg_dots<kRedRows><<<g, dim3(BX, BY), 0, s>>>(z, dprev, r, info, nx, ny, first, partials,
                                                  done);
  int n = g.x * g.y;
  const double* pp = folded(s, partials, &n, done, 0);
  k_fcg_coef<<<1, 1024, 0, s>>>(pp, n, state, first, done);
}

void launch_fcg_update(hipStream_t s, const float* z, const float* dprev, float* dnext, double* p,
                       double* r, const uint8_t* info, float* r32, int nx, int ny, int first,
                       const double* state, double* partials, const int* done) {
  k_fcg_update<kRedRows><<<reduce_grid_rows(nx, ny), dim3(BX, BY), 0, s>>>(
      z, dprev, dnext, p, r, info, r32, state, nx, ny, first, partials, done);
}

void launch_fcg_norm(hipStream_t s, const double* partials, int nx, int ny, double* sums,
                     const double* sums0, double* host_sums, int* done, double scale,
                     double tol) {
  const dim3 g = reduce_grid_rows(nx, ny);
  int n = g.x * g.y;
  partials = folded(s, partials, &n, done, 0);
  k_finalize_sum4<<<1, 1024, 0, s>>>(partials, n, sums, host_sums, done, scale, tol, 0, sums0);
}

void launch_norms(hipStream_t s, const float* p32, const double* p64, const double* div,
                  const uint8_t* lbl, int nx, int ny, double scale, double* partials,
                  double* sums) {
  dim3 g = reduce_grid(nx, ny);
  k_norms<<<g, dim3(BX, BY), 0, s>>>(p32, p64, div, lbl, nx, ny, scale, partials);
  launch_finalize_sum4(s, partials, g.x * g.y, sums);
}

// ---- batched launches ----------------------------------------------------------
// The defect and CG passes over every scene of the work list (batch_block),
// each scene at its own rows per thread, and the one-block reductions as one
// block per scene.  A scene's partials land in its own buffer at the slots
// of its single-grid launch, so its norms have the same bits.

__global__ __launch_bounds__(256) void kb_dc_update(const SceneDesc* __restrict__ D, BatchGrid g,
                                                    const BStep* __restrict__ st, int q,
                                                    int init) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const BStep& t = st[s];
  if (init == 2) {  // the initial pass with the divergence of velocity [q ^ 1]
    Extrapolation x = t.xp;
    x.h0 = x.h0 ? gm(x.h0) : nullptr;
    x.h1 = x.h1 ? gm(x.h1) : nullptr;
    double* ps = t.ps ? gm(t.ps) : nullptr;
    const int X = ps ? t.xo : 0;
    if (X == 1)
      dc_update_body<kRedRows, true, 1>(gm(t.p0), ps, nullptr, nullptr, gm(d.lv[0].info),
                                        gm(d.lv[0].b), nullptr, d.nx, d.ny, t.scale,
                                        gm(d.partials), nullptr, bx, by, gy, bid, gm(d.u[q ^ 1]),
                                        gm(d.v[q ^ 1]), t.inv_h, gm(d.div), x);
    else if (X == 2)
      dc_update_body<kRedRows, true, 2>(gm(t.p0), ps, nullptr, nullptr, gm(d.lv[0].info),
                                        gm(d.lv[0].b), nullptr, d.nx, d.ny, t.scale,
                                        gm(d.partials), nullptr, bx, by, gy, bid, gm(d.u[q ^ 1]),
                                        gm(d.v[q ^ 1]), t.inv_h, gm(d.div), x);
    else
      dc_update_body<kRedRows, true, 0>(gm(t.p0), ps, nullptr, nullptr, gm(d.lv[0].info),
                                        gm(d.lv[0].b), nullptr, d.nx, d.ny, t.scale,
                                        gm(d.partials), nullptr, bx, by, gy, bid, gm(d.u[q ^ 1]),
                                        gm(d.v[q ^ 1]), t.inv_h, gm(d.div), x);
    return;
  }
  const double* start = t.ps ? t.ps : t.p0;
  double* pn = init ? nullptr : gm(t.pr);
  const float* e = init ? nullptr : gm(d.lv[0].x);
  const int* done = init ? nullptr : gm(d.done);
  dc_update_body<kRedRows>(gm(start), pn, e, gm(d.div), gm(d.lv[0].info), gm(d.lv[0].b),
                           init ? nullptr : gm(d.rvec), d.nx, d.ny, t.scale, gm(d.partials),
                           done, bx, by, gy, bid);
}

__global__ __launch_bounds__(1024) void kb_finalize(const SceneDesc* __restrict__ D,
                                                    const BStep* __restrict__ st, int mode,
                                                    int slot, double tol) {
  const SceneDesc& d = D[blockIdx.x];
  const double scale = st[blockIdx.x].scale;
  if (mode == 0)
    finalize_sum4_body(gm(d.partials), d.red_blocks, gm(d.sums), gm(d.d_res) + 16, gm(d.done),
                       scale, 0.0, 1, nullptr);
  else
    finalize_sum4_body(gm(d.partials), d.red_blocks, gm(d.sums), gm(d.d_res) + 8 * slot, gm(d.done),
                       scale, tol, 0, mode == 2 ? gm(d.sums) : nullptr);
}

__global__ __launch_bounds__(256) void kb_fcg_dots(const SceneDesc* __restrict__ D, BatchGrid g,
                                                   int cp, int first) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  fcg_dots_body<kRedRows>(gm(d.lv[0].x), gm(d.dvec[cp]), gm(d.rvec), gm(d.lv[0].info), d.nx, d.ny,
                          first, gm(d.partials), gm(d.done), bx, by, gy, bid);
}

__global__ __launch_bounds__(1024) void kb_fcg_coef(const SceneDesc* __restrict__ D, int first) {
  const SceneDesc& d = D[blockIdx.x];
  fcg_coef_body(gm(d.partials), d.red_blocks, gm(d.fcg), first, gm(d.done));
}

__global__ __launch_bounds__(256) void kb_fcg_update(const SceneDesc* __restrict__ D,
                                                     BatchGrid g, const BStep* __restrict__ st,
                                                     int cp, int first) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  fcg_update_body<kRedRows>(gm(d.lv[0].x), gm(d.dvec[cp]), gm(d.dvec[cp ^ 1]), gm(st[s].pr),
                            gm(d.rvec), gm(d.lv[0].info), gm(d.lv[0].b), gm(d.fcg), d.nx, d.ny,
                            first, gm(d.partials), gm(d.done), bx, by, gy, bid);
}

dim3 batch_grid_reduce(int nx, int ny) { return reduce_grid_rows(nx, ny); }

void launch_batch_dc_update(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                            const BStep* st, int q, int init) {
  kb_dc_update<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q, init);
}

void launch_batch_finalize(hipStream_t s, const SceneDesc* D, int nscenes, const BStep* st,
                           int mode, int slot, double tol) {
  kb_finalize<<<nscenes, 1024, 0, s>>>(D, st, mode, slot, tol);
}

void launch_batch_fcg(hipStream_t s, const SceneDesc* D, const BatchGrid& g, const BStep* st,
                      int nscenes, int cp, int first) {
  kb_fcg_dots<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, cp, first);
  kb_fcg_coef<<<nscenes, 1024, 0, s>>>(D, first);
  kb_fcg_update<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, cp, first);
}

}  // namespace fh
