// Multigrid tile bodies shared by the single-grid kernels (fh_mgfused.hip)
// and the batched ones (fh_batch.hip): LDS-tiled pre-smooth + restrict,
// prolong + post-smooth, and the one-workgroup coarse sub-cycle of the
// V(2,2) cycle (pressureMultigrid.ts:253-275).  The tile shape is a
// template parameter so the autotuner can pick it per level size; temporal
// blocking makes every tiling give the same bits.
#pragma once
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {

constexpr int NT = 256;  // threads per tile workgroup unless the shape says otherwise

// Tile shapes (output cells per workgroup, both sides even, and the
// workgroup's threads) the tiled smoothers are instantiated for; the
// autotuners choose among them per level.  A larger tile recomputes less
// halo (the 16x16 tile sweeps 2.25 cells per output cell, 64x64 1.27) but
// holds fewer workgroups per CU; the wider workgroups split a large tile's
// sweeps over more waves.  FH_TILE_SHAPES lists them for the launchers'
// switches: (index, tx, ty, threads).
#define FH_TILE_SHAPES(X)                                                            \
  X(0, 64, 32, 256) X(1, 32, 32, 256) X(2, 64, 16, 256) X(3, 128, 16, 256)          \
  X(4, 32, 16, 256) X(5, 16, 16, 256) X(6, 32, 8, 256) X(7, 64, 32, 512)            \
  X(8, 32, 32, 512) X(9, 64, 64, 1024) X(10, 128, 32, 1024)
struct TileShape {
  int tx, ty, nt;
  const char* name;
};
constexpr TileShape kTileShapes[] = {
    {64, 32, 256, "64x32"},       {32, 32, 256, "32x32"},      {64, 16, 256, "64x16"},
    {128, 16, 256, "128x16"},     {32, 16, 256, "32x16"},      {16, 16, 256, "16x16"},
    {32, 8, 256, "32x8"},         {64, 32, 512, "64x32/512"},  {32, 32, 512, "32x32/512"},
    {64, 64, 1024, "64x64/1024"}, {128, 32, 1024, "128x32/1024"}};
constexpr int kNumTileShapes = sizeof(kTileShapes) / sizeof(kTileShapes[0]);
constexpr int HA = 5;    // halo of the pre-smooth+restrict tile (4 half-sweeps + residual)
constexpr int HB = 4;    // halo of the prolong+post-smooth tile (4 half-sweeps)

// info byte of a cell: bit 7 = fluid; bits 3..6 = which neighbours are
// non-solid and in the domain (+x, -x, +y, -y); bits 0..2 = how many (the
// equation's diagonal).  0 off the domain and on non-fluid cells.
__device__ __forceinline__ uint8_t cell_info(const uint8_t* __restrict__ lbl, int nx, int ny,
                                             int i, int j) {
  if (i < 0 || j < 0 || i >= nx || j >= ny) return 0;
  if (lbl[i + j * nx] != FLUID) return 0;
  int e = !solid_or_outside(lbl, nx, ny, i + 1, j);
  int w = !solid_or_outside(lbl, nx, ny, i - 1, j);
  int n = !solid_or_outside(lbl, nx, ny, i, j + 1);
  int s = !solid_or_outside(lbl, nx, ny, i, j - 1);
  return (uint8_t)(0x80 | (e << 3) | (w << 4) | (n << 5) | (s << 6) | (e + w + n + s));
}

// 1 / count for a sweep, count in 1..4: 1, 1/2 and 1/4 are exact powers of
// two (exponent 127 - count/2), 1/3 is the rounded constant (within an ulp of
// the division); (sum + b) * recip is the update.  One select, no branch.
__device__ __forceinline__ float gs_recip(int cnt) {
  const float p2 = __int_as_float((127 - (cnt >> 1)) << 23);
  return cnt == 3 ? (1.f / 3.f) : p2;
}

// The defect-correction update's arithmetic at one cell with info byte f,
// given the updated pressure pc and its four neighbours' and the divergence
// d: returns r = b - A p (b = -scale d; 0 off the fluid or without an open
// neighbour, neighbours summed +x, -x, +y, -y) and adds the norm terms
// {r^2, cells with such an r, d^2, fluid cells} to acc.  Contraction is off,
// so every pass that forms it (the update pass, and the finest prolongation
// tile fused with it) gives the same bits.
__device__ __forceinline__ double dc_cell(int f, double d, double pc, double pE, double pW,
                                          double pN, double pS, double scale,
                                          double (&acc)[4]) {
#pragma clang fp contract(off)
  double rr = 0.0;
  if (f & 0x80) {
    acc[2] += d * d;
    acc[3] += 1.0;
    const int cnt = f & 7;
    if (cnt > 0) {
      double sum = 0.0;
      sum += (f & 0x08) ? pE : 0.0;
      sum += (f & 0x10) ? pW : 0.0;
      sum += (f & 0x20) ? pN : 0.0;
      sum += (f & 0x40) ? pS : 0.0;
      rr = -scale * d - ((double)cnt * pc - sum);
      acc[0] += rr * rr;
      acc[1] += 1.0;
    }
  }
  return rr;
}

// Red-black Gauss-Seidel (omega = 1) on an LDS window of LX x LY cells.
// Labels and right-hand side do not change while a tile is smoothed, so each
// thread finds its cells once -- for each colour, K cells of the window's
// ring-1 interior, as LDS index, b and 1/count (0: not updated -- off the
// fluid, count 0, or a slot past the interior) -- and every half-sweep is
// then two neighbour reads, five flops and a write per cell.
//
// Every half-sweep updates the whole ring-1 interior.  A temporally blocked
// tile is exact only at LDS distance >= s from the border after s
// half-sweeps; cells nearer the border take values that differ from the
// global sweep, and the cells that read them are themselves nearer than s+1.
// The exact region is the same as when the updated region shrinks by one
// ring per half-sweep, so the tile interior gets the same bits.
template <int LX, int LY, int W = NT>
struct TileSweep {
  static constexpr int P = (LX + 1) / 2;  // pairs per row
  static constexpr int ITEMS = P * (LY - 2);
  static constexpr int K = (ITEMS + W - 1) / W;
  int l[2][K];
  float b[2][K], r[2][K];

  // par0 = parity of LDS (0,0) in global coordinates; colour c updates the
  // cells of global parity (i + j) & 1 == c.
  __device__ __forceinline__ void setup(const float* sb, const uint8_t* si, int par0, int tid) {
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int q = tid + k * W;
        const int lj = 1 + q / P;
        const int pc = q - (lj - 1) * P;
        const int li = 2 * pc + ((lj + par0 + c) & 1);
        const bool act = q < ITEMS && li >= 1 && li <= LX - 2;
        const int lq = act ? li + lj * LX : LX + 1;  // inactive: any interior cell
        const int cnt = act ? (si[lq] & 7) : 0;
        l[c][k] = lq;
        b[c][k] = sb[lq];
        r[c][k] = cnt ? gs_recip(cnt) : 0.f;
      }
  }

  // One half-sweep of colour C; zero: the unknown is zero on entry
  // (neighbour sum 0).  All reads are issued before any write.
  template <int C>
  __device__ __forceinline__ void sweep(float* sx, bool zero) const {
    float nv[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = l[C][k];
      const float sum = zero ? 0.f : ((sx[q + 1] + sx[q - 1]) + sx[q + LX]) + sx[q - LX];
      nv[k] = (sum + b[C][k]) * r[C][k];
    }
#pragma unroll
    for (int k = 0; k < K; k++)
      if (r[C][k] != 0.f) sx[l[C][k]] = nv[k];
  }
};

// LDS-typed pointers for the one-workgroup coarse levels, whose arrays sit at
// runtime offsets of one dynamic LDS block: with the address space explicit
// the accesses are ds_read/ds_write rather than flat ones.
typedef __attribute__((address_space(3))) float* lds_fp;
typedef __attribute__((address_space(3))) uint8_t* lds_bp;

// q / d for 0 <= q < 2^16 and 1 <= d < 2^16 by a multiply-high.  With
// m = ceil(2^32 / d) = d m' ... : m d = 2^32 + e, 0 <= e < d, so
// q m / 2^32 = q / d + q e / (d 2^32), and the second term is below
// q / 2^32 < 2^-16 < 1/d, too small to carry floor(q / d) (whose fractional
// part is at most 1 - 1/d) past the next integer.
struct FastDiv {
  unsigned m;
  int d;
  __device__ __forceinline__ explicit FastDiv(int dd) : m(fastdiv_magic(dd)), d(dd) {}
  // with the multiplier computed on the host (fastdiv_magic)
  __device__ __forceinline__ FastDiv(unsigned mm, int dd) : m(mm), d(dd) {}
  __device__ __forceinline__ int div(int q) const {
    return d > 1 ? (int)__umulhi((unsigned)q, m) : q;
  }
};

// The one-workgroup coarse levels' variant of TileSweep: runtime level sizes
// (arrays padded by a ring of zero cells, pitch LX), NC threads and up to KC
// cells per thread and colour, so a level holds at most NC * KC cells of a
// colour (the host keeps coarse levels within that).  A wave whose slots are
// all past the level's cells skips them, so on the small levels most of the
// workgroup does nothing.
template <int NC, int KC>
struct CoarseSweep {
  int l[2][KC];
  float b[2][KC], r[2][KC];
  int nk;  // slots holding a cell for this wave (wave-uniform)
  int LX;

  // dP: division by the pairs per row, (lx - 1) / 2
  __device__ __forceinline__ void setup(const lds_fp sb, const lds_bp si, int lx, int ly,
                                        FastDiv dP, int tid) {
    LX = lx;
    const int P = (lx - 1) >> 1;  // pairs per row of the nx = lx - 2 cells
    const int total = P * (ly - 2);
    const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63);
    nk = total > wbase ? min(KC, (total - wbase + NC - 1) / NC) : 0;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int k = 0; k < KC; k++) {
        l[c][k] = lx + 1;
        b[c][k] = 0.f;
        r[c][k] = 0.f;
        if (k >= nk) continue;
        const int q = tid + k * NC;
        const int rw = dP.div(q), pc = q - rw * P;
        const int lj = 1 + rw;
        // padded (0,0) is global (-1,-1): global parity of (li, lj) = li + lj
        const int li = 1 + 2 * pc + ((1 + lj + c) & 1);
        const bool act = q < total && li <= lx - 2;
        const int lq = act ? li + lj * lx : lx + 1;
        const int cnt = act ? (si[lq] & 7) : 0;
        l[c][k] = lq;
        b[c][k] = sb[lq];
        r[c][k] = cnt ? gs_recip(cnt) : 0.f;
      }
  }

  template <int C>
  __device__ __forceinline__ void sweep(lds_fp sx, bool zero) const {
    float nv[KC];
#pragma unroll
    for (int k = 0; k < KC; k++) {
      nv[k] = 0.f;
      if (k >= nk) continue;
      const int q = l[C][k];
      const float sum = zero ? 0.f : ((sx[q + 1] + sx[q - 1]) + sx[q + LX]) + sx[q - LX];
      nv[k] = (sum + b[C][k]) * r[C][k];
    }
#pragma unroll
    for (int k = 0; k < KC; k++)
      if (k < nk && r[C][k] != 0.f) sx[l[C][k]] = nv[k];
  }

  // Four half-sweeps (two red-black sweeps), colour 0 first, with a barrier
  // after each.
  __device__ __forceinline__ void two_sweeps(lds_fp sx, bool zero) const {
    sweep<0>(sx, zero);
    __syncthreads();
    sweep<1>(sx, false);
    __syncthreads();
    sweep<0>(sx, false);
    __syncthreads();
    sweep<1>(sx, false);
    __syncthreads();
  }
};

// Loads an LX x LY window with global origin (gx0, gy0) into LDS (zeros off
// the domain; x = 0 when xin is null).  Every thread first issues all of its
// global loads into registers and only then writes LDS, so one thread waits
// for memory once instead of once per element: with the trip count known at
// compile time the loop unrolls and the loads are in flight together.
//
// The defect-correction loop's `done` flag is read alongside: returns false
// (before any LDS write, uniformly for the workgroup) if it is set, so the
// flag costs no memory round trip of its own.
template <int LX, int LY, int W = NT>
__device__ __forceinline__ bool stage_tile(float* sx, float* sb, uint8_t* si,
                                           const float* __restrict__ xin,
                                           const float* __restrict__ b,
                                           const uint8_t* __restrict__ info, int nx, int ny,
                                           int gx0, int gy0, int tid,
                                           const int* __restrict__ done) {
  const int stop = done ? *done : 0;
  constexpr int N = LX * LY, K = (N + W - 1) / W;
  float rx[K], rb[K];
  uint8_t ri[K];
  if (gx0 >= 0 && gy0 >= 0 && gx0 + LX <= nx && gy0 + LY <= ny) {
    // the window lies inside the domain (most tiles): no cell needs a test,
    // and the last partial slot reads a clamped in-window index
    const int base = gx0 + gy0 * nx;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = min(tid + k * W, N - 1);
      const int lj = q / LX, li = q - lj * LX;
      const int g = base + li + lj * nx;
      rx[k] = xin ? xin[g] : 0.f;
      rb[k] = b[g];
      ri[k] = info[g];
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = tid + k * W;
      int lj = q / LX, li = q - lj * LX;
      int gi = gx0 + li, gj = gy0 + lj;
      bool in = q < N && gi >= 0 && gj >= 0 && gi < nx && gj < ny;
      int g = in ? gi + gj * nx : 0;
      rx[k] = (in && xin) ? xin[g] : 0.f;
      rb[k] = in ? b[g] : 0.f;
      ri[k] = in ? info[g] : 0;
    }
  }
  if (stop) return false;
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = tid + k * W;
    if (q < N) {
      sx[q] = rx[k];
      sb[q] = rb[k];
      si[q] = ri[k];
    }
  }
  return true;
}

// r = b - (count x - sum) at an LDS cell (0 off the fluid or if count = 0).
template <class XP, class BP, class IP>
__device__ __forceinline__ float lds_residual(XP sx, BP sb, IP si, int LX, int l) {
  int cnt = si[l] & 7;
  if (cnt == 0) return 0.f;
  float sum = ((sx[l + 1] + sx[l - 1]) + sx[l + LX]) + sx[l - LX];
  return sb[l] - ((float)cnt * sx[l] - sum);
}

// ---- k_smooth_restrict ------------------------------------------------------
// Pre-smoothing (2 sweeps) of level (nx,ny) and restriction of the residual
// to the coarse right-hand side bc (cnx,cny).  x is read from xin unless
// zero (then the unknown starts at 0), written to xout (xin != xout: halos
// are read from global memory while other tiles write their interiors).
// One TX x TY tile (tile index bx, by) per call; TX and TY even, so the
// children of a coarse cell never straddle tiles.  A caller that runs
// several tiles per workgroup puts a barrier between calls (LDS reuse).
template <int TX, int TY, int W = NT>
__device__ __forceinline__ void smooth_restrict_tile(const float* __restrict__ xin,
                                                     float* __restrict__ xout,
                                                     const float* __restrict__ b,
                                                     const uint8_t* __restrict__ info, int nx,
                                                     int ny, float* __restrict__ bc, int cnx,
                                                     int cny, int zero,
                                                     const int* __restrict__ done, int bx,
                                                     int by) {
  constexpr int AX = TX + 2 * HA, AY = TY + 2 * HA;
  __shared__ float sx[AX * AY];
  __shared__ float sb[AX * AY];
  __shared__ uint8_t si[AX * AY];
  // Opaque per call: keeps the compiler from hoisting tid-only index
  // arithmetic out of a caller's persistent task loop (which would hold
  // all of it in registers across tasks).
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  const int ox = bx * TX, oy = by * TY;
  const int gx0 = ox - HA, gy0 = oy - HA;
  if (!stage_tile<AX, AY, W>(sx, sb, si, zero ? nullptr : xin, b, info, nx, ny, gx0, gy0, tid,
                            done))
    return;
  __syncthreads();
  TileSweep<AX, AY, W> sw;
  sw.setup(sb, si, (gx0 + gy0) & 1, tid);
  sw.template sweep<0>(sx, zero);
  __syncthreads();
  sw.template sweep<1>(sx, false);
  __syncthreads();
  sw.template sweep<0>(sx, false);
  __syncthreads();
  sw.template sweep<1>(sx, false);
  __syncthreads();
  // coarse cells of this tile: children (2I+di, 2J+dj) are tile-interior
  const int CX = TX / 2, CY = TY / 2;
  for (int q = tid; q < CX * CY; q += W) {
    int cj = q / CX, ci = q - cj * CX;
    int I = ox / 2 + ci, J = oy / 2 + cj;
    if (I >= cnx || J >= cny) continue;
    float s = 0.f;
#pragma unroll
    for (int dj = 0; dj < 2; dj++)
#pragma unroll
      for (int di = 0; di < 2; di++) {
        int gi = 2 * I + di, gj = 2 * J + dj;
        if (gi < nx && gj < ny) s += lds_residual(sx, sb, si, AX, (gi - gx0) + (gj - gy0) * AX);
      }
    bc[I + J * cnx] = s;
  }
  for (int q = tid; q < TX * TY; q += W) {
    int lj = q / TX, li = q - lj * TX;
    int gi = ox + li, gj = oy + lj;
    if (gi < nx && gj < ny) xout[gi + gj * nx] = sx[(li + HA) + (lj + HA) * AX];
  }
}

// ---- k_prolong_smooth --------------------------------------------------------
// x += bilinear(e) on fluid cells (pressureMultigrid.ts:192-207), then 2
// post-smoothing sweeps, on a TX x TY tile (origin ox, oy) staged with a halo
// of H cells: the LDS windows sx (x), sb (b), si (info) of (TX + 2H) x
// (TY + 2H) cells and se of the coarse e around them.  Returns false
// (uniformly, before any LDS write) when `done` is set.  Afterwards every
// window cell at distance >= 4 from the border holds the global sweep's value
// (the tile interior at HB = 4, one more ring at H = 5).
template <int TX, int TY, int W, int H>
__device__ __forceinline__ bool prolong_smooth_core(float* sx, float* sb, uint8_t* si, float* se,
                                                    const float* __restrict__ xin,
                                                    const float* __restrict__ b,
                                                    const uint8_t* __restrict__ info, int nx,
                                                    int ny, const float* __restrict__ e, int cnx,
                                                    int cny, const int* __restrict__ done, int ox,
                                                    int oy, int tid) {
  constexpr int BXW = TX + 2 * H, BYH = TY + 2 * H;
  constexpr int EX = BXW / 2 + 3, EY = BYH / 2 + 3;
  const int gx0 = ox - H, gy0 = oy - H;
  // coarse window: parents I = gi>>1 and their +-1 neighbours
  const int ex0 = (gx0 >> 1) - 1, ey0 = (gy0 >> 1) - 1;
  {
    constexpr int K = (EX * EY + W - 1) / W;
    float re[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = tid + k * W;
      int ej = q / EX, ei = q - ej * EX;
      int I = ex0 + ei, J = ey0 + ej;
      bool in = q < EX * EY && I >= 0 && J >= 0 && I < cnx && J < cny;
      re[k] = in ? e[I + J * cnx] : 0.f;
    }
    if (!stage_tile<BXW, BYH, W>(sx, sb, si, xin, b, info, nx, ny, gx0, gy0, tid, done))
      return false;
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = tid + k * W;
      if (q < EX * EY) se[q] = re[k];
    }
  }
  __syncthreads();
  TileSweep<BXW, BYH, W> sw;
  sw.setup(sb, si, (gx0 + gy0) & 1, tid);
  for (int q = tid; q < BXW * BYH; q += W) {
    if (!(si[q] & 0x80)) continue;
    int lj = q / BXW, li = q - lj * BXW;
    int gi = gx0 + li, gj = gy0 + lj;
    int I = gi >> 1, J = gj >> 1;
    int I2 = (gi & 1) ? min(I + 1, cnx - 1) : max(I - 1, 0);
    int J2 = (gj & 1) ? min(J + 1, cny - 1) : max(J - 1, 0);
    float e00 = se[(I - ex0) + (J - ey0) * EX];
    float e10 = se[(I2 - ex0) + (J - ey0) * EX];
    float e01 = se[(I - ex0) + (J2 - ey0) * EX];
    float e11 = se[(I2 - ex0) + (J2 - ey0) * EX];
    sx[q] += 0.5625f * e00 + 0.1875f * (e10 + e01) + 0.0625f * e11;
  }
  __syncthreads();
  sw.template sweep<0>(sx, false);
  __syncthreads();
  sw.template sweep<1>(sx, false);
  __syncthreads();
  sw.template sweep<0>(sx, false);
  __syncthreads();
  sw.template sweep<1>(sx, false);
  __syncthreads();
  return true;
}

// Prolongation + post-smoothing of one tile; reads xin, writes xout (distinct
// buffers).
template <int TX, int TY, int W = NT>
__device__ __forceinline__ void prolong_smooth_tile(const float* __restrict__ xin,
                                                    float* __restrict__ xout,
                                                    const float* __restrict__ b,
                                                    const uint8_t* __restrict__ info, int nx,
                                                    int ny, const float* __restrict__ e, int cnx,
                                                    int cny, const int* __restrict__ done, int bx,
                                                    int by) {
  constexpr int BXW = TX + 2 * HB, BYH = TY + 2 * HB;
  constexpr int EX = BXW / 2 + 3, EY = BYH / 2 + 3;
  __shared__ float sx[BXW * BYH];
  __shared__ float sb[BXW * BYH];
  __shared__ uint8_t si[BXW * BYH];
  __shared__ float se[EX * EY];
  // Opaque per call: keeps the compiler from hoisting tid-only index
  // arithmetic out of a caller's persistent task loop (which would hold
  // all of it in registers across tasks).
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  const int ox = bx * TX, oy = by * TY;
  if (!prolong_smooth_core<TX, TY, W, HB>(sx, sb, si, se, xin, b, info, nx, ny, e, cnx, cny,
                                          done, ox, oy, tid))
    return;
  for (int q = tid; q < TX * TY; q += W) {
    int lj = q / TX, li = q - lj * TX;
    int gi = ox + li, gj = oy + lj;
    if (gi < nx && gj < ny) xout[gi + gj * nx] = sx[(li + HB) + (lj + HB) * BXW];
  }
}

// The finest level's prolongation + post-smoothing fused with the
// defect-correction update that follows it (dc_cell at every tile cell, the
// correction e taken from the tile one ring beyond its interior, hence the
// halo HB + 1): pn = po + e, r = b - A pn with b = -scale div, the FP32 r to
// rout (which must differ from b, whose halos other tiles still read), the
// FP64 r to r64 when given, and the tile's norm terms to partials[bid].  The
// correction itself is not stored: the next cycle starts from zero.
template <int TX, int TY, int W = NT>
__device__ __forceinline__ void prolong_smooth_dc_tile(
    const float* __restrict__ xin, const float* __restrict__ b, const uint8_t* __restrict__ info,
    int nx, int ny, const float* __restrict__ e, int cnx, int cny, const int* __restrict__ done,
    int bx, int by, const double* __restrict__ po, double* __restrict__ pn,
    const double* __restrict__ div, float* __restrict__ rout, double* __restrict__ r64,
    double scale, double* __restrict__ partials, int bid) {
  constexpr int H = HB + 1;
  constexpr int BXW = TX + 2 * H, BYH = TY + 2 * H;
  constexpr int EX = BXW / 2 + 3, EY = BYH / 2 + 3;
  __shared__ float sx[BXW * BYH];
  __shared__ float sb[BXW * BYH];
  __shared__ uint8_t si[BXW * BYH];
  __shared__ float se[EX * EY];
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  const int ox = bx * TX, oy = by * TY;
  if (!prolong_smooth_core<TX, TY, W, H>(sx, sb, si, se, xin, b, info, nx, ny, e, cnx, cny, done,
                                         ox, oy, tid))
    return;
  // each thread takes R consecutive rows of one tile column: the column's
  // pressure over rows -1 .. R (north and south shared between rows) and
  // its east and west neighbours, every load issued before the arithmetic;
  // indices clamped into the domain as the update pass clamps them (a closed
  // neighbour's value is not used)
  static_assert(W % TX == 0 && TY % (W / TX) == 0, "tile rows per thread");
  constexpr int R = TY / (W / TX);
  const int li = tid % TX, lj0 = (tid / TX) * R;
  const int gi = ox + li;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (gi < nx) {
    const int iE = gi + 1 < nx ? gi + 1 : gi, iW = gi > 0 ? gi - 1 : gi;
    double pc[R + 2], pe[R], pw[R], dv[R];
#pragma unroll
    for (int r = 0; r < R + 2; r++) pc[r] = po[gi + min(max(oy + lj0 - 1 + r, 0), ny - 1) * nx];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = min(oy + lj0 + r, ny - 1) * nx;
      pe[r] = po[iE + row];
      pw[r] = po[iW + row];
      dv[r] = div[gi + row];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int gj = oy + lj0 + r;
      if (gj >= ny) break;
      const int k = gi + gj * nx, l = (li + H) + (lj0 + r + H) * BXW;
      const double c = pc[r + 1] + (double)sx[l];
      const double pE = pe[r] + (double)sx[l + 1];
      const double pW = pw[r] + (double)sx[l - 1];
      const double pN = pc[r + 2] + (double)sx[l + BXW];
      const double pS = pc[r] + (double)sx[l - BXW];
      pn[k] = c;
      const double rr = dc_cell(si[l], dv[r], c, pE, pW, pN, pS, scale, acc);
      rout[k] = (float)rr;
      if (r64) r64[k] = rr;
    }
  }
  block_sum_store_at<4>(acc, partials, bid);
}

// ---- k_coarse_vcycle -----------------------------------------------------------
// The whole sub-cycle from level lc down, in one workgroup's LDS, laid out as
// CoarseLevels describes (each level padded by a one-cell ring of zeros, so
// the half-sweep's neighbour reads never leave the level).
// The workgroup size NC is a template parameter (the autotuner picks it): a
// larger group has more lanes per phase, a smaller one cheaper barriers.  KC
// = cells of a colour per thread, so NC*KC covers the usual coarse hierarchy
// in one pass.  With up to 16 waves on one CU every VALU instruction per
// thread costs ~16 cycles of each SIMD, so the per-level offsets and
// division multipliers come from the host and the staging is a straight copy.
constexpr int NC = 1024;  // the batch's coarse workgroup
constexpr int kCoarseGroups[] = {1024, 512, 256};
constexpr int kNumCoarseGroups = 3;

// One coarse level's LDS arrays and sizes.
struct CLevel {
  lds_fp x, b;
  lds_bp i;
  int nx, ny, px, py;
};

__device__ __forceinline__ CLevel clevel(lds_bp smem, const CoarseLevels& L, int l) {
  const int nx = L.nx[l], ny = L.ny[l];
  return {(lds_fp)smem + L.xo[l], (lds_fp)smem + L.bo[l], smem + L.io[l], nx, ny, nx + 2, ny + 2};
}

// Body of the sub-cycle for a workgroup of NC threads; smem is the dynamic
// LDS of coarse_vcycle_lds_bytes(L).
template <int NC, int KC = kCoarseSweepCells / NC>
__device__ __forceinline__ void coarse_vcycle_body(const CoarseLevels& L,
                                                   const float* __restrict__ b0,
                                                   const float* __restrict__ x0in,
                                                   float* __restrict__ x0out,
                                                   const double* __restrict__ M,
                                                   const int* __restrict__ cells, int ncoarse,
                                                   int direct, unsigned char* smem_generic) {
  const int tid = threadIdx.x;
  const lds_bp smem = (lds_bp)smem_generic;
  // The direct solve's operands for this thread -- its coarse unknown and
  // the first sixteen entries of its row segment of M (row tid / 8, columns
  // tid % 8 + 8m) -- loaded now: a barrier waits only for LDS, so the loads
  // arrive while the levels are smoothed.
  const int my_cell = (direct && tid < ncoarse) ? cells[tid] : 0;
  double mpre[16];
  {
    const bool has = direct && tid < 8 * ncoarse;
    const double* row = M + (size_t)(has ? tid >> 3 : 0) * ncoarse;
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int q = (tid & 7) + 8 * m;
      mpre[m] = (has && q < ncoarse) ? row[q] : 0.0;
    }
  }
  {
    // Staging: the info image in 16-byte chunks and level 0's right-hand
    // side into its padded b, each pass issuing all its loads before its
    // first LDS write (one pass for the usual hierarchy).  Every level's x is
    // zeroed (padding ring included); every other b is written by the
    // restriction before it is read.
    constexpr int KI = 4;  // 16-byte chunks per thread and pass
    typedef __attribute__((address_space(3))) unsigned* lds_up;
    const lds_up di = (lds_up)(smem + L.io[0]);
    const int nchunks = L.ibytes >> 4;
    const int nx0 = L.nx[0], n0 = nx0 * L.ny[0];
    const FastDiv d0(L.mnx[0], nx0);
    for (int pass = 0; pass * KI * NC < nchunks || pass * KC * NC < n0; pass++) {
      uint4 iv[KI];
      float bv0[KC];
      int bo0[KC];
#pragma unroll
      for (int k = 0; k < KI; k++) {
        const int c = tid + (pass * KI + k) * NC;
        iv[k] = c < nchunks ? reinterpret_cast<const uint4*>(gm(L.image))[c] : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int k = 0; k < KC; k++) {
        const int q = tid + (pass * KC + k) * NC;
        const int j = d0.div(q), i = q - j * nx0;
        bv0[k] = q < n0 ? b0[q] : 0.f;
        bo0[k] = q < n0 ? L.bo[0] + (i + 1) + (j + 1) * (nx0 + 2) : -1;
      }
      if (pass == 0)
        for (int q = tid; q < L.cells; q += NC) ((lds_fp)smem)[q] = 0.f;
#pragma unroll
      for (int k = 0; k < KI; k++) {
        const int c = tid + (pass * KI + k) * NC;
        if (c < nchunks) {
          di[4 * c] = iv[k].x;
          di[4 * c + 1] = iv[k].y;
          di[4 * c + 2] = iv[k].z;
          di[4 * c + 3] = iv[k].w;
        }
      }
#pragma unroll
      for (int k = 0; k < KC; k++)
        if (bo0[k] >= 0) ((lds_fp)smem)[bo0[k]] = bv0[k];
    }
    if (x0in) {  // warm start of level 0, over the zeros
      __syncthreads();
      for (int q = tid; q < n0; q += NC) {
        const int j = d0.div(q), i = q - j * nx0;
        ((lds_fp)smem)[L.xo[0] + (i + 1) + (j + 1) * (nx0 + 2)] = x0in[q];
      }
    }
  }
  __syncthreads();
  for (int l = 0; l < L.n - 1; l++) {
    const CLevel F = clevel(smem, L, l), G = clevel(smem, L, l + 1);
    bool zero = (l > 0) || !x0in;
    {
      CoarseSweep<NC, KC> sw;
      sw.setup(F.b, F.i, F.px, F.py, FastDiv(L.mpr[l], (F.nx + 1) >> 1), tid);
      sw.two_sweeps(F.x, zero);
    }
    const FastDiv dc(L.mnx[l + 1], G.nx);
    for (int q = tid; q < G.nx * G.ny; q += NC) {
      int J = dc.div(q), I = q - J * G.nx;
      float s = 0.f;
      for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
          int gi = 2 * I + di, gj = 2 * J + dj;
          if (gi < F.nx && gj < F.ny)
            s += lds_residual(F.x, F.b, F.i, F.px, (gi + 1) + (gj + 1) * F.px);
        }
      G.b[(I + 1) + (J + 1) * G.px] = s;
    }
    __syncthreads();
  }
  {
    const int l = L.n - 1;
    const CLevel C = clevel(smem, L, l);
    bool zero = (l > 0) || !x0in;
    if (direct) {
      // b at the coarse unknowns gathered into a dense LDS vector, with
      // their padded LDS indices beside it (both after the info region),
      // then x = M b in FP64, eight lanes per row; lane `part` takes columns
      // part + 8m, sixteen loads of M in flight at once (the first sixteen
      // of this thread's first row segment already in mpre).
      const lds_fp bv = (lds_fp)(smem + L.io[0] + L.ibytes);
      typedef __attribute__((address_space(3))) int* lds_ip;
      const lds_ip bi = (lds_ip)(smem + L.io[0] + L.ibytes + 4 * kMaxDirect);
      const FastDiv dn(L.mnx[l], C.nx);
      for (int q = tid; q < ncoarse; q += NC) {
        const int c = q == tid ? my_cell : cells[q], cj = dn.div(c), ci = c - cj * C.nx;
        const int k = (ci + 1) + (cj + 1) * C.px;
        bi[q] = k;
        bv[q] = C.b[k];
      }
      __syncthreads();
      for (int t = tid; t < 8 * ncoarse; t += NC) {
        int a = t >> 3, part = t & 7;
        const double* row = M + (size_t)a * ncoarse;
        double s = 0.0;
        for (int q0 = part; q0 < ncoarse; q0 += 128) {
          double mv[16];
          if (t == tid && q0 == part) {
#pragma unroll
            for (int m = 0; m < 16; m++) mv[m] = mpre[m];
          } else {
#pragma unroll
            for (int m = 0; m < 16; m++) {
              int q = q0 + 8 * m;
              mv[m] = q < ncoarse ? row[q] : 0.0;
            }
          }
#pragma unroll
          for (int m = 0; m < 16; m++)
            if (q0 + 8 * m < ncoarse) s += mv[m] * (double)bv[q0 + 8 * m];
        }
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        s += __shfl_xor(s, 4, 64);
        if (part == 0) C.x[bi[a]] = (float)s;
      }
    } else {
      CoarseSweep<NC, KC> sw;
      sw.setup(C.b, C.i, C.px, C.py, FastDiv(L.mpr[l], (C.nx + 1) >> 1), tid);
      sw.two_sweeps(C.x, zero);
      for (int s = 1; s < 4; s++) sw.two_sweeps(C.x, false);
    }
    __syncthreads();
  }
  for (int l = L.n - 2; l >= 0; l--) {
    const CLevel F = clevel(smem, L, l), G = clevel(smem, L, l + 1);
    const FastDiv dn(L.mnx[l], F.nx);
    for (int q = tid; q < F.nx * F.ny; q += NC) {
      int gj = dn.div(q), gi = q - gj * F.nx;
      int lq = (gi + 1) + (gj + 1) * F.px;
      if (!(F.i[lq] & 0x80)) continue;
      int I = gi >> 1, J = gj >> 1;
      int I2 = (gi & 1) ? min(I + 1, G.nx - 1) : max(I - 1, 0);
      int J2 = (gj & 1) ? min(J + 1, G.ny - 1) : max(J - 1, 0);
      const lds_fp E = G.x;
      F.x[lq] += 0.5625f * E[(I + 1) + (J + 1) * G.px] +
                 0.1875f * (E[(I2 + 1) + (J + 1) * G.px] + E[(I + 1) + (J2 + 1) * G.px]) +
                 0.0625f * E[(I2 + 1) + (J2 + 1) * G.px];
    }
    __syncthreads();
    {
      CoarseSweep<NC, KC> sw;
      sw.setup(F.b, F.i, F.px, F.py, FastDiv(L.mpr[l], (F.nx + 1) >> 1), tid);
      sw.two_sweeps(F.x, false);
    }
  }
  {
    const CLevel F = clevel(smem, L, 0);
    const FastDiv dn(L.mnx[0], F.nx);
    for (int q = tid; q < F.nx * F.ny; q += NC) {
      int gj = dn.div(q), gi = q - gj * F.nx;
      x0out[q] = F.x[(gi + 1) + (gj + 1) * F.px];
    }
  }
}

}  // namespace fh
