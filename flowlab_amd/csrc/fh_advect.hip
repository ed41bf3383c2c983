// Advection kernels: semi-Lagrangian and MacCormack for the MAC faces and
// for the three dye channels (src/core/advect.ts).
//
// One thread per output face/cell.  Every output is a gather of bilinear
// stencils around a CFL-bounded backtrace, so the reads are near-coalesced
// (a wavefront's 64 faces backtrace into a ~64+2c wide strip) and the kernels
// are bound by HBM traffic: MacCormack velocity moves ~68 B per cell
// (forward 2x13, correction 2x21), dye ~78 B per cell for 3 channels.
#include <algorithm>
#include "fh_batchdesc.hpp"
#include "fh_stencil.hpp"

namespace fh {

constexpr int BX = 64, BY = 4;

static inline dim3 grid2(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }

// One thread per output face/cell; the bodies are in fh_stencil.hpp.
__global__ __launch_bounds__(256) void k_advect_faces(
    const float* __restrict__ uc, const float* __restrict__ vc, const float* __restrict__ us,
    const float* __restrict__ vs, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c) {
  int bx, by;
  xcd_block(&bx, &by);
  advect_faces_at(uc, vc, us, vs, uo, vo, lbl, nx, ny, c,
                  bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

__global__ __launch_bounds__(256) void k_maccormack_faces(
    const float* __restrict__ un, const float* __restrict__ vn, const float* __restrict__ uh,
    const float* __restrict__ vh, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c) {
  int bx, by;
  xcd_block(&bx, &by);
  maccormack_faces_at(un, vn, uh, vh, uo, vo, lbl, nx, ny, c,
                      bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

__global__ __launch_bounds__(256) void k_advect_dye(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ q,
    float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx, int ny, int ds, float c,
    float k0, float k1, float k2) {
  int bx, by;
  xcd_block(&bx, &by);
  advect_dye_at(u, v, q, qo, lbl, nx, ny, ds, c, k0, k1, k2,
                bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

__global__ __launch_bounds__(256) void k_maccormack_dye(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ qn,
    const float* __restrict__ qh, float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx,
    int ny, int ds, float c, float k0, float k1, float k2) {
  int bx, by;
  xcd_block(&bx, &by);
  maccormack_dye_at(u, v, qn, qh, qo, lbl, nx, ny, ds, c, k0, k1, k2,
                    bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

// ---- LDS-staged variants ------------------------------------------------------
// A 64x4 workgroup covers a 64x16 tile of thread indices, four rows per
// thread.  The input fields are staged into LDS over the tile plus a halo of
// AH = 2 cells on each side and one more column and row (WW x WH).
//
// Reach: a backtrace moves a point by at most D = |c| * Vw cells, Vw the
// largest |u|, |v| in the staged velocity windows (the midpoint samples lie
// inside the window when the endpoint bound holds, and bilinear samples never
// exceed the values they blend).  For a thread index i in [x0, x0+63] the
// lowest stencil column read is floor(x0 - 0.5 - D) >= x0 - 2 and the highest
// floor(x0 + 63.5 + D) + 1 <= x0 + 66 when D <= 1.5, likewise for rows; the
// index clamping of the sampling only moves indices towards the tile.  So
// when D <= kReach (1.5 less a margin for rounding) every read of the tile is
// in its window and the kernel reads LDS unchecked; otherwise (a long
// backtrace, or a non-finite velocity) the whole workgroup reads global
// memory exactly as the plain kernels do.  Same bits either way.
constexpr int TX = 64, AH = 2;
constexpr int WW = TX + 2 * AH + 1;
constexpr float kReach = 1.25f;

// A tile of TX x TY outputs (TY a multiple of BY: each thread takes TY/BY
// rows) and its staged window of WW x WH elements per field.
template <int TY>
struct Tile {
  static_assert(TY % BY == 0, "rows per thread");
  static constexpr int WH = TY + 2 * AH + 1, WN = WW * WH;
  static constexpr int kStage = (WN + 255) / 256;
};

// Tile heights, measured per kernel at 8192^2 (us per launch for 8 / 16 /
// 32 rows): the MacCormack passes stage 4 and 8 fields and run more
// workgroups per CU on the smaller window (faces 664 / 690 / 815, dye
// 783 / 860 / 1033); the forward passes stage 2 and 5 and gain from the
// smaller halo share (faces 515 / 454 / 437, dye 520 / 496 / 579).
constexpr int kTyAdvectFaces = 32, kTyMacCormackFaces = 8, kTyAdvectDye = 16,
              kTyMacCormackDye = 8;

struct Win {
  const float* __restrict__ p;
  int w, h;  // field extent; row stride is w
};

// max(|a|, |b|) that keeps a NaN.
__device__ __forceinline__ float nan_max(float m, float x) {
  x = fabsf(x);
  return (x > m || x != x) ? x : m;
}

// Stages NF windows (all global loads first, then the LDS stores) and
// returns this thread's nan_max over what it staged of windows 0 and 1 (the
// velocity); elements outside a field are staged as 0.  speed (optional):
// the same maximum with fmaxf's treatment of NaN (ignored), as the max-speed
// reduction takes it.
//
// kClamped: every load unconditional, from the index clamped into the field,
// an element outside replaced by 0 after the load -- no branch per element.
// The batched kernels need it: their field pointers are read from the scene
// descriptor, and with a branch per element the compiler there waited for
// all loads in flight at almost every element (the batched MacCormack dye
// correction ran 1.8x the single-grid kernel's time on the same scene; 1.07x
// with kClamped).  The single-grid kernels keep the branches, which measured
// 4-8 % faster for them at 8192^2.  Same values either way.
//
// Windows that lie inside every field (all tiles but those along the domain
// edges; a workgroup-uniform test) take neither bounds tests nor branches:
// each load is the field's base plus a 32-bit byte offset, shared by the
// fields of one row stride, and the last slot's spare lanes load the
// window's last element (in bounds; not stored, and a window element for the
// maxima).  The edge tiles, and fields of 2^30 elements or more, stage as
// above.  W: the window's row length; NT: the workgroup's threads.
template <int NF, int WN, bool kClamped = false, int W = WW, int NT = 256>
__device__ __forceinline__ float stage_windows(float (*lds)[WN], const Win (&f)[NF], int wx0,
                                               int wy0, int tid, float* speed = nullptr) {
  constexpr int kStage = (WN + NT - 1) / NT, WH = WN / W;
  static_assert(WH * W == WN, "window shape");
  float v[NF][kStage];
  bool inside = wx0 >= 0 && wy0 >= 0;
#pragma unroll
  for (int q = 0; q < NF; q++)
    inside = inside && wx0 + W <= f[q].w && wy0 + WH <= f[q].h &&
             (long long)f[q].w * f[q].h < (1ll << 30);
  if (inside) {
#pragma unroll
    for (int r = 0; r < kStage; r++) {
      const int e = (r + 1) * NT <= WN ? tid + r * NT : min(tid + r * NT, WN - 1);
      const int b = e / W, a = e - b * W;
#pragma unroll
      for (int q = 0; q < NF; q++) {
        const uint32_t off = (uint32_t)((wx0 + a) + (wy0 + b) * f[q].w) * 4u;
        v[q][r] = *(const float*)((const char*)f[q].p + off);
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kStage; r++) {
      const int e = tid + r * NT;
      const int b = e / W, a = e - b * W;
      const int i = wx0 + a, j = wy0 + b;
#pragma unroll
      for (int q = 0; q < NF; q++) {
        const bool in = e < WN && i >= 0 && j >= 0 && i < f[q].w && j < f[q].h;
        if constexpr (kClamped) {
          const int ci = min(max(i, 0), f[q].w - 1), cj = min(max(j, 0), f[q].h - 1);
          const float x = f[q].p[ci + cj * f[q].w];
          v[q][r] = in ? x : 0.f;
        } else {
          v[q][r] = in ? f[q].p[i + j * f[q].w] : 0.f;
        }
      }
    }
  }
  float m = 0.f, sp = 0.f;
#pragma unroll
  for (int r = 0; r < kStage; r++) {
    const int e = tid + r * NT;
    if (e < WN) {
#pragma unroll
      for (int q = 0; q < NF; q++) lds[q][e] = v[q][r];
    }
    m = nan_max(nan_max(m, v[0][r]), v[1][r]);
    if (speed) sp = fmaxf(sp, fmaxf(fabsf(v[0][r]), fabsf(v[1][r])));
  }
  if (speed) *speed = sp;
  return m;
}

// Workgroup-uniform reach test: every thread passes its staged maximum, the
// barrier that publishes the windows also publishes the four waves' maxima.
template <int NW = 4>
__device__ __forceinline__ bool within_reach(float m, float c, int tid) {
  __shared__ float wm[NW];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nan_max(m, __shfl_xor(m, o, 64));
  if ((tid & 63) == 0) wm[tid >> 6] = m;
  __syncthreads();
  float vw = wm[0];
#pragma unroll
  for (int k = 1; k < NW; k++) vw = nan_max(vw, wm[k]);
  return fabsf(c) * vw <= kReach;
}

// within_reach that also stores the workgroup's largest |u|, |v| (speed,
// each thread's from stage_windows) to partials[bid], under the same barrier.
// The windows of all tiles of a launch cover every u and v element, so the
// largest of these partials is the fields' max face speed.
template <int NW = 4>
__device__ __forceinline__ bool within_reach_speed(float m, float sp, float c, int tid,
                                                   float* __restrict__ partials, int bid) {
  __shared__ float wm[NW], ws[NW];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    m = nan_max(m, __shfl_xor(m, o, 64));
    sp = fmaxf(sp, __shfl_xor(sp, o, 64));
  }
  if ((tid & 63) == 0) {
    wm[tid >> 6] = m;
    ws[tid >> 6] = sp;
  }
  __syncthreads();
  if (tid == 0) {
    float s = ws[0];
#pragma unroll
    for (int k = 1; k < NW; k++) s = fmaxf(s, ws[k]);
    partials[bid] = s;
  }
  float vw = wm[0];
#pragma unroll
  for (int k = 1; k < NW; k++) vw = nan_max(vw, wm[k]);
  return fabsf(c) * vw <= kReach;
}

template <class F, int WN>
__device__ __forceinline__ F window_as(const float (*lds)[WN], int q, int wx0, int wy0,
                                       int w = WW) {
  F f;
  f.l = (lds_cfp)lds[q];
  f.x0 = wx0;
  f.y0 = wy0;
  f.w = w;
  return f;
}

// Tiles clear of the domain edges.  With D <= kReach every position a face
// or cell (i, j) of the tile samples -- the backtraces both ways, their
// midpoints, the clamp stencil, with the half-cell offsets of the staggered
// fields -- has its stencil origin in [i-2, i+1] x [j-2, j+1], so when the
// tile's indices lie kEdge cells inside on every side no origin reaches the
// clamps (below 0, above count-2) of any field: sampling without the clamp
// (LFieldI) takes the same indices and weights.
constexpr int kEdge = 3;

template <int TY>
__device__ __forceinline__ bool interior_tile(int x0, int y0, int nx, int ny) {
  return x0 >= kEdge && y0 >= kEdge && x0 + TX - 1 <= nx - 1 - kEdge &&
         y0 + TY - 1 <= ny - 1 - kEdge;
}

// advect.ts:72-112 with carrier == source (the only use: u and v advect
// themselves), so two windows.
template <int TY, bool kBatch = false>
__device__ __forceinline__ void advect_faces_tile(
    const float* __restrict__ uc, const float* __restrict__ vc, float* __restrict__ uo,
    float* __restrict__ vo, const uint8_t* __restrict__ lbl, int nx, int ny, float c,
    uint8_t* __restrict__ sten, int bx, int by) {
  constexpr int WN = Tile<TY>::WN;
  __shared__ float lds[2][WN];
  const int x0 = bx * TX, y0 = by * TY, wx0 = x0 - AH, wy0 = y0 - AH;
  const int tid = threadIdx.x + threadIdx.y * TX;
  const Win f[2] = {{uc, nx + 1, ny}, {vc, nx, ny + 1}};
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  // every row's face solid tests, loaded with the windows
  int cl[TY / BY];
#pragma unroll
  for (int r = 0; r < TY / BY; r++) cl[r] = closed_faces(lbl, nx, ny, i, j + r * BY);
  const float m = stage_windows<2, Tile<TY>::WN, kBatch>(lds, f, wx0, wy0, tid);
  if (within_reach(m, c, tid)) {
    auto run = [&](auto tag) {
      using F = decltype(tag);
      const F u = window_as<F>(lds, 0, wx0, wy0), v = window_as<F>(lds, 1, wx0, wy0);
      for (int r = 0; r < TY / BY; r++)
        advect_faces_t(u, v, u, v, uo, vo, lbl, nx, ny, c, i, j + r * BY, sten, cl[r]);
    };
    if (interior_tile<TY>(x0, y0, nx, ny))
      run(LFieldI{});
    else
      run(LField{});
  } else {
    const GField u{uc, nx + 1}, v{vc, nx};
    for (int r = 0; r < TY / BY; r++)
      advect_faces_t(u, v, u, v, uo, vo, lbl, nx, ny, c, i, j + r * BY, sten);
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_advect_faces_lds(
    const float* __restrict__ uc, const float* __restrict__ vc, float* __restrict__ uo,
    float* __restrict__ vo, const uint8_t* __restrict__ lbl, int nx, int ny, float c,
    uint8_t* __restrict__ sten) {
  int bx, by;
  xcd_block(&bx, &by);
  advect_faces_tile<TY>(uc, vc, uo, vo, lbl, nx, ny, c, sten, bx, by);
}

// advect.ts:177-210 (fused reverse pass and clamp, as k_maccormack_faces).
template <int TY, bool kBatch = false>
__device__ __forceinline__ void maccormack_faces_tile(
    const float* __restrict__ un, const float* __restrict__ vn, const float* __restrict__ uh,
    const float* __restrict__ vh, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c, const uint8_t* __restrict__ sten,
    const float* __restrict__ heat, float gain, int bx, int by) {
  constexpr int WN = Tile<TY>::WN;
  __shared__ float lds[4][WN];
  const int x0 = bx * TX, y0 = by * TY, wx0 = x0 - AH, wy0 = y0 - AH;
  const int tid = threadIdx.x + threadIdx.y * TX;
  const Win f[4] = {{un, nx + 1, ny}, {vn, nx, ny + 1}, {uh, nx + 1, ny}, {vh, nx, ny + 1}};
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  // the recorded limiter stencils and every row's face solid tests, loaded
  // with the windows
  uint32_t st[TY / BY] = {};
  int cl[TY / BY];
#pragma unroll
  for (int r = 0; r < TY / BY; r++) {
    if (sten && i <= nx && j + r * BY <= ny) st[r] = sten[i + (j + r * BY) * (nx + 1)];
    cl[r] = closed_faces(lbl, nx, ny, i, j + r * BY);
  }
  const float m = stage_windows<4, Tile<TY>::WN, kBatch>(lds, f, wx0, wy0, tid);
  if (within_reach(m, c, tid)) {
    auto run = [&](auto tag) {
      using F = decltype(tag);
      const F a = window_as<F>(lds, 0, wx0, wy0), b = window_as<F>(lds, 1, wx0, wy0),
              ah = window_as<F>(lds, 2, wx0, wy0), bh = window_as<F>(lds, 3, wx0, wy0);
      for (int r = 0; r < TY / BY; r++)
        maccormack_faces_t(a, b, ah, bh, uo, vo, lbl, nx, ny, c, i, j + r * BY, sten != nullptr,
                           st[r], heat, gain, cl[r]);
    };
    if (interior_tile<TY>(x0, y0, nx, ny))
      run(LFieldI{});
    else
      run(LField{});
  } else {
    const GField a{un, nx + 1}, b{vn, nx}, ah{uh, nx + 1}, bh{vh, nx};
    for (int r = 0; r < TY / BY; r++)
      maccormack_faces_t(a, b, ah, bh, uo, vo, lbl, nx, ny, c, i, j + r * BY, false, 0, heat,
                         gain);
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_maccormack_faces_lds(
    const float* __restrict__ un, const float* __restrict__ vn, const float* __restrict__ uh,
    const float* __restrict__ vh, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c, const uint8_t* __restrict__ sten,
    const float* __restrict__ heat, float gain) {
  int bx, by;
  xcd_block(&bx, &by);
  maccormack_faces_tile<TY>(un, vn, uh, vh, uo, vo, lbl, nx, ny, c, sten, heat, gain, bx, by);
}

// advect.ts:241-270, dye on the velocity grid (scale 1): windows of u, v and
// the 3 channels.
template <int TY, bool kBatch = false>
__device__ __forceinline__ void advect_dye_tile(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ q,
    float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx, int ny, float c, float k0,
    float k1, float k2, uint8_t* __restrict__ sten, int bx, int by,
    float* __restrict__ speed = nullptr, int bid = 0) {
  constexpr int WN = Tile<TY>::WN;
  __shared__ float lds[5][WN];
  const int x0 = bx * TX, y0 = by * TY, wx0 = x0 - AH, wy0 = y0 - AH;
  const int tid = threadIdx.x + threadIdx.y * TX;
  const size_t plane = (size_t)nx * ny;
  const Win f[5] = {{u, nx + 1, ny}, {v, nx, ny + 1}, {q, nx, ny}, {q + plane, nx, ny},
                    {q + 2 * plane, nx, ny}};
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  // batched: each row's solid test loaded with the windows, so it is issued
  // before the first store even though the compiler cannot tell qo from lbl
  // (a batch's scene pointers come from memory, without the kernel
  // arguments' restrict); -1: the cell reads its label itself
  int sol[TY / BY];
#pragma unroll
  for (int r = 0; r < TY / BY; r++)
    sol[r] = !kBatch ? -1 : (i < nx && j + r * BY < ny) ? lbl[i + (j + r * BY) * nx] == SOLID : 1;
  float sp = 0.f;
  const float m =
      stage_windows<5, Tile<TY>::WN, kBatch>(lds, f, wx0, wy0, tid, &sp);
  const bool reach =
      speed ? within_reach_speed(m, sp, c, tid, speed, bid) : within_reach(m, c, tid);
  if (reach) {
    auto run = [&](auto tag) {
      using F = decltype(tag);
      const F a = window_as<F>(lds, 0, wx0, wy0), b = window_as<F>(lds, 1, wx0, wy0),
              q0 = window_as<F>(lds, 2, wx0, wy0), q1 = window_as<F>(lds, 3, wx0, wy0),
              q2 = window_as<F>(lds, 4, wx0, wy0);
      for (int r = 0; r < TY / BY; r++)
        advect_dye_t(a, b, q0, q1, q2, qo, lbl, nx, ny, 1, c, k0, k1, k2, i, j + r * BY, sten,
                     sol[r]);
    };
    if (interior_tile<TY>(x0, y0, nx, ny))
      run(LFieldI{});
    else
      run(LField{});
  } else {
    const GField a{u, nx + 1}, b{v, nx}, q0{q, nx}, q1{q + plane, nx}, q2{q + 2 * plane, nx};
    for (int r = 0; r < TY / BY; r++)
      advect_dye_t(a, b, q0, q1, q2, qo, lbl, nx, ny, 1, c, k0, k1, k2, i, j + r * BY, sten);
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_advect_dye_lds(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ q,
    float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx, int ny, float c, float k0,
    float k1, float k2, uint8_t* __restrict__ sten, float* __restrict__ speed) {
  int bx, by;
  xcd_block(&bx, &by);
  advect_dye_tile<TY>(u, v, q, qo, lbl, nx, ny, c, k0, k1, k2, sten, bx, by, speed,
                      blockIdx.x + blockIdx.y * gridDim.x);
}

// advect.ts:291-311, dye on the velocity grid: windows of u, v, the 3
// channels of q^n and of the forward pass.
template <int TY, bool kBatch = false>
__device__ __forceinline__ void maccormack_dye_tile(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ qn,
    const float* __restrict__ qh, float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx,
    int ny, float c, float k0, float k1, float k2, const uint8_t* __restrict__ sten, int bx,
    int by) {
  constexpr int WN = Tile<TY>::WN;
  __shared__ float lds[8][WN];
  const int x0 = bx * TX, y0 = by * TY, wx0 = x0 - AH, wy0 = y0 - AH;
  const int tid = threadIdx.x + threadIdx.y * TX;
  const size_t plane = (size_t)nx * ny;
  const Win f[8] = {{u, nx + 1, ny},      {v, nx, ny + 1},          {qn, nx, ny},
                    {qn + plane, nx, ny}, {qn + 2 * plane, nx, ny}, {qh, nx, ny},
                    {qh + plane, nx, ny}, {qh + 2 * plane, nx, ny}};
  const int i = x0 + threadIdx.x, j = y0 + threadIdx.y;
  // the recorded limiter stencils, loaded with the windows, and (batched)
  // each row's solid test, as in advect_dye_tile
  uint32_t st[TY / BY] = {};
  int sol[TY / BY];
#pragma unroll
  for (int r = 0; r < TY / BY; r++) {
    const bool in = i < nx && j + r * BY < ny;
    if (sten && in) st[r] = sten[i + (j + r * BY) * nx];
    sol[r] = !kBatch ? -1 : in ? lbl[i + (j + r * BY) * nx] == SOLID : 1;
  }
  const float m = stage_windows<8, Tile<TY>::WN, kBatch>(lds, f, wx0, wy0, tid);
  if (within_reach(m, c, tid)) {
    auto run = [&](auto tag) {
      using F = decltype(tag);
      const F a = window_as<F>(lds, 0, wx0, wy0), b = window_as<F>(lds, 1, wx0, wy0),
              n0 = window_as<F>(lds, 2, wx0, wy0), n1 = window_as<F>(lds, 3, wx0, wy0),
              n2 = window_as<F>(lds, 4, wx0, wy0), h0 = window_as<F>(lds, 5, wx0, wy0),
              h1 = window_as<F>(lds, 6, wx0, wy0), h2 = window_as<F>(lds, 7, wx0, wy0);
      for (int r = 0; r < TY / BY; r++)
        maccormack_dye_t(a, b, n0, n1, n2, h0, h1, h2, qo, lbl, nx, ny, 1, c, k0, k1, k2, i,
                         j + r * BY, sten != nullptr, st[r], sol[r]);
    };
    if (interior_tile<TY>(x0, y0, nx, ny))
      run(LFieldI{});
    else
      run(LField{});
  } else {
    const GField a{u, nx + 1}, b{v, nx}, n0{qn, nx}, n1{qn + plane, nx}, n2{qn + 2 * plane, nx},
        h0{qh, nx}, h1{qh + plane, nx}, h2{qh + 2 * plane, nx};
    for (int r = 0; r < TY / BY; r++)
      maccormack_dye_t(a, b, n0, n1, n2, h0, h1, h2, qo, lbl, nx, ny, 1, c, k0, k1, k2, i,
                       j + r * BY);
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_maccormack_dye_lds(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ qn,
    const float* __restrict__ qh, float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx,
    int ny, float c, float k0, float k1, float k2, const uint8_t* __restrict__ sten) {
  int bx, by;
  xcd_block(&bx, &by);
  maccormack_dye_tile<TY>(u, v, qn, qh, qo, lbl, nx, ny, c, k0, k1, k2, sten, bx, by);
}

// ---- MacCormack dye in one pass ----------------------------------------------------
// advect.ts:241-270 then 291-311 (dye scale 1) in one kernel.  The two-pass
// form moves 80 B per cell (34 + 46): the forward pass writes qHat and the
// limiter stencils, the correction reads them back with q^n and the velocity
// again.  Here a workgroup stages u, v and the three channels of q^n over
// its tile and a halo of AF = 4 cells, computes the forward pass itself on
// the tile plus a ring of two cells (every qHat the tile's reverse traces
// sample), keeps those in LDS (the tile's own centre velocities and stencil
// bytes in registers), and corrects its tile: 33 B per cell (u, v 8, q^n 12,
// label 1, out 12), at a third more forward-pass arithmetic for the ring.  Each value is computed by the
// same functions on the same staged values, so the outputs are the two
// passes' bits.
//
// Reach: with D <= kReach, a dye cell i's samples (the backtrace both ways,
// its midpoint, the clamp stencil, for u, v and q alike) read columns
// i - 2 .. i + 2 (u faces i - 1 .. i + 2), and the domain clamps only move
// them towards i; the ring's cells x0 - 2 .. x0 + TX + 1 so read
// x0 - 4 .. x0 + TX + 3, the window's TX + 2 AF columns, and likewise rows.
// The bound is the two-pass kernels' (kReach) on this window.  A workgroup over it writes nothing and marks the launch (fail =
// epoch); the two-pass kernels then run after it over every tile
// (k_dye_forward_rerun, k_dye_correct_rerun: each workgroup returns at once
// unless the mark is this launch's), so every output is the two passes'
// either way.  The max-speed partials come from the staged velocity windows
// of every tile, as in the forward kernel.
constexpr int AF = 4, WWF = TX + 2 * AF, BYF = 8, kTyDyeFused = 16, kTyFacesFused = 16;

// 5 windows and the ring's 3 channels: 50.9 KB, three workgroups of 512
// threads per CU.
template <int TY>
struct FusedTile {
  static_assert(TY % BYF == 0, "rows per thread");
  static constexpr int WH = TY + 2 * AF, WN = WWF * WH;
  // the forward-pass ring: cells (x0 - 2 .. x0 + TX + 1, y0 - 2 .. y0 + TY + 1)
  static constexpr int RW = TX + 4, RH = TY + 4, RN = RW * RH;
};

// The ring's forward-pass samples and the correction's reads all keep
// their stencil origins clear of the domain clamps (interior_tile's
// argument, for the ring's cells).
template <int TY>
__device__ __forceinline__ bool interior_ring(int x0, int y0, int nx, int ny) {
  return x0 - 2 >= kEdge && y0 - 2 >= kEdge && x0 + TX + 1 <= nx - 1 - kEdge &&
         y0 + TY + 1 <= ny - 1 - kEdge;
}

// Cells of the forward-pass ring of a TX x TY tile, by thread: k < R are the
// thread's own tile cells (threadIdx.x, threadIdx.y + k BYF), so that their
// centre velocity and stencil byte carry over to the correction in
// registers; k = R is one of the ring's 4 (RW + TY) cells outside the tile
// (threads below that count): two rows above, two below, two columns either
// side.  (ra, rb): the cell in ring coordinates; false when the thread has
// none.
template <int TY>
__device__ __forceinline__ bool ring_cell(int k, int* ra, int* rb) {
  using T = FusedTile<TY>;
  constexpr int R = TY / BYF;
  const int tid = threadIdx.x + threadIdx.y * TX;
  if (k < R) {
    *ra = (int)threadIdx.x + 2;
    *rb = (int)threadIdx.y + k * BYF + 2;
    return true;
  }
  if (tid < 2 * T::RW) {
    *rb = tid / T::RW;
    *ra = tid - *rb * T::RW;
    return true;
  }
  if (tid < 4 * T::RW) {
    const int t = tid - 2 * T::RW;
    *rb = TY + 2 + t / T::RW;
    *ra = t % T::RW;
    return true;
  }
  if (tid < 4 * T::RW + 4 * TY) {
    const int t = tid - 4 * T::RW, side = t & 3;
    *rb = 2 + (t >> 2);
    *ra = side < 2 ? side : TX + side;
    return true;
  }
  return false;
}

template <int TY>
__global__ __launch_bounds__(TX * BYF) void k_maccormack_dye_fused(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ q,
    float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx, int ny, float c, float k0,
    float k1, float k2, float* __restrict__ speed, int* __restrict__ fail, int epoch) {
  using T = FusedTile<TY>;
  constexpr int NT = TX * BYF, NW = NT / 64, R = TY / BYF;
  static_assert(4 * T::RW + 4 * TY <= NT, "one ring cell outside the tile per thread");
  __shared__ float lds[5][T::WN];
  __shared__ float hat[3][T::RN];
  int bx, by;
  xcd_block(&bx, &by);
  const int x0 = bx * TX, y0 = by * TY, wx0 = x0 - AF, wy0 = y0 - AF, rx0 = x0 - 2,
            ry0 = y0 - 2;
  const int tid = threadIdx.x + threadIdx.y * TX;
  const size_t plane = (size_t)nx * ny;
  const Win f[5] = {{u, nx + 1, ny}, {v, nx, ny + 1}, {q, nx, ny}, {q + plane, nx, ny},
                    {q + 2 * plane, nx, ny}};
  // this thread's ring cells (ring_cell) and their solid tests, loaded with
  // the windows (-1: none, or outside the grid)
  int ra[R + 1], rb[R + 1], sol[R + 1];
#pragma unroll
  for (int k = 0; k <= R; k++) {
    const bool has = ring_cell<TY>(k, &ra[k], &rb[k]);
    const int i = rx0 + ra[k], j = ry0 + rb[k];
    sol[k] = (has && i >= 0 && j >= 0 && i < nx && j < ny) ? lbl[i + j * nx] == SOLID : -1;
  }
  float sp = 0.f;
  const float m = stage_windows<5, T::WN, false, WWF, NT>(lds, f, wx0, wy0, tid, &sp);
  const int bid = (int)(blockIdx.x + blockIdx.y * gridDim.x);
  const bool reach = speed ? within_reach_speed<NW>(m, sp, c, tid, speed, bid)
                           : within_reach<NW>(m, c, tid);
  if (!reach) {
    if (tid == 0) atomicExch(fail, epoch);
    return;
  }
  auto run = [&](auto tag) {
    using F = decltype(tag);
    const F a = window_as<F>(lds, 0, wx0, wy0, WWF), b = window_as<F>(lds, 1, wx0, wy0, WWF),
            n0 = window_as<F>(lds, 2, wx0, wy0, WWF), n1 = window_as<F>(lds, 3, wx0, wy0, WWF),
            n2 = window_as<F>(lds, 4, wx0, wy0, WWF);
    // the forward pass (keep 1, as the two-pass forward kernel runs it) on
    // the ring's cells inside the grid; the others are never read
    uint32_t st[R] = {};
    float uv0[R][2] = {};
#pragma unroll
    for (int k = 0; k <= R; k++) {
      if (sol[k] < 0) continue;
      float o[3];
      const uint32_t s = advect_dye_cell(a, b, n0, n1, n2, lbl, nx, ny, 1, c, 1.f, 1.f, 1.f,
                                         rx0 + ra[k], ry0 + rb[k], sol[k], o,
                                         k < R ? uv0[k < R ? k : 0] : nullptr);
      if (k < R) st[k < R ? k : 0] = s;
      const int e = rb[k] * T::RW + ra[k];
      hat[0][e] = o[0];
      hat[1][e] = o[1];
      hat[2][e] = o[2];
    }
    __syncthreads();
    const F h0 = window_as<F>(hat, 0, rx0, ry0, T::RW), h1 = window_as<F>(hat, 1, rx0, ry0, T::RW),
            h2 = window_as<F>(hat, 2, rx0, ry0, T::RW);
    const int i = x0 + (int)threadIdx.x;
#pragma unroll
    for (int r = 0; r < R; r++)
      maccormack_dye_t(a, b, n0, n1, n2, h0, h1, h2, qo, lbl, nx, ny, 1, c, k0, k1, k2, i,
                       y0 + (int)threadIdx.y + r * BYF, true, st[r], sol[r] < 0 ? 1 : sol[r],
                       uv0[r]);
  };
  if (interior_ring<TY>(x0, y0, nx, ny))
    run(LFieldI{});
  else
    run(LField{});
}

// The two-pass kernels over every tile, for a launch that k_maccormack_dye_fused
// marked (fail == epoch); otherwise each workgroup returns at once.  A small
// grid strides over the tiles (n of them, gx per row).
template <int TY>
__global__ __launch_bounds__(256) void k_dye_forward_rerun(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ q,
    float* __restrict__ qh, const uint8_t* __restrict__ lbl, int nx, int ny, float c,
    uint8_t* __restrict__ sten, const int* __restrict__ fail, int epoch, int gx, int n) {
  if (*fail != epoch) return;
  for (int t = (int)blockIdx.x; t < n; t += (int)gridDim.x) {
    __syncthreads();  // the previous tile's reads of the windows are done
    advect_dye_tile<TY>(u, v, q, qh, lbl, nx, ny, c, 1.f, 1.f, 1.f, sten, t % gx, t / gx);
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_dye_correct_rerun(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ qn,
    const float* __restrict__ qh, float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx,
    int ny, float c, float k0, float k1, float k2, const uint8_t* __restrict__ sten,
    const int* __restrict__ fail, int epoch, int gx, int n) {
  if (*fail != epoch) return;
  for (int t = (int)blockIdx.x; t < n; t += (int)gridDim.x) {
    __syncthreads();
    maccormack_dye_tile<TY>(u, v, qn, qh, qo, lbl, nx, ny, c, k0, k1, k2, sten, t % gx, t / gx);
  }
}

// ---- MacCormack velocity in one pass -----------------------------------------------
// advect.ts:72-112 then 177-210 in one kernel, as k_maccormack_dye_fused does
// for the dye: a workgroup stages u^n and v^n over its thread indices and a
// halo of AF, runs the forward pass on them plus a ring of two (the tile's
// reverse traces sample uHat, vHat there), keeps those in LDS (its own
// indices' cross samples and stencil bytes in registers), and corrects its
// tile, applying buoyancy alone on the open v faces as
// k_maccormack_faces_lds does: 21 B per thread index (u, v 8 + out 8 +
// label 1 + heat 4) instead of the two passes' 48.  A face (i, j)'s samples
// read columns and rows i - 2 .. i + 2, j - 2 .. j + 2 when D <= kReach, so
// the window covers the ring's.  Same bits as the two passes; over the
// reach, the launch is marked and the two-pass kernels redo every tile
// (k_faces_forward_rerun, k_faces_correct_rerun).
template <int TY>
__global__ __launch_bounds__(TX * BYF) void k_maccormack_faces_fused(
    const float* __restrict__ un, const float* __restrict__ vn, float* __restrict__ uo,
    float* __restrict__ vo, const uint8_t* __restrict__ lbl, int nx, int ny, float c,
    const float* __restrict__ heat, float gain, int* __restrict__ fail, int epoch) {
  using T = FusedTile<TY>;
  constexpr int NT = TX * BYF, NW = NT / 64, R = TY / BYF;
  static_assert(4 * T::RW + 4 * TY <= NT, "one ring index outside the tile per thread");
  __shared__ float lds[2][T::WN];
  __shared__ float hat[2][T::RN];
  int bx, by;
  xcd_block(&bx, &by);
  const int x0 = bx * TX, y0 = by * TY, wx0 = x0 - AF, wy0 = y0 - AF, rx0 = x0 - 2,
            ry0 = y0 - 2;
  const int tid = threadIdx.x + threadIdx.y * TX;
  const Win f[2] = {{un, nx + 1, ny}, {vn, nx, ny + 1}};
  // this thread's ring indices (ring_cell) and their face solid tests,
  // loaded with the windows (-1: none, or outside (nx+1) x (ny+1))
  int ra[R + 1], rb[R + 1], cl[R + 1];
#pragma unroll
  for (int k = 0; k <= R; k++) {
    const bool has = ring_cell<TY>(k, &ra[k], &rb[k]);
    const int i = rx0 + ra[k], j = ry0 + rb[k];
    cl[k] = (has && i >= 0 && j >= 0 && i <= nx && j <= ny) ? closed_faces(lbl, nx, ny, i, j)
                                                            : -1;
  }
  const float m = stage_windows<2, T::WN, false, WWF, NT>(lds, f, wx0, wy0, tid);
  if (!within_reach<NW>(m, c, tid)) {
    if (tid == 0) atomicExch(fail, epoch);
    return;
  }
  auto run = [&](auto tag) {
    using F = decltype(tag);
    const F a = window_as<F>(lds, 0, wx0, wy0, WWF), b = window_as<F>(lds, 1, wx0, wy0, WWF);
    // the forward pass on the ring's indices; a face that does not exist
    // (u at j = ny, v at i = nx) stays 0 and is never sampled
    uint32_t st[R] = {};
    float cross[R][2] = {};
#pragma unroll
    for (int k = 0; k <= R; k++) {
      if (cl[k] < 0) continue;
      float uu = 0.f, vv = 0.f;
      const uint32_t s = advect_faces_cell(a, b, a, b, lbl, nx, ny, c, rx0 + ra[k], ry0 + rb[k],
                                           cl[k], &uu, &vv, k < R ? cross[k < R ? k : 0] : nullptr);
      if (k < R) st[k < R ? k : 0] = s;
      const int e = rb[k] * T::RW + ra[k];
      hat[0][e] = uu;
      hat[1][e] = vv;
    }
    __syncthreads();
    const F ha = window_as<F>(hat, 0, rx0, ry0, T::RW), hb = window_as<F>(hat, 1, rx0, ry0, T::RW);
    const int i = x0 + (int)threadIdx.x;
#pragma unroll
    for (int r = 0; r < R; r++)
      maccormack_faces_t(a, b, ha, hb, uo, vo, lbl, nx, ny, c, i, y0 + (int)threadIdx.y + r * BYF,
                         true, st[r], heat, gain, cl[r] < 0 ? 3 : cl[r], cross[r]);
  };
  if (interior_ring<TY>(x0, y0, nx, ny))
    run(LFieldI{});
  else
    run(LField{});
}

// The two-pass velocity kernels over every tile, for a launch that
// k_maccormack_faces_fused marked; otherwise each workgroup returns at once.
template <int TY>
__global__ __launch_bounds__(256) void k_faces_forward_rerun(
    const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ uh,
    float* __restrict__ vh, const uint8_t* __restrict__ lbl, int nx, int ny, float c,
    uint8_t* __restrict__ sten, const int* __restrict__ fail, int epoch, int gx, int n) {
  if (*fail != epoch) return;
  for (int t = (int)blockIdx.x; t < n; t += (int)gridDim.x) {
    __syncthreads();
    advect_faces_tile<TY>(u, v, uh, vh, lbl, nx, ny, c, sten, t % gx, t / gx);
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_faces_correct_rerun(
    const float* __restrict__ un, const float* __restrict__ vn, const float* __restrict__ uh,
    const float* __restrict__ vh, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c, const uint8_t* __restrict__ sten,
    const float* __restrict__ heat, float gain, const int* __restrict__ fail, int epoch, int gx,
    int n) {
  if (*fail != epoch) return;
  for (int t = (int)blockIdx.x; t < n; t += (int)gridDim.x) {
    __syncthreads();
    maccormack_faces_tile<TY>(un, vn, uh, vh, uo, vo, lbl, nx, ny, c, sten, heat, gain, t % gx,
                              t / gx);
  }
}

template <int TY>
static inline dim3 grid_t(int w, int h) {
  return dim3((w + TX - 1) / TX, (h + TY - 1) / TY);
}

void launch_advect_faces(hipStream_t s, const float* uc, const float* vc, const float* us,
                         const float* vs, float* uo, float* vo, const uint8_t* lbl, int nx, int ny,
                         float c, bool lds, uint8_t* sten) {
  if (lds && us == uc && vs == vc)
    k_advect_faces_lds<kTyAdvectFaces><<<grid_t<kTyAdvectFaces>(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(uc, vc, uo, vo, lbl, nx,
                                                                       ny, c, sten);
  else
    k_advect_faces<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(uc, vc, us, vs, uo, vo, lbl,
                                                                  nx, ny, c);
}

void launch_maccormack_faces(hipStream_t s, const float* un, const float* vn, const float* uh,
                             const float* vh, float* uo, float* vo, const uint8_t* lbl, int nx,
                             int ny, float c, bool lds, const uint8_t* sten, const float* heat,
                             float gain) {
  if (lds)
    k_maccormack_faces_lds<kTyMacCormackFaces><<<grid_t<kTyMacCormackFaces>(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(
        un, vn, uh, vh, uo, vo, lbl, nx, ny, c, sten, heat, gain);
  else
    k_maccormack_faces<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(un, vn, uh, vh, uo, vo,
                                                                      lbl, nx, ny, c);
}

int advect_dye_speed_blocks(int nx, int ny) {
  const dim3 g = grid_t<kTyAdvectDye>(nx, ny);
  return (int)(g.x * g.y);
}

void launch_advect_dye(hipStream_t s, const float* u, const float* v, const float* q, float* qo,
                       const uint8_t* lbl, int nx, int ny, int ds, float c, float k0, float k1,
                       float k2, bool lds, uint8_t* sten, float* speed) {
  if (lds && ds == 1)
    k_advect_dye_lds<kTyAdvectDye><<<grid_t<kTyAdvectDye>(nx, ny), dim3(BX, BY), 0, s>>>(u, v, q, qo, lbl, nx, ny, c, k0, k1,
                                                             k2, sten, speed);
  else
    k_advect_dye<<<grid2(nx * ds, ny * ds), dim3(BX, BY), 0, s>>>(u, v, q, qo, lbl, nx, ny, ds,
                                                                  c, k0, k1, k2);
}

void launch_maccormack_dye(hipStream_t s, const float* u, const float* v, const float* qn,
                           const float* qh, float* qo, const uint8_t* lbl, int nx, int ny, int ds,
                           float c, float k0, float k1, float k2, bool lds, const uint8_t* sten) {
  if (lds && ds == 1)
    k_maccormack_dye_lds<kTyMacCormackDye><<<grid_t<kTyMacCormackDye>(nx, ny), dim3(BX, BY), 0, s>>>(u, v, qn, qh, qo, lbl, nx, ny,
                                                                 c, k0, k1, k2, sten);
  else
    k_maccormack_dye<<<grid2(nx * ds, ny * ds), dim3(BX, BY), 0, s>>>(u, v, qn, qh, qo, lbl, nx,
                                                                      ny, ds, c, k0, k1, k2);
}

int maccormack_dye_fused_blocks(int nx, int ny) {
  const dim3 g = grid_t<kTyDyeFused>(nx, ny);
  return (int)(g.x * g.y);
}

// The two-pass kernels after a one-pass launch run with a small grid that
// strides over the tiles; unless that launch was marked, each of its
// workgroups returns at once.
constexpr int kRerunGroups = 1024;

void launch_maccormack_dye_fused(hipStream_t s, const float* u, const float* v, const float* qn,
                                 float* qo, const uint8_t* lbl, int nx, int ny, float c, float k0,
                                 float k1, float k2, float* speed, int* fail, int epoch) {
  k_maccormack_dye_fused<kTyDyeFused><<<grid_t<kTyDyeFused>(nx, ny), dim3(BX, BYF), 0, s>>>(
      u, v, qn, qo, lbl, nx, ny, c, k0, k1, k2, speed, fail, epoch);
}

void launch_maccormack_dye_rerun(hipStream_t s, const float* u, const float* v, const float* qn,
                                 float* qh, float* qo, const uint8_t* lbl, int nx, int ny, float c,
                                 float k0, float k1, float k2, uint8_t* sten, const int* fail,
                                 int epoch) {
  const dim3 gf = grid_t<kTyAdvectDye>(nx, ny), gc = grid_t<kTyMacCormackDye>(nx, ny);
  const int nf = (int)(gf.x * gf.y), nc = (int)(gc.x * gc.y);
  k_dye_forward_rerun<kTyAdvectDye><<<std::min(nf, kRerunGroups), dim3(BX, BY), 0, s>>>(
      u, v, qn, qh, lbl, nx, ny, c, sten, fail, epoch, (int)gf.x, nf);
  k_dye_correct_rerun<kTyMacCormackDye><<<std::min(nc, kRerunGroups), dim3(BX, BY), 0, s>>>(
      u, v, qn, qh, qo, lbl, nx, ny, c, k0, k1, k2, sten, fail, epoch, (int)gc.x, nc);
}

void launch_maccormack_faces_fused(hipStream_t s, const float* un, const float* vn, float* uo,
                                   float* vo, const uint8_t* lbl, int nx, int ny, float c,
                                   const float* heat, float gain, int* fail, int epoch) {
  const dim3 g = grid_t<kTyFacesFused>(nx + 1, ny + 1);
  k_maccormack_faces_fused<kTyFacesFused><<<g, dim3(BX, BYF), 0, s>>>(un, vn, uo, vo, lbl, nx, ny, c,
                                                                    heat, gain, fail, epoch);
}

void launch_maccormack_faces_rerun(hipStream_t s, const float* un, const float* vn, float* uh,
                                   float* vh, float* uo, float* vo, const uint8_t* lbl, int nx,
                                   int ny, float c, uint8_t* sten, const float* heat, float gain,
                                   const int* fail, int epoch) {
  const dim3 gf = grid_t<kTyAdvectFaces>(nx + 1, ny + 1),
             gc = grid_t<kTyMacCormackFaces>(nx + 1, ny + 1);
  const int nf = (int)(gf.x * gf.y), nc = (int)(gc.x * gc.y);
  k_faces_forward_rerun<kTyAdvectFaces><<<std::min(nf, kRerunGroups), dim3(BX, BY), 0, s>>>(
      un, vn, uh, vh, lbl, nx, ny, c, sten, fail, epoch, (int)gf.x, nf);
  k_faces_correct_rerun<kTyMacCormackFaces><<<std::min(nc, kRerunGroups), dim3(BX, BY), 0, s>>>(
      un, vn, uh, vh, uo, vo, lbl, nx, ny, c, sten, heat, gain, fail, epoch, (int)gc.x, nc);
}

// ---- batched launches ----------------------------------------------------------
// One launch over the tiles of every scene of the batch: each block finds its
// scene and tile (batch_block) and runs the single-grid kernel's tile body on
// that scene's buffers with that scene's c = dt / h.

template <int TY>
__global__ __launch_bounds__(256) void kb_advect_faces(const SceneDesc* __restrict__ D,
                                                       BatchGrid g, const BStep* __restrict__ st,
                                                       int q, int mac) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  advect_faces_tile<TY, true>(gm(d.u[q]), gm(d.v[q]), mac ? gm(d.uh) : gm(d.u[q ^ 1]),
                        mac ? gm(d.vh) : gm(d.v[q ^ 1]), gm(d.lbl), d.nx, d.ny, st[s].c,
                        mac ? gm(d.sten_f) : nullptr, bx, by);
}

template <int TY>
__global__ __launch_bounds__(256) void kb_maccormack_faces(const SceneDesc* __restrict__ D,
                                                           BatchGrid g,
                                                           const BStep* __restrict__ st, int q) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  // buoyancy alone (no confinement, gain != 0) is applied at the v-face
  // write with the dye not yet advected, as the single-grid step fuses it
  const BStep& t = st[s];
  const bool buoy = t.eps == 0.f && t.gain != 0.f;
  maccormack_faces_tile<TY, true>(gm(d.u[q]), gm(d.v[q]), gm(d.uh), gm(d.vh), gm(d.u[q ^ 1]),
                            gm(d.v[q ^ 1]), gm(d.lbl), d.nx, d.ny, t.c, gm(d.sten_f),
                            buoy ? gm(d.dye[q]) : nullptr, t.gain, bx, by);
}

// The forward MacCormack pass keeps the dye (keep 1 at its write, as
// dye_phase passes it); the semi-Lagrangian pass applies the step's keep.
template <int TY>
__global__ __launch_bounds__(256) void kb_advect_dye(const SceneDesc* __restrict__ D, BatchGrid g,
                                                     const BStep* __restrict__ st, int q,
                                                     int mac) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const BStep& t = st[s];
  advect_dye_tile<TY, true>(gm(d.u[q ^ 1]), gm(d.v[q ^ 1]), gm(d.dye[q]),
                      mac ? gm(d.dyeh) : gm(d.dye[q ^ 1]), gm(d.lbl), d.nx, d.ny, t.c,
                      mac ? 1.f : t.k0, mac ? 1.f : t.k1, mac ? 1.f : t.k2,
                      mac ? gm(d.sten_d) : nullptr, bx, by);
}

template <int TY>
__global__ __launch_bounds__(256) void kb_maccormack_dye(const SceneDesc* __restrict__ D,
                                                         BatchGrid g, const BStep* __restrict__ st,
                                                         int q) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const BStep& t = st[s];
  maccormack_dye_tile<TY, true>(gm(d.u[q ^ 1]), gm(d.v[q ^ 1]), gm(d.dye[q]), gm(d.dyeh),
                          gm(d.dye[q ^ 1]), gm(d.lbl), d.nx, d.ny, t.c, t.k0, t.k1, t.k2,
                          gm(d.sten_d), bx, by);
}

dim3 batch_grid_advect(int which, int nx, int ny) {
  switch (which) {
    case kBAdvectFaces: return grid_t<kTyAdvectFaces>(nx + 1, ny + 1);
    case kBMacCormackFaces: return grid_t<kTyMacCormackFaces>(nx + 1, ny + 1);
    case kBAdvectDye: return grid_t<kTyAdvectDye>(nx, ny);
    default: return grid_t<kTyMacCormackDye>(nx, ny);
  }
}

void launch_batch_advect_faces(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                               const BStep* st, int q, int mac) {
  kb_advect_faces<kTyAdvectFaces><<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q, mac);
}

void launch_batch_maccormack_faces(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                                   const BStep* st, int q) {
  kb_maccormack_faces<kTyMacCormackFaces><<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q);
}

void launch_batch_advect_dye(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                             const BStep* st, int q, int mac) {
  kb_advect_dye<kTyAdvectDye><<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q, mac);
}

void launch_batch_maccormack_dye(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                                 const BStep* st, int q) {
  kb_maccormack_dye<kTyMacCormackDye><<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q);
}

}  // namespace fh
