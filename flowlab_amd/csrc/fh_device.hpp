// Device-side helpers shared by every kernel of libflowhip (gfx950).
//
// MAC layout on the device is the reference's host layout
// (src/core/grid.ts:1-36, 118-141): p/label nx*ny stride nx, u (nx+1)*ny
// stride nx+1, v nx*(ny+1) stride nx, x fastest -- so a wavefront of 64
// consecutive threads along x reads 64 consecutive words.
//
// Positions are carried as (integer cell base, float offset) in CELL units
// (world / h).  A backtrace moves a point by at most a few cells, so the
// offset keeps ~2^-22 of a cell of precision at any grid size, where a
// plain float world coordinate loses it as the grid grows (2^-10 cell at
// 8192 cells).  Sampling semantics are grid.ts:173-250 exactly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

enum : uint8_t { FLUID = 0, AIR = 1, SOLID = 2 };

// A pointer read out of memory (a batch's scene descriptor, a struct of
// levels) is generic to the compiler -- only kernel arguments are known to
// be global memory -- so every access through it would be a flat
// instruction, which waits on the LDS and the vector-memory counters both.
// Every field buffer is global memory (device or mapped host), and gm()
// tells the compiler so: the address re-enters as a global pointer through
// an integer (a plain generic -> global -> generic cast pair is folded away
// again), and the accesses derived from it become global instructions.
// The conversion costs no instruction.
template <class T>
__device__ __forceinline__ T* gm(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

// grid.ts:159-161: off-grid counts as solid inside a stencil.
__device__ __forceinline__ bool solid_or_outside(const uint8_t* __restrict__ l, int nx, int ny,
                                                 int i, int j) {
  return i < 0 || j < 0 || i >= nx || j >= ny || l[i + j * nx] == SOLID;
}

// boundaries.ts:23-25: in-domain AND solid.
__device__ __forceinline__ bool is_solid_cell(const uint8_t* __restrict__ l, int nx, int ny,
                                              int i, int j) {
  return i >= 0 && j >= 0 && i < nx && j < ny && l[i + j * nx] == SOLID;
}

struct Pos {
  int i;    // integer cell base
  float f;  // offset in cells
};

struct Axis {
  int i0;
  float f;
};

// grid.ts:173-177 for pos = base + off: clamp pos to [0, count-1], then
// i0 = min(floor(pos), count-2), f = pos - i0.
__device__ __forceinline__ Axis clamped_axis(int base, float off, int count) {
  float fl = floorf(off);
  int i0 = base + (int)fl;
  float f = off - fl;
  if (i0 < 0) {
    i0 = 0;
    f = 0.f;
  } else if (i0 > count - 2) {
    i0 = count - 2;
    f = 1.f;
  }
  return {i0, f};
}

// clamped_axis where the caller has shown that the clamp cannot act (pos
// stays inside [0, count-1) by more than a cell): the same i0 and f.
template <bool CLAMP>
__device__ __forceinline__ Axis axis_t(int base, float off, int count) {
  if constexpr (CLAMP) {
    return clamped_axis(base, off, count);
  } else {
    const float fl = floorf(off);
    return {base + (int)fl, off - fl};
  }
}

// divergence.ts:17-25 of one cell in FP64 from its four FP32 faces (west and
// east u, south and north v), shared by every kernel that forms it.
__device__ __forceinline__ double div_value(float uw, float ue, float vs, float vn, double inv_h) {
  return inv_h * (((double)ue - (double)uw) + (double)vn - (double)vs);
}

// grid.ts:180-191
__device__ __forceinline__ float bilerp(float v00, float v10, float v01, float v11, float fx,
                                        float fy) {
  float lo = v00 + fx * (v10 - v00);
  float hi = v01 + fx * (v11 - v01);
  return lo + fy * (hi - lo);
}

// Gather of the four stencil values of a field with row stride `s`.
struct Quad {
  float a, b, c, d;
};

__device__ __forceinline__ Quad gather4(const float* __restrict__ q, int s, Axis ax, Axis ay) {
  const float* r0 = q + ax.i0 + ay.i0 * s;
  return {r0[0], r0[1], r0[s], r0[s + 1]};
}

// grid.ts:197-208: u at (i, j+1/2).
__device__ __forceinline__ float sample_u(const float* __restrict__ u, int nx, int ny, Pos x,
                                          Pos y) {
  Axis ax = clamped_axis(x.i, x.f, nx + 1);
  Axis ay = clamped_axis(y.i, y.f - 0.5f, ny);
  Quad g = gather4(u, nx + 1, ax, ay);
  return bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
}

// grid.ts:214-225: v at (i+1/2, j).
__device__ __forceinline__ float sample_v(const float* __restrict__ v, int nx, int ny, Pos x,
                                          Pos y) {
  Axis ax = clamped_axis(x.i, x.f - 0.5f, nx);
  Axis ay = clamped_axis(y.i, y.f, ny + 1);
  Quad g = gather4(v, nx, ax, ay);
  return bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
}

// grid.ts:239-250 on a cell-centred grid of nx*ny; x/y in that grid's cells.
__device__ __forceinline__ float sample_p(const float* __restrict__ q, int nx, int ny, Pos x,
                                          Pos y) {
  Axis ax = clamped_axis(x.i, x.f - 0.5f, nx);
  Axis ay = clamped_axis(y.i, y.f - 0.5f, ny);
  Quad g = gather4(q, nx, ax, ay);
  return bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
}

// grid.ts:268-270
__device__ __forceinline__ float clamp_to_quad(float q, Quad g) {
  float lo = fminf(fminf(g.a, g.b), fminf(g.c, g.d));
  float hi = fmaxf(fmaxf(g.a, g.b), fmaxf(g.c, g.d));
  return fminf(fmaxf(q, lo), hi);
}

__device__ __forceinline__ Pos shift(Pos p, float d) { return {p.i, p.f + d}; }

// advect.ts:41-55 in cell units: c = dt/h.
//   mid  = P - c/2 * (u0, v0);   prev = P - c * (u(mid), v(mid))
__device__ __forceinline__ void backtrace(const float* __restrict__ uc,
                                          const float* __restrict__ vc, int nx, int ny, Pos x,
                                          Pos y, float u0, float v0, float c, Pos* bx, Pos* by) {
  Pos mx = shift(x, -0.5f * c * u0);
  Pos my = shift(y, -0.5f * c * v0);
  *bx = shift(x, -c * sample_u(uc, nx, ny, mx, my));
  *by = shift(y, -c * sample_v(vc, nx, ny, mx, my));
}

// ---- field accessors ------------------------------------------------------
// The sampling above, generic over where a field's values come from: GField
// reads global memory (the same elements the pointer versions read); LField
// reads a window of the field staged in LDS, for callers that have shown
// every index they will read lies inside the window.
struct GField {
  static constexpr bool kClamp = true;  // sampling clamps to the domain (grid.ts:173-177)
  const float* __restrict__ p;
  int s;  // row stride
  __device__ __forceinline__ float at(int i, int j) const { return p[i + j * s]; }
  __device__ __forceinline__ Quad quad(int i, int j) const {
    const float* r = p + i + j * s;
    return {r[0], r[1], r[s], r[s + 1]};
  }
};

typedef __attribute__((address_space(3))) const float* lds_cfp;

// A window's row offset is small (every index read lies inside the window),
// so the 24-bit multiply gives the same product at full rate; a 32-bit
// integer multiply issues at a quarter of the VALU rate.
struct LField {
  static constexpr bool kClamp = true;
  lds_cfp l;    // window origin: field index (x0, y0)
  int x0, y0, w;  // w = window row stride
  __device__ __forceinline__ int off(int i, int j) const { return __mul24(j - y0, w) + (i - x0); }
  __device__ __forceinline__ float at(int i, int j) const { return l[off(i, j)]; }
  __device__ __forceinline__ Quad quad(int i, int j) const {
    lds_cfp r = l + off(i, j);
    return {r[0], r[1], r[w], r[w + 1]};
  }
};

// An LDS window whose every sample position is known to stay clear of the
// domain edges, so sampling skips the clamp (it would not act).
struct LFieldI : LField {
  static constexpr bool kClamp = false;
};

template <class F>
__device__ __forceinline__ Quad gather4f(const F& q, Axis ax, Axis ay) {
  return q.quad(ax.i0, ay.i0);
}

template <class F>
__device__ __forceinline__ float sample_uf(const F& u, int nx, int ny, Pos x, Pos y) {
  Axis ax = axis_t<F::kClamp>(x.i, x.f, nx + 1);
  Axis ay = axis_t<F::kClamp>(y.i, y.f - 0.5f, ny);
  Quad g = gather4f(u, ax, ay);
  return bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
}

template <class F>
__device__ __forceinline__ float sample_vf(const F& v, int nx, int ny, Pos x, Pos y) {
  Axis ax = axis_t<F::kClamp>(x.i, x.f - 0.5f, nx);
  Axis ay = axis_t<F::kClamp>(y.i, y.f, ny + 1);
  Quad g = gather4f(v, ax, ay);
  return bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
}

template <class FU, class FV>
__device__ __forceinline__ void backtracef(const FU& uc, const FV& vc, int nx, int ny, Pos x, Pos y,
                                           float u0, float v0, float c, Pos* bx, Pos* by) {
  Pos mx = shift(x, -0.5f * c * u0);
  Pos my = shift(y, -0.5f * c * v0);
  *bx = shift(x, -c * sample_uf(uc, nx, ny, mx, my));
  *by = shift(y, -c * sample_vf(vc, nx, ny, mx, my));
}

// Workgroups of a launch are dispatched round-robin over the chip's 8 XCDs,
// each with its own L2.  Remapping the dispatch index gives every XCD a
// contiguous run of logical blocks (row-major over the block grid), so the
// halos that neighbouring blocks share are fetched into one L2 instead of
// several.  With n = 8q + r blocks, XCD x runs the q + (x < r) dispatch
// indices x, x+8, ...; the k-th of them becomes logical block
// x q + min(x, r) + k, a bijection onto [0, n).
__device__ __forceinline__ void xcd_map(int b, int n, int gx, int* bx, int* by) {
  const int x = b & 7, q = n >> 3, r = n & 7;
  const int l = x * q + (x < r ? x : r) + (b >> 3);
  *by = l / gx;
  *bx = l - *by * gx;
}

__device__ __forceinline__ void xcd_block(int* bx, int* by) {
  const int gx = (int)gridDim.x;
  xcd_map((int)blockIdx.x + (int)blockIdx.y * gx, gx * (int)gridDim.y, gx, bx, by);
}

// ---- reductions ---------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace fh

namespace fh {

// Block-level reduction without atomics: each block reduces its threads'
// values (wavefront shuffles, then across waves in LDS) and thread 0 writes
// one partial per slot to partials[block*NS + s].  A one-block finalize
// kernel then sums the partials in block order, so the result is
// deterministic.  (A last-block-done atomic counter serialises thousands of
// device-scope atomics across the XCDs and was measured at ~25 ns each.)
// Block size must be a multiple of 64 and at most 1024.
// bid: the slot of this block's partials (its dispatch index in a
// single-grid launch, its index within its scene's grid in a batched one).
template <int NS>
__device__ void block_sum_store_at(double (&val)[NS], double* partials, int bid) {
  __shared__ double sh[NS][16];
  const int nthreads = blockDim.x * blockDim.y;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    double v = wave_sum(val[s]);
    if (lane == 0) sh[s][wid] = v;
  }
  __syncthreads();
  if (tid < NS) {
    double t = 0;
    for (int w = 0; w < nw; w++) t += sh[tid][w];
    partials[(size_t)bid * NS + tid] = t;
  }
}

template <int NS>
__device__ void block_sum_store(double (&val)[NS], double* partials) {
  block_sum_store_at<NS>(val, partials, blockIdx.x + blockIdx.y * gridDim.x);
}

__device__ inline void block_max_store_at(float val, float* partials, int bid) {
  __shared__ float sh[16];
  const int nthreads = blockDim.x * blockDim.y;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
  float v = wave_max(val);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (tid == 0) {
    float t = 0.f;
    for (int w = 0; w < nw; w++) t = fmaxf(t, sh[w]);
    partials[bid] = t;
  }
}

__device__ inline void block_max_store(float val, float* partials) {
  block_max_store_at(val, partials, blockIdx.x + blockIdx.y * gridDim.x);
}

}  // namespace fh
