// Body forces, divergence, gradient subtraction, boundaries and dye sources
// (src/core/{vorticity,buoyancy,stir,divergence,subtractGradient,boundaries,dye}.ts).
//
// All of these are single-pass streaming stencils: one thread per face or
// cell, x fastest, so each wavefront touches contiguous rows.
#include <algorithm>
#include <vector>

#include "fh_batchdesc.hpp"
#include "fh_stencil.hpp"

namespace fh {

constexpr int BX = 64, BY = 4;

static inline dim3 grid2(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }

// One thread per face/cell; the bodies are in fh_stencil.hpp.
__global__ __launch_bounds__(256) void k_vorticity(const float* __restrict__ u,
                                                   const float* __restrict__ v,
                                                   float* __restrict__ w, int nx, int ny,
                                                   float inv_h) {
  int bx, by;
  xcd_block(&bx, &by);
  vorticity_at(u, v, w, nx, ny, inv_h, bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

__global__ __launch_bounds__(256) void k_forces(float* __restrict__ u, float* __restrict__ v,
                                                const uint8_t* __restrict__ lbl,
                                                const float* __restrict__ heat,
                                                const float* __restrict__ w, int nx, int ny,
                                                int ds, float h, float dt, float gain, float eps,
                                                StirBox st) {
  int bx, by;
  xcd_block(&bx, &by);
  forces_at(u, v, lbl, heat, w, nx, ny, ds, h, dt, gain, eps, st,
            bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

// Rows per thread of the streaming kernels below: each thread issues the
// loads of kRows consecutive rows before any arithmetic or store, so every
// CU keeps enough bytes in flight to stream HBM.  Measured at 8192^2 (one
// row -> four): buoyancy 0.43 -> 0.71 of the HBM peak, divergence
// 0.67 -> 0.74.  Per element the arithmetic is the one-row bodies' (the same
// bits).
constexpr int kRows = 4;

// Thread grid covering w x h elements, kRows rows per thread.
static inline dim3 grid_rows(int w, int h) {
  return dim3((w + BX - 1) / BX, ((h + kRows - 1) / kRows + BY - 1) / BY);
}

// Buoyancy alone (buoyancy.ts:70-88: no stir, no confinement, gain != 0):
// u is unchanged, so only the interior v faces j in [1, ny-1] are touched,
// kRows faces of a column per thread; the stored value is forces_at's.
__device__ __forceinline__ void buoyancy_rows(float* __restrict__ v, const uint8_t* __restrict__ lbl,
                                              const float* __restrict__ heat, int nx, int ny,
                                              int ds, float gain, int bx, int by) {
  const int i = bx * BX + threadIdx.x, j0 = 1 + (by * BY + threadIdx.y) * kRows;
  if (i >= nx || j0 > ny - 1) return;
  const int dnx = nx * ds;
  const Axis ax = clamped_axis(i * ds, 0.5f * ds - 0.5f, dnx);
  uint8_t below[kRows], above[kRows];
  float val[kRows], fy[kRows];
  Quad g[kRows];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const int j = min(j0 + r, ny - 1), k = i + j * nx;
    below[r] = lbl[k - nx];
    above[r] = lbl[k];
    val[r] = v[k];
    const Axis ay = clamped_axis(j * ds, -0.5f, ny * ds);
    fy[r] = ay.f;
    g[r] = gather4(heat, dnx, ax, ay);
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const int j = j0 + r;
    if (j > ny - 1 || below[r] == SOLID || above[r] == SOLID) continue;
    v[i + j * nx] = val[r] + gain * bilerp(g[r].a, g[r].b, g[r].c, g[r].d, ax.f, fy[r]);
  }
}

__global__ __launch_bounds__(256) void k_buoyancy(float* __restrict__ v,
                                                  const uint8_t* __restrict__ lbl,
                                                  const float* __restrict__ heat, int nx, int ny,
                                                  int ds, float gain) {
  int bx, by;
  xcd_block(&bx, &by);
  buoyancy_rows(v, lbl, heat, nx, ny, ds, gain, bx, by);
}

// divergence.ts:17-25, evaluated in FP64 from the FP32 faces, kRows cells of
// a column per thread.
__device__ __forceinline__ void divergence_rows(const float* __restrict__ u,
                                                const float* __restrict__ v,
                                                double* __restrict__ div, int nx, int ny,
                                                double inv_h, int bx, int by) {
  const int i = bx * BX + threadIdx.x, j0 = (by * BY + threadIdx.y) * kRows;
  if (i >= nx || j0 >= ny) return;
  const int su = nx + 1;
  float ua[kRows], ub[kRows], vv[kRows + 1];
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const int j = min(j0 + r, ny - 1);
    ua[r] = u[i + j * su];
    ub[r] = u[i + 1 + j * su];
  }
#pragma unroll
  for (int r = 0; r <= kRows; r++) vv[r] = v[i + min(j0 + r, ny) * nx];
#pragma unroll
  for (int r = 0; r < kRows; r++)
    if (j0 + r < ny)
      div[i + (j0 + r) * nx] = div_value(ua[r], ub[r], vv[r], vv[r + 1], inv_h);
}

__global__ __launch_bounds__(256) void k_divergence(const float* __restrict__ u,
                                                    const float* __restrict__ v,
                                                    double* __restrict__ div, int nx, int ny,
                                                    double inv_h) {
  int bx, by;
  xcd_block(&bx, &by);
  divergence_rows(u, v, div, nx, ny, inv_h, bx, by);
}

__global__ __launch_bounds__(256) void k_subtract(float* __restrict__ u, float* __restrict__ v,
                                                  const double* __restrict__ p,
                                                  const uint8_t* __restrict__ lbl, int nx, int ny,
                                                  double gs) {
  int bx, by;
  xcd_block(&bx, &by);
  subtract_at(u, v, p, lbl, nx, ny, gs, bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

// Gradient subtraction with applyOutflow (boundaries.ts:115-143) after it in
// the same pass.  Outflow reads only the face next to an Air edge cell and
// writes that face and the edge face; the thread that subtracts at face 1
// (nx-1) of row j applies the row's low (high) edge right after, likewise
// face 1 (ny-1) of column i of v.  No other thread reads or writes those
// faces, so the values are those of the two passes in turn.  Needs nx, ny
// >= 3, where a row's two edges touch different faces.
// The outflow rule (boundaries.ts:115-143) at the faces next to Air edge
// cells, applied by the thread that subtracted at face 1 (nx-1) of row j and
// face 1 (ny-1) of column i.
__device__ __forceinline__ void outflow_edges_at(float* __restrict__ u, float* __restrict__ v,
                                                 const uint8_t* __restrict__ lbl, int nx, int ny,
                                                 int su, int i, int j) {
  if (j < ny) {
    if (i == 1 && lbl[j * nx] == AIR) {
      const float o = fminf(u[1 + j * su], 0.f);
      u[1 + j * su] = o;
      u[j * su] = o;
    } else if (i == nx - 1 && lbl[nx - 1 + j * nx] == AIR) {
      const float o = fmaxf(u[nx - 1 + j * su], 0.f);
      u[nx - 1 + j * su] = o;
      u[nx + j * su] = o;
    }
  }
  if (i < nx) {
    if (j == 1 && lbl[i] == AIR) {
      const float o = fminf(v[i + nx], 0.f);
      v[i + nx] = o;
      v[i] = o;
    } else if (j == ny - 1 && lbl[i + (ny - 1) * nx] == AIR) {
      const float o = fmaxf(v[i + (ny - 1) * nx], 0.f);
      v[i + (ny - 1) * nx] = o;
      v[i + ny * nx] = o;
    }
  }
}

// kRows face rows per thread: every row's loads are issued before any store;
// then the subtraction of each row, then each row's outflow (which reads back
// only faces this thread has just written).  The pressure and labels of a
// thread's column are loaded once for rows j0-1 .. j0+kRows-1 and those of
// the west column for rows j0 .. j0+kRows-1: row j's south cell is row j-1's
// cell.  These are the values subtract_load reads wherever a face is
// written (at j = ny, where the clamped south index differs, no face is).
__device__ __forceinline__ void subtract_outflow_rows(float* __restrict__ u, float* __restrict__ v,
                                                      const double* __restrict__ p,
                                                      const uint8_t* __restrict__ lbl, int nx,
                                                      int ny, double gs, int bx, int by) {
  const int i = bx * BX + threadIdx.x, j0 = (by * BY + threadIdx.y) * kRows, su = nx + 1;
  if (i > nx || j0 > ny) return;
  const int ic = min(i, nx - 1), iw = i >= 1 ? ic - 1 : ic;
  double pc[kRows + 1], pw[kRows];
  uint8_t lc[kRows + 1], lw[kRows];
  float u0[kRows], v0[kRows];
#pragma unroll
  for (int r = 0; r <= kRows; r++) {
    const int k = ic + min(max(j0 - 1 + r, 0), ny - 1) * nx;
    pc[r] = p[k];
    lc[r] = lbl[k];
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const int j = min(j0 + r, ny), k = iw + min(j, ny - 1) * nx;
    pw[r] = p[k];
    lw[r] = lbl[k];
    // the faces from indices clamped onto the interior faces, loaded
    // unconditionally: a branch per load made each one wait for the last
    // (the value is only used on the interior face itself)
    u0[r] = u[min(max(i, 1), nx - 1) + min(j, ny - 1) * su];
    v0[r] = v[min(i, nx - 1) + min(max(j, 1), ny - 1) * nx];
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const int j = j0 + r;
    if (j > ny) break;
    const bool uf = j < ny && i >= 1 && i <= nx - 1;
    const bool vf = i < nx && j >= 1 && j <= ny - 1;
    const uint8_t l0 = lc[r + 1];
    if (uf && lw[r] != SOLID && l0 != SOLID)
      u[i + j * su] = (float)((double)u0[r] - gs * (pc[r + 1] - pw[r]));
    if (vf && lc[r] != SOLID && l0 != SOLID)
      v[i + j * nx] = (float)((double)v0[r] - gs * (pc[r + 1] - pc[r]));
  }
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    const int j = j0 + r;
    if (j > ny) break;
    outflow_edges_at(u, v, lbl, nx, ny, su, i, j);
  }
}

__global__ __launch_bounds__(256) void k_subtract_outflow(float* __restrict__ u,
                                                          float* __restrict__ v,
                                                          const double* __restrict__ p,
                                                          const uint8_t* __restrict__ lbl, int nx,
                                                          int ny, double gs) {
  int bx, by;
  xcd_block(&bx, &by);
  subtract_outflow_rows(u, v, p, lbl, nx, ny, gs, bx, by);
}


__global__ __launch_bounds__(256) void k_outflow(float* __restrict__ u, float* __restrict__ v,
                                                 const uint8_t* __restrict__ lbl, int nx, int ny) {
  outflow_at(u, v, lbl, nx, ny, blockIdx.x * blockDim.x + threadIdx.x);
}

// dye.ts:70-87: dye += cov (val - dye) inside the patch, skipped at cov 0.
// One workgroup per BX x BY patch tile with any non-zero coverage (the list
// the host built when the patch was set), so a patch that spans the domain
// around a few small sources costs those sources' tiles.
__global__ __launch_bounds__(256) void k_dye_patch(float* __restrict__ dye,
                                                   const float* __restrict__ pd, int dnx, int dny,
                                                   int i0, int j0, int pnx, int pny,
                                                   const int* __restrict__ tiles) {
  const int gx = (pnx + BX - 1) / BX, t = tiles[blockIdx.x];
  const int ty = t / gx, tx = t - ty * gx;
  int i = tx * BX + threadIdx.x;
  int j = ty * BY + threadIdx.y;
  if (i >= pnx || j >= pny) return;
  int di = i0 + i, dj = j0 + j;
  if (di < 0 || dj < 0 || di >= dnx || dj >= dny) return;
  int cells = pnx * pny, s = i + j * pnx;
  float cov = pd[3 * cells + s];
  if (cov == 0.f) return;
  size_t plane = (size_t)dnx * dny;
  int k = di + dj * dnx;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float q = dye[k + c * plane];
    dye[k + c * plane] = q + cov * (pd[c * cells + s] - q);
  }
}

// boundaries.ts:45-64 on the device fields (dye indexed by the label index,
// as the reference does).
__global__ __launch_bounds__(256) void k_commit_labels(double* __restrict__ p,
                                                       float* __restrict__ u,
                                                       float* __restrict__ v,
                                                       float* __restrict__ dye,
                                                       const uint8_t* __restrict__ lbl, int nx,
                                                       int ny, int dye_plane) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i > nx || j > ny) return;
  if (i < nx && j < ny) {
    int k = i + j * nx;
    if (lbl[k] != FLUID) {
      p[k] = 0.0;
      if (lbl[k] == SOLID)
        for (int c = 0; c < 3; c++) dye[k + (size_t)c * dye_plane] = 0.f;
    }
  }
  if (j < ny && (is_solid_cell(lbl, nx, ny, i - 1, j) || is_solid_cell(lbl, nx, ny, i, j)))
    u[i + j * (nx + 1)] = 0.f;
  if (i < nx && (is_solid_cell(lbl, nx, ny, i, j - 1) || is_solid_cell(lbl, nx, ny, i, j)))
    v[i + j * nx] = 0.f;
}

// simulation.ts:451-459: max |face velocity|.
__global__ __launch_bounds__(256) void k_max_speed(const float* __restrict__ u,
                                                   const float* __restrict__ v, long nu, long nv,
                                                   float* partials, unsigned* counter,
                                                   float* out) {
  const long stride = (long)gridDim.x * blockDim.x;
  const long t0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
  float m = 0.f;
  // 16-byte loads over the whole float4s of each array (both start 256-byte
  // aligned), then the tail; the maximum does not depend on the order.
  auto scan = [&](const float* __restrict__ a, long n) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const long n4 = n >> 2;
    for (long k = t0; k < n4; k += stride) {
      const float4 x = a4[k];
      m = fmaxf(m, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))));
    }
    for (long k = (n4 << 2) + t0; k < n; k += stride) m = fmaxf(m, fabsf(a[k]));
  };
  scan(u, nu);
  scan(v, nv);
  block_max_store(m, partials);
}

// host_out (optional) is mapped pinned host memory that receives the same
// value, so the host reads it after an event with no copy on the stream.
__global__ __launch_bounds__(256) void k_finalize_max(const float* __restrict__ partials,
                                                      int nblocks, float* __restrict__ out,
                                                      float* __restrict__ host_out) {
  __shared__ float sh[4];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  float t = 0.f;
  for (int b = tid; b < nblocks; b += 256) t = fmaxf(t, partials[b]);
  t = wave_max(t);
  if (lane == 0) sh[wid] = t;
  __syncthreads();
  if (tid == 0) {
    const float m = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    out[0] = m;
    if (host_out) host_out[0] = m;
  }
}

// The maximum of many partials: 1024 threads, each taking float4s four deep
// (partials is a hipMalloc base, so 16-byte aligned), then the tail.
__global__ __launch_bounds__(1024) void k_finalize_max_wide(const float* __restrict__ partials,
                                                            int n, float* __restrict__ out,
                                                            float* __restrict__ host_out) {
  __shared__ float sh[16];
  const int tid = threadIdx.x;
  const float4* p4 = reinterpret_cast<const float4*>(partials);
  const int n4 = n >> 2;
  float t = 0.f;
  auto m4 = [](float4 x) { return fmaxf(fmaxf(x.x, x.y), fmaxf(x.z, x.w)); };
  int k = tid;
  for (; k + 3 * 1024 < n4; k += 4 * 1024) {
    const float4 a = p4[k], b = p4[k + 1024], c = p4[k + 2048], d = p4[k + 3072];
    t = fmaxf(t, fmaxf(fmaxf(m4(a), m4(b)), fmaxf(m4(c), m4(d))));
  }
  for (; k < n4; k += 1024) t = fmaxf(t, m4(p4[k]));
  for (int b = (n4 << 2) + tid; b < n; b += 1024) t = fmaxf(t, partials[b]);
  t = wave_max(t);
  if ((tid & 63) == 0) sh[tid >> 6] = t;
  __syncthreads();
  if (tid == 0) {
    float m = 0.f;
    for (int w = 0; w < 16; w++) m = fmaxf(m, sh[w]);
    out[0] = m;
    if (host_out) host_out[0] = m;
  }
}

void launch_finalize_max(hipStream_t s, const float* partials, int n, float* out,
                         float* host_out) {
  k_finalize_max_wide<<<1, 1024, 0, s>>>(partials, n, out, host_out);
}

void launch_vorticity(hipStream_t s, const float* u, const float* v, float* w, int nx, int ny,
                      float inv_h) {
  k_vorticity<<<grid2(nx, ny), dim3(BX, BY), 0, s>>>(u, v, w, nx, ny, inv_h);
}

void launch_forces(hipStream_t s, float* u, float* v, const uint8_t* lbl, const float* heat,
                   const float* w, int nx, int ny, int ds, float h, float dt, float gain,
                   float eps, StirBox stir) {
  if (!stir.active && eps == 0.f && gain != 0.f) {
    k_buoyancy<<<grid_rows(nx, ny - 1), dim3(BX, BY), 0, s>>>(v, lbl, heat, nx, ny, ds, gain);
    return;
  }
  k_forces<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(u, v, lbl, heat, w, nx, ny, ds, h, dt,
                                                          gain, eps, stir);
}

void launch_divergence(hipStream_t s, const float* u, const float* v, double* div, int nx, int ny,
                       double inv_h) {
  k_divergence<<<grid_rows(nx, ny), dim3(BX, BY), 0, s>>>(u, v, div, nx, ny, inv_h);
}

void launch_subtract_gradient(hipStream_t s, float* u, float* v, const double* p,
                              const uint8_t* lbl, int nx, int ny, double grad_scale) {
  k_subtract<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(u, v, p, lbl, nx, ny, grad_scale);
}

void launch_subtract_outflow(hipStream_t s, float* u, float* v, const double* p,
                             const uint8_t* lbl, int nx, int ny, double grad_scale) {
  if (nx < 3 || ny < 3) {
    launch_subtract_gradient(s, u, v, p, lbl, nx, ny, grad_scale);
    launch_outflow(s, u, v, lbl, nx, ny);
    return;
  }
  k_subtract_outflow<<<grid_rows(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(u, v, p, lbl, nx, ny,
                                                                     grad_scale);
}

void launch_outflow(hipStream_t s, float* u, float* v, const uint8_t* lbl, int nx, int ny) {
  int n = nx > ny ? nx : ny;
  k_outflow<<<(n + 255) / 256, 256, 0, s>>>(u, v, lbl, nx, ny);
}

void dye_patch_tiles(const float* data, int pnx, int pny, std::vector<int>& tiles) {
  const int gx = (pnx + BX - 1) / BX, gy = (pny + BY - 1) / BY;
  const float* cov = data + (size_t)3 * pnx * pny;
  tiles.clear();
  for (int ty = 0; ty < gy; ty++)
    for (int tx = 0; tx < gx; tx++) {
      bool any = false;
      for (int j = ty * BY; j < std::min(pny, (ty + 1) * BY) && !any; j++)
        for (int i = tx * BX; i < std::min(pnx, (tx + 1) * BX); i++)
          if (cov[i + (size_t)j * pnx] != 0.f) {
            any = true;
            break;
          }
      if (any) tiles.push_back(tx + ty * gx);
    }
}

void launch_dye_patch(hipStream_t s, float* dye, const float* patch, int dnx, int dny, int i0,
                      int j0, int pnx, int pny, const int* tiles, int ntiles) {
  if (ntiles > 0)
    k_dye_patch<<<ntiles, dim3(BX, BY), 0, s>>>(dye, patch, dnx, dny, i0, j0, pnx, pny, tiles);
}

void launch_commit_labels(hipStream_t s, double* p, float* u, float* v, float* dye,
                          const uint8_t* lbl, int nx, int ny, int dye_plane) {
  k_commit_labels<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(p, u, v, dye, lbl, nx, ny,
                                                                 dye_plane);
}

void launch_max_speed(hipStream_t s, const float* u, const float* v, int nx, int ny,
                      float* partials, unsigned* counter, float* out, float* host_out) {
  k_max_speed<<<1024, 256, 0, s>>>(u, v, (long)(nx + 1) * ny, (long)nx * (ny + 1), partials,
                                   counter, out);
  k_finalize_max<<<1, 256, 0, s>>>(partials, 1024, out, host_out);
}

// ---- batched launches ----------------------------------------------------------
// One launch over every scene of the work list; each block runs the
// single-grid kernel's body on its scene (batch_block).  The forces act on
// the advected velocity [q ^ 1] with the heat of the dye not yet advected
// [q], as the single-grid step orders them.

__global__ __launch_bounds__(256) void kb_vorticity(const SceneDesc* __restrict__ D, BatchGrid g,
                                                    const BStep* __restrict__ st, int q) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  vorticity_at(gm(d.u[q ^ 1]), gm(d.v[q ^ 1]), gm(d.w), d.nx, d.ny, st[s].inv_hf,
               bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

__global__ __launch_bounds__(256) void kb_forces(const SceneDesc* __restrict__ D, BatchGrid g,
                                                 const BStep* __restrict__ st, int q) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const BStep& t = st[s];
  const StirBox none{};
  forces_at(gm(d.u[q ^ 1]), gm(d.v[q ^ 1]), gm(d.lbl), gm(d.dye[q]), gm(d.w), d.nx, d.ny, 1, t.hf,
            t.dtf, t.gain, t.eps, none, bx * BX + threadIdx.x, by * BY + threadIdx.y);
}

// The single-grid step skips buoyancy when buoyancy * dt rounds to 0.
__global__ __launch_bounds__(256) void kb_buoyancy(const SceneDesc* __restrict__ D, BatchGrid g,
                                                   const BStep* __restrict__ st, int q) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const float gain = st[s].gain;
  if (gain == 0.f) return;
  buoyancy_rows(gm(d.v[q ^ 1]), gm(d.lbl), gm(d.dye[q]), d.nx, d.ny, 1, gain, bx, by);
}

// Gradient subtraction with the pressure the solve left in the step's pr.
__global__ __launch_bounds__(256) void kb_project(const SceneDesc* __restrict__ D, BatchGrid g,
                                                  const BStep* __restrict__ st, int q) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  subtract_outflow_rows(gm(d.u[q ^ 1]), gm(d.v[q ^ 1]), gm(st[s].pr), gm(d.lbl), d.nx, d.ny,
                        st[s].grad_scale, bx, by);
}

// kBatchSpeedBlocks blocks per scene stride over its u and v faces (the
// maximum does not depend on the order); one block per scene then reduces
// the scene's partials into host_out[s].
__global__ __launch_bounds__(256) void kb_max_speed(const SceneDesc* __restrict__ D, BatchGrid g,
                                                    int q, float* __restrict__ partials) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const long stride = (long)kBatchSpeedBlocks * 256;
  const long t0 = (long)bid * 256 + threadIdx.x;
  float m = 0.f;
  // float4 loads over the whole float4s of each array (hipMalloc bases, so
  // 16-byte aligned), two in flight per thread, then the tail
  auto scan = [&](const float* __restrict__ a, long n) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const long n4 = n >> 2;
    auto m4 = [](float4 x) {
      return fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w)));
    };
    long k = t0;
    for (; k + stride < n4; k += 2 * stride) {
      const float4 x = a4[k], y = a4[k + stride];
      m = fmaxf(m, fmaxf(m4(x), m4(y)));
    }
    if (k < n4) m = fmaxf(m, m4(a4[k]));
    for (long t = (n4 << 2) + t0; t < n; t += stride) m = fmaxf(m, fabsf(a[t]));
  };
  scan(gm(d.u[q ^ 1]), (long)(d.nx + 1) * d.ny);
  scan(gm(d.v[q ^ 1]), (long)d.nx * (d.ny + 1));
  block_max_store_at(m, partials + (size_t)s * kBatchSpeedBlocks, bid);
}

__global__ __launch_bounds__(64) void kb_finalize_max(const float* __restrict__ partials,
                                                      float* __restrict__ host_out) {
  const int s = blockIdx.x;
  float m = 0.f;
  for (int b = threadIdx.x; b < kBatchSpeedBlocks; b += 64)
    m = fmaxf(m, partials[(size_t)s * kBatchSpeedBlocks + b]);
  m = wave_max(m);
  if (threadIdx.x == 0) host_out[s] = m;
}

dim3 batch_grid_project(int which, int nx, int ny) {
  switch (which) {
    case kBVorticity: return grid2(nx, ny);
    case kBForces: return grid2(nx + 1, ny + 1);
    case kBBuoyancy: return grid_rows(nx, ny - 1);
    default: return grid_rows(nx + 1, ny + 1);
  }
}

void launch_batch_vorticity(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                            const BStep* st, int q) {
  kb_vorticity<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q);
}

void launch_batch_forces(hipStream_t s, const SceneDesc* D, const BatchGrid& g, const BStep* st,
                         int q) {
  kb_forces<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q);
}

void launch_batch_buoyancy(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                           const BStep* st, int q) {
  kb_buoyancy<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q);
}

void launch_batch_project(hipStream_t s, const SceneDesc* D, const BatchGrid& g, const BStep* st,
                          int q) {
  kb_project<<<g.blocks, dim3(BX, BY), 0, s>>>(D, g, st, q);
}

void launch_batch_max_speed(hipStream_t s, const SceneDesc* D, const BatchGrid& g, int q,
                            int nscenes, float* partials, float* host_out) {
  kb_max_speed<<<g.blocks, 256, 0, s>>>(D, g, q, partials);
  kb_finalize_max<<<nscenes, 64, 0, s>>>(partials, host_out);
}

}  // namespace fh
