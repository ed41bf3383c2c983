// Body forces, divergence, gradient subtraction, boundaries and dye sources
// (src/core/{vorticity,buoyancy,stir,divergence,subtractGradient,boundaries,dye}.ts).
//
// All of these are single-pass streaming stencils: one thread per face or
// cell, x fastest, so each wavefront touches contiguous rows.
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {

constexpr int BX = 64, BY = 4;

static inline dim3 grid2(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }

// vorticity.ts:9-16 at corner (i,j), times h.
__device__ __forceinline__ float curl_h(const float* __restrict__ u, const float* __restrict__ v,
                                        int nx, int i, int j) {
  return (v[i + j * nx] - v[i - 1 + j * nx]) - (u[i + j * (nx + 1)] - u[i + (j - 1) * (nx + 1)]);
}

// vorticity.ts:24-42: cell-centred vorticity from the four clamped corners.
__global__ __launch_bounds__(256) void k_vorticity(const float* __restrict__ u,
                                                   const float* __restrict__ v,
                                                   float* __restrict__ w, int nx, int ny,
                                                   float inv_h) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  int hiI = nx - 1, hiJ = ny - 1;
  int j0 = min(max(j, 1), hiJ), j1 = min(max(j + 1, 1), hiJ);
  int i0 = min(max(i, 1), hiI), i1 = min(max(i + 1, 1), hiI);
  float s = curl_h(u, v, nx, i0, j0) * inv_h + curl_h(u, v, nx, i1, j0) * inv_h +
            curl_h(u, v, nx, i0, j1) * inv_h + curl_h(u, v, nx, i1, j1) * inv_h;
  w[i + j * nx] = 0.25f * s;
}

// Confinement force at cell (i,j), in velocity per unit time:
//   a = |w|, g = grad a (clamped central differences),
//   N = g / (|g| + 1e-4 a/h + 1e-30),  f = eps h (N_y w, -N_x w).
// Evaluated in cell units (g*h), which is the same N.
__device__ __forceinline__ void conf_force(const float* __restrict__ w, int nx, int ny, float h,
                                           float eps, int i, int j, float* fx, float* fy) {
  int ip = min(i + 1, nx - 1), im = max(i - 1, 0);
  int jp = min(j + 1, ny - 1), jm = max(j - 1, 0);
  float om = w[i + j * nx];
  float a = fabsf(om);
  float gx = (fabsf(w[ip + j * nx]) - fabsf(w[im + j * nx])) / (float)(ip - im);
  float gy = (fabsf(w[i + jp * nx]) - fabsf(w[i + jm * nx])) / (float)(jp - jm);
  float len = sqrtf(gx * gx + gy * gy);
  float inv = 1.0f / (len + 1e-4f * a + 1e-30f * h);
  *fx = eps * h * (gy * inv * om);
  *fy = eps * h * (-(gx * inv) * om);
}

// Body forces on the faces, in the order the CPU step adds them: stir
// (stir.ts:49-83), buoyancy on v (buoyancy.ts:70-88), then confinement
// computed from the velocity before any force.  Face set: u i in [1,nx-1],
// v j in [1,ny-1], skipping faces that touch an in-domain Solid.
__global__ __launch_bounds__(256) void k_forces(float* __restrict__ u, float* __restrict__ v,
                                                const uint8_t* __restrict__ lbl,
                                                const float* __restrict__ heat,
                                                const float* __restrict__ w, int nx, int ny,
                                                int ds, float h, float dt, float gain, float eps,
                                                StirBox st) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i > nx || j > ny) return;
  float invR2 = st.active ? 1.0f / (st.r * st.r) : 0.f;
  if (j < ny && i >= 1 && i <= nx - 1 && !is_solid_cell(lbl, nx, ny, i - 1, j) &&
      !is_solid_cell(lbl, nx, ny, i, j)) {
    int k = i + j * (nx + 1);
    float val = u[k];
    if (st.active && i >= st.ui0 && i <= st.ui1 && j >= st.uj0 && j <= st.uj1) {
      float dx = (float)i - st.x, dy = ((float)j + 0.5f) - st.y;
      val += st.dx * expf(-(dx * dx + dy * dy) * invR2);
    }
    if (eps != 0.f) {
      float fa, ga, fb, gb;
      conf_force(w, nx, ny, h, eps, i - 1, j, &fa, &ga);
      conf_force(w, nx, ny, h, eps, i, j, &fb, &gb);
      val += dt * (0.5f * (fa + fb));
    }
    u[k] = val;
  }
  if (i < nx && j >= 1 && j <= ny - 1 && !is_solid_cell(lbl, nx, ny, i, j - 1) &&
      !is_solid_cell(lbl, nx, ny, i, j)) {
    int k = i + j * nx;
    float val = v[k];
    if (st.active && i >= st.vi0 && i <= st.vi1 && j >= st.vj0 && j <= st.vj1) {
      float dx = ((float)i + 0.5f) - st.x, dy = (float)j - st.y;
      val += st.dy * expf(-(dx * dx + dy * dy) * invR2);
    }
    if (gain != 0.f) {
      // sampleP on the dye grid at the face ((i+1/2) h, j h)
      Axis ax = clamped_axis(i * ds, 0.5f * ds - 0.5f, nx * ds);
      Axis ay = clamped_axis(j * ds, -0.5f, ny * ds);
      Quad g = gather4(heat, nx * ds, ax, ay);
      val += gain * bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
    }
    if (eps != 0.f) {
      float fa, ga, fb, gb;
      conf_force(w, nx, ny, h, eps, i, j - 1, &fa, &ga);
      conf_force(w, nx, ny, h, eps, i, j, &fb, &gb);
      val += dt * (0.5f * (ga + gb));
    }
    v[k] = val;
  }
}

// divergence.ts:17-25, evaluated in FP64 from the FP32 faces.
__global__ __launch_bounds__(256) void k_divergence(const float* __restrict__ u,
                                                    const float* __restrict__ v,
                                                    double* __restrict__ div, int nx, int ny,
                                                    double inv_h) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const float* ur = u + j * (nx + 1);
  double d = ((double)ur[i + 1] - (double)ur[i]) + (double)v[i + (j + 1) * nx] -
             (double)v[i + j * nx];
  div[i + j * nx] = inv_h * d;
}

// subtractGradient.ts:40-61 with an FP64 pressure.
__global__ __launch_bounds__(256) void k_subtract(float* __restrict__ u, float* __restrict__ v,
                                                  const double* __restrict__ p,
                                                  const uint8_t* __restrict__ lbl, int nx, int ny,
                                                  double gs) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i > nx || j > ny) return;
  if (j < ny && i >= 1 && i <= nx - 1 && !solid_or_outside(lbl, nx, ny, i - 1, j) &&
      !solid_or_outside(lbl, nx, ny, i, j)) {
    int k = i + j * (nx + 1);
    u[k] = (float)((double)u[k] - gs * (p[i + j * nx] - p[i - 1 + j * nx]));
  }
  if (i < nx && j >= 1 && j <= ny - 1 && !solid_or_outside(lbl, nx, ny, i, j - 1) &&
      !solid_or_outside(lbl, nx, ny, i, j)) {
    int k = i + j * nx;
    v[k] = (float)((double)v[k] - gs * (p[i + j * nx] - p[i + (j - 1) * nx]));
  }
}

// boundaries.ts:115-143: thread t owns u row t and v column t, replaying the
// reference's low-edge-then-high-edge order within each.
__global__ __launch_bounds__(256) void k_outflow(float* __restrict__ u, float* __restrict__ v,
                                                 const uint8_t* __restrict__ lbl, int nx, int ny) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int su = nx + 1;
  if (t < ny) {
    int j = t;
    if (lbl[j * nx] == AIR) {
      float o = fminf(u[1 + j * su], 0.f);
      u[1 + j * su] = o;
      u[j * su] = o;
    }
    if (lbl[nx - 1 + j * nx] == AIR) {
      float o = fmaxf(u[nx - 1 + j * su], 0.f);
      u[nx - 1 + j * su] = o;
      u[nx + j * su] = o;
    }
  }
  if (t < nx) {
    int i = t;
    if (lbl[i] == AIR) {
      float o = fminf(v[i + nx], 0.f);
      v[i + nx] = o;
      v[i] = o;
    }
    if (lbl[i + (ny - 1) * nx] == AIR) {
      float o = fmaxf(v[i + (ny - 1) * nx], 0.f);
      v[i + (ny - 1) * nx] = o;
      v[i + ny * nx] = o;
    }
  }
}

// dye.ts:70-87: dye += cov (val - dye) inside the patch, skipped at cov 0.
__global__ __launch_bounds__(256) void k_dye_patch(float* __restrict__ dye,
                                                   const float* __restrict__ pd, int dnx, int dny,
                                                   int i0, int j0, int pnx, int pny) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
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
  long stride = (long)gridDim.x * blockDim.x;
  float m = 0.f;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nu; k += stride)
    m = fmaxf(m, fabsf(u[k]));
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nv; k += stride)
    m = fmaxf(m, fabsf(v[k]));
  block_max_store(m, partials);
}

__global__ __launch_bounds__(256) void k_finalize_max(const float* __restrict__ partials,
                                                      int nblocks, float* __restrict__ out) {
  __shared__ float sh[4];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  float t = 0.f;
  for (int b = tid; b < nblocks; b += 256) t = fmaxf(t, partials[b]);
  t = wave_max(t);
  if (lane == 0) sh[wid] = t;
  __syncthreads();
  if (tid == 0) out[0] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

void launch_vorticity(hipStream_t s, const float* u, const float* v, float* w, int nx, int ny,
                      float inv_h) {
  k_vorticity<<<grid2(nx, ny), dim3(BX, BY), 0, s>>>(u, v, w, nx, ny, inv_h);
}

void launch_forces(hipStream_t s, float* u, float* v, const uint8_t* lbl, const float* heat,
                   const float* w, int nx, int ny, int ds, float h, float dt, float gain,
                   float eps, StirBox stir) {
  k_forces<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(u, v, lbl, heat, w, nx, ny, ds, h, dt,
                                                          gain, eps, stir);
}

void launch_divergence(hipStream_t s, const float* u, const float* v, double* div, int nx, int ny,
                       double inv_h) {
  k_divergence<<<grid2(nx, ny), dim3(BX, BY), 0, s>>>(u, v, div, nx, ny, inv_h);
}

void launch_subtract_gradient(hipStream_t s, float* u, float* v, const double* p,
                              const uint8_t* lbl, int nx, int ny, double grad_scale) {
  k_subtract<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(u, v, p, lbl, nx, ny, grad_scale);
}

void launch_outflow(hipStream_t s, float* u, float* v, const uint8_t* lbl, int nx, int ny) {
  int n = nx > ny ? nx : ny;
  k_outflow<<<(n + 255) / 256, 256, 0, s>>>(u, v, lbl, nx, ny);
}

void launch_dye_patch(hipStream_t s, float* dye, const float* patch, int dnx, int dny, int i0,
                      int j0, int pnx, int pny) {
  k_dye_patch<<<grid2(pnx, pny), dim3(BX, BY), 0, s>>>(dye, patch, dnx, dny, i0, j0, pnx, pny);
}

void launch_commit_labels(hipStream_t s, double* p, float* u, float* v, float* dye,
                          const uint8_t* lbl, int nx, int ny, int dye_plane) {
  k_commit_labels<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(p, u, v, dye, lbl, nx, ny,
                                                                 dye_plane);
}

void launch_max_speed(hipStream_t s, const float* u, const float* v, int nx, int ny,
                      float* partials, unsigned* counter, float* out) {
  k_max_speed<<<1024, 256, 0, s>>>(u, v, (long)(nx + 1) * ny, (long)nx * (ny + 1), partials,
                                   counter, out);
  k_finalize_max<<<1, 256, 0, s>>>(partials, 1024, out);
}

}  // namespace fh
