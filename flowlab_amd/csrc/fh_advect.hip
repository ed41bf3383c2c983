// Advection kernels: semi-Lagrangian and MacCormack for the MAC faces and
// for the three dye channels (src/core/advect.ts).
//
// One thread per output face/cell.  Every output is a gather of bilinear
// stencils around a CFL-bounded backtrace, so the reads are near-coalesced
// (a wavefront's 64 faces backtrace into a ~64+2c wide strip) and the kernels
// are bound by HBM traffic: MacCormack velocity moves ~68 B per cell
// (forward 2x13, correction 2x21), dye ~78 B per cell for 3 channels.
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {

constexpr int BX = 64, BY = 4;

static inline dim3 grid2(int w, int h) { return dim3((w + BX - 1) / BX, (h + BY - 1) / BY); }

// advect.ts:72-112 -- one launch covers the (nx+1)*ny u faces and the
// nx*(ny+1) v faces: thread (i,j) owns u(i,j) when j<ny and v(i,j) when i<nx.
__global__ __launch_bounds__(256) void k_advect_faces(
    const float* __restrict__ uc, const float* __restrict__ vc, const float* __restrict__ us,
    const float* __restrict__ vs, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i > nx || j > ny) return;
  Pos bx, by;
  if (j < ny) {
    int k = i + j * (nx + 1);
    if (solid_or_outside(lbl, nx, ny, i - 1, j) || solid_or_outside(lbl, nx, ny, i, j)) {
      uo[k] = us[k];
    } else {
      Pos X{i, 0.f}, Y{j, 0.5f};
      float v0 = sample_v(vc, nx, ny, X, Y);
      backtrace(uc, vc, nx, ny, X, Y, uc[k], v0, c, &bx, &by);
      uo[k] = sample_u(us, nx, ny, bx, by);
    }
  }
  if (i < nx) {
    int k = i + j * nx;
    if (solid_or_outside(lbl, nx, ny, i, j - 1) || solid_or_outside(lbl, nx, ny, i, j)) {
      vo[k] = vs[k];
    } else {
      Pos X{i, 0.5f}, Y{j, 0.f};
      float u0 = sample_u(uc, nx, ny, X, Y);
      backtrace(uc, vc, nx, ny, X, Y, u0, vc[k], c, &bx, &by);
      vo[k] = sample_v(vs, nx, ny, bx, by);
    }
  }
}

// advect.ts:177-210 fused: the reverse pass (carrier u^n, source uHat, -dt)
// is evaluated for this face only, then corrected = uHat + (u^n - uBar)/2 is
// clamped to the u^n stencil at the forward backtrace.  Each face reads only
// its own uBar, so fusing changes no value.
__global__ __launch_bounds__(256) void k_maccormack_faces(
    const float* __restrict__ un, const float* __restrict__ vn, const float* __restrict__ uh,
    const float* __restrict__ vh, float* __restrict__ uo, float* __restrict__ vo,
    const uint8_t* __restrict__ lbl, int nx, int ny, float c) {
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i > nx || j > ny) return;
  Pos rx, ry, fx, fy;
  if (j < ny) {
    int k = i + j * (nx + 1);
    if (solid_or_outside(lbl, nx, ny, i - 1, j) || solid_or_outside(lbl, nx, ny, i, j)) {
      uo[k] = un[k];
    } else {
      Pos X{i, 0.f}, Y{j, 0.5f};
      float u0 = un[k];
      float v0 = sample_v(vn, nx, ny, X, Y);
      backtrace(un, vn, nx, ny, X, Y, u0, v0, -c, &rx, &ry);
      float ubar = sample_u(uh, nx, ny, rx, ry);
      float corrected = uh[k] + 0.5f * (u0 - ubar);
      backtrace(un, vn, nx, ny, X, Y, u0, v0, c, &fx, &fy);
      Axis ax = clamped_axis(fx.i, fx.f, nx + 1);
      Axis ay = clamped_axis(fy.i, fy.f - 0.5f, ny);
      uo[k] = clamp_to_quad(corrected, gather4(un, nx + 1, ax, ay));
    }
  }
  if (i < nx) {
    int k = i + j * nx;
    if (solid_or_outside(lbl, nx, ny, i, j - 1) || solid_or_outside(lbl, nx, ny, i, j)) {
      vo[k] = vn[k];
    } else {
      Pos X{i, 0.5f}, Y{j, 0.f};
      float v0 = vn[k];
      float u0 = sample_u(un, nx, ny, X, Y);
      backtrace(un, vn, nx, ny, X, Y, u0, v0, -c, &rx, &ry);
      float vbar = sample_v(vh, nx, ny, rx, ry);
      float corrected = vh[k] + 0.5f * (v0 - vbar);
      backtrace(un, vn, nx, ny, X, Y, u0, v0, c, &fx, &fy);
      Axis ax = clamped_axis(fx.i, fx.f - 0.5f, nx);
      Axis ay = clamped_axis(fy.i, fy.f, ny + 1);
      vo[k] = clamp_to_quad(corrected, gather4(vn, nx, ax, ay));
    }
  }
}

// Cell (i,j) of the dye grid (ds x finer) in velocity-cell units.
__device__ __forceinline__ Pos dye_pos(int i, int ds) {
  return {i / ds, ((float)(i % ds) + 0.5f) / (float)ds};
}

// advect.ts:241-270 for three channels sharing one backtrace; out *= keep.
__global__ __launch_bounds__(256) void k_advect_dye(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ q,
    float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx, int ny, int ds, float c,
    float k0, float k1, float k2) {
  int dnx = nx * ds, dny = ny * ds;
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i >= dnx || j >= dny) return;
  size_t plane = (size_t)dnx * dny;
  int k = i + j * dnx;
  if (solid_or_outside(lbl, nx, ny, i / ds, j / ds)) {
    qo[k] = q[k] * k0;
    qo[k + plane] = q[k + plane] * k1;
    qo[k + 2 * plane] = q[k + 2 * plane] * k2;
    return;
  }
  Pos X = dye_pos(i, ds), Y = dye_pos(j, ds);
  float u0 = sample_u(u, nx, ny, X, Y);
  float v0 = sample_v(v, nx, ny, X, Y);
  Pos bx, by;
  backtrace(u, v, nx, ny, X, Y, u0, v0, c, &bx, &by);
  Pos dx{i, 0.5f + ds * (bx.f - X.f)}, dy{j, 0.5f + ds * (by.f - Y.f)};
  Axis ax = clamped_axis(dx.i, dx.f - 0.5f, dnx);
  Axis ay = clamped_axis(dy.i, dy.f - 0.5f, dny);
  Quad g0 = gather4(q, dnx, ax, ay);
  Quad g1 = gather4(q + plane, dnx, ax, ay);
  Quad g2 = gather4(q + 2 * plane, dnx, ax, ay);
  qo[k] = bilerp(g0.a, g0.b, g0.c, g0.d, ax.f, ay.f) * k0;
  qo[k + plane] = bilerp(g1.a, g1.b, g1.c, g1.d, ax.f, ay.f) * k1;
  qo[k + 2 * plane] = bilerp(g2.a, g2.b, g2.c, g2.d, ax.f, ay.f) * k2;
}

// advect.ts:291-311 for three channels, fused as for the faces; out *= keep.
__global__ __launch_bounds__(256) void k_maccormack_dye(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ qn,
    const float* __restrict__ qh, float* __restrict__ qo, const uint8_t* __restrict__ lbl, int nx,
    int ny, int ds, float c, float k0, float k1, float k2) {
  int dnx = nx * ds, dny = ny * ds;
  int i = blockIdx.x * BX + threadIdx.x;
  int j = blockIdx.y * BY + threadIdx.y;
  if (i >= dnx || j >= dny) return;
  size_t plane = (size_t)dnx * dny;
  int k = i + j * dnx;
  if (solid_or_outside(lbl, nx, ny, i / ds, j / ds)) {
    qo[k] = qn[k] * k0;
    qo[k + plane] = qn[k + plane] * k1;
    qo[k + 2 * plane] = qn[k + 2 * plane] * k2;
    return;
  }
  Pos X = dye_pos(i, ds), Y = dye_pos(j, ds);
  float u0 = sample_u(u, nx, ny, X, Y);
  float v0 = sample_v(v, nx, ny, X, Y);
  Pos bx, by;
  backtrace(u, v, nx, ny, X, Y, u0, v0, -c, &bx, &by);
  Axis ax = clamped_axis(i, ds * (bx.f - X.f), dnx);
  Axis ay = clamped_axis(j, ds * (by.f - Y.f), dny);
  Quad r0 = gather4(qh, dnx, ax, ay);
  Quad r1 = gather4(qh + plane, dnx, ax, ay);
  Quad r2 = gather4(qh + 2 * plane, dnx, ax, ay);
  float c0 = qh[k] + 0.5f * (qn[k] - bilerp(r0.a, r0.b, r0.c, r0.d, ax.f, ay.f));
  float c1 = qh[k + plane] + 0.5f * (qn[k + plane] - bilerp(r1.a, r1.b, r1.c, r1.d, ax.f, ay.f));
  float c2 = qh[k + 2 * plane] +
             0.5f * (qn[k + 2 * plane] - bilerp(r2.a, r2.b, r2.c, r2.d, ax.f, ay.f));
  backtrace(u, v, nx, ny, X, Y, u0, v0, c, &bx, &by);
  ax = clamped_axis(i, ds * (bx.f - X.f), dnx);
  ay = clamped_axis(j, ds * (by.f - Y.f), dny);
  qo[k] = clamp_to_quad(c0, gather4(qn, dnx, ax, ay)) * k0;
  qo[k + plane] = clamp_to_quad(c1, gather4(qn + plane, dnx, ax, ay)) * k1;
  qo[k + 2 * plane] = clamp_to_quad(c2, gather4(qn + 2 * plane, dnx, ax, ay)) * k2;
}

void launch_advect_faces(hipStream_t s, const float* uc, const float* vc, const float* us,
                         const float* vs, float* uo, float* vo, const uint8_t* lbl, int nx, int ny,
                         float c) {
  k_advect_faces<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(uc, vc, us, vs, uo, vo, lbl, nx,
                                                                ny, c);
}

void launch_maccormack_faces(hipStream_t s, const float* un, const float* vn, const float* uh,
                             const float* vh, float* uo, float* vo, const uint8_t* lbl, int nx,
                             int ny, float c) {
  k_maccormack_faces<<<grid2(nx + 1, ny + 1), dim3(BX, BY), 0, s>>>(un, vn, uh, vh, uo, vo, lbl,
                                                                    nx, ny, c);
}

void launch_advect_dye(hipStream_t s, const float* u, const float* v, const float* q, float* qo,
                       const uint8_t* lbl, int nx, int ny, int ds, float c, float k0, float k1,
                       float k2) {
  k_advect_dye<<<grid2(nx * ds, ny * ds), dim3(BX, BY), 0, s>>>(u, v, q, qo, lbl, nx, ny, ds, c,
                                                                k0, k1, k2);
}

void launch_maccormack_dye(hipStream_t s, const float* u, const float* v, const float* qn,
                           const float* qh, float* qo, const uint8_t* lbl, int nx, int ny, int ds,
                           float c, float k0, float k1, float k2) {
  k_maccormack_dye<<<grid2(nx * ds, ny * ds), dim3(BX, BY), 0, s>>>(u, v, qn, qh, qo, lbl, nx,
                                                                    ny, ds, c, k0, k1, k2);
}

}  // namespace fh
