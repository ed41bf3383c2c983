// Per-element bodies of the streaming stencil kernels: advection
// (src/core/advect.ts), forces, divergence, gradient subtraction, outflow,
// dye sources and label commits (src/core/{vorticity,buoyancy,stir,
// divergence,subtractGradient,boundaries,dye}.ts).
//
// Each function computes the outputs owned by one thread index (i, j) of one
// grid and returns at once outside its range.  The single-grid kernels map
// (blockIdx, threadIdx) to (i, j); the batched kernels map a task of their
// work list (scene, block) to the scene's fields and (i, j).  Both run this
// same code, so a scene gives the same bits either way.
#pragma once
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {

// ---- advection (advect.ts) ---------------------------------------------------
// Bodies are generic over field accessors (GField / LField, fh_device.hpp);
// the pointer versions read global memory and give the same bits.

// MacCormack's limiter clamps to the source stencil at the forward
// backtrace (advect.ts:200-209), the stencil the forward pass has just
// sampled.  The forward pass can record each face's (or dye cell's) stencil
// origin relative to its own index in 2 bits per axis (offset + 2), and the
// correction pass reads it instead of tracing forward again.  The offsets lie
// in [-2, 1] whenever the correction's workgroup passes its reach test (every
// backtrace of the tile then moves at most kReach = 1.25 cells, and the
// domain clamp only moves an origin towards the face), which is the only
// case that reads them.
__device__ __forceinline__ uint32_t sten2(int d0, int d1) {
  return (uint32_t)((d0 + 2) & 3) | ((uint32_t)((d1 + 2) & 3) << 2);
}
__device__ __forceinline__ int sten_d0(uint32_t s) { return (int)(s & 3) - 2; }
__device__ __forceinline__ int sten_d1(uint32_t s) { return (int)((s >> 2) & 3) - 2; }

// The solid tests of thread index (i, j)'s faces (bit 0: u face (i, j),
// solid_or_outside at (i-1, j) or (i, j); bit 1: v face, at (i, j-1) or
// (i, j)) from three label loads issued unconditionally, at indices clamped
// into the grid: tested one by one behind their short-circuit branches,
// each label load waited for the one before.
__device__ __forceinline__ int closed_faces(const uint8_t* __restrict__ lbl, int nx, int ny, int i,
                                            int j) {
  auto at = [&](int a, int b) {
    const uint8_t l = lbl[min(max(a, 0), nx - 1) + min(max(b, 0), ny - 1) * nx];
    return a < 0 || b < 0 || a >= nx || b >= ny || l == SOLID;
  };
  const bool c = at(i, j), w = at(i - 1, j), s = at(i, j - 1);
  return (int)(w || c) | ((int)(s || c) << 1);
}

// advect.ts:72-112 -- one thread index covers u(i,j) when j<ny and v(i,j)
// when i<nx, over (nx+1) x (ny+1).  uc/vc carry, us/vs are advected.  The
// values go to *uo_v and *vo_v; returns the stencil origins of both faces'
// samples as one byte (u face in bits 0-3, v face in bits 4-7; 0 for solid
// faces).  closed: the two faces' solid tests when the caller has loaded
// them (closed_faces; -1: read lbl here).  cross (may be null): the carrier
// sampled across at each open face, v at the u face then u at the v face
// (maccormack_faces_t samples the same).
template <class CU, class CV, class SU, class SV>
__device__ __forceinline__ uint32_t advect_faces_cell(const CU& uc, const CV& vc, const SU& us,
                                                      const SV& vs,
                                                      const uint8_t* __restrict__ lbl, int nx,
                                                      int ny, float c, int i, int j, int closed,
                                                      float* uo_v, float* vo_v,
                                                      float* cross = nullptr) {
  Pos bx, by;
  uint32_t st = 0;
  if (j < ny) {
    if (closed >= 0 ? (closed & 1) != 0
                    : solid_or_outside(lbl, nx, ny, i - 1, j) ||
                          solid_or_outside(lbl, nx, ny, i, j)) {
      *uo_v = us.at(i, j);
    } else {
      Pos X{i, 0.f}, Y{j, 0.5f};
      float v0 = sample_vf(vc, nx, ny, X, Y);
      if (cross) cross[0] = v0;
      backtracef(uc, vc, nx, ny, X, Y, uc.at(i, j), v0, c, &bx, &by);
      // sample_uf, with the stencil origin kept
      Axis ax = axis_t<SU::kClamp>(bx.i, bx.f, nx + 1);
      Axis ay = axis_t<SU::kClamp>(by.i, by.f - 0.5f, ny);
      Quad g = gather4f(us, ax, ay);
      *uo_v = bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
      st = sten2(ax.i0 - i, ay.i0 - j);
    }
  }
  if (i < nx) {
    if (closed >= 0 ? (closed & 2) != 0
                    : solid_or_outside(lbl, nx, ny, i, j - 1) ||
                          solid_or_outside(lbl, nx, ny, i, j)) {
      *vo_v = vs.at(i, j);
    } else {
      Pos X{i, 0.5f}, Y{j, 0.f};
      float u0 = sample_uf(uc, nx, ny, X, Y);
      if (cross) cross[1] = u0;
      backtracef(uc, vc, nx, ny, X, Y, u0, vc.at(i, j), c, &bx, &by);
      // sample_vf, with the stencil origin kept
      Axis ax = axis_t<SV::kClamp>(bx.i, bx.f - 0.5f, nx);
      Axis ay = axis_t<SV::kClamp>(by.i, by.f, ny + 1);
      Quad g = gather4f(vs, ax, ay);
      *vo_v = bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
      st |= sten2(ax.i0 - i, ay.i0 - j) << 4;
    }
  }
  return st;
}

// advect_faces_cell stored: u(i, j) into uo when j < ny, v(i, j) into vo
// when i < nx, the stencil byte into sten[i + j (nx + 1)] (sten may be null).
template <class CU, class CV, class SU, class SV>
__device__ __forceinline__ void advect_faces_t(const CU& uc, const CV& vc, const SU& us,
                                               const SV& vs, float* __restrict__ uo,
                                               float* __restrict__ vo,
                                               const uint8_t* __restrict__ lbl, int nx, int ny,
                                               float c, int i, int j,
                                               uint8_t* __restrict__ sten = nullptr,
                                               int closed = -1) {
  if (i > nx || j > ny) return;
  float a = 0.f, b = 0.f;
  const uint32_t st = advect_faces_cell(uc, vc, us, vs, lbl, nx, ny, c, i, j, closed, &a, &b);
  if (j < ny) uo[i + j * (nx + 1)] = a;
  if (i < nx) vo[i + j * nx] = b;
  if (sten) sten[i + j * (nx + 1)] = (uint8_t)st;
}

__device__ __forceinline__ void advect_faces_at(const float* __restrict__ uc,
                                                const float* __restrict__ vc,
                                                const float* __restrict__ us,
                                                const float* __restrict__ vs,
                                                float* __restrict__ uo, float* __restrict__ vo,
                                                const uint8_t* __restrict__ lbl, int nx, int ny,
                                                float c, int i, int j) {
  advect_faces_t(GField{uc, nx + 1}, GField{vc, nx}, GField{us, nx + 1}, GField{vs, nx}, uo, vo,
                 lbl, nx, ny, c, i, j);
}

// buoyancy.ts:70-88 on one open interior v face (i, j): val + gain times the
// heat (dye channel 0 on a grid ds x finer) at the face, as k_buoyancy
// applies it; shared so the fused and the separate pass give the same bits.
__device__ __forceinline__ float buoyancy_value(float val, const float* __restrict__ heat, int nx,
                                                int ny, int ds, float gain, int i, int j) {
  const int dnx = nx * ds;
  const Axis ax = clamped_axis(i * ds, 0.5f * ds - 0.5f, dnx);
  const Axis ay = clamped_axis(j * ds, -0.5f, ny * ds);
  const Quad g = gather4(heat, dnx, ax, ay);
  return val + gain * bilerp(g.a, g.b, g.c, g.d, ax.f, ay.f);
}

// advect.ts:177-210 fused: the reverse pass (carrier u^n, source uHat, -dt)
// is evaluated for this face only, then corrected = uHat + (u^n - uBar)/2 is
// clamped to the u^n stencil at the forward backtrace.  Each face reads only
// its own uBar, so fusing changes no value.  has_sten: st is this index's
// byte of the forward pass's stencil origins (advect_faces_t), used instead
// of tracing forward again; the caller sets it only where the offsets are
// known to fit.  cross (may be null): advect_faces_cell's cross samples of
// this index, used instead of sampling them again.
template <class NU, class NV, class HU, class HV>
__device__ __forceinline__ void maccormack_faces_t(const NU& un, const NV& vn, const HU& uh,
                                                   const HV& vh, float* __restrict__ uo,
                                                   float* __restrict__ vo,
                                                   const uint8_t* __restrict__ lbl, int nx, int ny,
                                                   float c, int i, int j, bool has_sten = false,
                                                   uint32_t st = 0,
                                                   const float* __restrict__ heat = nullptr,
                                                   float gain = 0.f, int closed = -1,
                                                   const float* cross = nullptr) {
  if (i > nx || j > ny) return;
  Pos rx, ry, fx, fy;
  if (j < ny) {
    int k = i + j * (nx + 1);
    if (closed >= 0 ? (closed & 1) != 0
                    : solid_or_outside(lbl, nx, ny, i - 1, j) ||
                          solid_or_outside(lbl, nx, ny, i, j)) {
      uo[k] = un.at(i, j);
    } else {
      Pos X{i, 0.f}, Y{j, 0.5f};
      float u0 = un.at(i, j);
      float v0 = cross ? cross[0] : sample_vf(vn, nx, ny, X, Y);
      backtracef(un, vn, nx, ny, X, Y, u0, v0, -c, &rx, &ry);
      float ubar = sample_uf(uh, nx, ny, rx, ry);
      float corrected = uh.at(i, j) + 0.5f * (u0 - ubar);
      Axis ax, ay;
      if (has_sten) {
        ax = {i + sten_d0(st), 0.f};
        ay = {j + sten_d1(st), 0.f};
      } else {
        backtracef(un, vn, nx, ny, X, Y, u0, v0, c, &fx, &fy);
        ax = axis_t<NU::kClamp>(fx.i, fx.f, nx + 1);
        ay = axis_t<NU::kClamp>(fy.i, fy.f - 0.5f, ny);
      }
      uo[k] = clamp_to_quad(corrected, gather4f(un, ax, ay));
    }
  }
  if (i < nx) {
    int k = i + j * nx;
    if (closed >= 0 ? (closed & 2) != 0
                    : solid_or_outside(lbl, nx, ny, i, j - 1) ||
                          solid_or_outside(lbl, nx, ny, i, j)) {
      vo[k] = vn.at(i, j);
    } else {
      Pos X{i, 0.5f}, Y{j, 0.f};
      float v0 = vn.at(i, j);
      float u0 = cross ? cross[1] : sample_uf(un, nx, ny, X, Y);
      backtracef(un, vn, nx, ny, X, Y, u0, v0, -c, &rx, &ry);
      float vbar = sample_vf(vh, nx, ny, rx, ry);
      float corrected = vh.at(i, j) + 0.5f * (v0 - vbar);
      Axis ax, ay;
      if (has_sten) {
        ax = {i + sten_d0(st >> 4), 0.f};
        ay = {j + sten_d1(st >> 4), 0.f};
      } else {
        backtracef(un, vn, nx, ny, X, Y, u0, v0, c, &fx, &fy);
        ax = axis_t<NV::kClamp>(fx.i, fx.f - 0.5f, nx);
        ay = axis_t<NV::kClamp>(fy.i, fy.f, ny + 1);
      }
      const float vc = clamp_to_quad(corrected, gather4f(vn, ax, ay));
      // buoyancy alone (k_buoyancy's value) on this open interior face
      vo[k] = heat ? buoyancy_value(vc, heat, nx, ny, 1, gain, i, j) : vc;
    }
  }
}

__device__ __forceinline__ void maccormack_faces_at(const float* __restrict__ un,
                                                    const float* __restrict__ vn,
                                                    const float* __restrict__ uh,
                                                    const float* __restrict__ vh,
                                                    float* __restrict__ uo, float* __restrict__ vo,
                                                    const uint8_t* __restrict__ lbl, int nx,
                                                    int ny, float c, int i, int j) {
  maccormack_faces_t(GField{un, nx + 1}, GField{vn, nx}, GField{uh, nx + 1}, GField{vh, nx}, uo,
                     vo, lbl, nx, ny, c, i, j);
}

// Cell (i,j) of the dye grid (ds x finer) in velocity-cell units.
__device__ __forceinline__ Pos dye_pos(int i, int ds) {
  return {i / ds, ((float)(i % ds) + 0.5f) / (float)ds};
}

// advect.ts:241-270 for three channels sharing one backtrace, at dye cell
// (i, j) inside the dye grid: out = the three samples times keep; returns
// the origin of the stencil MacCormack's limiter clamps to, relative to
// (i, j) (sten2; maccormack_dye_t's forward trace; 0 for solid cells).  q0..q2
// are the channels as accessors over the dye grid (dnx x dny).  solid: the
// cell's solid_or_outside test when the caller has already loaded it (-1:
// read lbl).  uv0 (may be null): the velocity sampled at the cell centre, u
// then v, for an open cell (maccormack_dye_t samples the same).
template <class CU, class CV, class Q>
__device__ __forceinline__ uint32_t advect_dye_cell(const CU& u, const CV& v, const Q& q0,
                                                    const Q& q1, const Q& q2,
                                                    const uint8_t* __restrict__ lbl, int nx,
                                                    int ny, int ds, float c, float k0, float k1,
                                                    float k2, int i, int j, int solid,
                                                    float (&out)[3], float* uv0 = nullptr) {
  const int dnx = nx * ds, dny = ny * ds;
  if (solid >= 0 ? solid != 0 : solid_or_outside(lbl, nx, ny, i / ds, j / ds)) {
    out[0] = q0.at(i, j) * k0;
    out[1] = q1.at(i, j) * k1;
    out[2] = q2.at(i, j) * k2;
    return 0;
  }
  Pos X = dye_pos(i, ds), Y = dye_pos(j, ds);
  float u0 = sample_uf(u, nx, ny, X, Y);
  float v0 = sample_vf(v, nx, ny, X, Y);
  if (uv0) {
    uv0[0] = u0;
    uv0[1] = v0;
  }
  Pos bx, by;
  backtracef(u, v, nx, ny, X, Y, u0, v0, c, &bx, &by);
  // the sample at the backtrace: dye cell i's centre sits at i + 1/2 in
  // dye-cell units, so the stencil axis is i + ds (bx - X) -- the axes the
  // limiter of maccormack_dye_t clamps with, returned as the stencil byte
  Axis ax = axis_t<Q::kClamp>(i, ds * (bx.f - X.f), dnx);
  Axis ay = axis_t<Q::kClamp>(j, ds * (by.f - Y.f), dny);
  Quad g0 = gather4f(q0, ax, ay);
  Quad g1 = gather4f(q1, ax, ay);
  Quad g2 = gather4f(q2, ax, ay);
  out[0] = bilerp(g0.a, g0.b, g0.c, g0.d, ax.f, ay.f) * k0;
  out[1] = bilerp(g1.a, g1.b, g1.c, g1.d, ax.f, ay.f) * k1;
  out[2] = bilerp(g2.a, g2.b, g2.c, g2.d, ax.f, ay.f) * k2;
  return sten2(ax.i0 - i, ay.i0 - j);
}

// advect_dye_cell stored: the output planes are qo, qo + plane, qo + 2 plane;
// sten (may be null): the stencil bytes, one per cell.
template <class CU, class CV, class Q>
__device__ __forceinline__ void advect_dye_t(const CU& u, const CV& v, const Q& q0, const Q& q1,
                                             const Q& q2, float* __restrict__ qo,
                                             const uint8_t* __restrict__ lbl, int nx, int ny,
                                             int ds, float c, float k0, float k1, float k2, int i,
                                             int j, uint8_t* __restrict__ sten = nullptr,
                                             int solid = -1) {
  const int dnx = nx * ds, dny = ny * ds;
  if (i >= dnx || j >= dny) return;
  const size_t plane = (size_t)dnx * dny;
  const int k = i + j * dnx;
  float o[3];
  const uint32_t st = advect_dye_cell(u, v, q0, q1, q2, lbl, nx, ny, ds, c, k0, k1, k2, i, j,
                                      solid, o);
  qo[k] = o[0];
  qo[k + plane] = o[1];
  qo[k + 2 * plane] = o[2];
  if (sten) sten[k] = (uint8_t)st;
}

__device__ __forceinline__ void advect_dye_at(const float* __restrict__ u,
                                              const float* __restrict__ v,
                                              const float* __restrict__ q, float* __restrict__ qo,
                                              const uint8_t* __restrict__ lbl, int nx, int ny,
                                              int ds, float c, float k0, float k1, float k2, int i,
                                              int j) {
  const int dnx = nx * ds;
  const size_t plane = (size_t)dnx * ny * ds;
  advect_dye_t(GField{u, nx + 1}, GField{v, nx}, GField{q, dnx}, GField{q + plane, dnx},
               GField{q + 2 * plane, dnx}, qo, lbl, nx, ny, ds, c, k0, k1, k2, i, j);
}

// advect.ts:291-311 for three channels, fused as for the faces; out *= keep.
// has_sten: st is this cell's byte of the forward pass's clamp stencil
// origins (advect_dye_t), used instead of tracing forward again.  solid as
// for advect_dye_t.  uv0 (may be null): the centre velocity advect_dye_cell
// sampled for this cell, used instead of sampling it again.
template <class CU, class CV, class Q>
__device__ __forceinline__ void maccormack_dye_t(const CU& u, const CV& v, const Q& n0,
                                                 const Q& n1, const Q& n2, const Q& h0,
                                                 const Q& h1, const Q& h2, float* __restrict__ qo,
                                                 const uint8_t* __restrict__ lbl, int nx, int ny,
                                                 int ds, float c, float k0, float k1, float k2,
                                                 int i, int j, bool has_sten = false,
                                                 uint32_t st = 0, int solid = -1,
                                                 const float* uv0 = nullptr) {
  int dnx = nx * ds, dny = ny * ds;
  if (i >= dnx || j >= dny) return;
  size_t plane = (size_t)dnx * dny;
  int k = i + j * dnx;
  if (solid >= 0 ? solid != 0 : solid_or_outside(lbl, nx, ny, i / ds, j / ds)) {
    qo[k] = n0.at(i, j) * k0;
    qo[k + plane] = n1.at(i, j) * k1;
    qo[k + 2 * plane] = n2.at(i, j) * k2;
    return;
  }
  Pos X = dye_pos(i, ds), Y = dye_pos(j, ds);
  const float u0 = uv0 ? uv0[0] : sample_uf(u, nx, ny, X, Y);
  const float v0 = uv0 ? uv0[1] : sample_vf(v, nx, ny, X, Y);
  Pos bx, by;
  backtracef(u, v, nx, ny, X, Y, u0, v0, -c, &bx, &by);
  Axis ax = axis_t<Q::kClamp>(i, ds * (bx.f - X.f), dnx);
  Axis ay = axis_t<Q::kClamp>(j, ds * (by.f - Y.f), dny);
  Quad r0 = gather4f(h0, ax, ay);
  Quad r1 = gather4f(h1, ax, ay);
  Quad r2 = gather4f(h2, ax, ay);
  float c0 = h0.at(i, j) + 0.5f * (n0.at(i, j) - bilerp(r0.a, r0.b, r0.c, r0.d, ax.f, ay.f));
  float c1 = h1.at(i, j) + 0.5f * (n1.at(i, j) - bilerp(r1.a, r1.b, r1.c, r1.d, ax.f, ay.f));
  float c2 = h2.at(i, j) + 0.5f * (n2.at(i, j) - bilerp(r2.a, r2.b, r2.c, r2.d, ax.f, ay.f));
  if (has_sten) {
    ax = {i + sten_d0(st), 0.f};
    ay = {j + sten_d1(st), 0.f};
  } else {
    backtracef(u, v, nx, ny, X, Y, u0, v0, c, &bx, &by);
    ax = axis_t<Q::kClamp>(i, ds * (bx.f - X.f), dnx);
    ay = axis_t<Q::kClamp>(j, ds * (by.f - Y.f), dny);
  }
  qo[k] = clamp_to_quad(c0, gather4f(n0, ax, ay)) * k0;
  qo[k + plane] = clamp_to_quad(c1, gather4f(n1, ax, ay)) * k1;
  qo[k + 2 * plane] = clamp_to_quad(c2, gather4f(n2, ax, ay)) * k2;
}

__device__ __forceinline__ void maccormack_dye_at(const float* __restrict__ u,
                                                  const float* __restrict__ v,
                                                  const float* __restrict__ qn,
                                                  const float* __restrict__ qh,
                                                  float* __restrict__ qo,
                                                  const uint8_t* __restrict__ lbl, int nx, int ny,
                                                  int ds, float c, float k0, float k1, float k2,
                                                  int i, int j) {
  const int dnx = nx * ds;
  const size_t plane = (size_t)dnx * ny * ds;
  maccormack_dye_t(GField{u, nx + 1}, GField{v, nx}, GField{qn, dnx}, GField{qn + plane, dnx},
                   GField{qn + 2 * plane, dnx}, GField{qh, dnx}, GField{qh + plane, dnx},
                   GField{qh + 2 * plane, dnx}, qo, lbl, nx, ny, ds, c, k0, k1, k2, i, j);
}

// ---- forces and projection -------------------------------------------------------

// vorticity.ts:9-16 at corner (i,j), times h.
__device__ __forceinline__ float curl_h(const float* __restrict__ u, const float* __restrict__ v,
                                        int nx, int i, int j) {
  return (v[i + j * nx] - v[i - 1 + j * nx]) - (u[i + j * (nx + 1)] - u[i + (j - 1) * (nx + 1)]);
}

// vorticity.ts:24-42: cell-centred vorticity from the four clamped corners.
__device__ __forceinline__ void vorticity_at(const float* __restrict__ u,
                                             const float* __restrict__ v, float* __restrict__ w,
                                             int nx, int ny, float inv_h, int i, int j) {
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
__device__ __forceinline__ void forces_at(float* __restrict__ u, float* __restrict__ v,
                                          const uint8_t* __restrict__ lbl,
                                          const float* __restrict__ heat,
                                          const float* __restrict__ w, int nx, int ny, int ds,
                                          float h, float dt, float gain, float eps,
                                          const StirBox& st, int i, int j) {
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

// subtractGradient.ts:40-61 with an FP64 pressure.
__device__ __forceinline__ void subtract_at(float* __restrict__ u, float* __restrict__ v,
                                            const double* __restrict__ p,
                                            const uint8_t* __restrict__ lbl, int nx, int ny,
                                            double gs, int i, int j) {
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

// subtract_at for an index inside the face grid (i <= nx, j <= ny) of a grid
// of at least 2 x 2 cells, split into its loads (the three cell labels and
// pressures it can need, u and v: issued before any store) and its stores;
// the same faces and values as subtract_at.
struct SubLoad {
  bool uw, vw;  // u / v face of this index is subtracted (interior, no solid side)
  int ku, kv;
  float u0, v0;
  double p0, pw, ps;
};

__device__ __forceinline__ SubLoad subtract_load(const float* __restrict__ u,
                                                 const float* __restrict__ v,
                                                 const double* __restrict__ p,
                                                 const uint8_t* __restrict__ lbl, int nx, int ny,
                                                 int i, int j) {
  SubLoad L;
  const bool uf = j < ny && i >= 1 && i <= nx - 1;
  const bool vf = i < nx && j >= 1 && j <= ny - 1;
  const int c = min(i, nx - 1) + min(j, ny - 1) * nx;  // cell (i, j), in the domain here
  const int cw = i >= 1 ? c - 1 : c, cs = j >= 1 ? c - nx : c;  // (i-1, j), (i, j-1)
  const uint8_t l0 = lbl[c], lw = lbl[cw], ls = lbl[cs];
  L.p0 = p[c];
  L.pw = p[cw];
  L.ps = p[cs];
  L.ku = i + j * (nx + 1);
  L.kv = i + j * nx;
  L.u0 = uf ? u[L.ku] : 0.f;
  L.v0 = vf ? v[L.kv] : 0.f;
  L.uw = uf && lw != SOLID && l0 != SOLID;
  L.vw = vf && ls != SOLID && l0 != SOLID;
  return L;
}

__device__ __forceinline__ void subtract_store(const SubLoad& L, float* __restrict__ u,
                                               float* __restrict__ v, double gs) {
  if (L.uw) u[L.ku] = (float)((double)L.u0 - gs * (L.p0 - L.pw));
  if (L.vw) v[L.kv] = (float)((double)L.v0 - gs * (L.p0 - L.ps));
}

// boundaries.ts:115-143: index t owns u row t and v column t, replaying the
// reference's low-edge-then-high-edge order within each.
__device__ __forceinline__ void outflow_at(float* __restrict__ u, float* __restrict__ v,
                                           const uint8_t* __restrict__ lbl, int nx, int ny,
                                           int t) {
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

}  // namespace fh
