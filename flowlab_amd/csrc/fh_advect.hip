// Advection kernels: semi-Lagrangian and MacCormack for the MAC faces and
// for the three dye channels (src/core/advect.ts).
//
// One thread per output face/cell.  Every output is a gather of bilinear
// stencils around a CFL-bounded backtrace, so the reads are near-coalesced
// (a wavefront's 64 faces backtrace into a ~64+2c wide strip) and the kernels
// are bound by HBM traffic: MacCormack velocity moves ~68 B per cell
// (forward 2x13, correction 2x21), dye ~78 B per cell for 3 channels.
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
