// Fused multigrid kernels for the V(2,2) cycle (pressureMultigrid.ts:253-275).
//
// A V-cycle launch per level used to be 8 colour half-sweeps + residual/
// restrict + prolong = 10 launches of a few microseconds each, so the cycle
// was bound by launch latency, not by HBM.  Here each level costs 2 launches
// and all levels small enough for one workgroup's LDS cost 1 in total:
//
//   k_smooth_restrict  (level l)   2 red-black sweeps + residual + restriction
//   k_prolong_smooth   (level l)   prolongation + 2 red-black sweeps
//   k_coarse_vcycle    (levels >= lc, one workgroup, all in LDS) the whole
//                                  sub-cycle including the coarsest solve
//
// Temporal blocking: a 64x32 tile is loaded with a halo of H cells into LDS
// and the 4 colour half-sweeps run there, the valid region shrinking by one
// cell per half-sweep; each cell's update depends only on its 4 neighbours'
// previous values, so the tile interior equals the global sweep.  Unknowns
// at non-fluid cells are zero, so a neighbour sum over all four neighbours in
// the reference order (+x, -x, +y, -y) equals the reference's sum over
// non-solid neighbours exactly.
//
// Every kernel takes `done`: once the defect-correction loop has converged
// (flag set on the device by the norm kernel), cycles the host enqueued
// ahead of that knowledge return at once and change nothing.
#include "fh_batchdesc.hpp"
#include "fh_mgtile.hpp"

namespace fh {

template <int TX, int TY, int W>
__global__ __launch_bounds__(W) void k_smooth_restrict(const float* __restrict__ xin,
                                                        float* __restrict__ xout,
                                                        const float* __restrict__ b,
                                                        const uint8_t* __restrict__ info, int nx,
                                                        int ny, float* __restrict__ bc, int cnx,
                                                        int cny, int zero,
                                                        const int* __restrict__ done) {
  int bx, by;
  xcd_block(&bx, &by);
  smooth_restrict_tile<TX, TY, W>(xin, xout, b, info, nx, ny, bc, cnx, cny, zero, done, bx, by);
}

template <int TX, int TY, int W>
__global__ __launch_bounds__(W) void k_prolong_smooth(const float* __restrict__ xin,
                                                       float* __restrict__ xout,
                                                       const float* __restrict__ b,
                                                       const uint8_t* __restrict__ info, int nx,
                                                       int ny, const float* __restrict__ e,
                                                       int cnx, int cny,
                                                       const int* __restrict__ done) {
  int bx, by;
  xcd_block(&bx, &by);
  prolong_smooth_tile<TX, TY, W>(xin, xout, b, info, nx, ny, e, cnx, cny, done, bx, by);
}

template <int TX, int TY, int W>
__global__ __launch_bounds__(W) void k_prolong_smooth_dc(
    const float* __restrict__ xin, const float* __restrict__ b, const uint8_t* __restrict__ info,
    int nx, int ny, const float* __restrict__ e, int cnx, int cny, const int* __restrict__ done,
    const double* __restrict__ po, double* __restrict__ pn, const double* __restrict__ div,
    float* __restrict__ rout, double* __restrict__ r64, double scale,
    double* __restrict__ partials) {
  int bx, by;
  xcd_block(&bx, &by);
  prolong_smooth_dc_tile<TX, TY, W>(xin, b, info, nx, ny, e, cnx, cny, done, bx, by, po, pn, div,
                                    rout, r64, scale, partials, bx + by * (int)gridDim.x);
}

template <int G>
__global__ __launch_bounds__(G) void k_coarse_vcycle(CoarseLevels L, const float* __restrict__ b0,
                                                     const float* __restrict__ x0in,
                                                     float* __restrict__ x0out,
                                                     const double* __restrict__ M,
                                                     const int* __restrict__ cells, int ncoarse,
                                                     int direct, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ unsigned char smem[];
  coarse_vcycle_body<G>(L, b0, x0in, x0out, M, cells, ncoarse, direct, smem);
}

__global__ void k_cell_info(const uint8_t* __restrict__ lbl, int nx, int ny,
                            uint8_t* __restrict__ info) {
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i < nx && j < ny) info[i + j * nx] = cell_info(lbl, nx, ny, i, j);
}

void coarse_levels_layout(CoarseLevels& L) {
  int cells = 0, ibytes = 0;
  for (int l = 0; l < L.n; l++) cells += (L.nx[l] + 2) * (L.ny[l] + 2);
  for (int l = 0, c = 0; l < L.n; l++) {
    L.xo[l] = c;
    L.bo[l] = cells + c;
    L.io[l] = 8 * cells + ibytes;
    L.mnx[l] = fastdiv_magic(L.nx[l]);
    L.mpr[l] = fastdiv_magic((L.nx[l] + 1) >> 1);
    c += (L.nx[l] + 2) * (L.ny[l] + 2);
    ibytes += padded_info_bytes(L.nx[l], L.ny[l]);
  }
  L.cells = cells;
  L.ibytes = ibytes;
}

size_t coarse_vcycle_lds_bytes(const CoarseLevels& L) {
  size_t b = 0;
  for (int l = 0; l < L.n; l++)
    b += 8 * (size_t)(L.nx[l] + 2) * (L.ny[l] + 2) + padded_info_bytes(L.nx[l], L.ny[l]);
  return b + 8 * (size_t)kMaxDirect;  // the direct solve's dense right-hand side and indices
}

__global__ void k_pad_info(const uint8_t* __restrict__ info, int nx, int ny,
                           uint8_t* __restrict__ out, int nbytes) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nbytes) return;
  const int px = nx + 2, j = q / px, i = q - j * px;
  const bool in = i >= 1 && i <= nx && j >= 1 && j <= ny;
  out[q] = in ? info[(i - 1) + (j - 1) * nx] : 0;
}

void launch_pad_info(hipStream_t s, const uint8_t* info, int nx, int ny, uint8_t* out) {
  const int n = padded_info_bytes(nx, ny);
  k_pad_info<<<(n + 255) / 256, 256, 0, s>>>(info, nx, ny, out, n);
}

void launch_cell_info(hipStream_t s, const uint8_t* lbl, int nx, int ny, uint8_t* info) {
  dim3 g((nx + 63) / 64, (ny + 3) / 4);
  k_cell_info<<<g, dim3(64, 4), 0, s>>>(lbl, nx, ny, info);
}

template <int TX, int TY, int W>
static void sr_launch(hipStream_t s, const float* xin, float* xout, const float* b,
                      const uint8_t* info, int nx, int ny, float* bc, int cnx, int cny, int zero,
                      const int* done) {
  dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY);
  k_smooth_restrict<TX, TY, W><<<g, W, 0, s>>>(xin, xout, b, info, nx, ny, bc, cnx, cny, zero,
                                               done);
}

template <int TX, int TY, int W>
static void ps_launch(hipStream_t s, const float* xin, float* xout, const float* b,
                      const uint8_t* info, int nx, int ny, const float* e, int cnx, int cny,
                      const int* done) {
  dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY);
  k_prolong_smooth<TX, TY, W><<<g, W, 0, s>>>(xin, xout, b, info, nx, ny, e, cnx, cny, done);
}

template <int TX, int TY, int W>
static int psdc_launch(hipStream_t s, const float* xin, const float* b, const uint8_t* info,
                       int nx, int ny, const float* e, int cnx, int cny, const int* done,
                       const double* po, double* pn, const double* div, float* rout, double* r64,
                       double scale, double* partials) {
  dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY);
  k_prolong_smooth_dc<TX, TY, W><<<g, W, 0, s>>>(xin, b, info, nx, ny, e, cnx, cny, done, po, pn,
                                                 div, rout, r64, scale, partials);
  return (int)(g.x * g.y);
}

int prolong_smooth_dc_tiles(int shape, int nx, int ny) {
  const TileShape& t = kTileShapes[shape];
  return ((nx + t.tx - 1) / t.tx) * ((ny + t.ty - 1) / t.ty);
}

int launch_prolong_smooth_dc(hipStream_t s, int shape, const float* xin, const float* b,
                             const uint8_t* info, int nx, int ny, const float* e, int cnx,
                             int cny, const int* done, const double* po, double* pn,
                             const double* div, float* rout, double* r64, double scale,
                             double* partials) {
  switch (shape) {
#define FH_PSDC_CASE(k, tx, ty, w)                                                             \
  case k:                                                                                      \
    return psdc_launch<tx, ty, w>(s, xin, b, info, nx, ny, e, cnx, cny, done, po, pn, div, rout, \
                                  r64, scale, partials);
    FH_TILE_SHAPES(FH_PSDC_CASE)
#undef FH_PSDC_CASE
  }
  return 0;
}

void launch_smooth_restrict(hipStream_t s, int shape, const float* xin, float* xout,
                            const float* b, const uint8_t* info, int nx, int ny, float* bc, int cnx,
                            int cny, int zero, const int* done) {
  switch (shape) {
#define FH_SR_CASE(k, tx, ty, w) \
  case k: sr_launch<tx, ty, w>(s, xin, xout, b, info, nx, ny, bc, cnx, cny, zero, done); break;
    FH_TILE_SHAPES(FH_SR_CASE)
#undef FH_SR_CASE
  }
}

void launch_prolong_smooth(hipStream_t s, int shape, const float* xin, float* xout,
                           const float* b, const uint8_t* info, int nx, int ny, const float* e,
                           int cnx, int cny, const int* done) {
  switch (shape) {
#define FH_PS_CASE(k, tx, ty, w) \
  case k: ps_launch<tx, ty, w>(s, xin, xout, b, info, nx, ny, e, cnx, cny, done); break;
    FH_TILE_SHAPES(FH_PS_CASE)
#undef FH_PS_CASE
  }
}

void launch_coarse_vcycle(hipStream_t s, int group, const CoarseLevels& L, const float* b0,
                          const float* x0in, float* x0out, const double* M, const int* cells,
                          int ncoarse, int direct, const int* done) {
  size_t lds = coarse_vcycle_lds_bytes(L);
  switch (group) {
    case 1:
      k_coarse_vcycle<512><<<1, 512, lds, s>>>(L, b0, x0in, x0out, M, cells, ncoarse, direct, done);
      break;
    case 2:
      k_coarse_vcycle<256><<<1, 256, lds, s>>>(L, b0, x0in, x0out, M, cells, ncoarse, direct, done);
      break;
    default:
      k_coarse_vcycle<1024><<<1, 1024, lds, s>>>(L, b0, x0in, x0out, M, cells, ncoarse, direct,
                                                 done);
      break;
  }
}

// ---- batched launches ----------------------------------------------------------
// Level l of every scene that tiles it (l < its lc; the work list holds only
// those), then each scene's sub-cycle as one workgroup per scene: the same
// tile bodies on each scene's level buffers, so the same bits as its own
// V-cycle.

template <int TX, int TY, int W>
__global__ __launch_bounds__(W) void kb_smooth_restrict(const SceneDesc* __restrict__ D,
                                                        BatchGrid g, int l) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const LevelDesc& L = d.lv[l];
  const LevelDesc& N = d.lv[l + 1];
  smooth_restrict_tile<TX, TY, W>(nullptr, gm(L.xa), gm(L.b), gm(L.info), L.nx, L.ny, gm(N.b),
                                  N.nx, N.ny, 1, gm(d.done), bx, by);
}

template <int TX, int TY, int W>
__global__ __launch_bounds__(W) void kb_prolong_smooth(const SceneDesc* __restrict__ D,
                                                       BatchGrid g, int l) {
  int s, bx, by, gy, bid;
  if (!batch_block(g, &s, &bx, &by, &gy, &bid)) return;
  const SceneDesc& d = D[s];
  const LevelDesc& L = d.lv[l];
  const LevelDesc& N = d.lv[l + 1];
  prolong_smooth_tile<TX, TY, W>(gm(L.xa), gm(L.x), gm(L.b), gm(L.info), L.nx, L.ny, gm(N.x),
                                 N.nx, N.ny, gm(d.done), bx, by);
}

__global__ __launch_bounds__(NC) void kb_coarse_vcycle(const SceneDesc* __restrict__ D) {
  const SceneDesc& d = D[blockIdx.x];
  if (*gm(d.done)) return;
  extern __shared__ unsigned char smem[];
  coarse_vcycle_body<NC>(d.cl, gm(d.lv[d.lc].b), nullptr, gm(d.lv[d.lc].x), gm(d.M), gm(d.cells),
                         d.ncoarse, 1, smem);
}

template <int TX, int TY, int W>
static void bsr(hipStream_t s, const SceneDesc* D, const BatchGrid& g, int l) {
  kb_smooth_restrict<TX, TY, W><<<g.blocks, W, 0, s>>>(D, g, l);
}

template <int TX, int TY, int W>
static void bps(hipStream_t s, const SceneDesc* D, const BatchGrid& g, int l) {
  kb_prolong_smooth<TX, TY, W><<<g.blocks, W, 0, s>>>(D, g, l);
}

void launch_batch_smooth_restrict(hipStream_t s, int shape, const SceneDesc* D,
                                  const BatchGrid& g, int l) {
  switch (shape) {
#define FH_BSR_CASE(k, tx, ty, w) \
  case k: bsr<tx, ty, w>(s, D, g, l); break;
    FH_TILE_SHAPES(FH_BSR_CASE)
#undef FH_BSR_CASE
  }
}

void launch_batch_prolong_smooth(hipStream_t s, int shape, const SceneDesc* D, const BatchGrid& g,
                                 int l) {
  switch (shape) {
#define FH_BPS_CASE(k, tx, ty, w) \
  case k: bps<tx, ty, w>(s, D, g, l); break;
    FH_TILE_SHAPES(FH_BPS_CASE)
#undef FH_BPS_CASE
  }
}

void launch_batch_coarse_vcycle(hipStream_t s, const SceneDesc* D, int nscenes, size_t lds) {
  kb_coarse_vcycle<<<nscenes, NC, lds, s>>>(D);
}

}  // namespace fh
