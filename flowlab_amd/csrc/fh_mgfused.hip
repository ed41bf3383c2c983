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
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {

constexpr int TX = 64, TY = 32;  // tile of output cells (even, so children
                                 // of a coarse cell never straddle tiles)
constexpr int NT = 256;          // threads per tile workgroup

// info byte of a cell: bit 7 = fluid, bits 0..2 = number of non-solid
// in-domain neighbours (the equation's diagonal), 0 off the domain.
__device__ __forceinline__ uint8_t cell_info(const uint8_t* __restrict__ lbl, int nx, int ny,
                                             int i, int j) {
  if (i < 0 || j < 0 || i >= nx || j >= ny) return 0;
  if (lbl[i + j * nx] != FLUID) return 0;
  int c = !solid_or_outside(lbl, nx, ny, i + 1, j) + !solid_or_outside(lbl, nx, ny, i - 1, j) +
          !solid_or_outside(lbl, nx, ny, i, j + 1) + !solid_or_outside(lbl, nx, ny, i, j - 1);
  return (uint8_t)(0x80 | c);
}

__global__ void k_cell_info(const uint8_t* __restrict__ lbl, int nx, int ny,
                            uint8_t* __restrict__ info) {
  int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i < nx && j < ny) info[i + j * nx] = cell_info(lbl, nx, ny, i, j);
}

// (sum + b) / count, with count in 1..4 (1/3 rounded: within an ulp of the
// division).
__device__ __forceinline__ float gs_update(float sum, float b, int cnt) {
  float r = cnt == 1 ? 1.f : cnt == 2 ? 0.5f : cnt == 3 ? (1.f / 3.f) : 0.25f;
  return (sum + b) * r;
}

// One colour half-sweep with omega = 1 on an LDS region of LX x LY cells,
// updating cells of the colour whose LDS distance to the region border is
// >= ring.  par0 = parity of LDS (0,0) in global coordinates.  zero: the
// unknown is zero on entry (neighbour sum = 0).  LX is a template constant so
// the pair index splits into (row, pair) with a multiply-shift.
template <int LX, int LY>
__device__ __forceinline__ void lds_half_sweep(float* sx, const float* sb, const uint8_t* si,
                                               int ring, int color, int par0, bool zero, int tid,
                                               int nthreads) {
  constexpr int P = (LX + 1) / 2;  // pairs per row
  const int r0 = ring, r1 = LY - 1 - ring;
  const int total = P * (r1 - r0 + 1);
  for (int q = tid; q < total; q += nthreads) {
    int lj = r0 + q / P;
    int pc = q - (lj - r0) * P;
    int li = 2 * pc + ((lj + par0 + color) & 1);
    if (li < ring || li > LX - 1 - ring) continue;
    int l = li + lj * LX;
    int cnt = si[l] & 7;
    if (cnt == 0) continue;
    float sum = zero ? 0.f : (((sx[l + 1] + sx[l - 1]) + sx[l + LX]) + sx[l - LX]);
    sx[l] = gs_update(sum, sb[l], cnt);
  }
}

// Runtime-width variant for the one-workgroup coarse levels.
__device__ __forceinline__ void lds_half_sweep_rt(float* sx, const float* sb, const uint8_t* si,
                                                  int LX, int LY, int color, bool zero, int tid,
                                                  int nthreads) {
  const int w = LX - 2, P = (w + 1) >> 1;
  const int total = P * (LY - 2);
  for (int q = tid; q < total; q += nthreads) {
    int r = q / P, pc = q - r * P;
    int lj = 1 + r;
    // padded (0,0) is global (-1,-1): global parity of (li, lj) = li + lj
    int li = 1 + 2 * pc + (((1 + 2 * pc + lj + color) & 1) ? 1 : 0);
    if (li > LX - 2) continue;
    int l = li + lj * LX;
    int cnt = si[l] & 7;
    if (cnt == 0) continue;
    float sum = zero ? 0.f : (((sx[l + 1] + sx[l - 1]) + sx[l + LX]) + sx[l - LX]);
    sx[l] = gs_update(sum, sb[l], cnt);
  }
}

// r = b - (count x - sum) at an LDS cell (0 off the fluid or if count = 0).
__device__ __forceinline__ float lds_residual(const float* sx, const float* sb, const uint8_t* si,
                                              int LX, int l) {
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
constexpr int HA = 5;
constexpr int AX = TX + 2 * HA, AY = TY + 2 * HA;

__global__ __launch_bounds__(NT) void k_smooth_restrict(const float* __restrict__ xin,
                                                        float* __restrict__ xout,
                                                        const float* __restrict__ b,
                                                        const uint8_t* __restrict__ info, int nx,
                                                        int ny, float* __restrict__ bc, int cnx,
                                                        int cny, int zero,
                                                        const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float sx[AX * AY];
  __shared__ float sb[AX * AY];
  __shared__ uint8_t si[AX * AY];
  const int tid = threadIdx.x;
  const int ox = blockIdx.x * TX, oy = blockIdx.y * TY;
  const int gx0 = ox - HA, gy0 = oy - HA;
  for (int q = tid; q < AX * AY; q += NT) {
    int lj = q / AX, li = q - lj * AX;
    int gi = gx0 + li, gj = gy0 + lj;
    bool in = gi >= 0 && gj >= 0 && gi < nx && gj < ny;
    int g = gi + gj * nx;
    sx[q] = (in && !zero) ? xin[g] : 0.f;
    sb[q] = in ? b[g] : 0.f;
    si[q] = in ? info[g] : 0;
  }
  __syncthreads();
  const int par0 = (gx0 + gy0) & 1;
  for (int s = 1; s <= 4; s++) {
    lds_half_sweep<AX, AY>(sx, sb, si, s, (s - 1) & 1, par0, zero && s == 1, tid, NT);
    __syncthreads();
  }
  // coarse cells of this tile: children (2I+di, 2J+dj) are tile-interior
  const int CX = TX / 2, CY = TY / 2;
  for (int q = tid; q < CX * CY; q += NT) {
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
  for (int q = tid; q < TX * TY; q += NT) {
    int lj = q / TX, li = q - lj * TX;
    int gi = ox + li, gj = oy + lj;
    if (gi < nx && gj < ny) xout[gi + gj * nx] = sx[(li + HA) + (lj + HA) * AX];
  }
}

// ---- k_prolong_smooth --------------------------------------------------------
// x += bilinear(e) on fluid cells (pressureMultigrid.ts:192-207), then 2
// post-smoothing sweeps.  Reads xin, writes xout (distinct buffers).
constexpr int HB = 4;
constexpr int BXW = TX + 2 * HB, BYH = TY + 2 * HB;
constexpr int EX = BXW / 2 + 3, EY = BYH / 2 + 3;

__global__ __launch_bounds__(NT) void k_prolong_smooth(const float* __restrict__ xin,
                                                       float* __restrict__ xout,
                                                       const float* __restrict__ b,
                                                       const uint8_t* __restrict__ info, int nx,
                                                       int ny, const float* __restrict__ e,
                                                       int cnx, int cny,
                                                       const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ float sx[BXW * BYH];
  __shared__ float sb[BXW * BYH];
  __shared__ uint8_t si[BXW * BYH];
  __shared__ float se[EX * EY];
  const int tid = threadIdx.x;
  const int ox = blockIdx.x * TX, oy = blockIdx.y * TY;
  const int gx0 = ox - HB, gy0 = oy - HB;
  // coarse window: parents I = gi>>1 and their +-1 neighbours
  const int ex0 = (gx0 >> 1) - 1, ey0 = (gy0 >> 1) - 1;
  for (int q = tid; q < EX * EY; q += NT) {
    int ej = q / EX, ei = q - ej * EX;
    int I = ex0 + ei, J = ey0 + ej;
    se[q] = (I >= 0 && J >= 0 && I < cnx && J < cny) ? e[I + J * cnx] : 0.f;
  }
  for (int q = tid; q < BXW * BYH; q += NT) {
    int lj = q / BXW, li = q - lj * BXW;
    int gi = gx0 + li, gj = gy0 + lj;
    bool in = gi >= 0 && gj >= 0 && gi < nx && gj < ny;
    int g = gi + gj * nx;
    sx[q] = in ? xin[g] : 0.f;
    sb[q] = in ? b[g] : 0.f;
    si[q] = in ? info[g] : 0;
  }
  __syncthreads();
  for (int q = tid; q < BXW * BYH; q += NT) {
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
  const int par0 = (gx0 + gy0) & 1;
  for (int s = 1; s <= 4; s++) {
    lds_half_sweep<BXW, BYH>(sx, sb, si, s, (s - 1) & 1, par0, false, tid, NT);
    __syncthreads();
  }
  for (int q = tid; q < TX * TY; q += NT) {
    int lj = q / TX, li = q - lj * TX;
    int gi = ox + li, gj = oy + lj;
    if (gi < nx && gj < ny) xout[gi + gj * nx] = sx[(li + HB) + (lj + HB) * BXW];
  }
}

// ---- k_coarse_vcycle -----------------------------------------------------------
// The whole sub-cycle from level lc down, in one workgroup's LDS.  Each
// level's arrays are padded by a one-cell ring of zeros (pitch nx+2) so the
// half-sweep's neighbour reads never leave the level.
constexpr int NC = 1024;

__device__ __forceinline__ int padded(int nx, int ny) { return (nx + 2) * (ny + 2); }

__global__ __launch_bounds__(NC) void k_coarse_vcycle(CoarseLevels L, const float* __restrict__ b0,
                                                      const float* __restrict__ x0in,
                                                      float* __restrict__ x0out,
                                                      const double* __restrict__ M,
                                                      const int* __restrict__ cells, int ncoarse,
                                                      int direct, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ unsigned char smem[];
  const int tid = threadIdx.x;
  float* X[kMaxCoarse];
  float* B[kMaxCoarse];
  uint8_t* I8[kMaxCoarse];
  {
    unsigned char* p = smem;
    for (int l = 0; l < L.n; l++) {
      int n = padded(L.nx[l], L.ny[l]);
      X[l] = (float*)p;
      p += n * sizeof(float);
      B[l] = (float*)p;
      p += n * sizeof(float);
      I8[l] = (uint8_t*)p;
      p += (n + 3) & ~3;
    }
  }
  for (int l = 0; l < L.n; l++) {
    int nx = L.nx[l], ny = L.ny[l], px = nx + 2, n = padded(nx, ny);
    for (int q = tid; q < n; q += NC) {
      int lj = q / px, li = q - lj * px;
      int gi = li - 1, gj = lj - 1;
      bool in = gi >= 0 && gj >= 0 && gi < nx && gj < ny;
      I8[l][q] = in ? L.info[l][gi + gj * nx] : 0;
      X[l][q] = (l == 0 && in && x0in) ? x0in[gi + gj * nx] : 0.f;
      B[l][q] = (l == 0 && in) ? b0[gi + gj * nx] : 0.f;
    }
  }
  __syncthreads();
  for (int l = 0; l < L.n - 1; l++) {
    int px = L.nx[l] + 2, py = L.ny[l] + 2;
    bool zero = (l > 0) || !x0in;
    for (int s = 1; s <= 4; s++) {
      lds_half_sweep_rt(X[l], B[l], I8[l], px, py, (s - 1) & 1, zero && s == 1, tid, NC);
      __syncthreads();
    }
    int cnx = L.nx[l + 1], cny = L.ny[l + 1], cpx = cnx + 2;
    for (int q = tid; q < cnx * cny; q += NC) {
      int J = q / cnx, I = q - J * cnx;
      float s = 0.f;
      for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
          int gi = 2 * I + di, gj = 2 * J + dj;
          if (gi < L.nx[l] && gj < L.ny[l])
            s += lds_residual(X[l], B[l], I8[l], px, (gi + 1) + (gj + 1) * px);
        }
      B[l + 1][(I + 1) + (J + 1) * cpx] = s;
    }
    __syncthreads();
  }
  {
    int l = L.n - 1;
    int nx = L.nx[l], ny = L.ny[l], px = nx + 2, py = ny + 2;
    bool zero = (l > 0) || !x0in;
    if (direct) {
      // x = M b in FP64, four lanes per row
      for (int t = tid; t < 4 * ncoarse; t += NC) {
        int a = t >> 2, part = t & 3;
        const double* row = M + (size_t)a * ncoarse;
        double s = 0.0;
        for (int q = part; q < ncoarse; q += 4) {
          int c = cells[q];
          s += row[q] * (double)B[l][(c % nx + 1) + (c / nx + 1) * px];
        }
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (part == 0) {
          int c = cells[a];
          X[l][(c % nx + 1) + (c / nx + 1) * px] = (float)s;
        }
      }
    } else {
      for (int s = 0; s < 16; s++) {
        lds_half_sweep_rt(X[l], B[l], I8[l], px, py, s & 1, zero && s == 0, tid, NC);
        __syncthreads();
      }
    }
    __syncthreads();
  }
  for (int l = L.n - 2; l >= 0; l--) {
    int nx = L.nx[l], ny = L.ny[l], px = nx + 2, py = ny + 2;
    int cnx = L.nx[l + 1], cny = L.ny[l + 1], cpx = cnx + 2;
    for (int q = tid; q < nx * ny; q += NC) {
      int gj = q / nx, gi = q - gj * nx;
      int lq = (gi + 1) + (gj + 1) * px;
      if (!(I8[l][lq] & 0x80)) continue;
      int I = gi >> 1, J = gj >> 1;
      int I2 = (gi & 1) ? min(I + 1, cnx - 1) : max(I - 1, 0);
      int J2 = (gj & 1) ? min(J + 1, cny - 1) : max(J - 1, 0);
      const float* E = X[l + 1];
      X[l][lq] += 0.5625f * E[(I + 1) + (J + 1) * cpx] +
                  0.1875f * (E[(I2 + 1) + (J + 1) * cpx] + E[(I + 1) + (J2 + 1) * cpx]) +
                  0.0625f * E[(I2 + 1) + (J2 + 1) * cpx];
    }
    __syncthreads();
    for (int s = 1; s <= 4; s++) {
      lds_half_sweep_rt(X[l], B[l], I8[l], px, py, (s - 1) & 1, false, tid, NC);
      __syncthreads();
    }
  }
  {
    int nx = L.nx[0], ny = L.ny[0], px = nx + 2;
    for (int q = tid; q < nx * ny; q += NC) {
      int gj = q / nx, gi = q - gj * nx;
      x0out[q] = X[0][(gi + 1) + (gj + 1) * px];
    }
  }
}

size_t coarse_vcycle_lds_bytes(const CoarseLevels& L) {
  size_t b = 0;
  for (int l = 0; l < L.n; l++) {
    size_t n = (size_t)(L.nx[l] + 2) * (L.ny[l] + 2);
    b += 8 * n + ((n + 3) & ~(size_t)3);
  }
  return b;
}

void launch_cell_info(hipStream_t s, const uint8_t* lbl, int nx, int ny, uint8_t* info) {
  dim3 g((nx + 63) / 64, (ny + 3) / 4);
  k_cell_info<<<g, dim3(64, 4), 0, s>>>(lbl, nx, ny, info);
}

void launch_smooth_restrict(hipStream_t s, const float* xin, float* xout, const float* b,
                            const uint8_t* info, int nx, int ny, float* bc, int cnx, int cny,
                            int zero, const int* done) {
  dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY);
  k_smooth_restrict<<<g, NT, 0, s>>>(xin, xout, b, info, nx, ny, bc, cnx, cny, zero, done);
}

void launch_prolong_smooth(hipStream_t s, const float* xin, float* xout, const float* b,
                           const uint8_t* info, int nx, int ny, const float* e, int cnx, int cny,
                           const int* done) {
  dim3 g((nx + TX - 1) / TX, (ny + TY - 1) / TY);
  k_prolong_smooth<<<g, NT, 0, s>>>(xin, xout, b, info, nx, ny, e, cnx, cny, done);
}

void launch_coarse_vcycle(hipStream_t s, const CoarseLevels& L, const float* b0,
                          const float* x0in, float* x0out, const double* M, const int* cells,
                          int ncoarse, int direct, const int* done) {
  size_t lds = coarse_vcycle_lds_bytes(L);
  k_coarse_vcycle<<<1, NC, lds, s>>>(L, b0, x0in, x0out, M, cells, ncoarse, direct, done);
}

}  // namespace fh
