// Kernels of the learned warm start: a 4-level convolutional U-Net with
// BF16 weights and activations, FP32 accumulation on MFMA.
//
// Tensors are channels-last, [N][H][W][C] with C a multiple of 8, so the 8
// consecutive K values one MFMA lane needs (8 channels of one 3x3 tap) are
// one 16-byte load.  A 3x3 "same" convolution is an implicit GEMM
//   D[co][px] = sum_k W[co][k] * X[px + tap(k)][ci(k)],  k = tap*Cin + ci,
// with M = output channels, N = pixels, K = 9*Cin padded to 32.  One
// V_MFMA_F32_16X16X32_BF16 multiplies a 16x32 weight block by a 32x16 patch
// block; the D layout (4 consecutive channels of one pixel per lane) makes
// the epilogue store 8 contiguous bytes per lane.
#include <algorithm>
#include <atomic>

#include "fh_device.hpp"
#include "fh_launch.hpp"
#include "fh_unet.hpp"

namespace fh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 load8(const bf16_t* p) {
  uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// Round-to-nearest-even float -> bf16 (finite inputs).
__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---- conv3x3 forward --------------------------------------------------------
// Workgroup = 4 waves; wave w computes row y = blockIdx.y*4 + w, pixels
// x0 .. x0+63 (four 16-pixel MFMA tiles), all COUT channels.
template <int COUT>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16_t* __restrict__ X, int cin,
                                                 const bf16_t* __restrict__ Wt, int kpad,
                                                 const float* __restrict__ bias, int N, int H,
                                                 int Wd, int relu, bf16_t* __restrict__ Y,
                                                 float* __restrict__ Yf) {
  constexpr int CT = COUT / 16;  // channel tiles
  constexpr int PT = 4;          // pixel tiles per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = blockIdx.z;
  const int y = blockIdx.y * 4 + wave;
  const int x0 = blockIdx.x * 64;
  if (y >= H) return;
  const int j = lane & 15, g = lane >> 4;
  f32x4 acc[PT][CT];
#pragma unroll
  for (int p = 0; p < PT; p++)
#pragma unroll
    for (int t = 0; t < CT; t++) acc[p][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kreal = 9 * cin;
  const size_t plane = (size_t)H * Wd;
  const bf16_t* Xn = X + (size_t)n * plane * cin;
  for (int k0 = 0; k0 < kpad; k0 += 32) {
    const int k = k0 + 8 * g;
    bf16x8 b[PT];
    if (k < kreal) {
      const int tap = k / cin, ci = k - tap * cin;
      const int yy = y + tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
      for (int p = 0; p < PT; p++) {
        int xx = x0 + 16 * p + j + dx;
        b[p] = (yy >= 0 && yy < H && xx >= 0 && xx < Wd)
                   ? load8(Xn + ((size_t)yy * Wd + xx) * cin + ci)
                   : zero8();
      }
    } else {
#pragma unroll
      for (int p = 0; p < PT; p++) b[p] = zero8();
    }
#pragma unroll
    for (int t = 0; t < CT; t++) {
      bf16x8 a = load8(Wt + (size_t)(16 * t + j) * kpad + k);
#pragma unroll
      for (int p = 0; p < PT; p++)
        acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[p], acc[p][t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int p = 0; p < PT; p++) {
    int x = x0 + 16 * p + j;
    if (x >= Wd) continue;
    size_t pix = (size_t)n * plane + (size_t)y * Wd + x;
#pragma unroll
    for (int t = 0; t < CT; t++) {
      int co = 16 * t + 4 * g;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        v[r] = acc[p][t][r] + (bias ? bias[co + r] : 0.f);
        if (relu) v[r] = fmaxf(v[r], 0.f);
      }
      if (Yf) {
        if (co == 0) Yf[pix] = v[0];
      } else {
        uint2 packed;
        packed.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        packed.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *reinterpret_cast<uint2*>(Y + pix * COUT + co) = packed;
      }
    }
  }
}

// The same convolution with the workgroup's input staged in LDS first: its
// four output rows need input rows y0-1 .. y0+4 and columns x0-1 .. x0+64
// (zero outside the image), all cin channels, [row][col][cin].  Each input
// pixel is then fetched from memory once per workgroup instead of once per
// tap; the MFMA operands and their order are those of k_conv3x3, so the
// same bits.  Dynamic LDS: kConvRows x kConvCols x cin bf16.
constexpr int kConvRows = 6, kConvCols = 66;

template <int COUT>
__global__ __launch_bounds__(256) void k_conv3x3_lds(const bf16_t* __restrict__ X, int cin,
                                                     const bf16_t* __restrict__ Wt, int kpad,
                                                     const float* __restrict__ bias, int N,
                                                     int H, int Wd, int relu,
                                                     bf16_t* __restrict__ Y,
                                                     float* __restrict__ Yf) {
  extern __shared__ unsigned char conv_smem[];
  bf16_t* sX = reinterpret_cast<bf16_t*>(conv_smem);
  constexpr int CT = COUT / 16;  // channel tiles
  constexpr int PT = 4;          // pixel tiles per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = blockIdx.z;
  const int y0 = blockIdx.y * 4, y = y0 + wave;
  const int x0 = blockIdx.x * 64;
  const size_t plane = (size_t)H * Wd;
  const bf16_t* Xn = X + (size_t)n * plane * cin;
  {
    const int c8 = cin >> 3, chunks = kConvRows * kConvCols * c8;
    for (int q = threadIdx.x; q < chunks; q += 256) {
      const int pix = q / c8, part = q - pix * c8;
      const int r = pix / kConvCols, c = pix - r * kConvCols;
      const int yy = y0 - 1 + r, xx = x0 - 1 + c;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (yy >= 0 && yy < H && xx >= 0 && xx < Wd)
        v = reinterpret_cast<const uint4*>(Xn + ((size_t)yy * Wd + xx) * cin)[part];
      reinterpret_cast<uint4*>(sX)[q] = v;
    }
  }
  __syncthreads();
  if (y >= H) return;
  const int j = lane & 15, g = lane >> 4;
  f32x4 acc[PT][CT];
#pragma unroll
  for (int p = 0; p < PT; p++)
#pragma unroll
    for (int t = 0; t < CT; t++) acc[p][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kreal = 9 * cin;
  for (int k0 = 0; k0 < kpad; k0 += 32) {
    const int k = k0 + 8 * g;
    bf16x8 b[PT];
    if (k < kreal) {
      const int tap = k / cin, ci = k - tap * cin;
      const int r = wave + tap / 3, dx = tap % 3 - 1;  // staged row of input row y + tap/3 - 1
#pragma unroll
      for (int p = 0; p < PT; p++) {
        const int c = 16 * p + j + dx + 1;  // staged column of x0 + 16p + j + dx
        b[p] = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const uint4*>(sX + (r * kConvCols + c) * cin + ci));
      }
    } else {
#pragma unroll
      for (int p = 0; p < PT; p++) b[p] = zero8();
    }
#pragma unroll
    for (int t = 0; t < CT; t++) {
      bf16x8 a = load8(Wt + (size_t)(16 * t + j) * kpad + k);
#pragma unroll
      for (int p = 0; p < PT; p++)
        acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[p], acc[p][t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int p = 0; p < PT; p++) {
    int x = x0 + 16 * p + j;
    if (x >= Wd) continue;
    size_t pix = (size_t)n * plane + (size_t)y * Wd + x;
#pragma unroll
    for (int t = 0; t < CT; t++) {
      int co = 16 * t + 4 * g;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        v[r] = acc[p][t][r] + (bias ? bias[co + r] : 0.f);
        if (relu) v[r] = fmaxf(v[r], 0.f);
      }
      if (Yf) {
        if (co == 0) Yf[pix] = v[0];
      } else {
        uint2 packed;
        packed.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        packed.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *reinterpret_cast<uint2*>(Y + pix * COUT + co) = packed;
      }
    }
  }
}

// The same convolution for images too small to fill the chip (the learned
// warm start's coarse grid and its deeper levels): a workgroup takes one
// output row of up to 16 * PT pixels and all COUT channels, and its four
// waves split the reduction over K (wave w takes the 32-wide k-steps w,
// w + 4, ...), each with its own loads in flight; waves 1..3 leave their
// partial sums in LDS and wave 0 adds them in wave order, then applies the
// bias and activation.  Four times the workgroups of k_conv3x3 and a quarter
// of each wave's chain of dependent weight loads, which on a grid of a few
// dozen workgroups (one or two per XCD, every weight load an L2 miss) is
// what bounds it.  Same operands as k_conv3x3; the partial sums are added in
// another order (FP32).  Dynamic LDS: 3 x PT x COUT/16 x 64 f32x4.
template <int COUT, int PT>
__global__ __launch_bounds__(256) void k_conv3x3_ksplit(const bf16_t* __restrict__ X, int cin,
                                                        const bf16_t* __restrict__ Wt, int kpad,
                                                        const float* __restrict__ bias, int N,
                                                        int H, int Wd, int relu,
                                                        bf16_t* __restrict__ Y,
                                                        float* __restrict__ Yf) {
  extern __shared__ unsigned char ks_smem[];
  f32x4* red = reinterpret_cast<f32x4*>(ks_smem);
  constexpr int CT = COUT / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = blockIdx.z, y = blockIdx.y, x0 = blockIdx.x * 16 * PT;
  const int j = lane & 15, g = lane >> 4;
  f32x4 acc[PT][CT];
#pragma unroll
  for (int p = 0; p < PT; p++)
#pragma unroll
    for (int t = 0; t < CT; t++) acc[p][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kreal = 9 * cin;
  const size_t plane = (size_t)H * Wd;
  const bf16_t* Xn = X + (size_t)n * plane * cin;
  // two k-steps per trip, every load of both issued before their MFMAs
  for (int k0 = 32 * wave; k0 < kpad; k0 += 256) {
    bf16x8 a[2][CT], b[2][PT];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = k0 + 128 * h + 8 * g;
      const bool live = k0 + 128 * h < kpad;
      if (live && k < kreal) {
        const int tap = k / cin, ci = k - tap * cin;
        const int yy = y + tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int p = 0; p < PT; p++) {
          const int xx = x0 + 16 * p + j + dx;
          b[h][p] = (yy >= 0 && yy < H && xx >= 0 && xx < Wd)
                        ? load8(Xn + ((size_t)yy * Wd + xx) * cin + ci)
                        : zero8();
        }
      } else {
#pragma unroll
        for (int p = 0; p < PT; p++) b[h][p] = zero8();
      }
#pragma unroll
      for (int t = 0; t < CT; t++)
        a[h][t] = live ? load8(Wt + (size_t)(16 * t + j) * kpad + k) : zero8();
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int t = 0; t < CT; t++)
#pragma unroll
        for (int p = 0; p < PT; p++)
          acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[h][t], b[h][p], acc[p][t], 0, 0, 0);
  }
  if (wave > 0) {
#pragma unroll
    for (int p = 0; p < PT; p++)
#pragma unroll
      for (int t = 0; t < CT; t++) red[(((wave - 1) * PT + p) * CT + t) * 64 + lane] = acc[p][t];
  }
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int w = 0; w < 3; w++)
#pragma unroll
    for (int p = 0; p < PT; p++)
#pragma unroll
      for (int t = 0; t < CT; t++) acc[p][t] += red[((w * PT + p) * CT + t) * 64 + lane];
#pragma unroll
  for (int p = 0; p < PT; p++) {
    const int x = x0 + 16 * p + j;
    if (x >= Wd) continue;
    const size_t pix = (size_t)n * plane + (size_t)y * Wd + x;
#pragma unroll
    for (int t = 0; t < CT; t++) {
      const int co = 16 * t + 4 * g;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        v[r] = acc[p][t][r] + (bias ? bias[co + r] : 0.f);
        if (relu) v[r] = fmaxf(v[r], 0.f);
      }
      if (Yf) {
        if (co == 0) Yf[pix] = v[0];
      } else {
        uint2 packed;
        packed.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        packed.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *reinterpret_cast<uint2*>(Y + pix * COUT + co) = packed;
      }
    }
  }
}

// Workgroups of k_conv3x3 below which the small-image variant runs.
constexpr int kConvSmallGrid = 512;

template <int COUT>
static bool launch_conv_small(hipStream_t s, const bf16_t* X, int cin, const bf16_t* Wt, int kpad,
                              const float* bias, int N, int H, int W, int relu, bf16_t* Y,
                              float* Yf) {
  constexpr int CT = COUT / 16;
  // pixel tiles per row: as many as the row needs, at most 16 accumulator
  // tiles per lane (PT * CT)
  constexpr int kMaxPT = 16 / CT >= 4 ? 4 : 16 / CT >= 2 ? 2 : 1;
  const int PT = std::min(W <= 16 ? 1 : W <= 32 ? 2 : 4, kMaxPT);
  const dim3 g((W + 16 * PT - 1) / (16 * PT), H, N);
  const size_t lds = (size_t)3 * PT * CT * 64 * sizeof(f32x4);
  if (lds > (size_t)device_max_lds()) return false;
  if (PT == 1) {
    k_conv3x3_ksplit<COUT, 1><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf);
  } else if (PT == 2) {
    if constexpr (kMaxPT >= 2)
      k_conv3x3_ksplit<COUT, 2><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y,
                                                    Yf);
  } else {
    if constexpr (kMaxPT >= 4)
      k_conv3x3_ksplit<COUT, 4><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y,
                                                    Yf);
  }
  return true;
}

void launch_conv3x3(hipStream_t s, const bf16_t* X, int cin, const bf16_t* Wt, int kpad,
                    const float* bias, int cout, int N, int H, int W, int relu, bf16_t* Y,
                    float* Yf) {
  dim3 g((W + 63) / 64, (H + 3) / 4, N);
  if ((long)g.x * g.y * g.z < kConvSmallGrid && cin % 8 == 0 &&
      (((uintptr_t)X) & 15) == 0) {
    bool done = false;
    switch (cout) {
      case 16: done = launch_conv_small<16>(s, X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
      case 32: done = launch_conv_small<32>(s, X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
      case 48: done = launch_conv_small<48>(s, X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
      case 64: done = launch_conv_small<64>(s, X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
      case 96: done = launch_conv_small<96>(s, X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
      case 128: done = launch_conv_small<128>(s, X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
      default: break;
    }
    if (done) return;
  }
  const int max_lds = device_max_lds();
  const size_t lds = (size_t)kConvRows * kConvCols * cin * sizeof(bf16_t);
  if (cin % 8 == 0 && lds <= (size_t)max_lds && (((uintptr_t)X) & 15) == 0) {
    switch (cout) {
      case 16: k_conv3x3_lds<16><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); return;
      case 32: k_conv3x3_lds<32><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); return;
      case 48: k_conv3x3_lds<48><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); return;
      case 64: k_conv3x3_lds<64><<<g, 256, lds, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); return;
      default: break;
    }
  }
  switch (cout) {
    case 16: k_conv3x3<16><<<g, 256, 0, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
    case 32: k_conv3x3<32><<<g, 256, 0, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
    case 48: k_conv3x3<48><<<g, 256, 0, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
    case 64: k_conv3x3<64><<<g, 256, 0, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
    case 96: k_conv3x3<96><<<g, 256, 0, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
    case 128: k_conv3x3<128><<<g, 256, 0, s>>>(X, cin, Wt, kpad, bias, N, H, W, relu, Y, Yf); break;
    default: break;
  }
}

// ---- weight-gradient ----------------------------------------------------------
// dW[co][k] = sum_px dY[px][co] * X[px + tap(k)][ci(k)],  db[co] = sum dY.
// One workgroup owns 16 channels x 64 K values and a slice of the pixels;
// thread t accumulates 4 consecutive k of one channel in FP32 and the slice
// partials are added into dW with atomics (FP32).
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const bf16_t* __restrict__ X, int cin,
                                                       const bf16_t* __restrict__ dY, int cout,
                                                       int N, int H, int Wd, int kstride,
                                                       float* __restrict__ dW,
                                                       float* __restrict__ db, int slices) {
  const int t = threadIdx.x;
  const int co = blockIdx.y * 16 + (t >> 4);
  const int k = blockIdx.x * 64 + (t & 15) * 4;
  const int kreal = 9 * cin;
  const long total = (long)N * H * Wd;
  const long per = (total + slices - 1) / slices;
  const long p0 = blockIdx.z * per, p1 = min(total, p0 + per);
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, ab = 0.f;
  const bool kin = k < kreal && co < cout;
  int tap = 0, ci = 0, dy = 0, dx = 0;
  if (kin) {
    tap = k / cin;
    ci = k - tap * cin;
    dy = tap / 3 - 1;
    dx = tap % 3 - 1;
  }
  const bool dob = (k == 0) && co < cout;
  if (kin || dob) {
    for (long p = p0; p < p1; p++) {
      int x = (int)(p % Wd);
      long q = p / Wd;
      int y = (int)(q % H);
      int n = (int)(q / H);
      float gy = bf2f(dY[p * cout + co]);
      if (dob) ab += gy;
      if (!kin) continue;
      int yy = y + dy, xx = x + dx;
      if (yy < 0 || yy >= H || xx < 0 || xx >= Wd) continue;
      const bf16_t* xp = X + (((size_t)n * H + yy) * Wd + xx) * cin + ci;
      uint2 v = *reinterpret_cast<const uint2*>(xp);
      a0 += gy * __uint_as_float(v.x << 16);
      a1 += gy * __uint_as_float(v.x & 0xffff0000u);
      a2 += gy * __uint_as_float(v.y << 16);
      a3 += gy * __uint_as_float(v.y & 0xffff0000u);
    }
  }
  if (kin) {
    float* w = dW + (size_t)co * kstride + k;
    atomicAdd(w + 0, a0);
    atomicAdd(w + 1, a1);
    atomicAdd(w + 2, a2);
    atomicAdd(w + 3, a3);
  }
  if (dob) atomicAdd(db + co, ab);
}

void launch_conv3x3_wgrad(hipStream_t s, const bf16_t* X, int cin, const bf16_t* dY, int cout,
                          int N, int H, int W, int kstride, float* dW, float* db) {
  long total = (long)N * H * W;
  int slices = (int)min(256L, max(1L, total / 2048));
  dim3 g((9 * cin + 63) / 64, (cout + 15) / 16, slices);
  k_conv3x3_wgrad<<<g, 256, 0, s>>>(X, cin, dY, cout, N, H, W, kstride, dW, db, slices);
}

// ---- elementwise ----------------------------------------------------------------

// 2x2 average pool, [N][H][W][C] -> [N][H/2][W/2][C]
__global__ void k_pool2(const bf16_t* __restrict__ X, bf16_t* __restrict__ Y, int N, int H, int W,
                        int C) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int h = H / 2, w = W / 2;
  long total = (long)N * h * w * C;
  if (i >= total) return;
  int c = (int)(i % C);
  long q = i / C;
  int x = (int)(q % w);
  q /= w;
  int y = (int)(q % h);
  int n = (int)(q / h);
  const bf16_t* b = X + (((size_t)n * H + 2 * y) * W + 2 * x) * C + c;
  float s = (bf2f(b[0]) + bf2f(b[C])) + (bf2f(b[(size_t)W * C]) + bf2f(b[(size_t)W * C + C]));
  Y[i] = f2bf(0.25f * s);
}

// d(pool): dX (+)= dY/4 broadcast over each 2x2 block.
__global__ void k_pool2_bwd(const bf16_t* __restrict__ dY, bf16_t* __restrict__ dX, int N, int H,
                            int W, int C, int accumulate) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long total = (long)N * H * W * C;
  if (i >= total) return;
  int c = (int)(i % C);
  long q = i / C;
  int x = (int)(q % W);
  q /= W;
  int y = (int)(q % H);
  int n = (int)(q / H);
  int h = H / 2, w = W / 2;
  float d = 0.25f * bf2f(dY[(((size_t)n * h + y / 2) * w + x / 2) * C + c]);
  if (accumulate) d += bf2f(dX[i]);
  dX[i] = f2bf(d);
}

// out[N][H][W][Ca+Cb] = concat(upsample2(a [N][H/2][W/2][Ca]), b [N][H][W][Cb])
__global__ void k_up_concat(const bf16_t* __restrict__ a, int Ca, const bf16_t* __restrict__ b,
                            int Cb, bf16_t* __restrict__ out, int N, int H, int W) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int C = Ca + Cb;
  long total = (long)N * H * W * C;
  if (i >= total) return;
  int c = (int)(i % C);
  long pix = i / C;
  int x = (int)(pix % W);
  long q = pix / W;
  int y = (int)(q % H);
  int n = (int)(q / H);
  if (c < Ca)
    out[i] = a[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * Ca + c];
  else
    out[i] = b[pix * Cb + (c - Ca)];
}

// The same with 8 channels (16 bytes) per thread, for Ca and Cb multiples of
// 8 and 16-byte aligned tensors: blockIdx.y is the output row (n, y), the
// threads of a row take (x, chunk) pairs.  Pure copies, so the same bits.
__global__ void k_up_concat8(const bf16_t* __restrict__ a, int Ca, const bf16_t* __restrict__ b,
                             int Cb, bf16_t* __restrict__ out, int H, int W) {
  const int C8 = (Ca + Cb) >> 3, Ca8 = Ca >> 3;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * C8) return;
  const int row = blockIdx.y, n = row / H, y = row - n * H;
  const int x = t / C8, c8 = t - x * C8;
  const uint4* src;
  if (c8 < Ca8)
    src = reinterpret_cast<const uint4*>(
              a + (((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * Ca) + c8;
  else
    src = reinterpret_cast<const uint4*>(b + ((size_t)row * W + x) * Cb) + (c8 - Ca8);
  reinterpret_cast<uint4*>(out + ((size_t)row * W + x) * (Ca + Cb))[c8] = *src;
}

// backward of k_up_concat: da = sum over each 2x2 block, db = slice.
__global__ void k_up_concat_bwd_a(const bf16_t* __restrict__ dout, int Ca, int Cb,
                                  bf16_t* __restrict__ da, int N, int H, int W) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int h = H / 2, w = W / 2, C = Ca + Cb;
  long total = (long)N * h * w * Ca;
  if (i >= total) return;
  int c = (int)(i % Ca);
  long q = i / Ca;
  int x = (int)(q % w);
  q /= w;
  int y = (int)(q % h);
  int n = (int)(q / h);
  const bf16_t* d = dout + (((size_t)n * H + 2 * y) * W + 2 * x) * C + c;
  float s = (bf2f(d[0]) + bf2f(d[C])) + (bf2f(d[(size_t)W * C]) + bf2f(d[(size_t)W * C + C]));
  da[i] = f2bf(s);
}

// db (the skip branch) accumulates: skip tensors also feed the next level's
// pool, so their gradient is the sum of both uses.
__global__ void k_up_concat_bwd_b(const bf16_t* __restrict__ dout, int Ca, int Cb,
                                  bf16_t* __restrict__ db, int N, int H, int W, int accumulate) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long total = (long)N * H * W * Cb;
  if (i >= total) return;
  int c = (int)(i % Cb);
  long pix = i / Cb;
  float v = bf2f(dout[pix * (Ca + Cb) + Ca + c]);
  if (accumulate) v += bf2f(db[i]);
  db[i] = f2bf(v);
}

// dY *= (Y > 0)
__global__ void k_relu_bwd(bf16_t* __restrict__ dY, const bf16_t* __restrict__ Y, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!(bf2f(Y[i]) > 0.f)) dY[i] = 0;
}

// Weight layout transforms.  Forward layout: Wt[co][kpad], k = tap*cin+ci.
// Backward-data layout: Wb[ci][kpadT], k' = tap'*cout+co, tap' = 8 - tap.
__global__ void k_pack_weights(const float* __restrict__ w, int cin, int cout, int kpad,
                               bf16_t* __restrict__ Wt, int kpadT, bf16_t* __restrict__ Wb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int total = cout * 9 * cin;
  if (i >= total) return;
  int ci = i % cin;
  int tap = (i / cin) % 9;
  int co = i / (9 * cin);
  bf16_t v = f2bf(w[i]);  // master layout: [co][tap][ci]
  Wt[(size_t)co * kpad + tap * cin + ci] = v;
  if (Wb) Wb[(size_t)ci * kpadT + (8 - tap) * cout + co] = v;
}

// Adam (Kingma & Ba 2015) on FP32 master weights with bias correction.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float grad_scale) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i] * grad_scale;
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static inline int nb(long n) { return (int)((n + 255) / 256); }

void launch_pool2(hipStream_t s, const bf16_t* X, bf16_t* Y, int N, int H, int W, int C) {
  long n = (long)N * (H / 2) * (W / 2) * C;
  k_pool2<<<nb(n), 256, 0, s>>>(X, Y, N, H, W, C);
}
void launch_pool2_bwd(hipStream_t s, const bf16_t* dY, bf16_t* dX, int N, int H, int W, int C,
                      int accumulate) {
  long n = (long)N * H * W * C;
  k_pool2_bwd<<<nb(n), 256, 0, s>>>(dY, dX, N, H, W, C, accumulate);
}
void launch_up_concat(hipStream_t s, const bf16_t* a, int Ca, const bf16_t* b, int Cb, bf16_t* out,
                      int N, int H, int W) {
  const bool vec = Ca % 8 == 0 && Cb % 8 == 0 &&
                   (((uintptr_t)a | (uintptr_t)b | (uintptr_t)out) & 15) == 0 &&
                   (long)W * ((Ca + Cb) / 8) < (1L << 30) && (long)N * H < 65536;
  if (vec) {
    const int per_row = W * ((Ca + Cb) / 8);
    k_up_concat8<<<dim3((per_row + 255) / 256, N * H), 256, 0, s>>>(a, Ca, b, Cb, out, H, W);
    return;
  }
  long n = (long)N * H * W * (Ca + Cb);
  k_up_concat<<<nb(n), 256, 0, s>>>(a, Ca, b, Cb, out, N, H, W);
}
void launch_up_concat_bwd(hipStream_t s, const bf16_t* dout, int Ca, int Cb, bf16_t* da,
                          bf16_t* db, int N, int H, int W, int accumulate_b) {
  long na = (long)N * (H / 2) * (W / 2) * Ca;
  k_up_concat_bwd_a<<<nb(na), 256, 0, s>>>(dout, Ca, Cb, da, N, H, W);
  long nbb = (long)N * H * W * Cb;
  k_up_concat_bwd_b<<<nb(nbb), 256, 0, s>>>(dout, Ca, Cb, db, N, H, W, accumulate_b);
}
void launch_relu_bwd(hipStream_t s, bf16_t* dY, const bf16_t* Y, long n) {
  k_relu_bwd<<<nb(n), 256, 0, s>>>(dY, Y, n);
}
void launch_pack_weights(hipStream_t s, const float* w, int cin, int cout, int kpad, bf16_t* Wt,
                         int kpadT, bf16_t* Wb) {
  int n = cout * 9 * cin;
  k_pack_weights<<<nb(n), 256, 0, s>>>(w, cin, cout, kpad, Wt, kpadT, Wb);
}
void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, long n, float lr,
                 float b1, float b2, float eps, float c1, float c2, float grad_scale) {
  k_adam<<<nb(n), 256, 0, s>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2, grad_scale);
}

}  // namespace fh
