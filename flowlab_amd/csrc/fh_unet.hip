// Learned warm start for the pressure solve: the U-Net's host side (layer
// table, parameters, forward/backward schedules, Adam), its training set
// captured from the engine's own multigrid solves, and the warm-start step
// used inside the defect-correction solve.
//
// Formulation.  The solve's start p (the previous pressure, or its
// extrapolation in time) leaves the defect r0 = b - A p.  The fine grid is
// gathered into F x F blocks (F a power of two, coarse sides at most 128):
// per block the sum of r0 (the coarse right-hand side: a smooth p with
// A p = r on the fine grid has A_c p_c = sum of r over the block on the
// coarse one), its solid and Air fractions and the fluid mean of p.  The
// network sees those four channels (the defect over its RMS s_r, the
// pressure over its RMS s_p) and predicts o, the coarse correction over
// s_r c0; the correction q = s_r c0 o is interpolated bilinearly onto the
// fine fluid cells and the start becomes p + m q with the energy-optimal
// step m >= 0 (k_warm_dots), so it never has a larger A-norm error than p.
// Training minimises the mean squared error of o against the fluid mean of
// (p* - p) over each block, over s_r c0, p* being the multigrid solution.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flowhip.h"
#include "fh_device.hpp"
#include "fh_launch.hpp"
#include "fh_unet.hpp"

namespace fh {

namespace {

__device__ __forceinline__ bf16_t f2bf_(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

// ---- the coarse grid the network runs on -------------------------------------
// Each coarse cell gathers an F x F block of fine cells as float4
// {sum of the defect r, solid cells, Air cells, sum of p over fluid cells}
// (fluid cells = F^2 - solid - Air), built by repeated 2 x 2 block sums.

// First level of the gather from the fine fields (G = 1: each fine cell as
// it is; G = 2: 2 x 2 blocks).  With p1 non-null the pressure summed is
// p1 - p (the captured correction), r unread.
template <int G>
__global__ void k_gather_first(const float* __restrict__ r, const uint8_t* __restrict__ lbl,
                               const double* __restrict__ p, const double* __restrict__ p1, int nx,
                               int cw, int ch, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= cw || j >= ch) return;
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int dy = 0; dy < G; dy++)
#pragma unroll
    for (int dx = 0; dx < G; dx++) {
      const long k = (long)(G * i + dx) + (long)(G * j + dy) * nx;
      const uint8_t l = lbl[k];
      if (!p1) a.x += r[k];
      a.y += l == SOLID ? 1.f : 0.f;
      a.z += l == AIR ? 1.f : 0.f;
      if (l == FLUID) a.w += p1 ? (float)(p1[k] - p[k]) : (float)p[k];
    }
  out[i + (long)j * cw] = a;
}

// One more 2 x 2 level: out (cw x ch) from in (2 cw x 2 ch).
__global__ void k_gather2(const float4* __restrict__ in, int cw, int ch, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= cw || j >= ch) return;
  const long w = 2L * cw, k = 2L * i + 2L * j * w;
  const float4 a = in[k], b = in[k + 1], c = in[k + w], d = in[k + w + 1];
  out[i + (long)j * cw] = make_float4((a.x + b.x) + (c.x + d.x), (a.y + b.y) + (c.y + d.y),
                                      (a.z + b.z) + (c.z + d.z), (a.w + b.w) + (c.w + d.w));
}

// The network's input scales of one coarse grid: s_r, s_p = RMS of the
// defect sum and of the fluid-mean pressure over coarse cells holding fluid
// (1 where zero).  One workgroup of 1024 threads.
// A coarse grid holds at most kUnetCoarseSide^2 = 16 x 1024 cells: each
// thread loads its (up to) 16 cells at once, then sums them.
__global__ __launch_bounds__(1024) void k_coarse_stats(const float4* __restrict__ c, int n,
                                                       float area, float* __restrict__ out2) {
  constexpr int kPer = kUnetCoarseSide * kUnetCoarseSide / 1024;
  __shared__ double sh[3][16];
  double a[3] = {0.0, 0.0, 0.0};
  float4 v[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int k = threadIdx.x + q * 1024;
    v[q] = k < n ? c[k] : make_float4(0.f, 0.f, area, 0.f);
  }
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const float fl = area - v[q].y - v[q].z;
    if (fl < 0.5f) continue;
    const double pm = (double)v[q].w / fl;
    a[0] += (double)v[q].x * v[q].x;
    a[1] += pm * pm;
    a[2] += 1.0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < 3; q++) {
    const double t = wave_sum(a[q]);
    if (lane == 0) sh[q][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 3; q++)
      for (int w = 0; w < 16; w++) t[q] += sh[q][w];
    const double sr = t[2] > 0 ? sqrt(t[0] / t[2]) : 0.0, sp = t[2] > 0 ? sqrt(t[1] / t[2]) : 0.0;
    out2[0] = (float)(sr > 0 ? sr : 1.0);
    out2[1] = (float)(sp > 0 ? sp : 1.0);
  }
}

// Normalised network input of a coarse grid: {r / s_r, solid fraction, Air
// fraction, fluid-mean p / s_p}; written as the float4 sample x (capture)
// and/or the bf16 input tensor (8 channels, 4 used).
__global__ void k_coarse_input(const float4* __restrict__ c, int n, float area,
                               const float* __restrict__ stats2, float4* __restrict__ x,
                               bf16_t* __restrict__ in) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float4 v = c[k];
  const float fl = area - v.y - v.z;
  const float inv = 1.f / area;
  float4 o;
  o.x = v.x / stats2[0];
  o.y = v.y * inv;
  o.z = v.z * inv;
  o.w = fl >= 0.5f ? (v.w / fl) / stats2[1] : 0.f;
  if (x) x[k] = o;
  if (in) {
    uint4 u;
    u.x = (uint32_t)f2bf_(o.x) | ((uint32_t)f2bf_(o.y) << 16);
    u.y = (uint32_t)f2bf_(o.z) | ((uint32_t)f2bf_(o.w) << 16);
    u.z = 0;
    u.w = 0;
    *reinterpret_cast<uint4*>(in + (long)k * kUnetIn) = u;
  }
}

// Captured target: the fluid-mean correction of each coarse cell over s_r.
__global__ void k_coarse_target(const float4* __restrict__ c, int n, float area,
                                const float* __restrict__ stats2, float* __restrict__ t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float4 v = c[k];
  const float fl = area - v.y - v.z;
  t[k] = fl >= 0.5f ? (v.w / fl) / stats2[0] : 0.f;
}

// Training batch b: the bf16 input from stored samples x, the target
// t / c0 and the mask of coarse cells holding fluid.
__global__ void k_train_prep(const float4* __restrict__ x, const float* __restrict__ t,
                             const int* __restrict__ idx, long hw, float inv_c0,
                             bf16_t* __restrict__ in, float* __restrict__ tgt,
                             float* __restrict__ mask) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= hw) return;
  const long src = (long)idx[b] * hw + i, o = (long)b * hw + i;
  const float4 v = x[src];
  uint4 u;
  u.x = (uint32_t)f2bf_(v.x) | ((uint32_t)f2bf_(v.y) << 16);
  u.y = (uint32_t)f2bf_(v.z) | ((uint32_t)f2bf_(v.w) << 16);
  u.z = 0;
  u.w = 0;
  *reinterpret_cast<uint4*>(in + o * kUnetIn) = u;
  const bool fluid = 1.f - v.y - v.z > 1e-6f;
  tgt[o] = fluid ? t[src] * inv_c0 : 0.f;
  mask[o] = fluid ? 1.f : 0.f;
}

// Masked MSE: partial sums {sum m (o-t)^2, sum m}.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ o,
                                              const float* __restrict__ t,
                                              const float* __restrict__ m, long n,
                                              double* partials) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double d = (double)o[i] - (double)t[i];
    acc[0] += m[i] * d * d;
    acc[1] += m[i];
  }
  block_sum_store<4>(acc, partials);
}

// dO = 2 m (o - t) / sum m, as the head's bf16 16-channel gradient (ch 0).
__global__ void k_loss_grad(const float* __restrict__ o, const float* __restrict__ t,
                            const float* __restrict__ m, const double* __restrict__ sums, long n,
                            bf16_t* __restrict__ dhead, float* __restrict__ dout) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float cnt = (float)sums[1];
  float d = cnt > 0 ? 2.f * m[i] * (o[i] - t[i]) / cnt : 0.f;
  if (dout) dout[i] = d;
  uint4 z = make_uint4(0, 0, 0, 0);
  uint4* h = reinterpret_cast<uint4*>(dhead + i * 16);
  h[0] = z;
  h[1] = z;
  dhead[i * 16] = f2bf_(d);
}

// FP32 per-pixel gradient -> head gradient tensor.
__global__ void k_head_grad(const float* __restrict__ dout, long n, bf16_t* __restrict__ dhead) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint4 z = make_uint4(0, 0, 0, 0);
  uint4* h = reinterpret_cast<uint4*>(dhead + i * 16);
  h[0] = z;
  h[1] = z;
  dhead[i * 16] = f2bf_(dout[i]);
}

// The network's correction on the fine grid: q = s_r c0 o at the cell's
// centre, o bilinear between coarse cell centres (clamped at the edges), on
// fluid cells; 0 elsewhere.
__global__ void k_prolong_out(const float* __restrict__ o, int cw, int ch, int F,
                              const uint8_t* __restrict__ info, int nx, int ny,
                              const float* __restrict__ stats2, float c0, float* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx || j >= ny) return;
  const long k = i + (long)j * nx;
  if (!(info[k] & 0x80)) {
    q[k] = 0.f;
    return;
  }
  const float invF = 1.f / (float)F;
  const float x = fminf(fmaxf(((float)i + 0.5f) * invF - 0.5f, 0.f), (float)(cw - 1));
  const float y = fminf(fmaxf(((float)j + 0.5f) * invF - 0.5f, 0.f), (float)(ch - 1));
  const int x0 = min((int)x, cw - 1), y0 = min((int)y, ch - 1);
  const int x1 = min(x0 + 1, cw - 1), y1 = min(y0 + 1, ch - 1);
  const float fx = x - (float)x0, fy = y - (float)y0;
  const float top = o[x0 + y0 * cw] + fx * (o[x1 + y0 * cw] - o[x0 + y0 * cw]);
  const float bot = o[x0 + y1 * cw] + fx * (o[x1 + y1 * cw] - o[x0 + y1 * cw]);
  q[k] = stats2[0] * c0 * (top + fy * (bot - top));
}

// The warm start is p + m q with the step m that minimises the energy norm
// of the error, |p* - p - m q|_A, over m >= 0: m = <r0, q> / <q, A q> (r0 = b
// - A p, so A (p* - p) = r0).  The learned start thus never has a larger
// A-norm error than the start it corrects, and the output scale drops out.
// Blocks of k_warm_dots: 16 per CU, so the pass streams at HBM rate.
constexpr int kWarmBlocks = 4096;

// k_warm_dots: block partial sums {<r0, q>, <q, A q>} over fluid cells with
// an open neighbour (info byte: bit 7 fluid, bits 3..6 open +x,-x,+y,-y,
// 0..2 count); A q = count q - sum of open fluid neighbours' q.
__global__ __launch_bounds__(256) void k_warm_dots(const float* __restrict__ r0,
                                                   const float* __restrict__ q,
                                                   const uint8_t* __restrict__ info, int nx,
                                                   long n, double* partials) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int f = info[k];
    if (!(f & 0x80) || (f & 7) == 0) continue;
    // an open neighbour is in the domain, so its index is valid
    float nb = 0.f;
    if (f & 0x08) nb += q[k + 1];
    if (f & 0x10) nb += q[k - 1];
    if (f & 0x20) nb += q[k + nx];
    if (f & 0x40) nb += q[k - nx];
    const double aq = (double)(f & 7) * (double)q[k] - (double)nb;
    acc[0] += (double)r0[k] * (double)q[k];
    acc[1] += (double)q[k] * aq;
  }
  block_sum_store<4>(acc, partials);
}

// p += m q on fluid cells, m from the finalized sums {<r0, q>, <q, A q>}.
__global__ void k_apply_warm(double* __restrict__ p, const float* __restrict__ q,
                             const double* __restrict__ sums, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = sums[1] > 0.0 ? fmax(sums[0] / sums[1], 0.0) : 0.0;
  p[i] += m * (double)q[i];
}

// Host input [N][H][W][8] f32 -> bf16
__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16_t* __restrict__ o, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) o[i] = f2bf_(a[i]);
}

inline int nblk(long n) { return (int)((n + 255) / 256); }

inline int roundup(int a, int b) { return (a + b - 1) / b * b; }

// SplitMix64 (Steele, Lea & Flood 2014) for the weight init.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

}  // namespace

// ---- network build, forward, backward -----------------------------------------

static const int kLayerSpec[kUnetLayers][3] = {
    // cin, cout, relu
    {8, 16, 1},   {16, 16, 1},  {16, 32, 1},  {32, 32, 1}, {32, 64, 1},
    {64, 64, 1},  {64, 64, 1},  {64, 64, 1},  {128, 64, 1}, {64, 64, 1},
    {96, 32, 1},  {32, 32, 1},  {48, 16, 1},  {16, 16, 1},  {16, 16, 0}};

// activation index of each layer's input and output (21 = the FP32 head out)
static const int kLayerIO[kUnetLayers][2] = {{0, 1},   {1, 2},   {3, 4},   {4, 5},   {6, 7},
                                             {7, 8},   {9, 10},  {10, 11}, {12, 13}, {13, 14},
                                             {15, 16}, {16, 17}, {18, 19}, {19, 20}, {20, 21}};
// activation tensors: level, channels
static const int kActSpec[21][2] = {{0, 8},  {0, 16}, {0, 16}, {1, 16}, {1, 32}, {1, 32}, {2, 32},
                                    {2, 64}, {2, 64}, {3, 64}, {3, 64}, {3, 64}, {2, 128}, {2, 64},
                                    {2, 64}, {1, 96}, {1, 32}, {1, 32}, {0, 48}, {0, 16}, {0, 16}};

size_t unet_param_count() {
  size_t n = 0;
  for (int l = 0; l < kUnetLayers; l++) n += (size_t)kLayerSpec[l][1] * (9 * kLayerSpec[l][0] + 1);
  return n;
}

hipError_t unet_init(UNet& net, uint64_t seed, hipStream_t s) {
  net.L.clear();
  size_t off = 0, wt = 0, wb = 0;
  for (int l = 0; l < kUnetLayers; l++) {
    ConvLayer c;
    c.cin = kLayerSpec[l][0];
    c.cout = kLayerSpec[l][1];
    c.relu = kLayerSpec[l][2];
    c.kpad = roundup(9 * c.cin, 32);
    c.kpadT = roundup(9 * c.cout, 32);
    c.woff = off;
    off += (size_t)c.cout * 9 * c.cin;
    c.boff = off;
    off += c.cout;
    c.wtoff = wt;
    wt += (size_t)c.cout * c.kpad;
    c.wboff = wb;
    wb += (size_t)c.cin * c.kpadT;
    net.L.push_back(c);
  }
  net.nparam = off;
  net.nwt = wt;
  net.nwb = wb;
  hipError_t e;
  if ((e = hipMalloc((void**)&net.w, off * 4)) || (e = hipMalloc((void**)&net.g, off * 4)) ||
      (e = hipMalloc((void**)&net.m, off * 4)) || (e = hipMalloc((void**)&net.v, off * 4)) ||
      (e = hipMalloc((void**)&net.wt, wt * 2)) || (e = hipMalloc((void**)&net.wb, wb * 2)))
    return e;
  if ((e = hipMemset(net.wt, 0, wt * 2)) || (e = hipMemset(net.wb, 0, wb * 2)) ||
      (e = hipMemset(net.m, 0, off * 4)) || (e = hipMemset(net.v, 0, off * 4)) ||
      (e = hipMemset(net.g, 0, off * 4)))
    return e;
  std::vector<float> h(off, 0.f);
  SplitMix rng{seed};
  for (int l = 0; l < kUnetLayers; l++) {
    const ConvLayer& c = net.L[l];
    double std_ = std::sqrt(2.0 / (9.0 * c.cin));
    int rows = c.cout;
    if (l == HEAD) {
      std_ *= 0.1;
      rows = 1;  // channels 1..15 of the head stay zero
    }
    for (int co = 0; co < rows; co++)
      for (int k = 0; k < 9 * c.cin; k++) h[c.woff + (size_t)co * 9 * c.cin + k] = (float)(std_ * rng.normal());
  }
  if ((e = hipMemcpy(net.w, h.data(), off * 4, hipMemcpyHostToDevice))) return e;
  net.step = 0;
  net.packed = false;
  return hipSuccess;
}

void unet_free(UNet& net) {
  for (void* p : {(void*)net.w, (void*)net.g, (void*)net.m, (void*)net.v, (void*)net.wt,
                  (void*)net.wb, (void*)net.out, (void*)net.dout, (void*)net.dhead})
    if (p) (void)hipFree(p);
  for (auto p : net.act)
    if (p) (void)hipFree(p);
  for (auto p : net.grad)
    if (p) (void)hipFree(p);
  net.act.clear();
  net.grad.clear();
}

void unet_pack(UNet& net, hipStream_t s) {
  for (auto& c : net.L)
    launch_pack_weights(s, net.w + c.woff, c.cin, c.cout, c.kpad, net.wt + c.wtoff, c.kpadT,
                        net.wb + c.wboff);
  net.packed = true;
}

// Activations for (N, H, W); with train, also their gradients and the head's
// gradient buffers (inference never touches them: at 8192^2 they would hold
// about as much memory again as the activations).
hipError_t unet_resize(UNet& net, int N, int H, int W, bool train) {
  if (net.N == N && net.H == H && net.W == W && !net.act.empty() && (net.trainable || !train))
    return hipSuccess;
  for (auto p : net.act)
    if (p) (void)hipFree(p);
  for (auto p : net.grad)
    if (p) (void)hipFree(p);
  if (net.out) (void)hipFree(net.out);
  if (net.dout) (void)hipFree(net.dout);
  if (net.dhead) (void)hipFree(net.dhead);
  net.out = net.dout = nullptr;
  net.dhead = nullptr;
  net.N = net.H = net.W = 0;
  net.trainable = false;
  net.act.assign(21, nullptr);
  net.grad.assign(21, nullptr);
  net.act_elems.assign(21, 0);
  hipError_t e;
  for (int a = 0; a < 21; a++) {
    int lv = kActSpec[a][0];
    size_t n = (size_t)N * (H >> lv) * (W >> lv) * kActSpec[a][1];
    net.act_elems[a] = n;
    if ((e = hipMalloc((void**)&net.act[a], n * 2 + 64))) return e;
    if (train && a > 0 && (e = hipMalloc((void**)&net.grad[a], n * 2 + 64))) return e;
  }
  size_t px = (size_t)N * H * W;
  if ((e = hipMalloc((void**)&net.out, px * 4 + 64))) return e;
  if (train && ((e = hipMalloc((void**)&net.dout, px * 4 + 64)) ||
                (e = hipMalloc((void**)&net.dhead, px * 16 * 2 + 64))))
    return e;
  net.trainable = train;
  net.N = N;
  net.H = H;
  net.W = W;
  return hipSuccess;
}

// act[0] must hold the input.  Result: net.out (FP32, N*H*W).
void unet_forward(UNet& net, hipStream_t s) {
  if (!net.packed) unet_pack(net, s);
  const int N = net.N, H = net.H, W = net.W;
  auto conv = [&](int l) {
    const ConvLayer& c = net.L[l];
    int ia = kLayerIO[l][0], oa = kLayerIO[l][1];
    int lv = kActSpec[ia][0];
    launch_conv3x3(s, net.act[ia], c.cin, net.wt + c.wtoff, c.kpad, net.w + c.boff, c.cout, N,
                   H >> lv, W >> lv, c.relu, oa < 21 ? net.act[oa] : nullptr,
                   oa < 21 ? nullptr : net.out);
  };
  conv(E0A);
  conv(E0B);
  launch_pool2(s, net.act[2], net.act[3], N, H, W, 16);
  conv(E1A);
  conv(E1B);
  launch_pool2(s, net.act[5], net.act[6], N, H >> 1, W >> 1, 32);
  conv(E2A);
  conv(E2B);
  launch_pool2(s, net.act[8], net.act[9], N, H >> 2, W >> 2, 64);
  conv(B3A);
  conv(B3B);
  launch_up_concat(s, net.act[11], 64, net.act[8], 64, net.act[12], N, H >> 2, W >> 2);
  conv(D2A);
  conv(D2B);
  launch_up_concat(s, net.act[14], 64, net.act[5], 32, net.act[15], N, H >> 1, W >> 1);
  conv(D1A);
  conv(D1B);
  launch_up_concat(s, net.act[17], 32, net.act[2], 16, net.act[18], N, H, W);
  conv(D0A);
  conv(D0B);
  conv(HEAD);
}

// Backward from net.dhead (the head's output gradient, bf16 16 channels with
// channel 0 live).  Accumulates parameter gradients into net.g (zero it first).
void unet_backward(UNet& net, hipStream_t s) {
  const int N = net.N, H = net.H, W = net.W;
  auto lvl = [&](int a) { return kActSpec[a][0]; };
  // gradient wrt a layer's weights, then (if needed) wrt its input
  auto back = [&](int l, const bf16_t* dy, bool need_dx) {
    const ConvLayer& c = net.L[l];
    int ia = kLayerIO[l][0];
    int lv = lvl(ia);
    launch_conv3x3_wgrad(s, net.act[ia], c.cin, dy, c.cout, N, H >> lv, W >> lv, 9 * c.cin,
                         net.g + c.woff, net.g + c.boff);
    if (need_dx)
      launch_conv3x3(s, dy, c.cout, net.wb + c.wboff, c.kpadT, nullptr, c.cin, N, H >> lv,
                     W >> lv, 0, net.grad[ia], nullptr);
  };
  auto relu = [&](int a) { launch_relu_bwd(s, net.grad[a], net.act[a], (long)net.act_elems[a]); };
  back(HEAD, net.dhead, true);
  relu(20);
  back(D0B, net.grad[20], true);
  relu(19);
  back(D0A, net.grad[19], true);
  launch_up_concat_bwd(s, net.grad[18], 32, 16, net.grad[17], net.grad[2], N, H, W, 0);
  relu(17);
  back(D1B, net.grad[17], true);
  relu(16);
  back(D1A, net.grad[16], true);
  launch_up_concat_bwd(s, net.grad[15], 64, 32, net.grad[14], net.grad[5], N, H >> 1, W >> 1, 0);
  relu(14);
  back(D2B, net.grad[14], true);
  relu(13);
  back(D2A, net.grad[13], true);
  launch_up_concat_bwd(s, net.grad[12], 64, 64, net.grad[11], net.grad[8], N, H >> 2, W >> 2, 0);
  relu(11);
  back(B3B, net.grad[11], true);
  relu(10);
  back(B3A, net.grad[10], true);
  launch_pool2_bwd(s, net.grad[9], net.grad[8], N, H >> 2, W >> 2, 64, 1);
  relu(8);
  back(E2B, net.grad[8], true);
  relu(7);
  back(E2A, net.grad[7], true);
  launch_pool2_bwd(s, net.grad[6], net.grad[5], N, H >> 1, W >> 1, 32, 1);
  relu(5);
  back(E1B, net.grad[5], true);
  relu(4);
  back(E1A, net.grad[4], true);
  launch_pool2_bwd(s, net.grad[3], net.grad[2], N, H, W, 16, 1);
  relu(2);
  back(E0B, net.grad[2], true);
  relu(1);
  back(E0A, net.grad[1], false);
}

}  // namespace fh

// ---- C ABI ------------------------------------------------------------------------

using namespace fh;

struct fh_unet {
  UNet net;
  hipStream_t s = nullptr;
  // training set (device), coarse samples of cw x ch cells: the normalised
  // input x (float4 per cell) and the target t (correction / s_r), and the
  // s_r of each sample
  int cap = 0, count = 0, cw = 0, ch = 0;
  float4* x = nullptr;
  float* t = nullptr;
  float* sr = nullptr;
  double* partials = nullptr;
  double* sums = nullptr;
  int* idx = nullptr;
  float *tgt = nullptr, *mask = nullptr;
  float c0 = 1.f;
  // gather scratch for the current fine grid (two float4 ping-pong levels),
  // the correction on the fine grid, and the capture's pending sample state
  float4* g[2] = {nullptr, nullptr};
  size_t g_cells = 0;
  float* q = nullptr;
  size_t q_cells = 0;
  float* cstats = nullptr;  // {s_r, s_p} of the coarse grid in use
  std::string err;
};

namespace fh {
// Engine-side hooks (fh_engine.hip).
int engine_capture_hook(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                        const double* p_prev, const double* p_final, long n, int H, int W,
                        bool begin);
int engine_warm_start(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                      const uint8_t* info, double* p, int H, int W, float* stats2);
}  // namespace fh

static thread_local std::string g_unet_err;

#define UCHK(call)                                                   \
  do {                                                               \
    hipError_t _e = (call);                                          \
    if (_e != hipSuccess) {                                          \
      g_unet_err = std::string(#call) + ": " + hipGetErrorString(_e); \
      return FH_ERR_HIP;                                             \
    }                                                                \
  } while (0)

namespace fh {

// The block size F of the network's coarse grid for an nx x ny fine grid:
// the smallest power of two with both nx / F and ny / F at most
// kUnetCoarseSide; 0 when F does not divide the sides or the coarse sides are
// not multiples of 8 (the network's three 2x poolings).
int unet_coarse_factor(int nx, int ny, int* cw, int* ch) {
  int F = 1;
  while (nx / F > kUnetCoarseSide || ny / F > kUnetCoarseSide) F *= 2;
  if (nx % F || ny % F || (nx / F) % 8 || (ny / F) % 8) return 0;
  *cw = nx / F;
  *ch = ny / F;
  return F;
}

// Gathers the fine grid (nx x ny) into F x F blocks: the result's address
// (in u->g) is returned through out.  p1 non-null: the pressure gathered is
// p1 - p (r unread).
static hipError_t gather(fh_unet* u, hipStream_t s, const float* r, const uint8_t* lbl,
                         const double* p, const double* p1, int nx, int ny, int F,
                         float4** out) {
  const size_t need = F == 1 ? (size_t)nx * ny : (size_t)(nx / 2) * (ny / 2);
  if (u->g_cells < need) {
    for (auto& b : u->g) {
      if (b) (void)hipFree(b);
      b = nullptr;
    }
    u->g_cells = 0;
    hipError_t e;
    if ((e = hipMalloc((void**)&u->g[0], need * sizeof(float4))) ||
        (e = hipMalloc((void**)&u->g[1], need * sizeof(float4))))
      return e;
    u->g_cells = need;
  }
  int cw = F == 1 ? nx : nx / 2, ch = F == 1 ? ny : ny / 2;
  dim3 blk(256), grd((cw + 255) / 256, ch);
  if (F == 1)
    k_gather_first<1><<<grd, blk, 0, s>>>(r, lbl, p, p1, nx, cw, ch, u->g[0]);
  else
    k_gather_first<2><<<grd, blk, 0, s>>>(r, lbl, p, p1, nx, cw, ch, u->g[0]);
  int cur = 0;
  for (int f = 2; f < F; f *= 2) {
    cw /= 2;
    ch /= 2;
    k_gather2<<<dim3((cw + 255) / 256, ch), blk, 0, s>>>(u->g[cur], cw, ch, u->g[cur ^ 1]);
    cur ^= 1;
  }
  *out = u->g[cur];
  return hipGetLastError();
}

int engine_capture_hook(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                        const double* p_prev, const double* p_final, long n, int H, int W,
                        bool begin) {
  (void)n;
  int cw = 0, ch = 0;
  const int F = unet_coarse_factor(W, H, &cw, &ch);
  if (!u || u->cap == 0 || F == 0 || cw != u->cw || ch != u->ch) return FH_OK;
  if (u->count >= u->cap) return FH_OK;
  // both calls gather over the labels; the first also over the defect, the
  // second over both pressures
  if (!lbl || !p_prev || (begin ? !r32 : !p_final)) {
    g_unet_err = "capture: missing field";
    return FH_ERR_BAD_ARG;
  }
  const long hw = (long)cw * ch, slot = u->count;
  const float area = (float)F * (float)F;
  float4* c = nullptr;
  if (begin) {
    UCHK(gather(u, s, r32, lbl, p_prev, nullptr, W, H, F, &c));
    k_coarse_stats<<<1, 1024, 0, s>>>(c, (int)hw, area, u->sr + 2 * slot);
    k_coarse_input<<<nblk(hw), 256, 0, s>>>(c, (int)hw, area, u->sr + 2 * slot,
                                            u->x + slot * hw, nullptr);
  } else {
    UCHK(gather(u, s, nullptr, lbl, p_prev, p_final, W, H, F, &c));
    k_coarse_target<<<nblk(hw), 256, 0, s>>>(c, (int)hw, area, u->sr + 2 * slot,
                                             u->t + slot * hw);
    u->count++;
  }
  UCHK(hipGetLastError());
  return FH_OK;
}

int engine_warm_start(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                      const uint8_t* info, double* p, int H, int W, float* stats2) {
  (void)stats2;
  UNet& net = u->net;
  int cw = 0, ch = 0;
  const int F = unet_coarse_factor(W, H, &cw, &ch);
  if (F == 0) return FH_OK;  // no coarse grid the network can take: the start stays
  if (!r32 || !lbl || !info || !p) {
    g_unet_err = "warm start: missing field";
    return FH_ERR_BAD_ARG;
  }
  const long n = (long)H * W, hw = (long)cw * ch;
  UCHK(unet_resize(net, 1, ch, cw, false));
  if (u->q_cells < (size_t)n) {
    if (u->q) (void)hipFree(u->q);
    u->q = nullptr;
    u->q_cells = 0;
    UCHK(hipMalloc((void**)&u->q, (size_t)n * sizeof(float)));
    u->q_cells = n;
  }
  float4* c = nullptr;
  UCHK(gather(u, s, r32, lbl, p, nullptr, W, H, F, &c));
  const float area = (float)F * (float)F;
  k_coarse_stats<<<1, 1024, 0, s>>>(c, (int)hw, area, u->cstats);
  k_coarse_input<<<nblk(hw), 256, 0, s>>>(c, (int)hw, area, u->cstats, nullptr, net.act[0]);
  unet_forward(net, s);
  k_prolong_out<<<dim3((W + 255) / 256, H), 256, 0, s>>>(net.out, cw, ch, F, info, W, H,
                                                         u->cstats, u->c0, u->q);
  // enough blocks to fill the chip, four cells per thread at least (the
  // partials buffer holds kWarmBlocks)
  const int dots_blocks = (int)std::min<long>(kWarmBlocks, std::max<long>(256, n / 1024));
  k_warm_dots<<<dots_blocks, 256, 0, s>>>(r32, u->q, info, W, n, u->partials);
  launch_finalize_sum4(s, u->partials, dots_blocks, u->sums);
  k_apply_warm<<<nblk(n), 256, 0, s>>>(p, u->q, u->sums, n);
  UCHK(hipGetLastError());
  return FH_OK;
}

}  // namespace fh

extern "C" {

const char* fh_unet_last_error(void) { return g_unet_err.c_str(); }

int fh_unet_num_params(void) { return (int)unet_param_count(); }

int fh_unet_coarse_grid(int nx, int ny, int* cw, int* ch) {
  if (!cw || !ch || nx < 1 || ny < 1) return 0;
  return unet_coarse_factor(nx, ny, cw, ch);
}

fh_unet* fh_unet_create(unsigned long long seed) {
  fh_unet* u = new fh_unet();
  if (hipStreamCreate(&u->s) != hipSuccess || unet_init(u->net, seed, u->s) != hipSuccess ||
      hipMalloc((void**)&u->partials, 4 * kWarmBlocks * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&u->sums, 64) != hipSuccess ||
      hipMalloc((void**)&u->cstats, 64) != hipSuccess) {
    g_unet_err = "fh_unet_create: allocation failed";
    unet_free(u->net);
    for (void* p : {(void*)u->partials, (void*)u->sums, (void*)u->cstats})
      if (p) (void)hipFree(p);
    if (u->s) (void)hipStreamDestroy(u->s);
    delete u;
    return nullptr;
  }
  return u;
}

void fh_unet_destroy(fh_unet* u) {
  if (!u) return;
  (void)hipStreamSynchronize(u->s);
  unet_free(u->net);
  for (void* p : {(void*)u->x, (void*)u->t, (void*)u->sr, (void*)u->partials, (void*)u->sums,
                  (void*)u->idx, (void*)u->tgt, (void*)u->mask, (void*)u->g[0], (void*)u->g[1],
                  (void*)u->q, (void*)u->cstats})
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(u->s);
  delete u;
}

int fh_unet_get_params(fh_unet* u, float* out, int n) {
  if (!u || !out || n != (int)u->net.nparam) return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  UCHK(hipMemcpy(out, u->net.w, (size_t)n * 4, hipMemcpyDeviceToHost));
  return FH_OK;
}

int fh_unet_set_params(fh_unet* u, const float* in, int n) {
  if (!u || !in || n != (int)u->net.nparam) return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  UCHK(hipMemcpy(u->net.w, in, (size_t)n * 4, hipMemcpyHostToDevice));
  u->net.packed = false;
  return FH_OK;
}

int fh_unet_set_scale(fh_unet* u, double c0) {
  if (!u || !(c0 > 0)) return FH_ERR_BAD_ARG;
  u->c0 = (float)c0;
  return FH_OK;
}

int fh_unet_get_scale(fh_unet* u, double* c0) {
  if (!u || !c0) return FH_ERR_BAD_ARG;
  *c0 = u->c0;
  return FH_OK;
}

int fh_unet_forward(fh_unet* u, int N, int H, int W, const float* in, float* out) {
  if (!u || !in || !out || N < 1 || H % 8 || W % 8 || H < 8 || W < 8) return FH_ERR_BAD_ARG;
  UCHK(unet_resize(u->net, N, H, W, false));
  long n = (long)N * H * W * kUnetIn;
  float* tmp = nullptr;
  UCHK(hipMalloc((void**)&tmp, n * 4));
  UCHK(hipMemcpyAsync(tmp, in, n * 4, hipMemcpyHostToDevice, u->s));
  k_f32_to_bf16<<<nblk(n), 256, 0, u->s>>>(tmp, u->net.act[0], n);
  unet_forward(u->net, u->s);
  UCHK(hipGetLastError());
  UCHK(hipMemcpyAsync(out, u->net.out, (size_t)N * H * W * 4, hipMemcpyDeviceToHost, u->s));
  UCHK(hipStreamSynchronize(u->s));
  UCHK(hipFree(tmp));
  return FH_OK;
}

int fh_unet_backward(fh_unet* u, int N, int H, int W, const float* in, const float* dout,
                     float* dparams) {
  if (!u || !in || !dout || !dparams || N < 1 || H % 8 || W % 8 || H < 8 || W < 8)
    return FH_ERR_BAD_ARG;
  UNet& net = u->net;
  UCHK(unet_resize(net, N, H, W, true));
  long n = (long)N * H * W * kUnetIn, px = (long)N * H * W;
  float* tmp = nullptr;
  UCHK(hipMalloc((void**)&tmp, n * 4));
  UCHK(hipMemcpyAsync(tmp, in, n * 4, hipMemcpyHostToDevice, u->s));
  k_f32_to_bf16<<<nblk(n), 256, 0, u->s>>>(tmp, net.act[0], n);
  unet_forward(net, u->s);
  UCHK(hipMemcpyAsync(net.dout, dout, px * 4, hipMemcpyHostToDevice, u->s));
  k_head_grad<<<nblk(px), 256, 0, u->s>>>(net.dout, px, net.dhead);
  UCHK(hipMemsetAsync(net.g, 0, net.nparam * 4, u->s));
  unet_backward(net, u->s);
  UCHK(hipGetLastError());
  UCHK(hipMemcpyAsync(dparams, net.g, net.nparam * 4, hipMemcpyDeviceToHost, u->s));
  UCHK(hipStreamSynchronize(u->s));
  UCHK(hipFree(tmp));
  return FH_OK;
}

int fh_unet_dataset_reserve(fh_unet* u, int capacity, int H, int W) {
  if (!u || capacity < 0 || H < 8 || W < 8 || H % 8 || W % 8 || H > kUnetCoarseSide ||
      W > kUnetCoarseSide)
    return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  for (void** p : {(void**)&u->x, (void**)&u->t, (void**)&u->sr})
    if (*p) {
      UCHK(hipFree(*p));
      *p = nullptr;
    }
  u->cap = capacity;
  u->count = 0;
  u->ch = H;
  u->cw = W;
  if (capacity == 0) return FH_OK;
  size_t n = (size_t)capacity * H * W;
  UCHK(hipMalloc((void**)&u->x, n * sizeof(float4)));
  UCHK(hipMalloc((void**)&u->t, n * 4));
  UCHK(hipMalloc((void**)&u->sr, (size_t)capacity * 2 * 4));
  return FH_OK;
}

int fh_unet_dataset_size(fh_unet* u) { return u ? u->count : 0; }

int fh_unet_dataset_get(fh_unet* u, int index, float* x, float* t, float* scales) {
  if (!u || index < 0 || index >= u->count) return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  UCHK(hipDeviceSynchronize());
  const size_t hw = (size_t)u->cw * u->ch;
  if (x) UCHK(hipMemcpy(x, u->x + index * hw, hw * sizeof(float4), hipMemcpyDeviceToHost));
  if (t) UCHK(hipMemcpy(t, u->t + index * hw, hw * 4, hipMemcpyDeviceToHost));
  if (scales) UCHK(hipMemcpy(scales, u->sr + 2 * (size_t)index, 8, hipMemcpyDeviceToHost));
  return FH_OK;
}

// c0 = RMS over the first samples' coarse cells holding fluid of the target
// (correction / s_r).
int fh_unet_calibrate(fh_unet* u, int samples) {
  if (!u || u->count == 0) return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  const long hw = (long)u->cw * u->ch;
  const int m = samples < u->count ? samples : u->count;
  std::vector<float4> x(hw);
  std::vector<float> t(hw);
  double acc = 0, cnt = 0;
  for (int k = 0; k < m; k++) {
    UCHK(hipMemcpy(x.data(), u->x + (size_t)k * hw, hw * sizeof(float4), hipMemcpyDeviceToHost));
    UCHK(hipMemcpy(t.data(), u->t + (size_t)k * hw, hw * 4, hipMemcpyDeviceToHost));
    for (long i = 0; i < hw; i++)
      if (1.f - x[i].y - x[i].z > 1e-6f) {
        acc += (double)t[i] * t[i];
        cnt += 1;
      }
  }
  if (cnt == 0 || acc == 0) return FH_ERR_BAD_ARG;
  u->c0 = (float)std::sqrt(acc / cnt);
  return FH_OK;
}

// `steps` Adam steps on batches of `batch` samples drawn (with a SplitMix64
// shuffle) from the captured set; loss_out[k] = masked MSE of step k.
int fh_unet_train(fh_unet* u, int steps, int batch, double lr, unsigned long long seed,
                  double* loss_out) {
  if (!u || u->count < 1 || batch < 1 || steps < 0) return FH_ERR_BAD_ARG;
  UNet& net = u->net;
  const int H = u->ch, W = u->cw;
  const long hw = (long)H * W, px = hw * batch;
  UCHK(unet_resize(net, batch, H, W, true));
  if (u->idx) (void)hipFree(u->idx);
  if (u->tgt) (void)hipFree(u->tgt);
  if (u->mask) (void)hipFree(u->mask);
  UCHK(hipMalloc((void**)&u->idx, batch * 4));
  UCHK(hipMalloc((void**)&u->tgt, px * 4));
  UCHK(hipMalloc((void**)&u->mask, px * 4));
  double* hsums = nullptr;
  UCHK(hipHostMalloc((void**)&hsums, 64, hipHostMallocDefault));
  SplitMix rng{seed};
  std::vector<int> order(u->count);
  for (int k = 0; k < u->count; k++) order[k] = k;
  size_t pos = order.size();
  std::vector<int> hidx(batch);
  const float inv_c0 = 1.f / u->c0;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int st = 0; st < steps; st++) {
    for (int b = 0; b < batch; b++) {
      if (pos >= order.size()) {
        for (size_t k = order.size() - 1; k > 0; k--) {
          size_t r = (size_t)(rng.next() % (k + 1));
          std::swap(order[k], order[r]);
        }
        pos = 0;
      }
      hidx[b] = order[pos++];
    }
    UCHK(hipMemcpyAsync(u->idx, hidx.data(), batch * 4, hipMemcpyHostToDevice, u->s));
    dim3 g(nblk(hw), batch);
    k_train_prep<<<g, 256, 0, u->s>>>(u->x, u->t, u->idx, hw, inv_c0, net.act[0], u->tgt,
                                      u->mask);
    unet_forward(net, u->s);
    k_loss<<<512, 256, 0, u->s>>>(net.out, u->tgt, u->mask, px, u->partials);
    launch_finalize_sum4(u->s, u->partials, 512, u->sums);
    k_loss_grad<<<nblk(px), 256, 0, u->s>>>(net.out, u->tgt, u->mask, u->sums, px, net.dhead,
                                            nullptr);
    UCHK(hipMemsetAsync(net.g, 0, net.nparam * 4, u->s));
    unet_backward(net, u->s);
    net.step++;
    float c1 = 1.f - std::pow(b1, (float)net.step), c2 = 1.f - std::pow(b2, (float)net.step);
    launch_adam(u->s, net.w, net.g, net.m, net.v, (long)net.nparam, (float)lr, b1, b2, eps, c1, c2,
                1.f);
    unet_pack(net, u->s);
    UCHK(hipGetLastError());
    if (loss_out) {
      UCHK(hipMemcpyAsync(hsums, u->sums, 16, hipMemcpyDeviceToHost, u->s));
      UCHK(hipStreamSynchronize(u->s));
      loss_out[st] = hsums[1] > 0 ? hsums[0] / hsums[1] : 0.0;
    }
  }
  UCHK(hipStreamSynchronize(u->s));
  (void)hipHostFree(hsums);
  return FH_OK;
}

}  // extern "C"
