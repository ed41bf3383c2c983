// Learned warm start for the pressure solve: the U-Net's host side (layer
// table, parameters, forward/backward schedules, Adam), its training set
// captured from the engine's own multigrid solves, and the warm-start step
// used inside the defect-correction solve.
//
// Formulation.  Given the previous pressure p_prev and the new right-hand
// side b, the network sees the defect r0 = b - A p_prev (normalised by its
// RMS s_r over the fluid), the solid and Air masks, and p_prev (normalised
// by its RMS s_p), and predicts the correction o; the warm start is
//   p0 = p_prev + s_r * c * o      on fluid cells,
// c = c0 * (H W / area0) scaling the trained magnitude with the grid area.
// Training minimises the mean squared error of o against
// (p* - p_prev) / (s_r c) over fluid cells, p* being the multigrid solution.
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

// Per-grid statistics for the normalisation: {sum r^2, sum p^2, n} over
// fluid cells, two-phase (block partials, then one block).
template <class PT>
__global__ __launch_bounds__(256) void k_stats(const float* __restrict__ r,
                                               const PT* __restrict__ p,
                                               const uint8_t* __restrict__ lbl, long n,
                                               double* partials) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    if (lbl[i] != FLUID) continue;
    double rv = r[i], pv = (double)p[i];
    acc[0] += rv * rv;
    acc[1] += pv * pv;
    acc[2] += 1.0;
  }
  block_sum_store<4>(acc, partials);
}

__global__ void k_stats_finalize(const double* __restrict__ partials, int nblocks,
                                 float* __restrict__ out2) {
  __shared__ double sh[3][4];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int s = 0; s < 3; s++) {
    double t = 0.0;
    for (int b = tid; b < nblocks; b += 256) t += partials[(size_t)b * 4 + s];
    t = wave_sum(t);
    if (lane == 0) sh[s][wid] = t;
  }
  __syncthreads();
  if (tid == 0) {
    double r2 = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
    double p2 = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
    double nn = ((sh[2][0] + sh[2][1]) + sh[2][2]) + sh[2][3];
    double sr = nn > 0 ? sqrt(r2 / nn) : 0.0;
    double sp = nn > 0 ? sqrt(p2 / nn) : 0.0;
    out2[0] = (float)(sr > 0 ? sr : 1.0);
    out2[1] = (float)(sp > 0 ? sp : 1.0);
  }
}

// Network input for sample b of a batch: channels {r/s_r, solid, air,
// p_prev/s_p, 0, 0, 0, 0}; optional target (p*-p_prev)/(s_r c) and mask.
template <class PT>
__global__ void k_prep(const float* __restrict__ r, const PT* __restrict__ pp,
                       const uint8_t* __restrict__ lbl, const float* __restrict__ dlt,
                       const float* __restrict__ stats, const int* __restrict__ idx, long hw,
                       int H, int W, float c, bf16_t* __restrict__ in, float* __restrict__ tgt,
                       float* __restrict__ mask) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (i >= hw) return;
  long slot = idx ? idx[b] : 0;
  long src = slot * hw + i;
  float sr = stats[2 * slot], sp = stats[2 * slot + 1];
  uint8_t l = lbl[src];
  float c0 = (l == FLUID) ? r[src] / sr : 0.f;
  float c3 = (float)((double)pp[src] / sp);
  uint4 v;
  v.x = (uint32_t)f2bf_(c0) | ((uint32_t)f2bf_(l == SOLID ? 1.f : 0.f) << 16);
  v.y = (uint32_t)f2bf_(l == AIR ? 1.f : 0.f) | ((uint32_t)f2bf_(c3) << 16);
  v.z = 0;
  v.w = 0;
  long o = (long)b * hw + i;
  *reinterpret_cast<uint4*>(in + o * kUnetIn) = v;
  if (tgt) {
    tgt[o] = (l == FLUID) ? dlt[src] / (sr * c) : 0.f;
    mask[o] = (l == FLUID) ? 1.f : 0.f;
  }
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

// The warm start is a line search along the network output o: with the
// defect r0 = b - A p_prev and q = A o (o counted as 0 off the fluid, where
// the correction is 0), the step m = <r0, q> / <q, q> (at least 0) minimises
// |b - A (p_prev + m o)| over fluid cells, so the warm-started residual is
// never larger than the previous pressure's, and the output scale drops out.
// Blocks of k_warm_dots: 16 per CU, so the pass streams at HBM rate (512
// left it at ~7 % of the HBM peak at 2048x1024).
constexpr int kWarmBlocks = 4096;

// k_warm_dots: block partial sums {r0.q, q.q} over fluid cells with an open
// neighbour (info byte: bit 7 fluid, bits 3..6 open +x,-x,+y,-y, 0..2 count).
__global__ __launch_bounds__(256) void k_warm_dots(const float* __restrict__ r0,
                                                   const float* __restrict__ o,
                                                   const uint8_t* __restrict__ info, int nx,
                                                   long n, double* partials) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int f = info[k];
    if (!(f & 0x80) || (f & 7) == 0) continue;
    // an open neighbour is in the domain, so its index is valid
    float nb = 0.f;
    if ((f & 0x08) && (info[k + 1] & 0x80)) nb += o[k + 1];
    if ((f & 0x10) && (info[k - 1] & 0x80)) nb += o[k - 1];
    if ((f & 0x20) && (info[k + nx] & 0x80)) nb += o[k + nx];
    if ((f & 0x40) && (info[k - nx] & 0x80)) nb += o[k - nx];
    const double q = (double)(f & 7) * (double)o[k] - (double)nb;
    acc[0] += (double)r0[k] * q;
    acc[1] += q * q;
  }
  block_sum_store<4>(acc, partials);
}

// p += m o on fluid cells, m from the finalized sums {r0.q, q.q}.
__global__ void k_apply_warm(double* __restrict__ p, const float* __restrict__ o,
                             const uint8_t* __restrict__ info, const double* __restrict__ sums,
                             long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = sums[1] > 0.0 ? fmax(sums[0] / sums[1], 0.0) : 0.0;
  if (info[i] & 0x80) p[i] += m * (double)o[i];
}

// Host input [N][H][W][8] f32 -> bf16
__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16_t* __restrict__ o, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) o[i] = f2bf_(a[i]);
}

template <class A, class B>
__global__ void k_convert(const A* __restrict__ a, B* __restrict__ b, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}

__global__ void k_delta(const double* __restrict__ p, const double* __restrict__ p0,
                        float* __restrict__ d, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)(p[i] - p0[i]);
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
  // training set (device): per sample r (f32), label (u8), p_prev (f32),
  // delta (f32), stats {s_r, s_p}
  int cap = 0, count = 0, H = 0, W = 0;
  float *r = nullptr, *pp = nullptr, *dl = nullptr, *stats = nullptr;
  uint8_t* lbl = nullptr;
  double* partials = nullptr;
  double* sums = nullptr;
  int* idx = nullptr;
  float *tgt = nullptr, *mask = nullptr;
  float c0 = 1.f;
  double area0 = 512.0 * 512.0;
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

static hipError_t stats_into(hipStream_t s, const float* r, const void* p, bool p64,
                             const uint8_t* lbl, long n, double* partials, float* out2) {
  const int blocks = 512;
  if (p64)
    k_stats<double><<<blocks, 256, 0, s>>>(r, (const double*)p, lbl, n, partials);
  else
    k_stats<float><<<blocks, 256, 0, s>>>(r, (const float*)p, lbl, n, partials);
  k_stats_finalize<<<1, 256, 0, s>>>(partials, blocks, out2);
  return hipGetLastError();
}

int engine_capture_hook(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                        const double* p_prev, const double* p_final, long n, int H, int W,
                        bool begin) {
  if (!u || u->cap == 0 || H != u->H || W != u->W) return FH_OK;
  if (u->count >= u->cap) return FH_OK;
  long slot = u->count;
  size_t o = (size_t)slot * n;
  if (begin) {
    UCHK(hipMemcpyAsync(u->r + o, r32, n * 4, hipMemcpyDeviceToDevice, s));
    UCHK(hipMemcpyAsync(u->lbl + o, lbl, n, hipMemcpyDeviceToDevice, s));
    k_convert<double, float><<<nblk(n), 256, 0, s>>>(p_prev, u->pp + o, n);
    UCHK(stats_into(s, r32, p_prev, true, lbl, n, u->partials, u->stats + 2 * slot));
  } else {
    k_delta<<<nblk(n), 256, 0, s>>>(p_final, p_prev, u->dl + o, n);
    UCHK(hipGetLastError());
    u->count++;
  }
  return FH_OK;
}

int engine_warm_start(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                      const uint8_t* info, double* p, int H, int W, float* stats2) {
  UNet& net = u->net;
  if ((H % 8) || (W % 8)) return FH_OK;  // three 2x poolings need multiples of 8
  UCHK(unet_resize(net, 1, H, W, false));
  long n = (long)H * W;
  UCHK(stats_into(s, r32, p, true, lbl, n, u->partials, stats2));
  float c = (float)(u->c0 * ((double)H * W / u->area0));
  dim3 g(nblk(n), 1);
  k_prep<double><<<g, 256, 0, s>>>(r32, p, lbl, nullptr, stats2, nullptr, n, H, W, c,
                                   net.act[0], nullptr, nullptr);
  unet_forward(net, s);
  // enough blocks to fill the chip (the partials buffer holds kWarmBlocks)
  k_warm_dots<<<kWarmBlocks, 256, 0, s>>>(r32, net.out, info, W, n, u->partials);
  launch_finalize_sum4(s, u->partials, kWarmBlocks, u->sums);
  k_apply_warm<<<nblk(n), 256, 0, s>>>(p, net.out, info, u->sums, n);
  UCHK(hipGetLastError());
  return FH_OK;
}

}  // namespace fh

extern "C" {

const char* fh_unet_last_error(void) { return g_unet_err.c_str(); }

int fh_unet_num_params(void) { return (int)unet_param_count(); }

fh_unet* fh_unet_create(unsigned long long seed) {
  fh_unet* u = new fh_unet();
  if (hipStreamCreate(&u->s) != hipSuccess || unet_init(u->net, seed, u->s) != hipSuccess) {
    g_unet_err = "fh_unet_create: allocation failed";
    unet_free(u->net);
    if (u->s) (void)hipStreamDestroy(u->s);
    delete u;
    return nullptr;
  }
  (void)hipMalloc((void**)&u->partials, 4 * kWarmBlocks * sizeof(double));
  (void)hipMalloc((void**)&u->sums, 64);
  return u;
}

void fh_unet_destroy(fh_unet* u) {
  if (!u) return;
  (void)hipStreamSynchronize(u->s);
  unet_free(u->net);
  for (void* p : {(void*)u->r, (void*)u->pp, (void*)u->dl, (void*)u->stats, (void*)u->lbl,
                  (void*)u->partials, (void*)u->sums, (void*)u->idx, (void*)u->tgt,
                  (void*)u->mask})
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

int fh_unet_set_scale(fh_unet* u, double c0, double area0) {
  if (!u || !(c0 > 0) || !(area0 > 0)) return FH_ERR_BAD_ARG;
  u->c0 = (float)c0;
  u->area0 = area0;
  return FH_OK;
}

int fh_unet_get_scale(fh_unet* u, double* c0, double* area0) {
  if (!u || !c0 || !area0) return FH_ERR_BAD_ARG;
  *c0 = u->c0;
  *area0 = u->area0;
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
  if (!u || capacity < 0 || H % 8 || W % 8) return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  for (void** p : {(void**)&u->r, (void**)&u->pp, (void**)&u->dl, (void**)&u->stats,
                   (void**)&u->lbl})
    if (*p) {
      UCHK(hipFree(*p));
      *p = nullptr;
    }
  u->cap = capacity;
  u->count = 0;
  u->H = H;
  u->W = W;
  if (capacity == 0) return FH_OK;
  size_t n = (size_t)capacity * H * W;
  UCHK(hipMalloc((void**)&u->r, n * 4));
  UCHK(hipMalloc((void**)&u->pp, n * 4));
  UCHK(hipMalloc((void**)&u->dl, n * 4));
  UCHK(hipMalloc((void**)&u->lbl, n));
  UCHK(hipMalloc((void**)&u->stats, (size_t)capacity * 2 * 4));
  return FH_OK;
}

int fh_unet_dataset_size(fh_unet* u) { return u ? u->count : 0; }

// RMS over the first samples of |p* - p_prev| / s_r: the output scale c0.
int fh_unet_calibrate(fh_unet* u, int samples) {
  if (!u || u->count == 0) return FH_ERR_BAD_ARG;
  UCHK(hipStreamSynchronize(u->s));
  long n = (long)u->H * u->W;
  int m = samples < u->count ? samples : u->count;
  std::vector<float> d(n), st(2);
  std::vector<uint8_t> l(n);
  double acc = 0;
  int used = 0;
  for (int k = 0; k < m; k++) {
    UCHK(hipMemcpy(d.data(), u->dl + (size_t)k * n, n * 4, hipMemcpyDeviceToHost));
    UCHK(hipMemcpy(l.data(), u->lbl + (size_t)k * n, n, hipMemcpyDeviceToHost));
    UCHK(hipMemcpy(st.data(), u->stats + 2 * k, 8, hipMemcpyDeviceToHost));
    double s2 = 0, cnt = 0;
    for (long i = 0; i < n; i++)
      if (l[i] == FLUID) {
        s2 += (double)d[i] * d[i];
        cnt += 1;
      }
    if (cnt > 0 && st[0] > 0) {
      acc += std::sqrt(s2 / cnt) / st[0];
      used++;
    }
  }
  if (used == 0) return FH_ERR_BAD_ARG;
  u->c0 = (float)(acc / used);
  u->area0 = (double)u->H * u->W;
  return FH_OK;
}

// `steps` Adam steps on batches of `batch` samples drawn (with a SplitMix64
// shuffle) from the captured set; loss_out[k] = masked MSE of step k.
int fh_unet_train(fh_unet* u, int steps, int batch, double lr, unsigned long long seed,
                  double* loss_out) {
  if (!u || u->count < 1 || batch < 1 || steps < 0) return FH_ERR_BAD_ARG;
  UNet& net = u->net;
  const int H = u->H, W = u->W;
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
  float c = (float)(u->c0 * ((double)hw / u->area0));
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
    k_prep<float><<<g, 256, 0, u->s>>>(u->r, u->pp, u->lbl, u->dl, u->stats, u->idx, hw, H, W, c,
                                       net.act[0], u->tgt, u->mask);
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
