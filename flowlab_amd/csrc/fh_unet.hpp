// Learned warm start: U-Net layer table, launchers and the host-side network.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace fh {

typedef uint16_t bf16_t;  // storage of a bfloat16

// ---- kernels (fh_unet_kernels.hip) -------------------------------------------
// Y = act(conv3x3(X) + bias), channels-last.  Wt: [cout][kpad] bf16 with
// k = tap*cin + ci (tap = 3*(dy+1) + (dx+1)), kpad = roundup(9 cin, 32).
// With Yf non-null only output channel 0 is written, in FP32.
void launch_conv3x3(hipStream_t s, const bf16_t* X, int cin, const bf16_t* Wt, int kpad,
                    const float* bias, int cout, int N, int H, int W, int relu, bf16_t* Y,
                    float* Yf);
// dW[co][k] += sum_px dY[px][co] X[px+tap][ci] (master layout [co][9 cin]),
// db[co] += sum_px dY[px][co].
void launch_conv3x3_wgrad(hipStream_t s, const bf16_t* X, int cin, const bf16_t* dY, int cout,
                          int N, int H, int W, int kstride, float* dW, float* db);
void launch_pool2(hipStream_t s, const bf16_t* X, bf16_t* Y, int N, int H, int W, int C);
void launch_pool2_bwd(hipStream_t s, const bf16_t* dY, bf16_t* dX, int N, int H, int W, int C,
                      int accumulate);
void launch_up_concat(hipStream_t s, const bf16_t* a, int Ca, const bf16_t* b, int Cb, bf16_t* out,
                      int N, int H, int W);
void launch_up_concat_bwd(hipStream_t s, const bf16_t* dout, int Ca, int Cb, bf16_t* da,
                          bf16_t* db, int N, int H, int W, int accumulate_b);
void launch_relu_bwd(hipStream_t s, bf16_t* dY, const bf16_t* Y, long n);
void launch_pack_weights(hipStream_t s, const float* w, int cin, int cout, int kpad, bf16_t* Wt,
                         int kpadT, bf16_t* Wb);
void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, long n, float lr,
                 float b1, float b2, float eps, float c1, float c2, float grad_scale);

// ---- the network ----------------------------------------------------------------
// 4 levels, channels 16-32-64-64, two 3x3 convs per level with ReLU, 2x2
// average pooling down, nearest upsampling + concat up, a 3x3 head to one
// FP32 channel.  Input: 8 channels (4 used).
struct ConvLayer {
  int cin, cout, relu;
  int kpad, kpadT;
  size_t woff, boff;     // offsets of weights [cout][9][cin] and bias in the master buffer
  size_t wtoff, wboff;   // offsets in the packed bf16 buffers
};

constexpr int kUnetIn = 8;
// The learned warm start runs the network on the fine grid gathered into
// blocks so that neither coarse side exceeds this (unet_coarse_factor).
constexpr int kUnetCoarseSide = 128;
constexpr int kUnetLayers = 15;
// layer indices
enum { E0A, E0B, E1A, E1B, E2A, E2B, B3A, B3B, D2A, D2B, D1A, D1B, D0A, D0B, HEAD };

struct UNet {
  std::vector<ConvLayer> L;
  size_t nparam = 0, nwt = 0, nwb = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  bf16_t *wt = nullptr, *wb = nullptr;
  int step = 0;
  bool packed = false;
  // activations for the current (N, H, W)
  int N = 0, H = 0, W = 0;
  bool trainable = false;     // grad, dout and dhead allocated
  std::vector<bf16_t*> act;   // forward tensors
  std::vector<bf16_t*> grad;  // their gradients
  std::vector<size_t> act_elems;
  float* out = nullptr;       // [N][H][W] FP32
  float* dout = nullptr;
  bf16_t* dhead = nullptr;    // [N][H][W][16]
};

}  // namespace fh
