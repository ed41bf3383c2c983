/*
 * FP64 CPU oracle of the learned warm start's U-Net (forward and parameter
 * gradient).  TEST INFRASTRUCTURE ONLY, like oracle.c.
 *
 * BUILD-DEFINED: the reference has no learned component.  This file states
 * the network exactly as libflowhip runs it (layer table, channels-last
 * tensors, parameter layout [cout][3][3][cin] + bias per layer) so the GPU's
 * forward and backward kernels can be diffed against it.  With emulate != 0
 * every value the GPU stores in BF16 (weights, activations, activation
 * gradients) is rounded to BF16 at the same point, leaving only FP32-vs-FP64
 * accumulation differences; with emulate == 0 it is plain FP64 math and is
 * pinned against torch float64 autograd and central differences in
 * tests/test_oracle_unet.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { NL = 15, NA = 22 };
static const int spec[NL][3] = {{8, 16, 1},  {16, 16, 1},  {16, 32, 1}, {32, 32, 1}, {32, 64, 1},
                                {64, 64, 1}, {64, 64, 1},  {64, 64, 1}, {128, 64, 1}, {64, 64, 1},
                                {96, 32, 1}, {32, 32, 1},  {48, 16, 1}, {16, 16, 1},  {16, 16, 0}};
static const int io[NL][2] = {{0, 1},   {1, 2},   {3, 4},   {4, 5},   {6, 7},
                              {7, 8},   {9, 10},  {10, 11}, {12, 13}, {13, 14},
                              {15, 16}, {16, 17}, {18, 19}, {19, 20}, {20, 21}};
/* level, channels; tensor 21 = the head output (16 channels, channel 0 used) */
static const int act[NA][2] = {{0, 8},  {0, 16}, {0, 16}, {1, 16}, {1, 32}, {1, 32},
                               {2, 32}, {2, 64}, {2, 64}, {3, 64}, {3, 64}, {3, 64},
                               {2, 128}, {2, 64}, {2, 64}, {1, 96}, {1, 32}, {1, 32},
                               {0, 48}, {0, 16}, {0, 16}, {0, 16}};

/* round-to-nearest-even to bfloat16, as the GPU's conversion does */
static double bf(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

static double rnd(double x, int emulate) { return emulate ? bf(x) : x; }

int fo_unet_num_params(void) {
  int n = 0;
  for (int l = 0; l < NL; l++) n += spec[l][1] * (9 * spec[l][0] + 1);
  return n;
}

static int woff(int l) {
  int n = 0;
  for (int k = 0; k < l; k++) n += spec[k][1] * (9 * spec[k][0] + 1);
  return n;
}

typedef struct {
  int N, H, W;
  double* a[NA];
  double* g[NA];
  size_t n[NA];
} net_t;

static void alloc_net(net_t* t, int N, int H, int W) {
  t->N = N;
  t->H = H;
  t->W = W;
  for (int i = 0; i < NA; i++) {
    int lv = act[i][0];
    t->n[i] = (size_t)N * (H >> lv) * (W >> lv) * act[i][1];
    t->a[i] = (double*)calloc(t->n[i], sizeof(double));
    t->g[i] = (double*)calloc(t->n[i], sizeof(double));
  }
}

static void free_net(net_t* t) {
  for (int i = 0; i < NA; i++) {
    free(t->a[i]);
    free(t->g[i]);
  }
}

/* y = act(conv3x3(x, w) + b); w [cout][3][3][cin] */
static void conv_fwd(const double* x, int cin, const double* w, const double* b, int cout, int N,
                     int H, int W, int relu, double* y, int round_out, int emulate) {
  for (int n = 0; n < N; n++)
    for (int yy = 0; yy < H; yy++)
      for (int xx = 0; xx < W; xx++)
        for (int co = 0; co < cout; co++) {
          double s = 0;
          for (int ky = 0; ky < 3; ky++) {
            int sy = yy + ky - 1;
            if (sy < 0 || sy >= H) continue;
            for (int kx = 0; kx < 3; kx++) {
              int sx = xx + kx - 1;
              if (sx < 0 || sx >= W) continue;
              const double* xp = x + (((size_t)n * H + sy) * W + sx) * cin;
              const double* wp = w + (((size_t)co * 3 + ky) * 3 + kx) * cin;
              for (int ci = 0; ci < cin; ci++) s += rnd(wp[ci], emulate) * xp[ci];
            }
          }
          s += b[co];
          if (relu && s < 0) s = 0;
          y[(((size_t)n * H + yy) * W + xx) * cout + co] = round_out ? rnd(s, emulate) : s;
        }
}

/* dw += sum dy x, db += sum dy, dx = conv^T(dy) (if dx) */
static void conv_bwd(const double* x, int cin, const double* w, int cout, int N, int H, int W,
                     const double* dy, double* dw, double* db, double* dx, int emulate) {
  if (dx) memset(dx, 0, sizeof(double) * (size_t)N * H * W * cin);
  for (int n = 0; n < N; n++)
    for (int yy = 0; yy < H; yy++)
      for (int xx = 0; xx < W; xx++)
        for (int co = 0; co < cout; co++) {
          double gy = dy[(((size_t)n * H + yy) * W + xx) * cout + co];
          db[co] += gy;
          for (int ky = 0; ky < 3; ky++) {
            int sy = yy + ky - 1;
            if (sy < 0 || sy >= H) continue;
            for (int kx = 0; kx < 3; kx++) {
              int sx = xx + kx - 1;
              if (sx < 0 || sx >= W) continue;
              size_t xo = (((size_t)n * H + sy) * W + sx) * cin;
              size_t wo = (((size_t)co * 3 + ky) * 3 + kx) * cin;
              for (int ci = 0; ci < cin; ci++) {
                dw[wo + ci] += gy * x[xo + ci];
                if (dx) dx[xo + ci] += gy * rnd(w[wo + ci], emulate);
              }
            }
          }
        }
  if (dx && emulate)
    for (size_t i = 0; i < (size_t)N * H * W * cin; i++) dx[i] = bf(dx[i]);
}

static void pool_fwd(const double* x, double* y, int N, int H, int W, int C, int emulate) {
  int h = H / 2, w = W / 2;
  for (int n = 0; n < N; n++)
    for (int yy = 0; yy < h; yy++)
      for (int xx = 0; xx < w; xx++)
        for (int c = 0; c < C; c++) {
          const double* b = x + (((size_t)n * H + 2 * yy) * W + 2 * xx) * C + c;
          double s = (b[0] + b[C]) + (b[(size_t)W * C] + b[(size_t)W * C + C]);
          y[(((size_t)n * h + yy) * w + xx) * C + c] = rnd(0.25 * s, emulate);
        }
}

static void pool_bwd_acc(const double* dy, double* dx, int N, int H, int W, int C, int emulate) {
  int h = H / 2, w = W / 2;
  for (int n = 0; n < N; n++)
    for (int yy = 0; yy < H; yy++)
      for (int xx = 0; xx < W; xx++)
        for (int c = 0; c < C; c++) {
          size_t i = (((size_t)n * H + yy) * W + xx) * C + c;
          dx[i] = rnd(dx[i] + 0.25 * dy[(((size_t)n * h + yy / 2) * w + xx / 2) * C + c], emulate);
        }
}

static void upcat_fwd(const double* a, int Ca, const double* b, int Cb, double* out, int N, int H,
                      int W) {
  int C = Ca + Cb;
  for (int n = 0; n < N; n++)
    for (int yy = 0; yy < H; yy++)
      for (int xx = 0; xx < W; xx++) {
        size_t p = ((size_t)n * H + yy) * W + xx;
        for (int c = 0; c < Ca; c++)
          out[p * C + c] = a[(((size_t)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * Ca + c];
        for (int c = 0; c < Cb; c++) out[p * C + Ca + c] = b[p * Cb + c];
      }
}

static void upcat_bwd(const double* dout, int Ca, int Cb, double* da, double* db, int N, int H,
                      int W, int emulate) {
  int C = Ca + Cb, h = H / 2, w = W / 2;
  for (int n = 0; n < N; n++)
    for (int yy = 0; yy < h; yy++)
      for (int xx = 0; xx < w; xx++)
        for (int c = 0; c < Ca; c++) {
          const double* d = dout + (((size_t)n * H + 2 * yy) * W + 2 * xx) * C + c;
          double s = (d[0] + d[C]) + (d[(size_t)W * C] + d[(size_t)W * C + C]);
          da[(((size_t)n * h + yy) * w + xx) * Ca + c] = rnd(s, emulate);
        }
  for (size_t p = 0; p < (size_t)N * H * W; p++)
    for (int c = 0; c < Cb; c++) db[p * Cb + c] = dout[p * C + Ca + c];
}

static void run_forward(net_t* t, const double* prm, int emulate) {
  int N = t->N, H = t->H, W = t->W;
  for (int l = 0; l < NL; l++) {
    int ia = io[l][0], oa = io[l][1];
    int lv = act[ia][0];
    const double* w = prm + woff(l);
    const double* b = w + (size_t)spec[l][1] * 9 * spec[l][0];
    conv_fwd(t->a[ia], spec[l][0], w, b, spec[l][1], N, H >> lv, W >> lv, spec[l][2], t->a[oa],
             l != NL - 1, emulate);
    /* the glue after each layer */
    if (l == 1) pool_fwd(t->a[2], t->a[3], N, H, W, 16, emulate);
    if (l == 3) pool_fwd(t->a[5], t->a[6], N, H >> 1, W >> 1, 32, emulate);
    if (l == 5) pool_fwd(t->a[8], t->a[9], N, H >> 2, W >> 2, 64, emulate);
    if (l == 7) upcat_fwd(t->a[11], 64, t->a[8], 64, t->a[12], N, H >> 2, W >> 2);
    if (l == 9) upcat_fwd(t->a[14], 64, t->a[5], 32, t->a[15], N, H >> 1, W >> 1);
    if (l == 11) upcat_fwd(t->a[17], 32, t->a[2], 16, t->a[18], N, H, W);
  }
}

static void load_input(net_t* t, const double* in, int emulate) {
  for (size_t i = 0; i < t->n[0]; i++) t->a[0][i] = rnd(in[i], emulate);
}

void fo_unet_forward(int N, int H, int W, const double* prm, const double* in, double* out,
                     int emulate) {
  net_t t;
  alloc_net(&t, N, H, W);
  load_input(&t, in, emulate);
  run_forward(&t, prm, emulate);
  for (size_t p = 0; p < (size_t)N * H * W; p++) out[p] = t.a[21][p * 16];
  free_net(&t);
}

static void relu_bwd(double* g, const double* a, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!(a[i] > 0)) g[i] = 0;
}

/* dparams = d/dprm sum(dout .* out) */
void fo_unet_backward(int N, int H, int W, const double* prm, const double* in,
                      const double* dout, double* dprm, int emulate) {
  net_t t;
  alloc_net(&t, N, H, W);
  load_input(&t, in, emulate);
  run_forward(&t, prm, emulate);
  memset(dprm, 0, sizeof(double) * fo_unet_num_params());
  memset(t.g[21], 0, sizeof(double) * t.n[21]);
  for (size_t p = 0; p < (size_t)N * H * W; p++) t.g[21][p * 16] = rnd(dout[p], emulate);
  for (int l = NL - 1; l >= 0; l--) {
    int ia = io[l][0], oa = io[l][1];
    int lv = act[ia][0];
    double* dw = dprm + woff(l);
    double* db = dw + (size_t)spec[l][1] * 9 * spec[l][0];
    const double* w = prm + woff(l);
    if (spec[l][2]) relu_bwd(t.g[oa], t.a[oa], t.n[oa]);
    conv_bwd(t.a[ia], spec[l][0], w, spec[l][1], N, H >> lv, W >> lv, t.g[oa], dw, db,
             l == 0 ? NULL : t.g[ia], emulate);
    /* glue before the layer, backwards */
    if (l == 12) upcat_bwd(t.g[18], 32, 16, t.g[17], t.g[2], N, H, W, emulate);
    if (l == 10) upcat_bwd(t.g[15], 64, 32, t.g[14], t.g[5], N, H >> 1, W >> 1, emulate);
    if (l == 8) upcat_bwd(t.g[12], 64, 64, t.g[11], t.g[8], N, H >> 2, W >> 2, emulate);
    if (l == 6) pool_bwd_acc(t.g[9], t.g[8], N, H >> 2, W >> 2, 64, emulate);
    if (l == 4) pool_bwd_acc(t.g[6], t.g[5], N, H >> 1, W >> 1, 32, emulate);
    if (l == 2) pool_bwd_acc(t.g[3], t.g[2], N, H, W, 16, emulate);
  }
  free_net(&t);
}
