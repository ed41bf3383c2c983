// CPU emulation of the LDS advection kernels (test only): a serial launch
// over blocks and threads, a window-bounds check on every unchecked LDS read,
// a check that every clamp-free sampling axis equals the clamped one, and the
// workgroup reach test evaluated serially.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

extern long emu_fast_blocks, emu_slow_blocks, emu_lds_reads, emu_axes;

// The height of the LDS windows of row length w the running kernel reads
// (each kernel's windows and ring have their own row lengths).
inline int& emu_win_h(int w) {
  static int h[256];
  if (w <= 0 || w >= 256) {
    fprintf(stderr, "window row length %d\n", w);
    abort();
  }
  return h[w];
}

inline void emu_inwin(int a, int b, int quad, int w) {
  emu_lds_reads++;
  if (a < 0 || b < 0 || a + quad >= w || b + quad >= emu_win_h(w)) {
    fprintf(stderr, "LDS read outside the window: a=%d b=%d quad=%d w=%d h=%d\n", a, b, quad, w,
            emu_win_h(w));
    abort();
  }
}

template <class A>
inline void emu_axis_check(const A& r, const A& c) {
  emu_axes++;
  if (r.i0 != c.i0 || !(r.f == c.f)) {
    fprintf(stderr, "clamp-free axis differs: i0 %d vs %d, f %g vs %g\n", r.i0, c.i0, r.f, c.f);
    abort();
  }
}

template <class K, class... A>
void emu_launch(K k, dim3 g, dim3 b, A... a) {
  gridDim = g;
  blockDim = b;
  for (unsigned by = 0; by < g.y; by++)
    for (unsigned bx = 0; bx < g.x; bx++) {
      blockIdx = dim3(bx, by);
      for (emu_phase = 0; emu_phase < 3; emu_phase++)
        for (unsigned ty = 0; ty < b.y; ty++)
          for (unsigned tx = 0; tx < b.x; tx++) {
            threadIdx = dim3(tx, ty);
            emu_barriers = 0;
            try {
              k(a...);
            } catch (EmuBarrier&) {
            }
          }
    }
}

// The workgroup's reach test over its two staged velocity windows of n
// elements each, of row length w; also sets the window height the LDS reads
// that follow are checked against (each kernel has its own tile height).
template <class L>
bool emu_within_reach(L lds, float c, float reach, int n, int w) {
  __syncthreads();
  emu_win_h(w) = n / w;
  float vw = 0.f;
  for (int q = 0; q < 2; q++)
    for (int e = 0; e < n; e++) {
      float x = fabsf(lds[q][e]);
      vw = (x > vw || x != x) ? x : vw;
    }
  bool fast = fabsf(c) * vw <= reach;
  if (threadIdx.x == 0 && threadIdx.y == 0) (fast ? emu_fast_blocks : emu_slow_blocks)++;
  return fast;
}
