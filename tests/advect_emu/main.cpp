// CPU emulation test driver (test only): each advection kernel on random
// fields and labels, LDS-staged against global gathers, over grid sizes that
// include tiles at the exact interior margin and CFL numbers below and above
// the reach limit.  Exits non-zero on any differing bit.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "emu.hpp"
#include "fh_launch.hpp"

dim3 threadIdx, blockIdx, gridDim, blockDim;
int emu_phase = 0;
long emu_fast_blocks = 0, emu_slow_blocks = 0, emu_lds_reads = 0, emu_axes = 0;
using namespace fh;

static std::mt19937 rng(7);

static std::vector<float> rnd(size_t n, float a) {
  std::uniform_real_distribution<float> d(-a, a);
  std::vector<float> v(n);
  for (auto& x : v) x = d(rng);
  return v;
}

static int cmp(const char* what, const float* a, const float* b, size_t n) {
  size_t bad = 0;
  for (size_t i = 0; i < n; i++) bad += memcmp(&a[i], &b[i], 4) != 0;
  if (bad) printf("  %-20s %zu / %zu differ\n", what, bad, n);
  return bad != 0;
}

int main() {
  int fails = 0;
  const int sizes[][2] = {{67, 45}, {130, 70}, {64, 16}, {200, 33}, {300, 100}, {199, 52}, {195, 51}};
  const float cs[] = {0.3f, 1.0f, 1.249f, 1.3f, 3.7f};  // speeds up to 1: displacement up to c cells
  for (auto& sz : sizes)
    for (float c : cs) {
      const int nx = sz[0], ny = sz[1];
      const size_t nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1), nc = (size_t)nx * ny;
      auto u = rnd(nu, 1.f), v = rnd(nv, 1.f), uh = rnd(nu, 1.f), vh = rnd(nv, 1.f);
      auto q = rnd(3 * nc, 1.f), qh = rnd(3 * nc, 1.f);
      std::vector<uint8_t> lbl(nc);
      std::uniform_int_distribution<int> dl(0, 9);
      for (auto& l : lbl) {
        int r = dl(rng);
        l = r == 0 ? 2 : (r == 1 ? 1 : 0);
      }
      std::vector<float> a0(nu), b0(nv), a1(nu), b1(nv), o0(3 * nc), o1(3 * nc);
      launch_advect_faces(nullptr, u.data(), v.data(), u.data(), v.data(), a0.data(), b0.data(),
                          lbl.data(), nx, ny, c, false);
      launch_advect_faces(nullptr, u.data(), v.data(), u.data(), v.data(), a1.data(), b1.data(),
                          lbl.data(), nx, ny, c, true);
      fails += cmp("advect_faces u", a0.data(), a1.data(), nu);
      fails += cmp("advect_faces v", b0.data(), b1.data(), nv);
      launch_maccormack_faces(nullptr, u.data(), v.data(), uh.data(), vh.data(), a0.data(),
                              b0.data(), lbl.data(), nx, ny, c, false);
      launch_maccormack_faces(nullptr, u.data(), v.data(), uh.data(), vh.data(), a1.data(),
                              b1.data(), lbl.data(), nx, ny, c, true);
      fails += cmp("maccormack_faces u", a0.data(), a1.data(), nu);
      fails += cmp("maccormack_faces v", b0.data(), b1.data(), nv);
      // the correction reading the forward pass's recorded limiter stencils
      std::vector<uint8_t> sf((size_t)(nx + 1) * (ny + 1), 0xff);
      std::vector<float> fa(nu), fb(nv);
      launch_advect_faces(nullptr, u.data(), v.data(), u.data(), v.data(), fa.data(), fb.data(),
                          lbl.data(), nx, ny, c, true, sf.data());
      launch_maccormack_faces(nullptr, u.data(), v.data(), uh.data(), vh.data(), a1.data(),
                              b1.data(), lbl.data(), nx, ny, c, true, sf.data());
      fails += cmp("maccormack_faces u sten", a0.data(), a1.data(), nu);
      fails += cmp("maccormack_faces v sten", b0.data(), b1.data(), nv);
      launch_advect_dye(nullptr, u.data(), v.data(), q.data(), o0.data(), lbl.data(), nx, ny, 1, c,
                        0.9f, 1.f, 0.7f, false);
      launch_advect_dye(nullptr, u.data(), v.data(), q.data(), o1.data(), lbl.data(), nx, ny, 1, c,
                        0.9f, 1.f, 0.7f, true);
      fails += cmp("advect_dye", o0.data(), o1.data(), 3 * nc);
      launch_maccormack_dye(nullptr, u.data(), v.data(), q.data(), qh.data(), o0.data(),
                            lbl.data(), nx, ny, 1, c, 0.9f, 1.f, 0.7f, false);
      launch_maccormack_dye(nullptr, u.data(), v.data(), q.data(), qh.data(), o1.data(),
                            lbl.data(), nx, ny, 1, c, 0.9f, 1.f, 0.7f, true);
      fails += cmp("maccormack_dye", o0.data(), o1.data(), 3 * nc);
      std::vector<uint8_t> sd(nc, 0xff);
      std::vector<float> fo(3 * nc);
      launch_advect_dye(nullptr, u.data(), v.data(), q.data(), fo.data(), lbl.data(), nx, ny, 1, c,
                        1.f, 1.f, 1.f, true, sd.data());
      launch_maccormack_dye(nullptr, u.data(), v.data(), q.data(), qh.data(), o1.data(),
                            lbl.data(), nx, ny, 1, c, 0.9f, 1.f, 0.7f, true, sd.data());
      fails += cmp("maccormack_dye sten", o0.data(), o1.data(), 3 * nc);
    }
  printf("staged blocks %ld, global blocks %ld, checked LDS reads %ld, clamp-free axes %ld\n",
         emu_fast_blocks, emu_slow_blocks, emu_lds_reads, emu_axes);
  printf(fails ? "FAIL\n" : "ALL BIT-IDENTICAL\n");
  return fails != 0;
}
