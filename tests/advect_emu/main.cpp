// CPU emulation test driver (test only): each advection kernel on random
// fields and labels, LDS-staged against global gathers, over grid sizes that
// include tiles at the exact interior margin and CFL numbers below and above
// the reach limit.  Exits non-zero on any differing bit.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "emu.hpp"
#include "fh_batchdesc.hpp"
#include "fh_launch.hpp"

dim3 threadIdx, blockIdx, gridDim, blockDim;
int emu_phase = 0, emu_barriers = 0;
long emu_fast_blocks = 0, emu_slow_blocks = 0, emu_lds_reads = 0, emu_axes = 0;
static long one_pass_marked = 0;
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

// The batched launches over three scenes of uneven size (work list built
// largest scene last, offsets padded to multiples of 8) against each scene's
// own single-grid launch: the same bits.
static int batch_check() {
  int fails = 0;
  const int sizes[][2] = {{67, 45}, {130, 70}, {200, 33}};
  const float cs[] = {0.8f, 1.3f, 0.4f};
  constexpr int NS = 3;
  struct Scene {
    std::vector<float> u[2], v[2], uh, vh, q[2], qh;
    std::vector<uint8_t> lbl, sf, sd;
  } sc[NS];
  std::vector<SceneDesc> D(NS);
  std::vector<BStep> st(NS);
  for (int s = 0; s < NS; s++) {
    const int nx = sizes[s][0], ny = sizes[s][1];
    const size_t nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1), nc = (size_t)nx * ny;
    Scene& S = sc[s];
    S.u[0] = rnd(nu, 1.f), S.v[0] = rnd(nv, 1.f), S.u[1].assign(nu, 0.f), S.v[1].assign(nv, 0.f);
    S.uh.assign(nu, 0.f), S.vh.assign(nv, 0.f);
    S.q[0] = rnd(3 * nc, 1.f), S.q[1].assign(3 * nc, 0.f), S.qh.assign(3 * nc, 0.f);
    S.lbl.resize(nc);
    std::uniform_int_distribution<int> dl(0, 9);
    for (auto& l : S.lbl) {
      int r = dl(rng);
      l = r == 0 ? 2 : (r == 1 ? 1 : 0);
    }
    S.sf.assign((size_t)(nx + 1) * (ny + 1), 0xff), S.sd.assign(nc, 0xff);
    SceneDesc& d = D[s];
    memset(&d, 0, sizeof d);
    d.nx = nx, d.ny = ny;
    for (int k = 0; k < 2; k++)
      d.u[k] = S.u[k].data(), d.v[k] = S.v[k].data(), d.dye[k] = S.q[k].data();
    d.uh = S.uh.data(), d.vh = S.vh.data(), d.dyeh = S.qh.data();
    d.lbl = S.lbl.data(), d.sten_f = S.sf.data(), d.sten_d = S.sd.data();
    st[s] = BStep{};
    st[s].c = cs[s], st[s].k0 = 0.9f, st[s].k1 = 1.f, st[s].k2 = 0.7f;
  }
  auto list = [&](int which, std::vector<BSpan>& span, std::vector<uint8_t>& chunk) {
    span.assign(NS, BSpan{0, 0, 0, 0});
    chunk.clear();
    int off = 0;
    for (int s = NS - 1; s >= 0; s--) {
      const dim3 g = batch_grid_advect(which, sizes[s][0], sizes[s][1]);
      const int n = (int)(g.x * g.y), padded = (n + 7) & ~7;
      span[s] = BSpan{off, (int)g.x, (int)g.y, 0};
      chunk.insert(chunk.end(), padded / 8, (uint8_t)s);
      off += padded;
    }
    return BatchGrid{span.data(), chunk.data(), off};
  };
  std::vector<BSpan> sp[4];
  std::vector<uint8_t> ch[4];
  BatchGrid g[4];
  for (int k = 0; k < 4; k++) g[k] = list(k, sp[k], ch[k]);
  for (int mac = 0; mac < 2; mac++) {
    // batched: velocity [0] -> [1], then dye [0] -> [1] along velocity [1]
    launch_batch_advect_faces(nullptr, D.data(), g[kBAdvectFaces], st.data(), 0, mac);
    if (mac) launch_batch_maccormack_faces(nullptr, D.data(), g[kBMacCormackFaces], st.data(), 0);
    launch_batch_advect_dye(nullptr, D.data(), g[kBAdvectDye], st.data(), 0, mac);
    if (mac) launch_batch_maccormack_dye(nullptr, D.data(), g[kBMacCormackDye], st.data(), 0);
    for (int s = 0; s < NS; s++) {
      const int nx = sizes[s][0], ny = sizes[s][1];
      const size_t nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1), nc = (size_t)nx * ny;
      Scene& S = sc[s];
      const float c = cs[s];
      std::vector<float> a(nu), b(nv), ah(nu), bh(nv), o(3 * nc), oh(3 * nc);
      std::vector<uint8_t> sf((size_t)(nx + 1) * (ny + 1)), sd(nc);
      if (mac) {
        launch_advect_faces(nullptr, S.u[0].data(), S.v[0].data(), S.u[0].data(), S.v[0].data(),
                            ah.data(), bh.data(), S.lbl.data(), nx, ny, c, true, sf.data());
        launch_maccormack_faces(nullptr, S.u[0].data(), S.v[0].data(), ah.data(), bh.data(),
                                a.data(), b.data(), S.lbl.data(), nx, ny, c, true, sf.data());
        launch_advect_dye(nullptr, S.u[1].data(), S.v[1].data(), S.q[0].data(), oh.data(),
                          S.lbl.data(), nx, ny, 1, c, 1.f, 1.f, 1.f, true, sd.data());
        launch_maccormack_dye(nullptr, S.u[1].data(), S.v[1].data(), S.q[0].data(), oh.data(),
                              o.data(), S.lbl.data(), nx, ny, 1, c, 0.9f, 1.f, 0.7f, true,
                              sd.data());
      } else {
        launch_advect_faces(nullptr, S.u[0].data(), S.v[0].data(), S.u[0].data(), S.v[0].data(),
                            a.data(), b.data(), S.lbl.data(), nx, ny, c, true);
        launch_advect_dye(nullptr, S.u[1].data(), S.v[1].data(), S.q[0].data(), o.data(),
                          S.lbl.data(), nx, ny, 1, c, 0.9f, 1.f, 0.7f, true);
      }
      fails += cmp(mac ? "batch mac u" : "batch sl u", a.data(), S.u[1].data(), nu);
      fails += cmp(mac ? "batch mac v" : "batch sl v", b.data(), S.v[1].data(), nv);
      fails += cmp(mac ? "batch mac dye" : "batch sl dye", o.data(), S.q[1].data(), 3 * nc);
    }
  }
  return fails;
}

int main() {
  int fails = batch_check();
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
      // the one-pass dye kernel (and the two-pass kernels after it when a
      // tile exceeds the reach) against the two global-gather passes
      std::vector<float> g1(3 * nc), g2(3 * nc);
      launch_advect_dye(nullptr, u.data(), v.data(), q.data(), g1.data(), lbl.data(), nx, ny, 1,
                        c, 1.f, 1.f, 1.f, false);
      launch_maccormack_dye(nullptr, u.data(), v.data(), q.data(), g1.data(), g2.data(),
                            lbl.data(), nx, ny, 1, c, 0.9f, 1.f, 0.7f, false);
      static int fail = 0, epoch = 0;
      ++epoch;
      std::vector<float> h1(3 * nc), f1(3 * nc);
      std::vector<uint8_t> sd1(nc, 0xff);
      launch_maccormack_dye_fused(nullptr, u.data(), v.data(), q.data(), f1.data(), lbl.data(), nx,
                                  ny, c, 0.9f, 1.f, 0.7f, nullptr, &fail, epoch);
      launch_maccormack_dye_rerun(nullptr, u.data(), v.data(), q.data(), h1.data(), f1.data(),
                                  lbl.data(), nx, ny, c, 0.9f, 1.f, 0.7f, sd1.data(), &fail, epoch);
      fails += cmp("maccormack_dye one pass", g2.data(), f1.data(), 3 * nc);
      if (fail == epoch) one_pass_marked++;
      // the one-pass velocity kernel against the two passes, without
      // buoyancy (against the global gathers) and with buoyancy from the
      // first dye channel (against the staged passes, the kernels that
      // apply it)
      for (int buoy = 0; buoy < 2; buoy++) {
      const float* heat = buoy ? q.data() : nullptr;
      std::vector<float> ga(nu), gb(nv), gha(nu), ghb(nv), fa1(nu), fb1(nv), ha1(nu), hb1(nv);
      std::vector<uint8_t> gsf((size_t)(nx + 1) * (ny + 1), 0xff);
      launch_advect_faces(nullptr, u.data(), v.data(), u.data(), v.data(), gha.data(), ghb.data(),
                          lbl.data(), nx, ny, c, buoy != 0, buoy ? gsf.data() : nullptr);
      launch_maccormack_faces(nullptr, u.data(), v.data(), gha.data(), ghb.data(), ga.data(),
                              gb.data(), lbl.data(), nx, ny, c, buoy != 0,
                              buoy ? gsf.data() : nullptr, heat, 0.3f);
      ++epoch;
      std::vector<uint8_t> sf1((size_t)(nx + 1) * (ny + 1), 0xff);
      launch_maccormack_faces_fused(nullptr, u.data(), v.data(), fa1.data(), fb1.data(),
                                    lbl.data(), nx, ny, c, heat, 0.3f, &fail, epoch);
      launch_maccormack_faces_rerun(nullptr, u.data(), v.data(), ha1.data(), hb1.data(),
                                    fa1.data(), fb1.data(), lbl.data(), nx, ny, c, sf1.data(),
                                    heat, 0.3f, &fail, epoch);
      fails += cmp("maccormack_faces one pass u", ga.data(), fa1.data(), nu);
      fails += cmp("maccormack_faces one pass v", gb.data(), fb1.data(), nv);
      }
    }
  printf("staged blocks %ld, global blocks %ld, checked LDS reads %ld, clamp-free axes %ld\n",
         emu_fast_blocks, emu_slow_blocks, emu_lds_reads, emu_axes);
  printf("one-pass dye launches marked over the reach %ld\n", one_pass_marked);
  printf(fails ? "FAIL\n" : "ALL BIT-IDENTICAL\n");
  return fails != 0;
}
