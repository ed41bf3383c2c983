// libflowhip engine: device state of one MAC grid and the C ABI of
// include/flowhip.h.  One engine = one grid on one HIP stream; nothing here
// synchronises except where a value must reach the host (solver
// convergence checks, readbacks, timings).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flowhip.h"
#include "fh_device.hpp"
#include "fh_engine_impl.hpp"
#include "fh_launch.hpp"
#include "fh_mgtile.hpp"

namespace fh {
// fh_unet.hip
int engine_capture_hook(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                        const double* p_prev, const double* p_final, long n, int H, int W,
                        bool begin);
int engine_warm_start(fh_unet* u, hipStream_t s, const float* r32, const uint8_t* lbl,
                      const uint8_t* info, double* p, int H, int W, float* stats2);
}  // namespace fh

using namespace fh;

namespace {

thread_local std::string g_err;

void set_err(const std::string& s) { g_err = s; }

// pressureMultigrid.ts:65-78 with a configurable coarsest side.
int level_count(int nx, int ny, int coarsest) {
  int n = 1;
  while (nx > coarsest || ny > coarsest) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
    n++;
  }
  return n;
}

int fail_hip(fh_engine* e, hipError_t err, const char* what) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(err));
  set_err(buf);
  if (e && (err == hipErrorIllegalAddress || err == hipErrorLaunchFailure ||
            err == hipErrorECCNotCorrectable || err == hipErrorAssert))
    e->lost = true;
  return e && e->lost ? FH_ERR_DEVICE_LOST : FH_ERR_HIP;
}

#define HIPCHK(call)                                    \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return fail_hip(e, _e, #call); \
  } while (0)

// Reentrancy / device-loss guard around every engine entry point.
struct Guard {
  fh_engine* e;
  int status = FH_OK;
  Guard(fh_engine* eng, const char* what) : e(eng) {
    if (!e) {
      set_err(std::string(what) + ": null engine");
      status = FH_ERR_BAD_ARG;
    } else if (e->lost) {
      set_err("GPU device lost");
      status = FH_ERR_DEVICE_LOST;
    } else {
      bool expected = false;
      if (!e->busy.compare_exchange_strong(expected, true)) {
        set_err(std::string(what) + " is not re-entrant");
        status = FH_ERR_BUSY;
      }
    }
  }
  ~Guard() {
    if (status == FH_OK) e->busy.store(false);
  }
};

#define GUARD(name)            \
  Guard _g(e, name);           \
  if (_g.status != FH_OK) return _g.status

int check_launch(fh_engine* e, const char* what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail_hip(e, err, what);
  return FH_OK;
}

#define LAUNCHED(what)                       \
  do {                                       \
    int _st = check_launch(e, what);         \
    if (_st != FH_OK) return _st;            \
  } while (0)

size_t n_u(const fh_engine* e) { return (size_t)(e->nx + 1) * e->ny; }
size_t n_v(const fh_engine* e) { return (size_t)e->nx * (e->ny + 1); }
size_t n_c(const fh_engine* e) { return (size_t)e->nx * e->ny; }
size_t n_d(const fh_engine* e) { return (size_t)e->dnx * e->dny; }

constexpr int kTraceMax = 8192;

// Brackets one launch of the traced kernel with a pair of events.
struct Trace {
  fh_engine* e;
  bool on;
  Trace(fh_engine* eng, int kid) : e(eng), on(eng->trace_kid == kid && eng->trace_n < kTraceMax) {
    if (!on) return;
    size_t need = 2 * (size_t)(e->trace_n + 1);
    while (e->trace_ev.size() < need) {
      hipEvent_t ev;
      if (hipEventCreate(&ev) != hipSuccess) {
        on = false;
        return;
      }
      e->trace_ev.push_back(ev);
    }
    (void)hipEventRecord(e->trace_ev[2 * e->trace_n], e->s);
  }
  ~Trace() {
    if (!on) return;
    (void)hipEventRecord(e->trace_ev[2 * e->trace_n + 1], e->s);
    e->trace_n++;
  }
};


template <class T>
int dalloc(fh_engine* e, T** p, size_t n) {
  hipError_t err = hipMalloc((void**)p, n * sizeof(T) + 256);
  if (err != hipSuccess) return fail_hip(e, err, "hipMalloc");
  err = hipMemset(*p, 0, n * sizeof(T) + 256);
  if (err != hipSuccess) return fail_hip(e, err, "hipMemset");
  return FH_OK;
}

// The pinned staging buffer, sized for the largest field (u, v or 3 dye planes).
int ensure_stage(fh_engine* e, size_t n) {
  if (e->stage && n <= e->stage_n) return FH_OK;
  size_t want = std::max(std::max(n_u(e), n_v(e)), 3 * n_d(e));
  want = std::max(want, n);
  if (e->stage) HIPCHK(hipHostFree(e->stage));
  e->stage = nullptr;
  e->stage_n = 0;
  HIPCHK(hipHostMalloc((void**)&e->stage, want * sizeof(float), hipHostMallocDefault));
  e->stage_n = want;
  return FH_OK;
}

// f64 host -> f32 device
int up32(fh_engine* e, float* dst, const double* src, size_t n) {
  int st = ensure_stage(e, n);
  if (st != FH_OK) return st;
  for (size_t k = 0; k < n; k++) e->stage[k] = (float)src[k];
  HIPCHK(hipMemcpyAsync(dst, e->stage, n * sizeof(float), hipMemcpyHostToDevice, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  return FH_OK;
}

int down32(fh_engine* e, double* dst, const float* src, size_t n) {
  int st = ensure_stage(e, n);
  if (st != FH_OK) return st;
  HIPCHK(hipMemcpyAsync(e->stage, src, n * sizeof(float), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  for (size_t k = 0; k < n; k++) dst[k] = (double)e->stage[k];
  return FH_OK;
}

template <class T>
void swap_ptr(T*& a, T*& b) {
  T* t = a;
  a = b;
  b = t;
}

}  // namespace

namespace fh {

// ---- coarsest-level pseudo-inverse (host, FP64) ---------------------------
// Unknowns: fluid cells with at least one non-solid neighbour.  A fluid
// component that touches no Air is pure Neumann; it is regularised by
// (1/n) 1 1^T and the component mean is projected out of the result, so
// M b is the minimum-norm least-squares solution A^+ b on every component.
void build_pseudo_inverse(int nx, int ny, const std::vector<uint8_t>& lab, std::vector<double>& M,
                          std::vector<int>& cells) {
  auto so = [&](int i, int j) {
    return i < 0 || j < 0 || i >= nx || j >= ny || lab[i + j * nx] == SOLID;
  };
  const int di[4] = {1, -1, 0, 0}, dj[4] = {0, 0, 1, -1};
  std::vector<int> map(nx * ny, -1);
  cells.clear();
  for (int k = 0; k < nx * ny; k++) {
    int i = k % nx, j = k / nx;
    if (lab[k] != FLUID) continue;
    int cnt = 0;
    for (int t = 0; t < 4; t++) cnt += !so(i + di[t], j + dj[t]);
    if (cnt > 0) {
      map[k] = (int)cells.size();
      cells.push_back(k);
    }
  }
  int n = (int)cells.size();
  std::vector<int> comp(n, -1), stack;
  std::vector<int> neumann, size;
  for (int a = 0; a < n; a++) {
    if (comp[a] >= 0) continue;
    int c = (int)neumann.size();
    int anchored = 0, sz = 0;
    stack.assign(1, a);
    comp[a] = c;
    while (!stack.empty()) {
      int b = stack.back();
      stack.pop_back();
      sz++;
      int k = cells[b], i = k % nx, j = k / nx;
      for (int t = 0; t < 4; t++) {
        int ii = i + di[t], jj = j + dj[t];
        if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
        int kk = ii + jj * nx;
        if (lab[kk] == AIR) anchored = 1;
        int m = map[kk];
        if (m >= 0 && comp[m] < 0) {
          comp[m] = c;
          stack.push_back(m);
        }
      }
    }
    neumann.push_back(!anchored);
    size.push_back(sz);
  }
  std::vector<double> A((size_t)n * n, 0.0);
  for (int a = 0; a < n; a++) {
    int k = cells[a], i = k % nx, j = k / nx;
    for (int t = 0; t < 4; t++) {
      int ii = i + di[t], jj = j + dj[t];
      if (so(ii, jj)) continue;
      A[(size_t)a * n + a] += 1.0;
      int m = map[ii + jj * nx];
      if (m >= 0) A[(size_t)a * n + m] -= 1.0;
    }
  }
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++)
      if (comp[a] == comp[b] && neumann[comp[a]]) A[(size_t)a * n + b] += 1.0 / size[comp[a]];
  // Cholesky A = L L^T (lower triangle in place)
  for (int c = 0; c < n; c++) {
    double sdiag = A[(size_t)c * n + c];
    for (int t = 0; t < c; t++) sdiag -= A[(size_t)c * n + t] * A[(size_t)c * n + t];
    double d = std::sqrt(sdiag);
    A[(size_t)c * n + c] = d;
    for (int r = c + 1; r < n; r++) {
      double q = A[(size_t)r * n + c];
      for (int t = 0; t < c; t++) q -= A[(size_t)r * n + t] * A[(size_t)c * n + t];
      A[(size_t)r * n + c] = q / d;
    }
  }
  // inverse column by column
  M.assign((size_t)n * n, 0.0);
  std::vector<double> y(n);
  for (int col = 0; col < n; col++) {
    for (int r = 0; r < n; r++) {
      double sacc = (r == col) ? 1.0 : 0.0;
      for (int t = 0; t < r; t++) sacc -= A[(size_t)r * n + t] * y[t];
      y[r] = sacc / A[(size_t)r * n + r];
    }
    for (int r = n - 1; r >= 0; r--) {
      double sacc = y[r];
      for (int t = r + 1; t < n; t++) sacc -= A[(size_t)t * n + r] * y[t];
      y[r] = sacc / A[(size_t)r * n + r];
    }
    for (int r = 0; r < n; r++) M[(size_t)r * n + col] = y[r];
  }
  // project out component means on Neumann components
  int ncomp = (int)neumann.size();
  std::vector<double> mean(ncomp);
  for (int col = 0; col < n; col++) {
    std::fill(mean.begin(), mean.end(), 0.0);
    for (int a = 0; a < n; a++)
      if (neumann[comp[a]]) mean[comp[a]] += M[(size_t)a * n + col];
    for (int a = 0; a < n; a++)
      if (neumann[comp[a]]) M[(size_t)a * n + col] -= mean[comp[a]] / size[comp[a]];
  }
}

// Forgets the pressure history of the extrapolated warm start: the resident
// p is no longer the last fh_step solve's result.
void drop_pressure_history(fh_engine* e) {
  e->p_dt = 0.0;
  e->ph_n = 0;
}

// Lagrange weights of the extrapolation to tau after the newest pressure:
// nodes 0 (p), -a (h0, a = dt of the step that produced p) and -(a + b)
// (h1, b = dt of the step that produced h0); order = nodes used - 1.
void extrapolation_weights(int order, double tau, double a, double b, double w[3]) {
  w[0] = 1.0;
  w[1] = w[2] = 0.0;
  if (order == 1) {
    w[0] = 1.0 + tau / a;
    w[1] = -tau / a;
  } else if (order == 2) {
    w[0] = (tau + a) * (tau + a + b) / (a * (a + b));
    w[1] = -tau * (tau + a + b) / (a * b);
    w[2] = tau * (tau + a) / (b * (a + b));
  }
}

StartPlan plan_start(fh_engine* e, double dt) {
  StartPlan sp;
  if (e->extrap <= 0 || !(dt > 0.0) || !(e->p_dt > 0.0)) return sp;
  sp.on = true;
  sp.order = std::min(e->ph_n, e->extrap);
  double w[3];
  extrapolation_weights(sp.order, dt, e->p_dt, e->ph_dt[0], w);
  sp.xp = Extrapolation{e->ph[0], e->ph[1], w[0], w[1], w[2]};
  return sp;
}

// After the start is in p2: it becomes p, the old p (the last solution)
// becomes the newest history entry, and the buffer of the entry that falls
// out of the history becomes p2.
void rotate_history(fh_engine* e) {
  double* last = e->p;
  e->p = e->p2;
  if (e->extrap >= 2) {
    e->p2 = e->ph[1];
    e->ph[1] = e->ph[0];
    e->ph_dt[1] = e->ph_dt[0];
  } else {
    e->p2 = e->ph[0];
  }
  e->ph[0] = last;
  e->ph_dt[0] = e->p_dt;
  e->ph_n = std::min(e->ph_n + 1, e->extrap);
}

}  // namespace fh

namespace {

// Coarsen the label chain and rebuild the coarsest-level operator; zero
// every multigrid unknown so no stale value from an old geometry survives.
int refresh_levels(fh_engine* e) {
  for (size_t l = 1; l < e->lv.size(); l++) {
    launch_coarsen_labels(e->s, e->lv[l - 1].lbl, e->lv[l - 1].nx, e->lv[l - 1].ny, e->lv[l].lbl,
                          e->lv[l].nx, e->lv[l].ny);
    LAUNCHED("coarsen_labels");
  }
  for (size_t l = 0; l < e->lv.size(); l++) {
    Level& L = e->lv[l];
    launch_cell_info(e->s, L.lbl, L.nx, L.ny, L.info);
    LAUNCHED("cell_info");
    if ((int)l >= e->cinfo_from) {
      launch_pad_info(e->s, L.info, L.nx, L.ny,
                      e->cinfo + (e->cinfo_off[l] - e->cinfo_off[e->cinfo_from]));
      LAUNCHED("pad_info");
    }
    HIPCHK(hipMemsetAsync(L.x, 0, (size_t)L.nx * L.ny * sizeof(float), e->s));
    HIPCHK(hipMemsetAsync(L.xa, 0, (size_t)L.nx * L.ny * sizeof(float), e->s));
  }
  Level& C = e->lv[e->nlev_dc - 1];
  std::vector<uint8_t> lab((size_t)C.nx * C.ny);
  HIPCHK(hipMemcpyAsync(lab.data(), C.lbl, lab.size(), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  std::vector<double> M;
  std::vector<int> cells;
  build_pseudo_inverse(C.nx, C.ny, lab, M, cells);
  int n = (int)cells.size();
  if (n > kMaxDirect) {
    set_err("coarsest level too large for the direct solve");
    return FH_ERR_BAD_ARG;
  }
  if (e->M) HIPCHK(hipFree(e->M));
  if (e->cells) HIPCHK(hipFree(e->cells));
  e->M = nullptr;
  e->cells = nullptr;
  e->ncoarse = n;
  if (n > 0) {
    HIPCHK(hipMalloc((void**)&e->M, M.size() * sizeof(double)));
    HIPCHK(hipMalloc((void**)&e->cells, cells.size() * sizeof(int)));
    HIPCHK(hipMemcpy(e->M, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  e->labels_dirty = false;
  return FH_OK;
}

int read_sums(fh_engine* e) {
  HIPCHK(hipMemcpyAsync(e->h_sums, e->sums, 6 * sizeof(double), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  return FH_OK;
}

void rb(fh_engine* e, int l, int color, float omega, float scale, int first) {
  Level& L = e->lv[l];
  Trace t(e, l == 0 ? FH_K_RB_SWEEP_FINE : -1);
  launch_rb_sweep(e->s, L.x, L.b, L.lbl, L.nx, L.ny, color, omega, scale, first);
}

CoarseLevels coarse_levels(const fh_engine* e, int lc, int nlev) {
  CoarseLevels C{};
  C.n = nlev - lc;
  for (int k = 0; k < C.n; k++) {
    C.nx[k] = e->lv[lc + k].nx;
    C.ny[k] = e->lv[lc + k].ny;
  }
  coarse_levels_layout(C);
  // sizing before the padded info exists (coarse_start) needs no image
  C.image = (e->cinfo && lc >= e->cinfo_from)
                ? e->cinfo + (e->cinfo_off[lc] - e->cinfo_off[e->cinfo_from])
                : nullptr;
  return C;
}

// First level from which the rest of the hierarchy fits one workgroup's LDS
// (and its sweeps, at most kCoarseSweepCells cells of a colour per level).
int coarse_start(const fh_engine* e, int nlev) {
  for (int l = 0; l < nlev; l++)
    if (nlev - l <= kMaxCoarse && (long)((e->lv[l].nx + 1) / 2) * e->lv[l].ny <= kCoarseSweepCells &&
        coarse_vcycle_lds_bytes(coarse_levels(e, l, nlev)) <= kCoarseLds)
      return l;
  return nlev - 1;
}

// One V(2,2) cycle (pressureMultigrid.ts:253-275) on levels [0, nlev):
// levels below lc with the LDS-tiled pre-smooth+restrict and
// prolong+post-smooth kernels, levels from lc in one workgroup.  warm: level
// 0 starts from its stored x (the reference mode's pressure warm start);
// otherwise every level's unknown starts at zero.
int vcycle(fh_engine* e, int nlev, bool dc, bool warm, const int* done,
           bool skip_fine_prolong = false) {
  int lc = dc ? e->lc_dc : e->lc_ref;
  for (int l = 0; l < lc; l++) {
    Level& L = e->lv[l];
    Level& N = e->lv[l + 1];
    bool zero = !(l == 0 && warm);
    Trace t(e, l == 0 ? FH_K_SMOOTH_RESTRICT_FINE : -1);
    launch_smooth_restrict(e->s, e->tile[l], zero ? nullptr : L.x, L.xa, L.b, L.info, L.nx, L.ny,
                           N.b, N.nx, N.ny, zero ? 1 : 0, done);
  }
  LAUNCHED("smooth_restrict");
  {
    Trace t(e, FH_K_COARSE_VCYCLE);
    launch_coarse_vcycle(e->s, e->coarse_group, coarse_levels(e, lc, nlev), e->lv[lc].b,
                         (lc == 0 && warm) ? e->lv[0].x : nullptr, e->lv[lc].x,
                         dc ? e->M : nullptr, dc ? e->cells : nullptr, dc ? e->ncoarse : 0,
                         dc ? 1 : 0, done);
  }
  LAUNCHED("coarse_vcycle");
  for (int l = lc - 1; l >= (skip_fine_prolong ? 1 : 0); l--) {
    Level& L = e->lv[l];
    Level& N = e->lv[l + 1];
    Trace t(e, l == 0 ? FH_K_PROLONG_SMOOTH_FINE : -1);
    launch_prolong_smooth(e->s, e->tile[l], L.xa, L.x, L.b, L.info, L.nx, L.ny, N.x, N.nx, N.ny,
                          done);
  }
  LAUNCHED("prolong_smooth");
  return FH_OK;
}

// One defect-correction cycle c: V-cycle on the correction, p += e with the
// new residual (also to r64 when given) and its norm (to host slot c & 1),
// then cycle event c & 1.
int enqueue_dc_cycle(fh_engine* e, int c, double scale, double tol, double* r64 = nullptr) {
  Level& L0 = e->lv[0];
  // the finest level's prolongation fused with the update when it is tiled
  // and its tiles' partials fit the reduction buffer
  const bool fused = e->fuse_update && e->lc_dc >= 1 &&
                     (size_t)prolong_smooth_dc_tiles(e->tile[0], e->nx, e->ny) <= e->partials_n;
  int st = vcycle(e, e->nlev_dc, true, false, e->done, fused);
  if (st != FH_OK) return st;
  if (fused) {
    const Level& L1 = e->lv[1];
    int n;
    {
      Trace t(e, FH_K_DC_UPDATE);
      n = launch_prolong_smooth_dc(e->s, e->tile[0], L0.xa, L0.b, L0.info, L0.nx, L0.ny, L1.x,
                                   L1.nx, L1.ny, e->done, e->p, e->p2, e->div, e->b2, r64, scale,
                                   e->partials);
    }
    launch_dc_norm_n(e->s, e->partials, n, e->sums, e->d_res + 8 * (c & 1), e->done, scale, tol);
    LAUNCHED("prolong_smooth_dc");
    // the new residual is the next cycle's right-hand side
    swap_ptr(L0.b, e->b2);
  }
  e->last_dc_fused = fused;
  if (!fused) {
    {
      Trace t(e, FH_K_DC_UPDATE);
      launch_dc_correct(e->s, e->p, e->p2, L0.x, e->div, L0.info, L0.b, e->nx, e->ny, scale,
                        e->partials, e->done, r64);
    }
    launch_dc_norm(e->s, e->partials, e->nx, e->ny, e->sums, e->d_res + 8 * (c & 1), e->done,
                   scale, tol);
    LAUNCHED("dc_update");
  }
  swap_ptr(e->p, e->p2);
  HIPCHK(hipEventRecord(e->cyc_ev[c & 1], e->s));
  return FH_OK;
}

// One flexible-CG cycle c >= 1 (cycle 0 of the solve is a defect-correction
// cycle): z = V-cycle(r32), then the CG step (p and r in place) and its norm
// (to host slot c & 1), then cycle event c & 1.
int enqueue_fcg_cycle(fh_engine* e, int c, double scale, double tol) {
  Level& L0 = e->lv[0];
  int st = vcycle(e, e->nlev_dc, true, false, e->done);
  if (st != FH_OK) return st;
  const int first = c == 1 ? 1 : 0;
  {
    Trace t(e, FH_K_FCG_DOTS);
    launch_fcg_dots(e->s, L0.x, e->dvec[c & 1], e->rvec, L0.info, e->nx, e->ny, first,
                    e->partials, e->fcg, e->done);
  }
  {
    Trace t(e, FH_K_FCG_UPDATE);
    launch_fcg_update(e->s, L0.x, e->dvec[c & 1], e->dvec[(c + 1) & 1], e->p, e->rvec, L0.info,
                      L0.b, e->nx, e->ny, first, e->fcg, e->partials, e->done);
  }
  launch_fcg_norm(e->s, e->partials, e->nx, e->ny, e->sums, e->sums, e->d_res + 8 * (c & 1),
                  e->done, scale, tol);
  LAUNCHED("fcg_step");
  HIPCHK(hipEventRecord(e->cyc_ev[c & 1], e->s));
  return FH_OK;
}

int ensure_fcg(fh_engine* e) {
  if (e->rvec) return FH_OK;
  const size_t n = (size_t)e->nx * e->ny;
  int st;
  if ((st = dalloc(e, &e->rvec, n)) || (st = dalloc(e, &e->dvec[0], n)) ||
      (st = dalloc(e, &e->dvec[1], n)) || (st = dalloc(e, &e->fcg, (size_t)16)))
    return st;
  return FH_OK;
}

double rel_resid(const double* s, double scale, bool scaled_r) {
  // sums = {sum r^2, n_r, sum div^2, n_fluid}; the defect-correction kernel
  // reports r in code units (b - A p), the norms kernel in divergence units.
  if (s[1] <= 0 || s[2] <= 0) return 0.0;
  double rr = std::sqrt(s[0] / s[1]);
  if (scaled_r) rr /= std::fabs(scale);
  return rr / std::sqrt(s[2] / s[3]);
}

// PressureSolver.solve on the resident p (FP64), div (FP64), label.
// div_fused: the divergence has not been formed; the initial defect pass of
// the FP64 solvers forms it from the resident velocity.  sp: the start.
int solve_impl(fh_engine* e, int solver, double scale, int iterations, double omega,
               double tol, int* out_iters, bool div_fused, const StartPlan& sp);

// Every solve is bracketed by two events: fh_last_solve_ms reports its
// device time (label refresh, warm start and all cycles included).
// step_dt > 0: the solve of an fh_step of that step size (its start may be
// extrapolated in time, and its result joins the pressure history); 0: any
// other solve, which ends the history.
int solve(fh_engine* e, int solver, double scale, int iterations, double omega, double tol,
          int* out_iters, bool div_fused = false, double step_dt = 0.0) {
  e->solve_ev_valid = false;
  if (!e->solve_ev[0]) {
    HIPCHK(hipEventCreate(&e->solve_ev[0]));
    HIPCHK(hipEventCreate(&e->solve_ev[1]));
  }
  HIPCHK(hipEventRecord(e->solve_ev[0], e->s));
  const StartPlan sp = plan_start(e, step_dt);
  if (!sp.on) e->ph_n = 0;
  int st = solve_impl(e, solver, scale, iterations, omega, tol, out_iters, div_fused, sp);
  if (st != FH_OK) {
    drop_pressure_history(e);
    return st;
  }
  e->p_dt = step_dt;
  HIPCHK(hipEventRecord(e->solve_ev[1], e->s));
  e->solve_ev_valid = true;
  return FH_OK;
}

int solve_impl(fh_engine* e, int solver, double scale, int iterations, double omega,
               double tol, int* out_iters, bool div_fused, const StartPlan& sp) {
  int N = e->nx * e->ny;
  if (e->labels_dirty) {
    int st = refresh_levels(e);
    if (st != FH_OK) return st;
  }
  e->hist.clear();
  Level& L0 = e->lv[0];
  const bool fp64 = solver == FH_SOLVER_MG_DC || solver == FH_SOLVER_MG_FCG;
  // a training capture reads the initial pass on the host; everything else
  // (the learned start included) runs on the device behind the norm kernel
  const bool sync_init = fp64 && e->capture;
  // The extrapolated start: formed by the FP64 solvers' fused initial pass,
  // otherwise by a pass of its own before anything reads p.
  const bool start_in_pass = sp.on && fp64 && div_fused && !sync_init;
  if (sp.on && !start_in_pass) {
    launch_extrapolate(e->s, e->p, sp.xp, sp.order, e->p2, N);
    LAUNCHED("extrapolate");
    rotate_history(e);
  }
  if (fp64) {
    const bool fcg = solver == FH_SOLVER_MG_FCG;
    if (fcg) {
      int st0 = ensure_fcg(e);
      if (st0 != FH_OK) return st0;
    }
    // CG keeps the FP64 residual; the passes before the first cycle need not
    // store it, since that cycle is a defect-correction cycle, whose update
    // pass stores the residual of its result before any CG step reads it
    double* r64 = fcg ? e->rvec : nullptr;
    // r = b - A p (no update) + norms.  The zero-divergence guard of
    // pressureMultigrid.ts:294-300 reads the same pass: on the device (the
    // norm sets `done`, so the cycles enqueued behind it skip themselves)
    // unless a training capture needs the norm on the host now.
    int st;
    if (sync_init) {
      launch_set_int(e->s, e->done, 0);
      launch_dc_correct(e->s, e->p, nullptr, nullptr, e->div, L0.info, L0.b, e->nx, e->ny, scale,
                        e->partials, nullptr, nullptr);
      launch_dc_norm(e->s, e->partials, e->nx, e->ny, e->sums, nullptr, nullptr, scale, 0.0);
      LAUNCHED("dc_update");
      st = read_sums(e);
      if (st != FH_OK) return st;
      if (e->h_sums[2] == 0.0) {
        HIPCHK(hipMemsetAsync(e->p, 0, N * sizeof(double), e->s));
        *out_iters = 0;
        return FH_OK;
      }
      if (e->capture) {
        HIPCHK(hipMemcpyAsync(e->psave, e->p, N * sizeof(double), hipMemcpyDeviceToDevice, e->s));
        st = engine_capture_hook(e->capture, e->s, L0.b, e->lbl, e->psave, nullptr, N, e->ny,
                                 e->nx, true);
        if (st != FH_OK) return st;
      }
      if (e->warm) {
        // learned warm start: p += net correction, then the defect again
        st = engine_warm_start(e->warm, e->s, L0.b, e->lbl, L0.info, e->p, e->ny, e->nx,
                               e->wstats);
        if (st != FH_OK) return st;
        launch_dc_correct(e->s, e->p, nullptr, nullptr, e->div, L0.info, L0.b, e->nx, e->ny,
                          scale, e->partials, nullptr, nullptr);
        launch_dc_norm(e->s, e->partials, e->nx, e->ny, e->sums, nullptr, nullptr, scale, 0.0);
        LAUNCHED("dc_update");
        st = read_sums(e);
        if (st != FH_OK) return st;
      }
      e->hist.push_back(rel_resid(e->h_sums, scale, true));
    } else {
      if (div_fused) {
        {
          Trace t(e, FH_K_DIVERGENCE);
          launch_div_defect(e->s, e->u, e->v, 1.0 / e->h, e->div, e->p, L0.info, L0.b, nullptr,
                            e->nx, e->ny, scale, e->partials, start_in_pass ? e->p2 : nullptr,
                            &sp.xp, sp.order);
        }
        if (start_in_pass) rotate_history(e);
      } else {
        launch_dc_correct(e->s, e->p, nullptr, nullptr, e->div, L0.info, L0.b, e->nx, e->ny,
                          scale, e->partials, nullptr, nullptr);
      }
      launch_dc_norm_init(e->s, e->partials, e->nx, e->ny, e->sums, e->d_res + 16, e->done,
                          scale);
      LAUNCHED("dc_update");
      if (e->warm) {
        // learned warm start: p += its correction of the start, then the
        // defect and its norm again (the norm the loop reads is the start's
        // after the correction; a zero divergence leaves p unchanged here)
        st = engine_warm_start(e->warm, e->s, L0.b, e->lbl, L0.info, e->p, e->ny, e->nx,
                               e->wstats);
        if (st != FH_OK) return st;
        launch_dc_correct(e->s, e->p, nullptr, nullptr, e->div, L0.info, L0.b, e->nx, e->ny,
                          scale, e->partials, nullptr, nullptr);
        launch_dc_norm_init(e->s, e->partials, e->nx, e->ny, e->sums, e->d_res + 16, e->done,
                            scale);
        LAUNCHED("dc_update");
      }
      HIPCHK(hipEventRecord(e->cyc_ev[2], e->s));
    }
    // Cycles are enqueued one ahead of the host's convergence check: cycle
    // c goes in before the host waits for cycle c-1's norm, so the GPU never
    // idles on the host.  The norm kernel sets `done` on the device once the
    // tolerance is met, and every kernel of a cycle enqueued after that
    // returns at entry, so no extra cycle is applied.  Such a skipped cycle
    // still costs its launches, so when cycle c-1's residual, predicted from
    // the contraction measured so far (for the first cycle, the last solve's
    // first-cycle contraction), is already within the tolerance, the host
    // reads cycle c-1's norm before enqueuing cycle c instead.  A wrong
    // guess costs either one skipped cycle or one host round trip, never a
    // different result.
    //
    // At the top of iteration c, cycles 0..c-1 are enqueued and the norms of
    // cycles 0..c-2 have been read.
    int used = iterations;
    bool stopped = false, init_pending = !sync_init;
    const int trace_c0 = e->trace_n;
    double r_init = sync_init ? e->hist.back() : -1.0;
    double r_prev2 = -1.0, r_prev = r_init;  // the last two norms read
    // The initial pass's norm, once its event (or any later one) has been
    // waited for.  Returns true on zero divergence.
    auto consume_init = [&]() -> bool {
      if (!init_pending) return false;
      init_pending = false;
      const double* init = e->h_res + 16;
      if (init[2] == 0.0) return true;
      e->hist.push_back(rel_resid(init, scale, true));
      r_init = r_prev = e->hist.back();
      return false;
    };
    // Zero divergence: every enqueued cycle skipped itself; p = 0.
    auto zero_solution = [&]() -> int {
      e->trace_n = trace_c0;
      e->hist.clear();
      HIPCHK(hipMemsetAsync(e->p, 0, N * sizeof(double), e->s));
      *out_iters = 0;
      return FH_OK;
    };
    auto read_cycle = [&](int c) -> bool {  // true: cycle c met the tolerance
      const double* r = e->h_res + 8 * (c & 1);
      e->hist.push_back(r[4]);
      r_prev2 = r_prev;
      r_prev = r[4];
      if (c == 0 && r_init > 0.0) e->rho_first = r[4] / r_init;
      return r[5] != 0.0;
    };
    for (int c = 0; c < iterations; c++) {
      bool prev_read = false;
      if (c >= 1 && tol > 0) {
        if (init_pending) {  // c == 1: cycle 0 keeps the GPU busy meanwhile
          HIPCHK(hipEventSynchronize(e->cyc_ev[2]));
          if (consume_init()) return zero_solution();
        }
        const double rho = c >= 2 ? r_prev / r_prev2 : e->rho_first;
        if (r_prev > 0.0 && rho > 0.0 && r_prev * rho <= tol) {
          HIPCHK(hipEventSynchronize(e->cyc_ev[(c - 1) & 1]));
          if (read_cycle(c - 1)) {
            used = c;
            stopped = true;
            break;
          }
          prev_read = true;
        }
      }
      const int trace_mark = e->trace_n;
      st = (fcg && c > 0) ? enqueue_fcg_cycle(e, c, scale, tol)
                          : enqueue_dc_cycle(e, c, scale, tol, r64);
      if (st != FH_OK) return st;
      if (c > 0 && tol > 0 && !prev_read) {
        HIPCHK(hipEventSynchronize(e->cyc_ev[(c - 1) & 1]));
        if (read_cycle(c - 1)) {
          used = c;  // cycle c was skipped on the device
          // p/p2 were swapped for the skipped cycle's no-op update (a CG
          // step, c > 0 of FH_SOLVER_MG_FCG, updates p in place), and its
          // kernels' trace records time only their early return
          if (!(fcg && c > 0)) {
            swap_ptr(e->p, e->p2);
            if (e->last_dc_fused) swap_ptr(L0.b, e->b2);
          }
          e->trace_n = trace_mark;
          stopped = true;
          break;
        }
      }
    }
    if (!stopped && iterations > 0) {
      int c = iterations - 1;
      HIPCHK(hipEventSynchronize(e->cyc_ev[c & 1]));
      if (consume_init()) return zero_solution();
      e->hist.push_back(e->h_res[8 * (c & 1) + 4]);
    }
    if (init_pending) {  // no cycle ran
      HIPCHK(hipStreamSynchronize(e->s));
      if (consume_init()) return zero_solution();
    }
    *out_iters = used;
    if (e->capture) {
      st = engine_capture_hook(e->capture, e->s, nullptr, e->lbl, e->psave, e->p, N, e->ny,
                               e->nx, false);
      if (st != FH_OK) return st;
    }
    return FH_OK;
  }
  // FP32 solvers on a single-precision copy of p (level-0 unknown).
  launch_norms(e->s, nullptr, e->p, e->div, e->lbl, e->nx, e->ny, scale, e->partials, e->sums);
  LAUNCHED("norms");
  int st = read_sums(e);
  if (st != FH_OK) return st;
  if (e->h_sums[2] == 0.0) {
    HIPCHK(hipMemsetAsync(e->p, 0, N * sizeof(double), e->s));
    *out_iters = 0;
    return FH_OK;
  }
  double divrms = std::sqrt(e->h_sums[2] / e->h_sums[3]);
  double target = tol > 0 ? tol * divrms : -1;
  launch_f64_to_f32(e->s, e->p, L0.x, N);
  int used = iterations;
  if (solver == FH_SOLVER_RBSOR && rb_sor_lds_fits(e->nx, e->ny)) {
    // the whole grid in one workgroup's LDS: the iterations between two
    // convergence checks (every 12th iteration) run as one launch
    launch_rhs32(e->s, e->div, L0.b, N, 1.0);
    for (int k = 0; k < iterations;) {
      const int chunk = target >= 0 ? std::min(12, iterations - k) : iterations - k;
      {
        Trace t(e, FH_K_RB_SWEEP_FINE);
        launch_rb_sor_lds(e->s, L0.x, L0.b, L0.lbl, L0.nx, L0.ny, chunk, (float)omega,
                          (float)scale);
      }
      LAUNCHED("rb_sor_lds");
      k += chunk;
      if (target >= 0 && chunk == 12) {
        launch_norms(e->s, L0.x, nullptr, e->div, e->lbl, e->nx, e->ny, scale, e->partials,
                     e->sums);
        st = read_sums(e);
        if (st != FH_OK) return st;
        if (e->h_sums[1] > 0 && std::sqrt(e->h_sums[0] / e->h_sums[1]) <= target) {
          used = k;
          break;
        }
      }
    }
  } else if (solver == FH_SOLVER_RBSOR) {
    launch_rhs32(e->s, e->div, L0.b, N, 1.0);
    for (int k = 0; k < iterations; k++) {
      rb(e, 0, 0, (float)omega, (float)scale, 0);
      rb(e, 0, 1, (float)omega, (float)scale, 0);
      LAUNCHED("rb_sweep");
      if (target >= 0 && k % 12 == 11) {
        launch_norms(e->s, L0.x, nullptr, e->div, e->lbl, e->nx, e->ny, scale, e->partials,
                     e->sums);
        st = read_sums(e);
        if (st != FH_OK) return st;
        if (e->h_sums[1] > 0 && std::sqrt(e->h_sums[0] / e->h_sums[1]) <= target) {
          used = k + 1;
          break;
        }
      }
    }
  } else if (solver == FH_SOLVER_MG) {
    launch_rhs32(e->s, e->div, L0.b, N, -scale);
    for (int c = 0; c < iterations; c++) {
      st = vcycle(e, e->nlev_ref, false, true, nullptr);
      if (st != FH_OK) return st;
      if (target >= 0) {
        launch_norms(e->s, L0.x, nullptr, e->div, e->lbl, e->nx, e->ny, scale, e->partials,
                     e->sums);
        st = read_sums(e);
        if (st != FH_OK) return st;
        if (e->h_sums[1] > 0 && std::sqrt(e->h_sums[0] / e->h_sums[1]) <= target) {
          used = c + 1;
          break;
        }
      }
    }
  } else {
    set_err("unknown solver");
    return FH_ERR_BAD_ARG;
  }
  launch_f32_to_f64(e->s, L0.x, e->p, N);
  LAUNCHED("solve");
  *out_iters = used;
  return FH_OK;
}

// A step whose body forces are buoyancy alone (no stir, no confinement,
// dye scale 1) applies it inside the LDS MacCormack face pass, at each open
// interior v face's write: the value the separate buoyancy pass would give,
// without reading and writing v again.
bool buoyancy_fused(const fh_engine* e, int scheme, double dt) {
  return scheme == FH_MACCORMACK && e->advect_lds && e->ds == 1 && !e->stir.active &&
         (float)e->confinement == 0.f && (float)(e->buoyancy * dt) != 0.f;
}

// The epoch of the next one-pass MacCormack dye launch: every launch's
// differs from the mark any earlier launch can have left (epochs start at 1).
int next_one_pass_epoch(fh_engine* e) {
  e->one_pass_epoch = e->one_pass_epoch == INT_MAX ? 1 : e->one_pass_epoch + 1;
  return e->one_pass_epoch;
}

int advect_velocity(fh_engine* e, int scheme, double dt, bool with_buoyancy = false) {
  float c = (float)(dt / e->h);
  if (scheme == FH_MACCORMACK && e->advect_lds && e->faces_one_pass) {
    const int epoch = next_one_pass_epoch(e);
    const float* heat = with_buoyancy ? e->dye : nullptr;
    const float gain = (float)(e->buoyancy * dt);
    {
      Trace t(e, FH_K_MACCORMACK_FACES);
      launch_maccormack_faces_fused(e->s, e->u, e->v, e->u2, e->v2, e->lbl, e->nx, e->ny, c, heat,
                                    gain, e->one_pass_fail, epoch);
    }
    launch_maccormack_faces_rerun(e->s, e->u, e->v, e->uh, e->vh, e->u2, e->v2, e->lbl, e->nx,
                                  e->ny, c, e->sten_f, heat, gain, e->one_pass_fail, epoch);
  } else if (scheme == FH_MACCORMACK) {
    {
      Trace t(e, FH_K_ADVECT_FACES);
      launch_advect_faces(e->s, e->u, e->v, e->u, e->v, e->uh, e->vh, e->lbl, e->nx, e->ny, c,
                          e->advect_lds, e->advect_lds ? e->sten_f : nullptr);
    }
    Trace t(e, FH_K_MACCORMACK_FACES);
    launch_maccormack_faces(e->s, e->u, e->v, e->uh, e->vh, e->u2, e->v2, e->lbl, e->nx, e->ny,
                            c, e->advect_lds, e->advect_lds ? e->sten_f : nullptr,
                            with_buoyancy ? e->dye : nullptr, (float)(e->buoyancy * dt));
  } else {
    Trace t(e, FH_K_ADVECT_FACES);
    launch_advect_faces(e->s, e->u, e->v, e->u, e->v, e->u2, e->v2, e->lbl, e->nx, e->ny, c,
                        e->advect_lds);
  }
  LAUNCHED("advect_velocity");
  swap_ptr(e->u, e->u2);
  swap_ptr(e->v, e->v2);
  e->vel_epoch++;
  return FH_OK;
}

// buoyancy_done: the face pass has applied buoyancy alone (buoyancy_fused).
int forces(fh_engine* e, double dt, bool buoyancy_done = false) {
  float eps = (float)e->confinement;
  float gain = (float)(e->buoyancy * dt);
  if (eps != 0.f) launch_vorticity(e->s, e->u, e->v, e->w, e->nx, e->ny, (float)(1.0 / e->h));
  if (!buoyancy_done && (e->stir.active || gain != 0.f || eps != 0.f)) {
    Trace t(e, FH_K_FORCES);
    launch_forces(e->s, e->u, e->v, e->lbl, e->dye, e->w, e->nx, e->ny, e->ds, (float)e->h,
                  (float)dt, gain, eps, e->stir);
  }
  e->stir.active = 0;  // one frame's displacement (simulation.ts:69-72)
  e->vel_epoch++;
  LAUNCHED("forces");
  return FH_OK;
}

int project(fh_engine* e, double grad_scale) {
  {
    Trace t(e, FH_K_PROJECT);
    launch_subtract_outflow(e->s, e->u, e->v, e->p, e->lbl, e->nx, e->ny, grad_scale);
  }
  e->vel_epoch++;
  LAUNCHED("project");
  return FH_OK;
}

// The max face speed of the (unchanged) velocity from the forward dye pass's
// staged windows: each workgroup's maximum, then their maximum into the
// mapped h_fmax, as the separate reduction would leave it.
bool speed_in_dye(const fh_engine* e) { return e->advect_lds && e->ds == 1; }

int dye_phase(fh_engine* e, int scheme, double dt, const double keep[3], bool speed = false) {
  float c = (float)(dt / e->h);
  float k0 = keep ? (float)keep[0] : 1.f, k1 = keep ? (float)keep[1] : 1.f,
        k2 = keep ? (float)keep[2] : 1.f;
  float* sp = (speed && speed_in_dye(e)) ? e->fpartials : nullptr;
  const bool fused = scheme == FH_MACCORMACK && e->advect_lds && e->ds == 1 && e->dye_one_pass;
  auto speed_done = [&]() -> int {
    if (!sp) return FH_OK;
    launch_finalize_max(e->s, e->fpartials,
                        fused ? maccormack_dye_fused_blocks(e->nx, e->ny)
                              : advect_dye_speed_blocks(e->nx, e->ny),
                        e->fmax, e->d_fmax);
    LAUNCHED("max_speed");
    HIPCHK(hipEventRecord(e->fmax_ev, e->s));
    e->fmax_epoch = e->vel_epoch;
    return FH_OK;
  };
  int st;
  if (fused) {
    const int epoch = next_one_pass_epoch(e);
    {
      Trace t(e, FH_K_MACCORMACK_DYE);
      launch_maccormack_dye_fused(e->s, e->u, e->v, e->dye, e->dye2, e->lbl, e->nx, e->ny, c, k0,
                                  k1, k2, sp, e->one_pass_fail, epoch);
    }
    launch_maccormack_dye_rerun(e->s, e->u, e->v, e->dye, e->dyeh, e->dye2, e->lbl, e->nx, e->ny,
                                c, k0, k1, k2, e->sten_d, e->one_pass_fail, epoch);
    if ((st = speed_done()) != FH_OK) return st;
  } else if (scheme == FH_MACCORMACK) {
    // the dye LDS kernels (dye scale 1) record and read the limiter stencils
    uint8_t* sten_d = (e->advect_lds && e->ds == 1) ? e->sten_d : nullptr;
    {
      Trace t(e, FH_K_ADVECT_DYE);
      launch_advect_dye(e->s, e->u, e->v, e->dye, e->dyeh, e->lbl, e->nx, e->ny, e->ds, c, 1.f,
                        1.f, 1.f, e->advect_lds, sten_d, sp);
    }
    if ((st = speed_done()) != FH_OK) return st;
    Trace t(e, FH_K_MACCORMACK_DYE);
    launch_maccormack_dye(e->s, e->u, e->v, e->dye, e->dyeh, e->dye2, e->lbl, e->nx, e->ny, e->ds,
                          c, k0, k1, k2, e->advect_lds, sten_d);
  } else {
    {
      Trace t(e, FH_K_ADVECT_DYE);
      launch_advect_dye(e->s, e->u, e->v, e->dye, e->dye2, e->lbl, e->nx, e->ny, e->ds, c, k0,
                        k1, k2, e->advect_lds, nullptr, sp);
    }
    if ((st = speed_done()) != FH_OK) return st;
  }
  LAUNCHED("advect_dye");
  swap_ptr(e->dye, e->dye2);
  if (e->has_patch) {
    launch_dye_patch(e->s, e->dye, e->patch, e->dnx, e->dny, e->pi0, e->pj0, e->pnx, e->pny,
                     e->patch_tiles, e->patch_ntiles);
    LAUNCHED("dye_patch");
  }
  return FH_OK;
}

int mark(fh_engine* e, int phase) {
  if (e->timing) HIPCHK(hipEventRecord(e->ev[phase], e->s));
  return FH_OK;
}

}  // namespace

fh::CoarseLevels fh::engine_coarse_levels(const fh_engine* e, int lc, int nlev) {
  return coarse_levels(e, lc, nlev);
}

int fh::engine_batch_prepare(fh_engine* e) {
  GUARD("HipBatch.prepare()");
  int st;
  if (e->labels_dirty && (st = refresh_levels(e)) != FH_OK) return st;
  return ensure_fcg(e);
}

extern "C" {

const char* fh_last_error(void) { return g_err.c_str(); }

int fh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

fh_engine* fh_create(int nx, int ny, int dye_scale, double h) {
  if (nx < 2 || ny < 2 || dye_scale < 1 || !(h > 0)) {
    set_err("fh_create: need nx, ny >= 2, dye_scale >= 1, h > 0");
    return nullptr;
  }
  if (fh_device_count() < 1) {
    set_err("fh_create: no HIP device");
    return nullptr;
  }
  fh_engine* e = new fh_engine();
  e->nx = nx;
  e->ny = ny;
  e->ds = dye_scale;
  e->dnx = nx * dye_scale;
  e->dny = ny * dye_scale;
  e->h = h;
  auto bail = [&]() {
    fh_destroy(e);
    return (fh_engine*)nullptr;
  };
  if (hipStreamCreate(&e->s) != hipSuccess) {
    set_err("fh_create: hipStreamCreate failed");
    return bail();
  }
  size_t nu = n_u(e), nv = n_v(e), nc = n_c(e), nd = n_d(e);
  if (dalloc(e, &e->u, nu) || dalloc(e, &e->v, nv) || dalloc(e, &e->u2, nu) ||
      dalloc(e, &e->v2, nv) || dalloc(e, &e->uh, nu) || dalloc(e, &e->vh, nv) ||
      dalloc(e, &e->dye, 3 * nd) || dalloc(e, &e->dye2, 3 * nd) || dalloc(e, &e->dyeh, 3 * nd) ||
      dalloc(e, &e->w, nc) || dalloc(e, &e->lbl, nc) || dalloc(e, &e->p, nc) ||
      dalloc(e, &e->sten_f, (size_t)(nx + 1) * (ny + 1)) || dalloc(e, &e->sten_d, nd) ||
      dalloc(e, &e->one_pass_fail, 1) || dalloc(e, &e->p2, nc) || dalloc(e, &e->div, nc))
    return bail();
  e->nlev_ref = level_count(nx, ny, 4);
  e->nlev_dc = level_count(nx, ny, 16);
  int lnx = nx, lny = ny;
  for (int l = 0; l < e->nlev_ref; l++) {
    Level L{lnx, lny, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t n = (size_t)lnx * lny;
    if (l == 0)
      L.lbl = e->lbl;
    else if (dalloc(e, &L.lbl, n))
      return bail();
    if (dalloc(e, &L.x, n) || dalloc(e, &L.b, n) || dalloc(e, &L.xa, n) || dalloc(e, &L.info, n))
      return bail();
    e->lv.push_back(L);
    lnx = (lnx + 1) / 2;
    lny = (lny + 1) / 2;
  }
  e->lc_ref = coarse_start(e, e->nlev_ref);
  e->lc_dc = coarse_start(e, e->nlev_dc);
  // padded info of every level a sub-cycle may stage
  e->cinfo_from = std::min(e->lc_ref, e->lc_dc);
  e->cinfo_off.assign(e->lv.size() + 1, 0);
  for (size_t l = 0; l < e->lv.size(); l++)
    e->cinfo_off[l + 1] = e->cinfo_off[l] +
                          ((int)l >= e->cinfo_from ? padded_info_bytes(e->lv[l].nx, e->lv[l].ny) : 0);
  if (dalloc(e, &e->cinfo, (size_t)(e->cinfo_off[e->lv.size()] - e->cinfo_off[e->cinfo_from])))
    return bail();
  size_t nblk = (size_t)((nx + 63) / 64) * ((ny + 3) / 4);
  e->partials_n = nblk;
  // the max-speed partials of either forward dye pass (two-pass or one-pass)
  const size_t nspeed = (size_t)std::max(advect_dye_speed_blocks(nx, ny),
                                         maccormack_dye_fused_blocks(nx, ny));
  if (dalloc(e, &e->partials, 4 * nblk + 4 * kFoldBlocks) || dalloc(e, &e->b2, nc) ||
      dalloc(e, &e->counter, 1) || dalloc(e, &e->sums, 8) ||
      dalloc(e, &e->fpartials, std::max<size_t>(1024, nspeed)) ||
      dalloc(e, &e->fmax, 1) || dalloc(e, &e->done, 1))
    return bail();
  if (hipHostMalloc((void**)&e->h_sums, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&e->h_res, 24 * sizeof(double),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&e->d_res, e->h_res, 0) != hipSuccess ||
      hipHostMalloc((void**)&e->h_fmax, sizeof(float),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&e->d_fmax, e->h_fmax, 0) != hipSuccess) {
    set_err("fh_create: hipHostMalloc failed");
    return bail();
  }
  for (int k = 0; k <= FH_NUM_PHASES; k++)
    if (hipEventCreate(&e->ev[k]) != hipSuccess) {
      set_err("fh_create: hipEventCreate failed");
      return bail();
    }
  for (hipEvent_t& ev : e->cyc_ev)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      set_err("fh_create: hipEventCreate failed");
      return bail();
    }
  if (hipEventCreateWithFlags(&e->fmax_ev, hipEventDisableTiming) != hipSuccess) {
    set_err("fh_create: hipEventCreate failed");
    return bail();
  }
  return e;
}

void fh_destroy(fh_engine* e) {
  if (!e) return;
  if (e->s) (void)hipStreamSynchronize(e->s);
  float* fbufs[] = {e->u, e->v, e->u2, e->v2, e->uh, e->vh, e->dye, e->dye2, e->dyeh, e->w, e->b2,
                    e->patch, e->fpartials, e->fmax, e->dvec[0], e->dvec[1]};
  for (float* p : fbufs)
    if (p) (void)hipFree(p);
  double* dbufs[] = {e->p,   e->p2,  e->div,   e->M,    e->partials,
                     e->sums, e->rvec, e->fcg, e->ph[0], e->ph[1]};
  for (double* p : dbufs)
    if (p) (void)hipFree(p);
  if (e->lbl) (void)hipFree(e->lbl);
  if (e->sten_f) (void)hipFree(e->sten_f);
  if (e->sten_d) (void)hipFree(e->sten_d);
  if (e->one_pass_fail) (void)hipFree(e->one_pass_fail);
  if (e->cells) (void)hipFree(e->cells);
  if (e->cinfo) (void)hipFree(e->cinfo);
  if (e->patch_tiles) (void)hipFree(e->patch_tiles);
  if (e->counter) (void)hipFree(e->counter);
  if (e->done) (void)hipFree(e->done);
  if (e->psave) (void)hipFree(e->psave);
  if (e->wstats) (void)hipFree(e->wstats);
  for (hipEvent_t ev : e->solve_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : e->cyc_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fmax_ev) (void)hipEventDestroy(e->fmax_ev);
  for (size_t l = 0; l < e->lv.size(); l++) {
    if (l > 0 && e->lv[l].lbl) (void)hipFree(e->lv[l].lbl);
    if (e->lv[l].x) (void)hipFree(e->lv[l].x);
    if (e->lv[l].b) (void)hipFree(e->lv[l].b);
    if (e->lv[l].xa) (void)hipFree(e->lv[l].xa);
    if (e->lv[l].info) (void)hipFree(e->lv[l].info);
  }
  if (e->h_res) (void)hipHostFree(e->h_res);
  if (e->h_sums) (void)hipHostFree(e->h_sums);
  if (e->h_fmax) (void)hipHostFree(e->h_fmax);
  if (e->stage) (void)hipHostFree(e->stage);
  for (int k = 0; k <= FH_NUM_PHASES; k++)
    if (e->ev[k]) (void)hipEventDestroy(e->ev[k]);
  for (hipEvent_t ev : e->trace_ev) (void)hipEventDestroy(ev);
  if (e->s) (void)hipStreamDestroy(e->s);
  delete e;
}

int fh_upload(fh_engine* e, const double* u, const double* v, const double* p,
              const uint8_t* label, const double* dye) {
  GUARD("HipStepper.upload()");
  int st;
  e->vel_epoch++;
  drop_pressure_history(e);
  if (u && (st = up32(e, e->u, u, n_u(e)))) return st;
  if (v && (st = up32(e, e->v, v, n_v(e)))) return st;
  if (dye && (st = up32(e, e->dye, dye, 3 * n_d(e)))) return st;
  if (p) HIPCHK(hipMemcpy(e->p, p, n_c(e) * sizeof(double), hipMemcpyHostToDevice));
  if (label) {
    HIPCHK(hipMemcpy(e->lbl, label, n_c(e), hipMemcpyHostToDevice));
    e->labels_dirty = true;
    // The label chain and the coarsest operator depend on the labels alone:
    // rebuilt here, once per upload, rather than inside the next solve.
    if ((st = refresh_levels(e))) return st;
  }
  return FH_OK;
}

int fh_set_labels(fh_engine* e, const uint8_t* label) {
  GUARD("HipStepper.setLabels()");
  if (!label) {
    set_err("setLabels: null label");
    return FH_ERR_BAD_ARG;
  }
  drop_pressure_history(e);
  HIPCHK(hipMemcpy(e->lbl, label, n_c(e), hipMemcpyHostToDevice));
  launch_commit_labels(e->s, e->p, e->u, e->v, e->dye, e->lbl, e->nx, e->ny, (int)n_d(e));
  e->vel_epoch++;
  LAUNCHED("commit_labels");
  HIPCHK(hipStreamSynchronize(e->s));
  e->labels_dirty = true;
  return FH_OK;
}

int fh_set_dye_patch(fh_engine* e, int i0, int j0, int pnx, int pny, const float* data) {
  GUARD("HipStepper.setDyePatch()");
  if (e->patch) HIPCHK(hipFree(e->patch));
  if (e->patch_tiles) HIPCHK(hipFree(e->patch_tiles));
  e->patch = nullptr;
  e->patch_tiles = nullptr;
  e->patch_ntiles = 0;
  e->has_patch = 0;
  if (!data) return FH_OK;
  if (pnx <= 0 || pny <= 0) {
    set_err("setDyePatch: empty patch");
    return FH_ERR_BAD_ARG;
  }
  size_t n = (size_t)4 * pnx * pny;
  HIPCHK(hipMalloc((void**)&e->patch, n * sizeof(float)));
  HIPCHK(hipMemcpy(e->patch, data, n * sizeof(float), hipMemcpyHostToDevice));
  std::vector<int> tiles;
  dye_patch_tiles(data, pnx, pny, tiles);
  if (!tiles.empty()) {
    HIPCHK(hipMalloc((void**)&e->patch_tiles, tiles.size() * sizeof(int)));
    HIPCHK(hipMemcpy(e->patch_tiles, tiles.data(), tiles.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  }
  e->patch_ntiles = (int)tiles.size();
  e->pi0 = i0;
  e->pj0 = j0;
  e->pnx = pnx;
  e->pny = pny;
  e->has_patch = 1;
  return FH_OK;
}

int fh_set_buoyancy(fh_engine* e, double w) {
  GUARD("HipStepper.setBuoyancy()");
  e->buoyancy = w;
  return FH_OK;
}

int fh_set_confinement(fh_engine* e, double eps) {
  GUARD("HipStepper.setConfinement()");
  e->confinement = eps;
  return FH_OK;
}

int fh_set_stir(fh_engine* e, int active, double x, double y, double r, double dx, double dy) {
  GUARD("HipStepper.setStir()");
  StirBox& s = e->stir;
  s.active = active ? 1 : 0;
  if (!active) return FH_OK;
  if (!(r > 0)) {
    set_err("setStir: radius must be positive");
    return FH_ERR_BAD_ARG;
  }
  // stir.ts:49-62 box, then cell units for the kernel
  double h = e->h, reach = 3 * r;
  int c0 = (int)std::floor((x - reach) / h), c1 = (int)std::ceil((x + reach) / h);
  int r0 = (int)std::floor((y - reach) / h), r1 = (int)std::ceil((y + reach) / h);
  auto cl = [](int a, int lo, int hi) { return a < lo ? lo : a > hi ? hi : a; };
  int nx = e->nx, ny = e->ny;
  s.ui0 = cl(c0, 1, nx - 1);
  s.ui1 = cl(c1, 1, nx - 1);
  s.uj0 = cl(r0, 0, ny - 1);
  s.uj1 = cl(r1, 0, ny - 1);
  s.vi0 = cl(c0, 0, nx - 1);
  s.vi1 = cl(c1, 0, nx - 1);
  s.vj0 = cl(r0, 1, ny - 1);
  s.vj1 = cl(r1, 1, ny - 1);
  s.x = (float)(x / h);
  s.y = (float)(y / h);
  s.r = (float)(r / h);
  s.dx = (float)dx;
  s.dy = (float)dy;
  return FH_OK;
}

int fh_step(fh_engine* e, int scheme, double dt, double scale, double grad_scale,
            const double keep[3], int solver, int iterations, double omega, double tol,
            int* out_iters) {
  GUARD("HipStepper.step()");
  if (!(dt > 0) || !out_iters) {
    set_err("step: dt must be positive and out_iters non-null");
    return FH_ERR_BAD_ARG;
  }
  int st;
  if ((st = mark(e, 0))) return st;
  const bool buoy = buoyancy_fused(e, scheme, dt);
  if ((st = advect_velocity(e, scheme, dt, buoy))) return st;
  if ((st = mark(e, 1))) return st;
  if ((st = forces(e, dt, buoy))) return st;
  if ((st = mark(e, 2))) return st;
  // The FP64 solvers' initial defect pass forms the divergence itself
  // (unless a training capture needs the defect on the host first); the
  // other solvers read it from its own pass.
  const bool fuse = (solver == FH_SOLVER_MG_DC || solver == FH_SOLVER_MG_FCG) && !e->capture;
  if (!fuse) {
    {
      Trace t(e, FH_K_DIVERGENCE);
      launch_divergence(e->s, e->u, e->v, e->div, e->nx, e->ny, 1.0 / e->h);
    }
    LAUNCHED("divergence");
  }
  if ((st = mark(e, 3))) return st;
  if ((st = solve(e, solver, scale, iterations, omega, tol, out_iters, fuse, dt))) return st;
  if ((st = mark(e, 4))) return st;
  if ((st = project(e, grad_scale))) return st;
  // The next step's CFL reads this velocity (simulation.ts:474-477); the dye
  // phase below leaves it unchanged, so its maximum is taken from the forward
  // dye pass's staged velocity windows (or, without the LDS dye pass, by its
  // own reduction here), the result lands in the mapped h_fmax and the host's
  // wait for it overlaps the remaining dye kernels.
  const bool speed = speed_in_dye(e);
  if (!speed) {
    launch_max_speed(e->s, e->u, e->v, e->nx, e->ny, e->fpartials, e->counter, e->fmax,
                     e->d_fmax);
    LAUNCHED("max_speed");
    HIPCHK(hipEventRecord(e->fmax_ev, e->s));
    e->fmax_epoch = e->vel_epoch;
  }
  if ((st = mark(e, 5))) return st;
  if ((st = dye_phase(e, scheme, dt, keep, speed))) return st;
  if ((st = mark(e, 6))) return st;
  e->ev_valid = e->timing;
  return FH_OK;
}

int fh_read_velocity(fh_engine* e, double* u, double* v) {
  GUARD("HipStepper.readVelocity()");
  int st;
  if (u && (st = down32(e, u, e->u, n_u(e)))) return st;
  if (v && (st = down32(e, v, e->v, n_v(e)))) return st;
  return FH_OK;
}

int fh_read_dye(fh_engine* e, double* dye) {
  GUARD("HipStepper.readDye()");
  if (!dye) return FH_ERR_BAD_ARG;
  return down32(e, dye, e->dye, 3 * n_d(e));
}

int fh_read_pressure(fh_engine* e, double* p) {
  GUARD("HipStepper.readPressure()");
  if (!p) return FH_ERR_BAD_ARG;
  HIPCHK(hipMemcpyAsync(p, e->p, n_c(e) * sizeof(double), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  return FH_OK;
}

int fh_read_divergence(fh_engine* e, double* div) {
  GUARD("HipStepper.readDivergence()");
  if (!div) return FH_ERR_BAD_ARG;
  HIPCHK(hipMemcpyAsync(div, e->div, n_c(e) * sizeof(double), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  return FH_OK;
}

int fh_max_face_speed(fh_engine* e, double* out) {
  GUARD("HipStepper.maxFaceSpeed()");
  if (!out) return FH_ERR_BAD_ARG;
  if (e->fmax_epoch == e->vel_epoch) {  // reduced by the last step, velocity unchanged since
    HIPCHK(hipEventSynchronize(e->fmax_ev));
    *out = (double)*e->h_fmax;
    return FH_OK;
  }
  launch_max_speed(e->s, e->u, e->v, e->nx, e->ny, e->fpartials, e->counter, e->fmax);
  LAUNCHED("max_speed");
  HIPCHK(hipMemcpyAsync(e->h_fmax, e->fmax, sizeof(float), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  *out = (double)*e->h_fmax;
  return FH_OK;
}

int fh_solve_pressure(fh_engine* e, int solver, double* p, const double* div,
                      const uint8_t* label, double scale, int iterations, double omega,
                      double tol, int* out_iters) {
  GUARD("HipPressureSolver.solve()");
  if (!p || !div || !label || !out_iters) {
    set_err("solve: null argument");
    return FH_ERR_BAD_ARG;
  }
  size_t nc = n_c(e);
  HIPCHK(hipMemcpy(e->p, p, nc * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->div, div, nc * sizeof(double), hipMemcpyHostToDevice));
  std::vector<uint8_t> cur(nc);
  HIPCHK(hipMemcpy(cur.data(), e->lbl, nc, hipMemcpyDeviceToHost));
  if (memcmp(cur.data(), label, nc) != 0) {
    HIPCHK(hipMemcpy(e->lbl, label, nc, hipMemcpyHostToDevice));
    e->labels_dirty = true;
  }
  int st = solve(e, solver, scale, iterations, omega, tol, out_iters);
  if (st != FH_OK) return st;
  HIPCHK(hipMemcpyAsync(p, e->p, nc * sizeof(double), hipMemcpyDeviceToHost, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  return FH_OK;
}

int fh_advect_velocity(fh_engine* e, int scheme, const double* u_in, const double* v_in,
                       double* u_out, double* v_out, const uint8_t* label, double dt) {
  GUARD("HipAdvector.velocity()");
  if (!u_in || !v_in || !u_out || !v_out || !label) return FH_ERR_BAD_ARG;
  int st;
  e->vel_epoch++;
  if ((st = up32(e, e->u, u_in, n_u(e))) || (st = up32(e, e->v, v_in, n_v(e)))) return st;
  HIPCHK(hipMemcpy(e->lbl, label, n_c(e), hipMemcpyHostToDevice));
  e->labels_dirty = true;
  if ((st = advect_velocity(e, scheme, dt))) return st;
  if ((st = down32(e, u_out, e->u, n_u(e))) || (st = down32(e, v_out, e->v, n_v(e)))) return st;
  return FH_OK;
}

int fh_advect_dye(fh_engine* e, int scheme, const double* u, const double* v, const double* q_in,
                  double* q_out, const uint8_t* label, double dt) {
  GUARD("HipAdvector.dye()");
  if (!u || !v || !q_in || !q_out || !label) return FH_ERR_BAD_ARG;
  int st;
  e->vel_epoch++;
  if ((st = up32(e, e->u, u, n_u(e))) || (st = up32(e, e->v, v, n_v(e))) ||
      (st = up32(e, e->dye, q_in, 3 * n_d(e))))
    return st;
  HIPCHK(hipMemcpy(e->lbl, label, n_c(e), hipMemcpyHostToDevice));
  e->labels_dirty = true;
  int had_patch = e->has_patch;
  e->has_patch = 0;
  st = dye_phase(e, scheme, dt, nullptr);
  e->has_patch = had_patch;
  if (st) return st;
  return down32(e, q_out, e->dye, 3 * n_d(e));
}

int fh_phase_advect_velocity(fh_engine* e, int scheme, double dt) {
  GUARD("phase.advectVelocity()");
  return advect_velocity(e, scheme, dt);
}

int fh_phase_forces(fh_engine* e, double dt) {
  GUARD("phase.forces()");
  return forces(e, dt);
}

int fh_phase_divergence(fh_engine* e) {
  GUARD("phase.divergence()");
  launch_divergence(e->s, e->u, e->v, e->div, e->nx, e->ny, 1.0 / e->h);
  LAUNCHED("divergence");
  return FH_OK;
}

int fh_phase_solve(fh_engine* e, int solver, double scale, int iterations, double omega,
                   double tol, int* out_iters) {
  GUARD("phase.solve()");
  if (!out_iters) return FH_ERR_BAD_ARG;
  return solve(e, solver, scale, iterations, omega, tol, out_iters);
}

int fh_phase_project(fh_engine* e, double grad_scale) {
  GUARD("phase.project()");
  return project(e, grad_scale);
}

int fh_phase_dye(fh_engine* e, int scheme, double dt, const double keep[3]) {
  GUARD("phase.dye()");
  return dye_phase(e, scheme, dt, keep);
}

int fh_solve_history(fh_engine* e, double* out, int max) {
  if (!e || !out) return 0;
  int n = (int)e->hist.size() < max ? (int)e->hist.size() : max;
  for (int k = 0; k < n; k++) out[k] = e->hist[k];
  return n;
}

int fh_set_timing(fh_engine* e, int on) {
  GUARD("setTiming()");
  e->timing = on != 0;
  e->ev_valid = false;
  return FH_OK;
}

int fh_get_timings(fh_engine* e, double* ms) {
  GUARD("getTimings()");
  if (!ms) return FH_ERR_BAD_ARG;
  for (int k = 0; k < FH_NUM_PHASES; k++) ms[k] = 0;
  if (!e->ev_valid) return FH_OK;
  HIPCHK(hipEventSynchronize(e->ev[6]));
  for (int k = 0; k < 6; k++) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e->ev[k], e->ev[k + 1]));
    ms[k] = t;
  }
  float tot = 0;
  HIPCHK(hipEventElapsedTime(&tot, e->ev[0], e->ev[6]));
  ms[FH_PHASE_TOTAL] = tot;
  return FH_OK;
}

int fh_last_solve_ms(fh_engine* e, double* ms) {
  GUARD("lastSolveMs()");
  if (!ms) return FH_ERR_BAD_ARG;
  *ms = 0;
  if (!e->solve_ev_valid) return FH_OK;
  HIPCHK(hipEventSynchronize(e->solve_ev[1]));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e->solve_ev[0], e->solve_ev[1]));
  *ms = t;
  return FH_OK;
}

int fh_set_capture(fh_engine* e, fh_unet* u) {
  GUARD("setCapture()");
  if (u && !e->psave) HIPCHK(hipMalloc((void**)&e->psave, n_c(e) * sizeof(double)));
  e->capture = u;
  return FH_OK;
}

int fh_set_warm_start(fh_engine* e, fh_unet* u) {
  GUARD("setWarmStart()");
  if (u && !e->wstats) HIPCHK(hipMalloc((void**)&e->wstats, 64));
  e->warm = u;
  return FH_OK;
}

int fh_set_fused_update(fh_engine* e, int on) {
  GUARD("setFusedUpdate()");
  e->fuse_update = on != 0;
  return FH_OK;
}

int fh_set_warm_extrapolation(fh_engine* e, int order) {
  GUARD("setWarmExtrapolation()");
  if (order < 0 || order > 2) {
    set_err("setWarmExtrapolation: order must be 0, 1 or 2");
    return FH_ERR_BAD_ARG;
  }
  int st;
  for (int k = 0; k < order; k++)
    if (!e->ph[k] && (st = dalloc(e, &e->ph[k], n_c(e)))) return st;
  if (order != e->extrap) e->ph_n = 0;
  e->extrap = order;
  return FH_OK;
}

int fh_trace_kernel(fh_engine* e, int kernel) {
  GUARD("traceKernel()");
  HIPCHK(hipStreamSynchronize(e->s));
  e->trace_kid = kernel;
  e->trace_n = 0;
  return FH_OK;
}

int fh_trace_stats(fh_engine* e, int* count, double* total_ms) {
  GUARD("traceStats()");
  if (!count || !total_ms) return FH_ERR_BAD_ARG;
  HIPCHK(hipStreamSynchronize(e->s));
  double tot = 0;
  for (int k = 0; k < e->trace_n; k++) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e->trace_ev[2 * k], e->trace_ev[2 * k + 1]));
    tot += t;
  }
  *count = e->trace_n;
  *total_ms = tot;
  return FH_OK;
}

// Device ms of an event pair with nothing between it, averaged over `reps`
// pairs recorded back to back on the engine's stream: the bracket cost that
// fh_trace_stats' per-launch times include.
int fh_trace_overhead(fh_engine* e, int reps, double* ms) {
  GUARD("traceOverhead()");
  if (!ms || reps < 1) return FH_ERR_BAD_ARG;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  double tot = 0;
  for (int r = 0; r < reps; r++) {
    HIPCHK(hipEventRecord(a, e->s));
    HIPCHK(hipEventRecord(b, e->s));
    HIPCHK(hipEventSynchronize(b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    tot += t;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = tot / reps;
  return FH_OK;
}

int fh_num_tile_shapes(void) { return kNumTileShapes; }

const char* fh_tile_shape_name(int shape) {
  return (shape >= 0 && shape < kNumTileShapes) ? kTileShapes[shape].name : "";
}

int fh_set_tile(fh_engine* e, int level, int shape) {
  GUARD("setTile()");
  if (shape < 0 || shape >= kNumTileShapes || level < -1 || level >= kMaxTiledLevels) {
    set_err("fh_set_tile: bad level or shape");
    return FH_ERR_BAD_ARG;
  }
  for (int l = 0; l < kMaxTiledLevels; l++)
    if (level < 0 || level == l) e->tile[l] = shape;
  return FH_OK;
}

// Tiled levels of the defect-correction hierarchy and their shapes.
int fh_get_tiles(fh_engine* e, int* shapes, int* nlevels) {
  GUARD("getTiles()");
  if (nlevels) *nlevels = e->lc_dc;
  if (shapes)
    for (int l = 0; l < e->lc_dc; l++) shapes[l] = e->tile[l];
  return FH_OK;
}

// Autotuner: for every tiled level of the defect-correction hierarchy, times
// `reps` pre-smooth+restrict and prolong+post-smooth launches of each tile
// shape and keeps the fastest.  Only multigrid scratch (level unknowns and
// coarse right-hand sides) is touched; labels must be set (a solve or step
// has run).  ms_out (optional): nlevels x fh_num_tile_shapes() ms per pair.
constexpr int kTuneRounds = 3;

int fh_tune_tiles(fh_engine* e, int reps, float* ms_out) {
  GUARD("tuneTiles()");
  if (reps < 1) reps = 1;
  if (e->labels_dirty) {
    int st = refresh_levels(e);
    if (st != FH_OK) return st;
  }
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  for (int l = 0; l < e->lc_dc && l < kMaxTiledLevels; l++) {
    Level& L = e->lv[l];
    Level& N = e->lv[l + 1];
    // kTuneRounds interleaved rounds over the shapes, each shape's best
    // round kept: one round's clock or queue hiccup cannot pick a shape.
    float msr[kNumTileShapes];
    for (int t = 0; t < kNumTileShapes; t++) msr[t] = 1e30f;
    for (int round = 0; round < kTuneRounds; round++)
      for (int t = 0; t < kNumTileShapes; t++) {
        for (int r = -2; r < reps; r++) {
          if (r == 0) HIPCHK(hipEventRecord(a, e->s));
          launch_smooth_restrict(e->s, t, nullptr, L.xa, L.b, L.info, L.nx, L.ny, N.b, N.nx, N.ny,
                                 1, nullptr);
          // the finest level's defect-correction cycles end in the fused
          // pass (its outputs, p2, b2 and the partials, are scratch here)
          if (l == 0 && e->fuse_update &&
              (size_t)prolong_smooth_dc_tiles(t, L.nx, L.ny) <= e->partials_n)
            launch_prolong_smooth_dc(e->s, t, L.xa, L.b, L.info, L.nx, L.ny, N.x, N.nx, N.ny,
                                     nullptr, e->p, e->p2, e->div, e->b2, nullptr, 1.0,
                                     e->partials);
          else
            launch_prolong_smooth(e->s, t, L.xa, L.x, L.b, L.info, L.nx, L.ny, N.x, N.nx, N.ny,
                                  nullptr);
        }
        HIPCHK(hipEventRecord(b, e->s));
        HIPCHK(hipEventSynchronize(b));
        LAUNCHED("tune_tiles");
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        msr[t] = std::min(msr[t], ms / reps);
      }
    int best_t = 0;
    for (int t = 0; t < kNumTileShapes; t++) {
      if (ms_out) ms_out[l * kNumTileShapes + t] = msr[t];
      if (msr[t] < msr[best_t]) best_t = t;
    }
    e->tile[l] = best_t;
  }
  // the one-workgroup coarse sub-cycle: its workgroup size
  {
    const int lc = e->lc_dc, nlev = e->nlev_dc;
    float msg[kNumCoarseGroups];
    for (int g = 0; g < kNumCoarseGroups; g++) msg[g] = 1e30f;
    for (int round = 0; round < kTuneRounds; round++)
      for (int g = 0; g < kNumCoarseGroups; g++) {
        for (int r = -2; r < reps; r++) {
          if (r == 0) HIPCHK(hipEventRecord(a, e->s));
          launch_coarse_vcycle(e->s, g, coarse_levels(e, lc, nlev), e->lv[lc].b, nullptr,
                               e->lv[lc].x, e->M, e->cells, e->ncoarse, 1, nullptr);
        }
        HIPCHK(hipEventRecord(b, e->s));
        HIPCHK(hipEventSynchronize(b));
        LAUNCHED("tune_coarse");
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        msg[g] = std::min(msg[g], ms / reps);
      }
    e->coarse_group = 0;
    for (int g = 1; g < kNumCoarseGroups; g++)
      if (msg[g] < msg[e->coarse_group]) e->coarse_group = g;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return FH_OK;
}

// Workgroup size of the coarse sub-cycle: threads of kCoarseGroups[group].
int fh_set_coarse_group(fh_engine* e, int group) {
  GUARD("setCoarseGroup()");
  if (group < 0 || group >= kNumCoarseGroups) {
    set_err("fh_set_coarse_group: bad group");
    return FH_ERR_BAD_ARG;
  }
  e->coarse_group = group;
  return FH_OK;
}

// Advection kernels: LDS-staged windows (1, the default) or plain gathers from
// global memory (0).  Same results either way.
int fh_set_advect_staging(fh_engine* e, int on) {
  GUARD("setAdvectStaging()");
  e->advect_lds = on != 0;
  return FH_OK;
}

int fh_set_faces_one_pass(fh_engine* e, int on) {
  GUARD("setFacesOnePass()");
  e->faces_one_pass = on != 0;
  return FH_OK;
}

int fh_set_dye_one_pass(fh_engine* e, int on) {
  GUARD("setDyeOnePass()");
  e->dye_one_pass = on != 0;
  return FH_OK;
}

int fh_coarse_group_threads(fh_engine* e) {
  return e ? kCoarseGroups[e->coarse_group] : 0;
}

int fh_synchronize(fh_engine* e) {
  GUARD("synchronize()");
  HIPCHK(hipStreamSynchronize(e->s));
  return FH_OK;
}

}  // extern "C"
