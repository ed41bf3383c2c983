// Batched scenes: many independent grids of uneven sizes advanced together on
// one GPU (include/flowhip.h, fh_batch_*).
//
// No counterpart in the reference, which steps one grid per Simulation
// (simulation.ts:467-587).  Each scene keeps a single-grid engine (fh_engine:
// its buffers, multigrid hierarchy and coarsest-level operator), but a batch
// step launches each phase once for all scenes: every batched kernel runs
// over a work list of (scene, block) tasks built from the scenes' own
// single-grid launch grids (fh_batchdesc.hpp), on one stream.  A scene's
// blocks run the single-grid kernels' bodies at the same block indices and
// partials slots, so every scene gets exactly the bits of fh_step on it
// alone: dt from its own CFL (simulation.ts:474-479) and its pressure by
// FH_SOLVER_MG_FCG to its own convergence.  Convergence is per scene on the
// device (each scene's norm sets its own `done`, and its blocks of later
// cycles return at once); the host enqueues cycles for all scenes, one ahead
// of reading their norms, until every scene has converged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flowhip.h"
#include "fh_batchdesc.hpp"
#include "fh_engine_impl.hpp"
#include "fh_mgtile.hpp"

using namespace fh;

namespace {

thread_local std::string b_err;

struct SceneParams {
  int nx = 0, ny = 0;
  double h = 0.0;
  double decay[3] = {0.0, 0.0, 0.0};
  // After an upload of both velocity components the next step's CFL takes
  // the largest |u|, |v| of the uploaded FP64 host fields, as Simulation.step
  // scans its host fields (simulation.ts:451-459); later steps read the
  // device's reduction of the stepped velocity.
  bool host_umax_valid = false;
  double host_umax = 0.0;
};

}  // namespace

struct fh_batch {
  std::vector<fh_engine*> e;
  std::vector<SceneParams> prm;
  std::vector<int> order;  // scenes by decreasing cell count
  std::vector<double> dt, resid;
  std::vector<int> iters;
  std::atomic<bool> busy{false};
  float last_ms = 0.f;
  long long cells = 0;
  hipStream_t s = nullptr;
  // Descriptors and work lists, rebuilt before a step after anything that
  // changes them (uploads, forces, tiles).  q: the step parity since.
  bool dirty = true;
  int q = 0;
  int tile[kBatchMaxLevels] = {};  // tile shape per level index
  int max_lc = 0;                  // tiled levels of the deepest scene
  size_t coarse_lds = 0;
  std::vector<SceneDesc> desc;
  SceneDesc* d_desc = nullptr;
  BStep* d_step = nullptr;
  BStep* h_step = nullptr;  // pinned
  unsigned char* d_lists = nullptr;
  size_t lists_bytes = 0;
  BatchGrid g_adv[4] = {}, g_proj[kNumBatchProject] = {}, g_red = {}, g_speed = {};
  BatchGrid g_lv[kBatchMaxLevels][kNumTileShapes] = {};
  // largest face speed of each scene's velocity: partials, then mapped host floats
  float* d_fpart = nullptr;
  float* h_umax = nullptr;
  float* d_umax = nullptr;
  bool umax_valid = false;
  hipEvent_t ev_umax = nullptr, ev_init = nullptr, ev_cyc[2] = {};
};

namespace {

struct BGuard {
  fh_batch* b;
  int status = FH_OK;
  BGuard(fh_batch* bb, const char* what) : b(bb) {
    if (!b) {
      b_err = std::string(what) + ": null batch";
      status = FH_ERR_BAD_ARG;
      return;
    }
    bool expected = false;
    if (!b->busy.compare_exchange_strong(expected, true)) {
      b_err = std::string(what) + " is not re-entrant";
      status = FH_ERR_BUSY;
    }
  }
  ~BGuard() {
    if (status == FH_OK) b->busy.store(false);
  }
};

#define BGUARD(name)        \
  BGuard _bg(b, name);      \
  if (_bg.status != FH_OK) return _bg.status

int fail(hipError_t err, const char* what) {
  b_err = std::string(what) + ": " + hipGetErrorString(err);
  return (err == hipErrorIllegalAddress || err == hipErrorLaunchFailure) ? FH_ERR_DEVICE_LOST
                                                                         : FH_ERR_HIP;
}

#define BCHK(call)                                 \
  do {                                             \
    hipError_t _e = (call);                        \
    if (_e != hipSuccess) return fail(_e, #call); \
  } while (0)

bool bad_scene(fh_batch* b, int scene, const char* what) {
  if (scene >= 0 && scene < (int)b->e.size()) return false;
  b_err = std::string(what) + ": scene out of range";
  return true;
}

// Work list of one batched launch: scene s (in `order`) takes the next
// multiple of 8 flat blocks for its grid; scenes with an empty grid are left
// out.  Appends the spans (one per scene, by scene index) and the chunk
// table to `host` and returns the grid with offsets into it (patched to
// device addresses once the buffer is placed).
struct ListBuild {
  std::vector<unsigned char> host;
  struct Pending {
    BatchGrid* g;
    size_t span_off, chunk_off;
  };
  std::vector<Pending> pending;

  void add(BatchGrid* g, const std::vector<int>& order, const std::vector<dim3>& grids) {
    const int ns = (int)grids.size();
    std::vector<BSpan> span(ns, BSpan{0, 0, 0, 0});
    std::vector<uint8_t> chunk;
    int off = 0;
    for (int s : order) {
      const int n = (int)(grids[s].x * grids[s].y);
      if (n == 0) continue;
      span[s] = BSpan{off, (int)grids[s].x, (int)grids[s].y, 0};
      const int padded = (n + 7) & ~7;
      chunk.insert(chunk.end(), padded / 8, (uint8_t)s);
      off += padded;
    }
    g->blocks = off;
    auto place = [&](const void* p, size_t n) {
      size_t at = (host.size() + 255) & ~(size_t)255;
      host.resize(at + n);
      if (n) memcpy(host.data() + at, p, n);
      return at;
    };
    Pending pd{g, place(span.data(), span.size() * sizeof(BSpan)),
               place(chunk.data(), chunk.size())};
    pending.push_back(pd);
  }
};

int prepare(fh_batch* b) {
  const int ns = (int)b->e.size();
  BCHK(hipStreamSynchronize(b->s));
  b->desc.assign(ns, SceneDesc{});
  b->max_lc = 0;
  b->coarse_lds = 0;
  for (int s = 0; s < ns; s++) {
    fh_engine* e = b->e[s];
    int st = engine_batch_prepare(e);
    if (st != FH_OK) {
      b_err = std::string("scene ") + std::to_string(s) + ": " + fh_last_error();
      return st;
    }
    SceneDesc& d = b->desc[s];
    d.nx = e->nx;
    d.ny = e->ny;
    d.u[0] = e->u, d.u[1] = e->u2, d.v[0] = e->v, d.v[1] = e->v2;
    d.uh = e->uh, d.vh = e->vh;
    d.dye[0] = e->dye, d.dye[1] = e->dye2, d.dyeh = e->dyeh, d.w = e->w;
    d.lbl = e->lbl, d.sten_f = e->sten_f, d.sten_d = e->sten_d;
    d.div = e->div, d.rvec = e->rvec, d.fcg = e->fcg;
    d.dvec[0] = e->dvec[0], d.dvec[1] = e->dvec[1];
    d.partials = e->partials, d.sums = e->sums, d.d_res = e->d_res, d.done = e->done;
    const dim3 rg = batch_grid_reduce(e->nx, e->ny);
    d.red_blocks = (int)(rg.x * rg.y);
    if (d.red_blocks > kFoldAbove || e->nlev_dc > kBatchMaxLevels) {
      b_err = "fh_batch_step: scene " + std::to_string(s) +
              " is too large for the batched step (step it with fh_step)";
      return FH_ERR_BAD_ARG;
    }
    d.lc = e->lc_dc;
    for (int l = 0; l < e->nlev_dc; l++) {
      const Level& L = e->lv[l];
      d.lv[l] = LevelDesc{L.x, L.xa, L.b, L.info, L.nx, L.ny};
    }
    d.cl = engine_coarse_levels(e, e->lc_dc, e->nlev_dc);
    const CoarseLevels& C = d.cl;
    d.M = e->M, d.cells = e->cells, d.ncoarse = e->ncoarse;
    b->max_lc = std::max(b->max_lc, d.lc);
    b->coarse_lds = std::max(b->coarse_lds, coarse_vcycle_lds_bytes(C));
  }
  // work lists
  ListBuild lb;
  std::vector<dim3> grids(ns);
  for (int k = 0; k < 4; k++) {
    for (int s = 0; s < ns; s++) grids[s] = batch_grid_advect(k, b->desc[s].nx, b->desc[s].ny);
    lb.add(&b->g_adv[k], b->order, grids);
  }
  for (int k = 0; k < kNumBatchProject; k++) {
    for (int s = 0; s < ns; s++) {
      const fh_engine* e = b->e[s];
      const bool conf = (float)e->confinement != 0.f;
      const bool in = (k == kBVorticity || k == kBForces) ? conf
                      : k == kBBuoyancy                    ? (!conf && e->buoyancy != 0.0)
                                                           : true;
      grids[s] = in ? batch_grid_project(k, e->nx, e->ny) : dim3(0, 0);
    }
    lb.add(&b->g_proj[k], b->order, grids);
  }
  for (int s = 0; s < ns; s++) grids[s] = batch_grid_reduce(b->desc[s].nx, b->desc[s].ny);
  lb.add(&b->g_red, b->order, grids);
  for (int s = 0; s < ns; s++) grids[s] = dim3(kBatchSpeedBlocks, 1);
  lb.add(&b->g_speed, b->order, grids);
  for (int l = 0; l < b->max_lc; l++)
    for (int t = 0; t < kNumTileShapes; t++) {
      const int tx = kTileShapes[t].tx, ty = kTileShapes[t].ty;
      for (int s = 0; s < ns; s++) {
        const SceneDesc& d = b->desc[s];
        grids[s] = l < d.lc ? dim3((d.lv[l].nx + tx - 1) / tx, (d.lv[l].ny + ty - 1) / ty)
                            : dim3(0, 0);
      }
      lb.add(&b->g_lv[l][t], b->order, grids);
    }
  if (lb.host.size() > b->lists_bytes) {
    if (b->d_lists) BCHK(hipFree(b->d_lists));
    b->d_lists = nullptr;
    b->lists_bytes = 0;
    BCHK(hipMalloc((void**)&b->d_lists, lb.host.size()));
    b->lists_bytes = lb.host.size();
  }
  BCHK(hipMemcpy(b->d_lists, lb.host.data(), lb.host.size(), hipMemcpyHostToDevice));
  for (auto& pd : lb.pending) {
    pd.g->span = (const BSpan*)(b->d_lists + pd.span_off);
    pd.g->chunk = b->d_lists + pd.chunk_off;
  }
  BCHK(hipMemcpy(b->d_desc, b->desc.data(), ns * sizeof(SceneDesc), hipMemcpyHostToDevice));
  b->q = 0;
  b->dirty = false;
  return FH_OK;
}

int check_launch(const char* what) {
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? FH_OK : fail(err, what);
}

#define BLAUNCHED(what)              \
  do {                               \
    int _st = check_launch(what);    \
    if (_st != FH_OK) return _st;    \
  } while (0)

// One V(2,2) cycle of every scene that has not converged.
int vcycle(fh_batch* b) {
  for (int l = 0; l < b->max_lc; l++) {
    const BatchGrid& g = b->g_lv[l][b->tile[l]];
    if (g.blocks) launch_batch_smooth_restrict(b->s, b->tile[l], b->d_desc, g, l);
  }
  launch_batch_coarse_vcycle(b->s, b->d_desc, (int)b->e.size(), b->coarse_lds);
  for (int l = b->max_lc - 1; l >= 0; l--) {
    const BatchGrid& g = b->g_lv[l][b->tile[l]];
    if (g.blocks) launch_batch_prolong_smooth(b->s, b->tile[l], b->d_desc, g, l);
  }
  BLAUNCHED("batched V-cycle");
  return FH_OK;
}

// The pressure solve of every scene: the initial defect pass, then cycles
// (the first by defect correction, the rest flexible CG) until each scene's
// relative residual is <= tol, at most max_cycles; per-scene cycle counts
// and final residuals into b->iters / b->resid as fh_step reports them.
int solve(fh_batch* b, double tol, int max_cycles) {
  const int ns = (int)b->e.size(), q = b->q;
  // the initial defect pass, forming each scene's divergence (as fh_step does)
  launch_batch_dc_update(b->s, b->d_desc, b->g_red, b->d_step, q, 2);
  launch_batch_finalize(b->s, b->d_desc, ns, b->d_step, 0, 0, 0.0);
  BLAUNCHED("batched initial defect");
  BCHK(hipEventRecord(b->ev_init, b->s));
  std::vector<char> fin(ns, 0), zero(ns, 0);
  int left = ns;
  bool init_read = false;
  auto read_init = [&]() {
    init_read = true;
    for (int s = 0; s < ns; s++)
      if (b->e[s]->h_res[16 + 2] == 0.0) {  // zero divergence: p = 0, no cycle
        zero[s] = fin[s] = 1;
        b->iters[s] = 0;
        b->resid[s] = 0.0;
        left--;
      }
  };
  int c = 0;
  for (; c < max_cycles; c++) {
    int st = vcycle(b);
    if (st != FH_OK) return st;
    if (c == 0) {
      launch_batch_dc_update(b->s, b->d_desc, b->g_red, b->d_step, q, 0);
      launch_batch_finalize(b->s, b->d_desc, ns, b->d_step, 1, 0, tol);
    } else {
      launch_batch_fcg(b->s, b->d_desc, b->g_red, b->d_step, ns, c & 1, c == 1 ? 1 : 0);
      launch_batch_finalize(b->s, b->d_desc, ns, b->d_step, 2, c & 1, tol);
    }
    BLAUNCHED("batched cycle");
    BCHK(hipEventRecord(b->ev_cyc[c & 1], b->s));
    if (c == 0) continue;
    // cycle c is in flight; read cycle c - 1's norms
    if (!init_read) {
      BCHK(hipEventSynchronize(b->ev_init));
      read_init();
    }
    BCHK(hipEventSynchronize(b->ev_cyc[(c - 1) & 1]));
    for (int s = 0; s < ns; s++) {
      if (fin[s]) continue;
      const double* r = b->e[s]->h_res + 8 * ((c - 1) & 1);
      if (r[5] != 0.0) {
        fin[s] = 1;
        b->iters[s] = c;
        b->resid[s] = r[4];
        left--;
      }
    }
    if (left == 0) break;  // cycle c skipped itself on every scene
  }
  if (left > 0) {  // max_cycles ran: the last cycle's norms
    const int last = max_cycles - 1;
    BCHK(hipEventSynchronize(b->ev_cyc[last & 1]));
    if (!init_read) read_init();
    for (int s = 0; s < ns; s++) {
      if (fin[s]) continue;
      b->iters[s] = max_cycles;
      b->resid[s] = b->e[s]->h_res[8 * (last & 1) + 4];
    }
  }
  for (int s = 0; s < ns; s++)
    if (zero[s])
      BCHK(hipMemsetAsync(b->h_step[s].pr, 0, (size_t)b->desc[s].nx * b->desc[s].ny * 8,
                          b->s));
  return FH_OK;
}

template <class T>
void swap_ptr(T*& a, T*& c) {
  T* t = a;
  a = c;
  c = t;
}

}  // namespace

extern "C" {

const char* fh_batch_last_error(void) { return b_err.c_str(); }

int fh_batch_num_tiles(void) { return fh_num_tile_shapes(); }

const char* fh_batch_tile_name(int t) { return fh_tile_shape_name(t); }

fh_batch* fh_batch_create(int nscenes, const int* nx, const int* ny, const double* h) {
  if (nscenes < 1 || nscenes > kBatchMaxScenes || !nx || !ny || !h) {
    b_err = "fh_batch_create: need 1..255 scenes and nx, ny, h arrays";
    return nullptr;
  }
  for (int s = 0; s < nscenes; s++)
    if (nx[s] < 3 || ny[s] < 3 || !(h[s] > 0)) {
      b_err = "fh_batch_create: scene sizes must be at least 3 x 3 with h > 0";
      return nullptr;
    }
  fh_batch* b = new fh_batch();
  b->e.assign(nscenes, nullptr);
  b->prm.resize(nscenes);
  b->dt.assign(nscenes, 0.0);
  b->resid.assign(nscenes, 0.0);
  b->iters.assign(nscenes, 0);
  auto bail = [&](const std::string& why) {
    b_err = why;
    fh_batch_destroy(b);
    return (fh_batch*)nullptr;
  };
  for (int s = 0; s < nscenes; s++) {
    b->e[s] = fh_create(nx[s], ny[s], 1, h[s]);
    if (!b->e[s])
      return bail(std::string("fh_batch_create: scene ") + std::to_string(s) + ": " +
                  fh_last_error());
    b->prm[s].nx = nx[s];
    b->prm[s].ny = ny[s];
    b->prm[s].h = h[s];
    b->cells += (long long)nx[s] * ny[s];
  }
  b->order.resize(nscenes);
  for (int s = 0; s < nscenes; s++) b->order[s] = s;
  std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
    return (long long)nx[a] * ny[a] > (long long)nx[c] * ny[c];
  });
  // the largest scene's default tile shapes per level
  for (int l = 0; l < kBatchMaxLevels && l < kMaxTiledLevels; l++)
    b->tile[l] = b->e[b->order[0]]->tile[l];
  if (hipStreamCreate(&b->s) != hipSuccess ||
      hipMalloc((void**)&b->d_desc, nscenes * sizeof(SceneDesc)) != hipSuccess ||
      hipMalloc((void**)&b->d_step, nscenes * sizeof(BStep)) != hipSuccess ||
      hipHostMalloc((void**)&b->h_step, nscenes * sizeof(BStep), hipHostMallocDefault) !=
          hipSuccess ||
      hipMalloc((void**)&b->d_fpart, (size_t)nscenes * kBatchSpeedBlocks * sizeof(float)) !=
          hipSuccess ||
      hipHostMalloc((void**)&b->h_umax, nscenes * sizeof(float),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&b->d_umax, b->h_umax, 0) != hipSuccess)
    return bail("fh_batch_create: device allocation failed");
  for (hipEvent_t* ev : {&b->ev_umax, &b->ev_init, &b->ev_cyc[0], &b->ev_cyc[1]})
    if (hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess)
      return bail("fh_batch_create: hipEventCreate failed");
  return b;
}

void fh_batch_destroy(fh_batch* b) {
  if (!b) return;
  if (b->s) (void)hipStreamSynchronize(b->s);
  for (fh_engine* e : b->e)
    if (e) fh_destroy(e);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_step) (void)hipFree(b->d_step);
  if (b->h_step) (void)hipHostFree(b->h_step);
  if (b->d_lists) (void)hipFree(b->d_lists);
  if (b->d_fpart) (void)hipFree(b->d_fpart);
  if (b->h_umax) (void)hipHostFree(b->h_umax);
  for (hipEvent_t ev : {b->ev_umax, b->ev_init, b->ev_cyc[0], b->ev_cyc[1]})
    if (ev) (void)hipEventDestroy(ev);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

int fh_batch_upload(fh_batch* b, int scene, const double* u, const double* v, const double* p,
                    const uint8_t* label, const double* dye) {
  BGUARD("HipBatch.upload()");
  if (bad_scene(b, scene, "fh_batch_upload")) return FH_ERR_BAD_ARG;
  BCHK(hipStreamSynchronize(b->s));
  int st = fh_upload(b->e[scene], u, v, p, label, dye);
  if (st != FH_OK) {
    b_err = fh_last_error();
    return st;
  }
  b->dirty = true;
  SceneParams& P = b->prm[scene];
  P.host_umax_valid = false;
  if (u || v) b->umax_valid = false;
  if (u && v) {
    double m = 0.0;
    const size_t nu = (size_t)(P.nx + 1) * P.ny, nv = (size_t)P.nx * (P.ny + 1);
    for (size_t k = 0; k < nu; k++) m = std::max(m, std::fabs(u[k]));
    for (size_t k = 0; k < nv; k++) m = std::max(m, std::fabs(v[k]));
    P.host_umax = m;
    P.host_umax_valid = true;
  }
  return FH_OK;
}

int fh_batch_set_forces(fh_batch* b, int scene, double buoyancy, double confinement,
                        const double* decay) {
  BGUARD("HipBatch.setForces()");
  if (bad_scene(b, scene, "fh_batch_set_forces")) return FH_ERR_BAD_ARG;
  BCHK(hipStreamSynchronize(b->s));
  int st = fh_set_buoyancy(b->e[scene], buoyancy);
  if (st == FH_OK) st = fh_set_confinement(b->e[scene], confinement);
  if (st != FH_OK) {
    b_err = fh_last_error();
    return st;
  }
  for (int c = 0; c < 3; c++) b->prm[scene].decay[c] = decay ? decay[c] : 0.0;
  b->dirty = true;
  return FH_OK;
}

int fh_batch_step(fh_batch* b, int scheme, double cfl, double dt_max, double rho, double tol,
                  int max_cycles) {
  BGUARD("HipBatch.step()");
  if (!(cfl > 0) || !(dt_max > 0) || !(rho > 0) || max_cycles < 1 ||
      (scheme != FH_SEMI_LAGRANGIAN && scheme != FH_MACCORMACK)) {
    b_err = "fh_batch_step: need cfl, dt_max, rho > 0, max_cycles >= 1 and a known scheme";
    return FH_ERR_BAD_ARG;
  }
  const auto t0 = std::chrono::steady_clock::now();
  int st;
  if (b->dirty && (st = prepare(b)) != FH_OK) return st;
  const int ns = (int)b->e.size(), q = b->q;
  const SceneDesc* D = b->d_desc;
  // dt per scene from its largest face speed (simulation.ts:474-479)
  bool need_dev = false;
  for (int s = 0; s < ns; s++) need_dev |= !b->prm[s].host_umax_valid;
  if (need_dev) {
    if (!b->umax_valid) {  // the velocity [q] as it stands
      launch_batch_max_speed(b->s, D, b->g_speed, q ^ 1, ns, b->d_fpart, b->d_umax);
      BLAUNCHED("batched max speed");
      BCHK(hipEventRecord(b->ev_umax, b->s));
    }
    BCHK(hipEventSynchronize(b->ev_umax));
  }
  for (int s = 0; s < ns; s++) {
    SceneParams& P = b->prm[s];
    const fh_engine* e = b->e[s];
    double umax;
    if (P.host_umax_valid) {
      umax = P.host_umax;
      P.host_umax_valid = false;
    } else {
      umax = (double)b->h_umax[s];
    }
    const double dt = umax > 1e-9 ? std::min(dt_max, cfl * P.h / umax) : dt_max;
    BStep& t = b->h_step[s];
    t.c = (float)(dt / e->h);
    float k[3];
    for (int c = 0; c < 3; c++) k[c] = P.decay[c] > 0 ? (float)std::exp(-P.decay[c] * dt) : 1.f;
    t.k0 = k[0], t.k1 = k[1], t.k2 = k[2];
    t.gain = (float)(e->buoyancy * dt);
    t.eps = (float)e->confinement;
    t.hf = (float)e->h;
    t.dtf = (float)dt;
    t.inv_hf = (float)(1.0 / e->h);
    t.scale = rho * P.h * P.h / dt;
    t.grad_scale = dt / (rho * P.h);
    t.inv_h = 1.0 / e->h;
    // the pressure buffers as fh_step would use them (the start p or its
    // extrapolation in time into p2, the result where fh_step's first
    // cycle puts it)
    const StartPlan sp = plan_start(b->e[s], dt);
    fh_engine* es = b->e[s];
    t.p0 = es->p;
    t.xp = sp.xp;
    t.xo = sp.order;
    if (sp.on) {
      t.ps = es->p2;
      t.pr = es->extrap >= 2 ? es->ph[1] : es->ph[0];
    } else {
      t.ps = nullptr;
      t.pr = es->p2;
      es->ph_n = 0;
    }
    b->dt[s] = dt;
  }
  BCHK(hipMemcpyAsync(b->d_step, b->h_step, ns * sizeof(BStep), hipMemcpyHostToDevice, b->s));
  const BStep* S = b->d_step;
  const bool mac = scheme == FH_MACCORMACK;
  // velocity: [q] -> [q ^ 1]
  launch_batch_advect_faces(b->s, D, b->g_adv[kBAdvectFaces], S, q, mac ? 1 : 0);
  if (mac) launch_batch_maccormack_faces(b->s, D, b->g_adv[kBMacCormackFaces], S, q);
  if (b->g_proj[kBVorticity].blocks) {
    launch_batch_vorticity(b->s, D, b->g_proj[kBVorticity], S, q);
    launch_batch_forces(b->s, D, b->g_proj[kBForces], S, q);
  }
  // MacCormack applies buoyancy alone inside its face pass (kb_maccormack_faces)
  if (!mac && b->g_proj[kBBuoyancy].blocks)
    launch_batch_buoyancy(b->s, D, b->g_proj[kBBuoyancy], S, q);
  BLAUNCHED("batched advection and forces");
  if ((st = solve(b, tol, max_cycles)) != FH_OK) {
    for (fh_engine* e : b->e) drop_pressure_history(e);
    return st;
  }
  launch_batch_project(b->s, D, b->g_proj[kBProject], S, q);
  // the next step's CFL reads this velocity; the dye phase leaves it alone
  launch_batch_max_speed(b->s, D, b->g_speed, q, ns, b->d_fpart, b->d_umax);
  BLAUNCHED("batched projection");
  BCHK(hipEventRecord(b->ev_umax, b->s));
  b->umax_valid = true;
  launch_batch_advect_dye(b->s, D, b->g_adv[kBAdvectDye], S, q, mac ? 1 : 0);
  if (mac) launch_batch_maccormack_dye(b->s, D, b->g_adv[kBMacCormackDye], S, q);
  BLAUNCHED("batched dye");
  // the engines' own view: the step's results are their current fields,
  // and with the extrapolated start the history rotates as fh_step's does
  for (int s = 0; s < ns; s++) {
    fh_engine* e = b->e[s];
    const BStep& t = b->h_step[s];
    swap_ptr(e->u, e->u2);
    swap_ptr(e->v, e->v2);
    swap_ptr(e->dye, e->dye2);
    if (t.ps) {
      double* last = e->p;
      e->p = t.pr;
      e->p2 = t.ps;
      if (e->extrap >= 2) {
        e->ph[1] = e->ph[0];
        e->ph_dt[1] = e->ph_dt[0];
      }
      e->ph[0] = last;
      e->ph_dt[0] = e->p_dt;
      e->ph_n = std::min(e->ph_n + 1, e->extrap);
    } else {
      swap_ptr(e->p, e->p2);
    }
    e->p_dt = b->dt[s];
    e->vel_epoch++;
  }
  b->q ^= 1;
  b->last_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  return FH_OK;
}

int fh_batch_set_warm_extrapolation(fh_batch* b, int order) {
  BGUARD("HipBatch.setWarmExtrapolation()");
  BCHK(hipStreamSynchronize(b->s));
  for (fh_engine* e : b->e) {
    const int st = fh_set_warm_extrapolation(e, order);
    if (st != FH_OK) {
      b_err = fh_last_error();
      return st;
    }
  }
  return FH_OK;
}

int fh_batch_results(fh_batch* b, double* dt, int* iters, double* resid) {
  BGUARD("HipBatch.results()");
  for (size_t s = 0; s < b->e.size(); s++) {
    if (dt) dt[s] = b->dt[s];
    if (iters) iters[s] = b->iters[s];
    if (resid) resid[s] = b->resid[s];
  }
  return FH_OK;
}

int fh_batch_last_step_ms(fh_batch* b, float* ms) {
  BGUARD("HipBatch.lastStepMs()");
  if (!ms) return FH_ERR_BAD_ARG;
  *ms = b->last_ms;
  return FH_OK;
}

long long fh_batch_cells(fh_batch* b) { return b ? b->cells : 0; }

int fh_batch_read(fh_batch* b, int scene, double* u, double* v, double* p, double* dye) {
  BGUARD("HipBatch.read()");
  if (bad_scene(b, scene, "fh_batch_read")) return FH_ERR_BAD_ARG;
  BCHK(hipStreamSynchronize(b->s));
  fh_engine* e = b->e[scene];
  int st = FH_OK;
  if ((u || v) && (st = fh_read_velocity(e, u, v)) != FH_OK) return b_err = fh_last_error(), st;
  if (p && (st = fh_read_pressure(e, p)) != FH_OK) return b_err = fh_last_error(), st;
  if (dye && (st = fh_read_dye(e, dye)) != FH_OK) return b_err = fh_last_error(), st;
  return FH_OK;
}

int fh_batch_set_tile(fh_batch* b, int level, int shape) {
  BGUARD("HipBatch.setTile()");
  if (shape < 0 || shape >= kNumTileShapes || level < -1 || level >= kBatchMaxLevels) {
    b_err = "fh_batch_set_tile: unknown shape or level";
    return FH_ERR_BAD_ARG;
  }
  for (int l = 0; l < kBatchMaxLevels; l++)
    if (level < 0 || l == level) b->tile[l] = shape;
  return FH_OK;
}

// The tile shape of each level index of the batched launches.
int fh_batch_get_tiles(fh_batch* b, int* shapes, int* nlevels) {
  BGUARD("HipBatch.getTiles()");
  if (!shapes || !nlevels) return FH_ERR_BAD_ARG;
  int st;
  if (b->dirty && (st = prepare(b)) != FH_OK) return st;
  for (int l = 0; l < b->max_lc; l++) shapes[l] = b->tile[l];
  *nlevels = b->max_lc;
  return FH_OK;
}

// Autotuner of the batched multigrid launches: per level index, every tile
// shape's pre-smooth+restrict and prolong+post-smooth launch pair over all
// scenes, timed with events (kTuneRounds interleaved rounds, each shape's
// best kept).  Every scene's `done` is cleared first so no block skips its
// work; the next solve's initial pass sets it again.
int fh_batch_tune(fh_batch* b, int reps, float* ms_out) {
  BGUARD("HipBatch.tune()");
  if (reps < 1) reps = 1;
  int st;
  if (b->dirty && (st = prepare(b)) != FH_OK) return st;
  for (const SceneDesc& d : b->desc) BCHK(hipMemsetAsync(d.done, 0, sizeof(int), b->s));
  hipEvent_t a, z;
  BCHK(hipEventCreate(&a));
  BCHK(hipEventCreate(&z));
  constexpr int kRounds = 3;
  for (int l = 0; l < b->max_lc; l++) {
    float msr[kNumTileShapes];
    for (int t = 0; t < kNumTileShapes; t++) msr[t] = 1e30f;
    for (int round = 0; round < kRounds; round++)
      for (int t = 0; t < kNumTileShapes; t++) {
        const BatchGrid& g = b->g_lv[l][t];
        for (int r = -2; r < reps; r++) {
          if (r == 0) BCHK(hipEventRecord(a, b->s));
          launch_batch_smooth_restrict(b->s, t, b->d_desc, g, l);
          launch_batch_prolong_smooth(b->s, t, b->d_desc, g, l);
        }
        BCHK(hipEventRecord(z, b->s));
        BCHK(hipEventSynchronize(z));
        BLAUNCHED("batched tile tuning");
        float ms = 0.f;
        BCHK(hipEventElapsedTime(&ms, a, z));
        msr[t] = std::min(msr[t], ms / reps);
      }
    int best = 0;
    for (int t = 0; t < kNumTileShapes; t++) {
      if (ms_out) ms_out[l * kNumTileShapes + t] = msr[t];
      if (msr[t] < msr[best]) best = t;
    }
    b->tile[l] = best;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(z);
  return FH_OK;
}

}  // extern "C"
