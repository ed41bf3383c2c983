// Batched scenes: many independent grids of uneven sizes advanced together on
// one GPU (include/flowhip.h, fh_batch_*).
//
// No counterpart in the reference, which steps one grid per Simulation
// (simulation.ts:467-587).  Each scene is a single-grid engine (fh_engine:
// its own buffers and HIP stream) and takes the engine's own step, with dt
// from its own CFL as Simulation.step computes it (simulation.ts:474-479).
// A batch step hands the scenes, largest first, to a pool of host threads;
// each thread steps one scene at a time, so the scenes' kernels -- on
// grids this small, mostly latency-bound multigrid phases and host
// convergence checks -- interleave on the device across the engines'
// streams instead of one grid running at a time.  Every scene computes
// exactly what fh_step computes for it alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flowhip.h"

namespace {

thread_local std::string b_err;

// Host threads stepping scenes.  Measured on 64 scenes of 256^2..1024^2:
// 1 thread 40.1 ms/step, 4 17.2, 8 16.1, 16 15.1; raising the process's
// hardware queues from 4 to 8 or 16 made it slower (23 ms), so the streams
// share the default queues.
constexpr int kBatchThreads = 16;

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

bool bad_scene(fh_batch* b, int scene, const char* what) {
  if (scene >= 0 && scene < (int)b->e.size()) return false;
  b_err = std::string(what) + ": scene out of range";
  return true;
}

// One scene's step: Simulation.step's dt and factors (simulation.ts:474-479,
// dyeKeepFor simulation.ts:243-253), then the engine's fused step.
int step_scene(fh_batch* b, int s, int scheme, double cfl, double dt_max, double rho, double tol,
               int max_cycles) {
  fh_engine* e = b->e[s];
  const SceneParams& P = b->prm[s];
  double umax = 0.0;
  int st = FH_OK;
  if (b->prm[s].host_umax_valid) {
    umax = b->prm[s].host_umax;
    b->prm[s].host_umax_valid = false;
  } else if ((st = fh_max_face_speed(e, &umax)) != FH_OK) {
    return st;
  }
  const double dt = umax > 1e-9 ? std::min(dt_max, cfl * P.h / umax) : dt_max;
  const double scale = rho * P.h * P.h / dt, grad_scale = dt / (rho * P.h);
  double keep[3];
  for (int c = 0; c < 3; c++) keep[c] = P.decay[c] > 0 ? std::exp(-P.decay[c] * dt) : 1.0;
  int it = 0;
  st = fh_step(e, scheme, dt, scale, grad_scale, keep, FH_SOLVER_MG_FCG, max_cycles, 1.0, tol,
               &it);
  if (st != FH_OK) return st;
  double hist[512];
  const int n = fh_solve_history(e, hist, 512);
  b->dt[s] = dt;
  b->iters[s] = it;
  b->resid[s] = n > 0 ? hist[n - 1] : 0.0;
  return FH_OK;
}

}  // namespace

extern "C" {

const char* fh_batch_last_error(void) { return b_err.c_str(); }

int fh_batch_num_tiles(void) { return fh_num_tile_shapes(); }

const char* fh_batch_tile_name(int t) { return fh_tile_shape_name(t); }

fh_batch* fh_batch_create(int nscenes, const int* nx, const int* ny, const double* h) {
  if (nscenes < 1 || nscenes > 255 || !nx || !ny || !h) {
    b_err = "fh_batch_create: need 1..255 scenes and nx, ny, h arrays";
    return nullptr;
  }
  for (int s = 0; s < nscenes; s++)
    if (nx[s] < 2 || ny[s] < 2 || !(h[s] > 0)) {
      b_err = "fh_batch_create: scene sizes must be at least 2 x 2 with h > 0";
      return nullptr;
    }
  fh_batch* b = new fh_batch();
  b->e.assign(nscenes, nullptr);
  b->prm.resize(nscenes);
  b->dt.assign(nscenes, 0.0);
  b->resid.assign(nscenes, 0.0);
  b->iters.assign(nscenes, 0);
  for (int s = 0; s < nscenes; s++) {
    b->e[s] = fh_create(nx[s], ny[s], 1, h[s]);
    if (!b->e[s]) {
      b_err = std::string("fh_batch_create: scene ") + std::to_string(s) + ": " + fh_last_error();
      fh_batch_destroy(b);
      return nullptr;
    }
    b->prm[s].nx = nx[s];
    b->prm[s].ny = ny[s];
    b->prm[s].h = h[s];
    b->cells += (long long)nx[s] * ny[s];
  }
  b->order.resize(nscenes);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
    return (long long)nx[a] * ny[a] > (long long)nx[c] * ny[c];
  });
  return b;
}

void fh_batch_destroy(fh_batch* b) {
  if (!b) return;
  for (fh_engine* e : b->e)
    if (e) fh_destroy(e);
  delete b;
}

int fh_batch_upload(fh_batch* b, int scene, const double* u, const double* v, const double* p,
                    const uint8_t* label, const double* dye) {
  BGUARD("HipBatch.upload()");
  if (bad_scene(b, scene, "fh_batch_upload")) return FH_ERR_BAD_ARG;
  int st = fh_upload(b->e[scene], u, v, p, label, dye);
  if (st != FH_OK) {
    b_err = fh_last_error();
    return st;
  }
  SceneParams& P = b->prm[scene];
  P.host_umax_valid = false;
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
  int st = fh_set_buoyancy(b->e[scene], buoyancy);
  if (st == FH_OK) st = fh_set_confinement(b->e[scene], confinement);
  if (st != FH_OK) {
    b_err = fh_last_error();
    return st;
  }
  for (int c = 0; c < 3; c++) b->prm[scene].decay[c] = decay ? decay[c] : 0.0;
  return FH_OK;
}

int fh_batch_step(fh_batch* b, int scheme, double cfl, double dt_max, double rho, double tol,
                  int max_cycles, int graph_cycles) {
  BGUARD("HipBatch.step()");
  (void)graph_cycles;  // accepted for the interface; the engines launch directly
  if (!(cfl > 0) || !(dt_max > 0) || !(rho > 0) || max_cycles < 1 ||
      (scheme != FH_SEMI_LAGRANGIAN && scheme != FH_MACCORMACK)) {
    b_err = "fh_batch_step: need cfl, dt_max, rho > 0, max_cycles >= 1 and a known scheme";
    return FH_ERR_BAD_ARG;
  }
  const int ns = (int)b->e.size();
  const auto t0 = std::chrono::steady_clock::now();
  std::atomic<int> next{0};
  std::vector<int> status(ns, FH_OK);
  std::vector<std::string> msg(ns);
  auto work = [&]() {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= ns) return;
      const int s = b->order[k];
      status[s] = step_scene(b, s, scheme, cfl, dt_max, rho, tol, max_cycles);
      if (status[s] != FH_OK) msg[s] = fh_last_error();  // this thread's error text
    }
  };
  const int nt = std::min(ns, kBatchThreads);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++) pool.emplace_back(work);
  work();
  for (std::thread& t : pool) t.join();
  b->last_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  for (int s = 0; s < ns; s++)
    if (status[s] != FH_OK) {
      b_err = "scene " + std::to_string(s) + ": " + msg[s];
      return status[s];
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
  fh_engine* e = b->e[scene];
  int st = FH_OK;
  if ((u || v) && (st = fh_read_velocity(e, u, v)) != FH_OK) return b_err = fh_last_error(), st;
  if (p && (st = fh_read_pressure(e, p)) != FH_OK) return b_err = fh_last_error(), st;
  if (dye && (st = fh_read_dye(e, dye)) != FH_OK) return b_err = fh_last_error(), st;
  return FH_OK;
}

int fh_batch_set_tile(fh_batch* b, int level, int shape) {
  BGUARD("HipBatch.setTile()");
  for (fh_engine* e : b->e) {
    int st = fh_set_tile(e, level, shape);
    if (st != FH_OK) {
      b_err = fh_last_error();
      return st;
    }
  }
  return FH_OK;
}

// The tiled levels of the largest scene and their shapes.
int fh_batch_get_tiles(fh_batch* b, int* shapes, int* nlevels) {
  BGUARD("HipBatch.getTiles()");
  return fh_get_tiles(b->e[b->order[0]], shapes, nlevels);
}

// Every scene tunes its own levels (fh_tune_tiles); ms_out (optional) gets
// the largest scene's timings, nlevels x fh_batch_num_tiles() ms.
int fh_batch_tune(fh_batch* b, int reps, float* ms_out) {
  BGUARD("HipBatch.tune()");
  for (size_t k = 0; k < b->order.size(); k++) {
    fh_engine* e = b->e[b->order[k]];
    int st = fh_tune_tiles(e, reps, k == 0 ? ms_out : nullptr);
    if (st != FH_OK) {
      b_err = fh_last_error();
      return st;
    }
  }
  return FH_OK;
}

}  // extern "C"
