// Batched scenes: many independent MAC grids of uneven size advanced together,
// one launch per phase for the whole batch.
//
// The reference steps one grid per Simulation (simulation.ts:467-587); its
// closest relative of this is the one-submit recording of a fused step
// (stepGpu.ts:441-542).  Here every phase kernel is persistent: a grid of at
// most `cap` workgroups walks a task list built once at creation, one task
// being (scene, block) for the streaming stencils or (scene, tile) for the
// multigrid tiles of one level, so a 256^2 and a 1024^2 scene share each
// launch.  Per-scene state lives in a device table (BScene) that the kernels
// read and update: dt from the scene's own CFL, the p/p2 and u/u2 ping-pong
// pointers, the defect-correction norms, iteration count and `done` flag.
// With nothing left on the host between phases, a whole step (max speed ->
// dt -> advection -> forces -> divergence -> K defect-correction V-cycles)
// is captured once into a hipGraph and replayed; scenes that converge early
// skip the remaining cycles on the device.
//
// Every per-cell body is the single-grid engine's (fh_stencil.hpp,
// fh_mgtile.hpp, the defect-correction update of fh_mg.hip), so a scene
// steps to the same bits batched or alone, whatever tile shape the
// autotuner picked (temporal blocking is tiling-invariant).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flowhip.h"
#include "fh_mgtile.hpp"
#include "fh_stencil.hpp"

using namespace fh;

namespace {

thread_local std::string b_err;

constexpr int kMaxLev = 16;
constexpr int SBX = 64, SBY = 4;  // streaming block: 64 x 4 cells, 256 threads
constexpr int kCoarseLdsB = 65536;

struct BLevel {
  int nx, ny;
  float *x, *xa, *b;
  uint8_t *lbl, *info;
};

// Per-scene device descriptor.  Fields below "step values" and the
// pointers are rewritten by kernels during a step.
struct BScene {
  int nx, ny;
  double h;
  float *u, *v, *u2, *v2, *uh, *vh, *w, *dye, *dye2, *dyeh;
  uint8_t* lbl;
  double *p, *p2, *div;
  double buoyancy, confinement, decay[3];
  // step values (written by k_b_dt)
  double dt, scale, gs;
  float c, gain, eps, k0, k1, k2;
  // multigrid (defect-correction hierarchy, coarsest side <= 16)
  int nlev, lc;
  BLevel lv[kMaxLev];
  CoarseLevels coarse;
  double* M;
  int* cells;
  int ncoarse;
  // solver state
  int done, zero_p, iters;
  double rel;
  // this scene's slice of the cell and face task lists (reduction ranges)
  int cell_t0, cell_nt, face_t0, face_nt;
};

// Result slot per scene, in mapped host memory.
struct BRes {
  double dt, rel;
  int iters, done;
};

// Task: scene (8 bits) | by (12 bits) | bx (12 bits).
__host__ __device__ inline uint32_t task_pack(int s, int bx, int by) {
  return ((uint32_t)s << 24) | ((uint32_t)by << 12) | (uint32_t)bx;
}
__device__ inline void task_unpack(uint32_t t, int* s, int* bx, int* by) {
  *s = (int)(t >> 24);
  *by = (int)((t >> 12) & 0xfff);
  *bx = (int)(t & 0xfff);
}

#define FOR_TASKS(tasks, n)                                          \
  for (int t_ = blockIdx.x; t_ < (n); t_ += gridDim.x)               \
    if (int s_, bx_, by_; (task_unpack((tasks)[t_], &s_, &bx_, &by_), true))

// ---- streaming kernels ---------------------------------------------------------

// Streaming kernels run 64 x 4 blocks (threadIdx.x < 64, threadIdx.y < 4),
// as the single-grid kernels do.
__device__ __forceinline__ int sx_of(int bx) { return bx * SBX + (int)threadIdx.x; }
__device__ __forceinline__ int sy_of(int by) { return by * SBY + (int)threadIdx.y; }
__device__ __forceinline__ int flat_tid() { return (int)(threadIdx.x + threadIdx.y * SBX); }

// simulation.ts:441-459 per scene: block maxima of |u|, |v| over face tasks.
__global__ __launch_bounds__(256) void k_b_max_speed(BScene* sc, const uint32_t* tasks, int n,
                                                     float* partials) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    int i = sx_of(bx_), j = sy_of(by_);
    float m = 0.f;
    if (j < S.ny && i <= S.nx) m = fabsf(S.u[i + j * (S.nx + 1)]);
    if (i < S.nx && j <= S.ny) m = fmaxf(m, fabsf(S.v[i + j * S.nx]));
    m = wave_max(m);
    __shared__ float sh[4];
    if (threadIdx.x == 0) sh[threadIdx.y] = m;
    __syncthreads();
    if (flat_tid() == 0) partials[t_] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
    __syncthreads();
  }
}

// Per scene (one workgroup each): dt = min(dt_max, cfl h / uMax) as
// simulation.ts:477-479, the derived factors, and a fresh solver state.
__global__ __launch_bounds__(256) void k_b_dt(BScene* sc, const float* partials, double cfl,
                                              double dt_max, double rho, BRes* res) {
  BScene& S = sc[blockIdx.x];
  float m = 0.f;
#pragma unroll 4
  for (int t = threadIdx.x; t < S.face_nt; t += 256) m = fmaxf(m, partials[S.face_t0 + t]);
  m = wave_max(m);
  __shared__ float sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double umax = (double)fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
  double dt = umax > 1e-9 ? fmin(dt_max, cfl * S.h / umax) : dt_max;
  S.dt = dt;
  S.scale = rho * S.h * S.h / dt;
  S.gs = dt / (rho * S.h);
  S.c = (float)(dt / S.h);
  S.gain = (float)(S.buoyancy * dt);
  S.eps = (float)S.confinement;
  S.k0 = S.decay[0] > 0 ? (float)exp(-S.decay[0] * dt) : 1.f;
  S.k1 = S.decay[1] > 0 ? (float)exp(-S.decay[1] * dt) : 1.f;
  S.k2 = S.decay[2] > 0 ? (float)exp(-S.decay[2] * dt) : 1.f;
  S.done = 0;
  S.zero_p = 0;
  S.iters = 0;
  S.rel = 0.0;
  res[blockIdx.x].dt = dt;
}

// which: 1 = swap u/u2 and v/v2; 2 = swap dye/dye2.
__global__ void k_b_swap(BScene* sc, int n, int which) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  BScene& S = sc[s];
  if (which == 1) {
    float* t = S.u;
    S.u = S.u2;
    S.u2 = t;
    t = S.v;
    S.v = S.v2;
    S.v2 = t;
  } else {
    float* t = S.dye;
    S.dye = S.dye2;
    S.dye2 = t;
  }
}

// mode 0: semi-Lagrangian into u2/v2; 1: MacCormack forward into uh/vh.
__global__ __launch_bounds__(256) void k_b_advect_faces(const BScene* sc, const uint32_t* tasks,
                                                        int n, int mode) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    float* uo = mode ? S.uh : S.u2;
    float* vo = mode ? S.vh : S.v2;
    advect_faces_at(S.u, S.v, S.u, S.v, uo, vo, S.lbl, S.nx, S.ny, S.c, sx_of(bx_), sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_maccormack_faces(const BScene* sc,
                                                            const uint32_t* tasks, int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    maccormack_faces_at(S.u, S.v, S.uh, S.vh, S.u2, S.v2, S.lbl, S.nx, S.ny, S.c, sx_of(bx_),
                        sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_vorticity(const BScene* sc, const uint32_t* tasks,
                                                     int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    if (S.eps == 0.f) continue;
    vorticity_at(S.u, S.v, S.w, S.nx, S.ny, (float)(1.0 / S.h), sx_of(bx_), sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_forces(const BScene* sc, const uint32_t* tasks, int n) {
  const StirBox none{};
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    if (S.gain == 0.f && S.eps == 0.f) continue;
    forces_at(S.u, S.v, S.lbl, S.dye, S.w, S.nx, S.ny, 1, (float)S.h, (float)S.dt, S.gain, S.eps,
              none, sx_of(bx_), sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_divergence(const BScene* sc, const uint32_t* tasks,
                                                      int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    divergence_at(S.u, S.v, S.div, S.nx, S.ny, 1.0 / S.h, sx_of(bx_), sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_subtract(const BScene* sc, const uint32_t* tasks,
                                                    int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    subtract_at(S.u, S.v, S.p, S.lbl, S.nx, S.ny, S.gs, sx_of(bx_), sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_outflow(const BScene* sc, const uint32_t* tasks,
                                                   int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    outflow_at(S.u, S.v, S.lbl, S.nx, S.ny, bx_ * 256 + flat_tid());
  }
}

// mode 0: semi-Lagrangian into dye2 with keep; 1: MacCormack forward into dyeh.
__global__ __launch_bounds__(256) void k_b_advect_dye(const BScene* sc, const uint32_t* tasks,
                                                      int n, int mode) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    if (mode)
      advect_dye_at(S.u, S.v, S.dye, S.dyeh, S.lbl, S.nx, S.ny, 1, S.c, 1.f, 1.f, 1.f,
                    sx_of(bx_), sy_of(by_));
    else
      advect_dye_at(S.u, S.v, S.dye, S.dye2, S.lbl, S.nx, S.ny, 1, S.c, S.k0, S.k1, S.k2,
                    sx_of(bx_), sy_of(by_));
  }
}

__global__ __launch_bounds__(256) void k_b_maccormack_dye(const BScene* sc, const uint32_t* tasks,
                                                          int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    maccormack_dye_at(S.u, S.v, S.dye, S.dyeh, S.dye2, S.lbl, S.nx, S.ny, 1, S.c, S.k0, S.k1,
                      S.k2, sx_of(bx_), sy_of(by_));
  }
}

// ---- pressure: defect correction -----------------------------------------------

// The single-grid k_dc_update body (fh_mg.hip) for one 64x4 block of a
// scene; partial sums go to slot t of `partials`.  initial: no correction
// (e = 0), the residual of the warm start.
__global__ __launch_bounds__(256) void k_b_dc_update(const BScene* sc, const uint32_t* tasks,
                                                     int n, double* partials, int initial) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    if (S.done) continue;
    const int nx = S.nx, ny = S.ny;
    const int i = sx_of(bx_), j = sy_of(by_);
    const double* po = S.p;
    double* pn = S.p2;
    const float* e = initial ? nullptr : S.lv[0].x;
    const uint8_t* info = S.lv[0].info;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < nx && j < ny) {
      const int k = i + j * nx;
      const int kE = i + 1 < nx ? k + 1 : k, kW = i > 0 ? k - 1 : k;
      const int kN = j + 1 < ny ? k + nx : k, kS = j > 0 ? k - nx : k;
      const int f = info[k];
      const double d = S.div[k];
      double pc = po[k], pE = po[kE], pW = po[kW], pN = po[kN], pS = po[kS];
      if (e) {
        pc += (double)e[k];
        pE += (double)e[kE];
        pW += (double)e[kW];
        pN += (double)e[kN];
        pS += (double)e[kS];
      }
      pn[k] = pc;
      double r = 0.0;
      if (f & 0x80) {
        acc[2] += d * d;
        acc[3] += 1.0;
        const int cnt = f & 7;
        if (cnt > 0) {
          double sum = 0.0;
          sum += (f & 0x08) ? pE : 0.0;
          sum += (f & 0x10) ? pW : 0.0;
          sum += (f & 0x20) ? pN : 0.0;
          sum += (f & 0x40) ? pS : 0.0;
          r = -S.scale * d - ((double)cnt * pc - sum);
          acc[0] += r * r;
          acc[1] += 1.0;
        }
      }
      S.lv[0].b[k] = (float)r;
    }
    // block_sum_store indexes partials by block id; here by task
    __shared__ double sh[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double v = wave_sum(acc[q]);
      if (threadIdx.x == 0) sh[q][threadIdx.y] = v;
    }
    __syncthreads();
    const int ft = flat_tid();
    if (ft < 4) partials[(size_t)t_ * 4 + ft] = ((sh[ft][0] + sh[ft][1]) + sh[ft][2]) + sh[ft][3];
    __syncthreads();
  }
}

// Per scene: relative residual from the task partials, p/p2 swap, counters,
// convergence.  initial: the warm-start residual (a zero divergence gives
// p = 0 and no cycles, as the single-grid solve does).
__global__ __launch_bounds__(256) void k_b_dc_norm(BScene* sc, const double* partials, double tol,
                                                   int initial, BRes* res) {
  BScene& S = sc[blockIdx.x];
  if (S.done) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int t = threadIdx.x; t < S.cell_nt; t += 256) {
    const double* p = partials + (size_t)(S.cell_t0 + t) * 4;
    acc[0] += p[0];
    acc[1] += p[1];
    acc[2] += p[2];
    acc[3] += p[3];
  }
  __shared__ double sh[4][4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    double v = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[4];
  for (int q = 0; q < 4; q++) v[q] = ((sh[q][0] + sh[q][1]) + sh[q][2]) + sh[q][3];
  double* t = S.p;
  S.p = S.p2;
  S.p2 = t;
  double rel = 0.0;
  if (v[1] > 0 && v[2] > 0) rel = sqrt(v[0] / v[1]) / (fabs(S.scale) * sqrt(v[2] / v[3]));
  if (initial) {
    if (v[2] == 0.0) {
      S.zero_p = 1;
      S.done = 1;
    }
  } else {
    S.iters++;
  }
  S.rel = rel;
  if (tol > 0 && rel <= tol) S.done = 1;
  BRes& R = res[blockIdx.x];
  R.rel = rel;
  R.iters = S.iters;
  R.done = S.done;
}

// p = 0 on scenes whose divergence was zero (single-grid solve's guard).
__global__ __launch_bounds__(256) void k_b_zero_p(const BScene* sc, const uint32_t* tasks, int n) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    if (!S.zero_p) continue;
    int i = sx_of(bx_), j = sy_of(by_);
    if (i < S.nx && j < S.ny) S.p[i + j * S.nx] = 0.0;
  }
}

// Scenes that hit the cycle cap stop where they are; mark them so the
// host's check sees every scene settled.
__global__ void k_b_set_done(BScene* sc, int n, int v) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) sc[s].done = v;
}

// ---- pressure: multigrid tiles -------------------------------------------------

template <int TX, int TY>
__global__ __launch_bounds__(NT) void k_b_smooth_restrict(const BScene* sc, const uint32_t* tasks,
                                                          int n, int l) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    const BLevel& L = S.lv[l];
    const BLevel& N = S.lv[l + 1];
    smooth_restrict_tile<TX, TY>(nullptr, L.xa, L.b, L.info, L.nx, L.ny, N.b, N.nx, N.ny, 1,
                                 &S.done, bx_, by_);
    __syncthreads();
  }
}

template <int TX, int TY>
__global__ __launch_bounds__(NT) void k_b_prolong_smooth(const BScene* sc, const uint32_t* tasks,
                                                         int n, int l) {
  FOR_TASKS(tasks, n) {
    const BScene& S = sc[s_];
    const BLevel& L = S.lv[l];
    const BLevel& N = S.lv[l + 1];
    prolong_smooth_tile<TX, TY>(L.xa, L.x, L.b, L.info, L.nx, L.ny, N.x, N.nx, N.ny, &S.done,
                                bx_, by_);
    __syncthreads();
  }
}

// One workgroup per scene: the scene's sub-cycle from its level lc down.
__global__ __launch_bounds__(NC) void k_b_coarse(const BScene* sc) {
  const BScene& S = sc[blockIdx.x];
  if (S.done) return;
  extern __shared__ unsigned char smem[];
  coarse_vcycle_body<NC>(S.coarse, S.lv[S.lc].b, nullptr, S.lv[S.lc].x, S.M, S.cells, S.ncoarse, 1,
                     smem);
}

void launch_b_smooth_restrict(int shape, dim3 g, hipStream_t s, const BScene* sc,
                              const uint32_t* tasks, int n, int l) {
  switch (shape) {
    case 0: k_b_smooth_restrict<64, 32><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 1: k_b_smooth_restrict<32, 32><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 2: k_b_smooth_restrict<64, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 3: k_b_smooth_restrict<128, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 5: k_b_smooth_restrict<16, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 6: k_b_smooth_restrict<32, 8><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    default: k_b_smooth_restrict<32, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
  }
}

void launch_b_prolong_smooth(int shape, dim3 g, hipStream_t s, const BScene* sc,
                             const uint32_t* tasks, int n, int l) {
  switch (shape) {
    case 0: k_b_prolong_smooth<64, 32><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 1: k_b_prolong_smooth<32, 32><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 2: k_b_prolong_smooth<64, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 3: k_b_prolong_smooth<128, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 5: k_b_prolong_smooth<16, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    case 6: k_b_prolong_smooth<32, 8><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
    default: k_b_prolong_smooth<32, 16><<<g, NT, 0, s>>>(sc, tasks, n, l); break;
  }
}

int level_count_b(int nx, int ny, int coarsest) {
  int n = 1;
  while (nx > coarsest || ny > coarsest) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
    n++;
  }
  return n;
}

// Stepping is switched off: the first device run of k_b_* took a GPU memory
// aperture fault (a flat access outside every aperture) whose cause is not
// yet shown; the host-side tables were checked and hold valid pointers and
// tasks.  Until the cause is found the kernels are not launched.
constexpr bool kBatchKernelsEnabled = false;
constexpr const char* kBatchOff =
    "batched stepping is switched off: its kernels took a GPU memory fault whose cause is "
    "not yet found";

}  // namespace

// ---- host object ---------------------------------------------------------------------

struct DevList {
  uint32_t* d = nullptr;
  int n = 0;
};

struct fh_batch {
  int ns = 0;
  hipStream_t s = nullptr;
  std::vector<BScene> h;  // host mirror (pointers as created; device copy is authoritative)
  BScene* d = nullptr;
  BRes* res = nullptr;    // mapped host
  BRes* d_res = nullptr;
  std::vector<void*> allocs;
  DevList faces, cells, outflow;
  int max_lc = 0;
  int tile[kMaxLev] = {};
  DevList tiles[kMaxLev];
  float* fpart = nullptr;
  double* dpart = nullptr;
  size_t coarse_lds = 0;
  int cap = 2048;  // persistent grid: at most this many workgroups
  std::atomic<bool> busy{false};
  bool lost = false;
  // captured step
  hipGraphExec_t g_main = nullptr, g_cycle = nullptr, g_post = nullptr;
  double key[8] = {};
  bool graph_ok = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evs = nullptr;
  float last_ms = 0.f;
  long long cells_total = 0;
};

namespace {

int bfail(fh_batch* b, hipError_t err, const char* what) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(err));
  b_err = buf;
  if (b && (err == hipErrorIllegalAddress || err == hipErrorLaunchFailure)) b->lost = true;
  return b && b->lost ? FH_ERR_DEVICE_LOST : FH_ERR_HIP;
}

#define BCHK(call)                                     \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return bfail(b, _e, #call);  \
  } while (0)

struct BGuard {
  fh_batch* b;
  int status = FH_OK;
  BGuard(fh_batch* bb, const char* what) : b(bb) {
    if (!b) {
      b_err = std::string(what) + ": null batch";
      status = FH_ERR_BAD_ARG;
    } else if (b->lost) {
      b_err = "GPU device lost";
      status = FH_ERR_DEVICE_LOST;
    } else {
      bool expected = false;
      if (!b->busy.compare_exchange_strong(expected, true)) {
        b_err = std::string(what) + " is not re-entrant";
        status = FH_ERR_BUSY;
      }
    }
  }
  ~BGuard() {
    if (status == FH_OK) b->busy.store(false);
  }
};

#define BGUARD(name)          \
  BGuard _g(b, name);         \
  if (_g.status != FH_OK) return _g.status

template <class T>
int balloc(fh_batch* b, T** p, size_t n) {
  hipError_t err = hipMalloc((void**)p, n * sizeof(T) + 256);
  if (err != hipSuccess) return bfail(b, err, "hipMalloc");
  b->allocs.push_back((void*)*p);
  err = hipMemset(*p, 0, n * sizeof(T) + 256);
  if (err != hipSuccess) return bfail(b, err, "hipMemset");
  return FH_OK;
}

int upload_list(fh_batch* b, DevList& L, const std::vector<uint32_t>& v) {
  if (L.d) {
    BCHK(hipFree(L.d));
    L.d = nullptr;
  }
  L.n = (int)v.size();
  if (L.n == 0) return FH_OK;
  BCHK(hipMalloc((void**)&L.d, v.size() * sizeof(uint32_t)));
  BCHK(hipMemcpy(L.d, v.data(), v.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  return FH_OK;
}

dim3 pgrid(const fh_batch* b, int n) { return dim3(std::max(1, std::min(n, b->cap))); }

// Tile task lists of level l for the level's chosen shape.
int build_tile_list(fh_batch* b, int l) {
  std::vector<uint32_t> v;
  const TileShape& T = kTileShapes[b->tile[l]];
  for (int s = 0; s < b->ns; s++) {
    const BScene& S = b->h[s];
    if (l >= S.lc) continue;
    const BLevel& L = S.lv[l];
    int gx = (L.nx + T.tx - 1) / T.tx, gy = (L.ny + T.ty - 1) / T.ty;
    for (int by = 0; by < gy; by++)
      for (int bx = 0; bx < gx; bx++) v.push_back(task_pack(s, bx, by));
  }
  return upload_list(b, b->tiles[l], v);
}

void drop_graphs(fh_batch* b) {
  for (hipGraphExec_t* g : {&b->g_main, &b->g_cycle, &b->g_post})
    if (*g) {
      (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  b->graph_ok = false;
}

struct StepArgs {
  int scheme;
  double cfl, dt_max, rho, tol;
};

// Enqueue helpers.  None synchronises; all read per-scene values from the
// device table, so they can be captured into a graph.
int enq_pre(fh_batch* b, const StepArgs& a) {
  hipStream_t s = b->s;
  k_b_max_speed<<<pgrid(b, b->faces.n), dim3(SBX, SBY), 0, s>>>(b->d, b->faces.d, b->faces.n, b->fpart);
  k_b_dt<<<b->ns, 256, 0, s>>>(b->d, b->fpart, a.cfl, a.dt_max, a.rho, b->d_res);
  if (a.scheme == FH_MACCORMACK) {
    k_b_advect_faces<<<pgrid(b, b->faces.n), dim3(SBX, SBY), 0, s>>>(b->d, b->faces.d, b->faces.n, 1);
    k_b_maccormack_faces<<<pgrid(b, b->faces.n), dim3(SBX, SBY), 0, s>>>(b->d, b->faces.d, b->faces.n);
  } else {
    k_b_advect_faces<<<pgrid(b, b->faces.n), dim3(SBX, SBY), 0, s>>>(b->d, b->faces.d, b->faces.n, 0);
  }
  k_b_swap<<<(b->ns + 63) / 64, 64, 0, s>>>(b->d, b->ns, 1);
  k_b_vorticity<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n);
  k_b_forces<<<pgrid(b, b->faces.n), dim3(SBX, SBY), 0, s>>>(b->d, b->faces.d, b->faces.n);
  k_b_divergence<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n);
  k_b_dc_update<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n, b->dpart, 1);
  k_b_dc_norm<<<b->ns, 256, 0, s>>>(b->d, b->dpart, a.tol, 1, b->d_res);
  k_b_zero_p<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, e, "batch step (pre)");
  return FH_OK;
}

// The V-cycle's tile and coarse phases (no update of p).
void enq_vcycle(fh_batch* b) {
  hipStream_t s = b->s;
  for (int l = 0; l < b->max_lc; l++)
    launch_b_smooth_restrict(b->tile[l], pgrid(b, b->tiles[l].n), s, b->d, b->tiles[l].d,
                             b->tiles[l].n, l);
  k_b_coarse<<<b->ns, NC, b->coarse_lds, s>>>(b->d);
  for (int l = b->max_lc - 1; l >= 0; l--)
    launch_b_prolong_smooth(b->tile[l], pgrid(b, b->tiles[l].n), s, b->d, b->tiles[l].d,
                            b->tiles[l].n, l);
}

int enq_cycle(fh_batch* b, const StepArgs& a) {
  enq_vcycle(b);
  k_b_dc_update<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, b->s>>>(b->d, b->cells.d, b->cells.n, b->dpart,
                                                        0);
  k_b_dc_norm<<<b->ns, 256, 0, b->s>>>(b->d, b->dpart, a.tol, 0, b->d_res);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, e, "batch step (cycle)");
  return FH_OK;
}

int enq_post(fh_batch* b, const StepArgs& a) {
  hipStream_t s = b->s;
  k_b_subtract<<<pgrid(b, b->faces.n), dim3(SBX, SBY), 0, s>>>(b->d, b->faces.d, b->faces.n);
  k_b_outflow<<<pgrid(b, b->outflow.n), dim3(SBX, SBY), 0, s>>>(b->d, b->outflow.d, b->outflow.n);
  if (a.scheme == FH_MACCORMACK) {
    k_b_advect_dye<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n, 1);
    k_b_maccormack_dye<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n);
  } else {
    k_b_advect_dye<<<pgrid(b, b->cells.n), dim3(SBX, SBY), 0, s>>>(b->d, b->cells.d, b->cells.n, 0);
  }
  k_b_swap<<<(b->ns + 63) / 64, 64, 0, s>>>(b->d, b->ns, 2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, e, "batch step (post)");
  return FH_OK;
}

bool all_done(const fh_batch* b) {
  for (int s = 0; s < b->ns; s++)
    if (!((volatile const BRes*)b->res)[s].done) return false;
  return true;
}

int capture(fh_batch* b, hipGraphExec_t* out, int (*body)(fh_batch*, const StepArgs&, int),
            const StepArgs& a, int k) {
  hipGraph_t g = nullptr;
  BCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  int st = body(b, a, k);
  hipError_t e = hipStreamEndCapture(b->s, &g);
  if (st != FH_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return bfail(b, e, "hipStreamEndCapture");
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(b, e, "hipGraphInstantiate");
  return FH_OK;
}

int body_main(fh_batch* b, const StepArgs& a, int k) {
  int st = enq_pre(b, a);
  for (int c = 0; c < k && st == FH_OK; c++) st = enq_cycle(b, a);
  return st;
}
int body_cycle(fh_batch* b, const StepArgs& a, int) { return enq_cycle(b, a); }
int body_post(fh_batch* b, const StepArgs& a, int) { return enq_post(b, a); }

}  // namespace

extern "C" {

const char* fh_batch_last_error(void) { return b_err.c_str(); }

int fh_batch_num_tiles(void) { return kNumTileShapes; }

const char* fh_batch_tile_name(int t) { return (t >= 0 && t < kNumTileShapes) ? kTileShapes[t].name : ""; }

fh_batch* fh_batch_create(int nscenes, const int* nx, const int* ny, const double* h) {
  if (nscenes < 1 || nscenes > 255 || !nx || !ny || !h) {
    b_err = "fh_batch_create: need 1..255 scenes and nx, ny, h arrays";
    return nullptr;
  }
  for (int s = 0; s < nscenes; s++)
    if (nx[s] < 2 || ny[s] < 2 || nx[s] > 64 * 4095 || ny[s] > 4 * 4095 || !(h[s] > 0)) {
      b_err = "fh_batch_create: scene sizes must be 2..262080 x 2..16380 with h > 0";
      return nullptr;
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    b_err = "fh_batch_create: no HIP device";
    return nullptr;
  }
  fh_batch* b = new fh_batch();
  b->ns = nscenes;
  auto bail = [&]() {
    fh_batch_destroy(b);
    return (fh_batch*)nullptr;
  };
  if (hipStreamCreate(&b->s) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&b->evs, hipEventDisableTiming) != hipSuccess) {
    b_err = "fh_batch_create: stream/event creation failed";
    return bail();
  }
  b->h.resize(nscenes);
  std::vector<uint32_t> faces, cells, outflow;
  for (int s = 0; s < nscenes; s++) {
    BScene& S = b->h[s];
    memset(&S, 0, sizeof S);
    S.nx = nx[s];
    S.ny = ny[s];
    S.h = h[s];
    size_t nu = (size_t)(S.nx + 1) * S.ny, nv = (size_t)S.nx * (S.ny + 1),
           nc = (size_t)S.nx * S.ny;
    if (balloc(b, &S.u, nu) || balloc(b, &S.v, nv) || balloc(b, &S.u2, nu) ||
        balloc(b, &S.v2, nv) || balloc(b, &S.uh, nu) || balloc(b, &S.vh, nv) ||
        balloc(b, &S.w, nc) || balloc(b, &S.dye, 3 * nc) || balloc(b, &S.dye2, 3 * nc) ||
        balloc(b, &S.dyeh, 3 * nc) || balloc(b, &S.lbl, nc) || balloc(b, &S.p, nc) ||
        balloc(b, &S.p2, nc) || balloc(b, &S.div, nc))
      return bail();
    S.nlev = level_count_b(S.nx, S.ny, 16);
    if (S.nlev > kMaxLev) {
      b_err = "fh_batch_create: scene too large";
      return bail();
    }
    int lnx = S.nx, lny = S.ny;
    for (int l = 0; l < S.nlev; l++) {
      BLevel& L = S.lv[l];
      L.nx = lnx;
      L.ny = lny;
      size_t n = (size_t)lnx * lny;
      if (l == 0)
        L.lbl = S.lbl;
      else if (balloc(b, &L.lbl, n))
        return bail();
      if (balloc(b, &L.x, n) || balloc(b, &L.xa, n) || balloc(b, &L.b, n) ||
          balloc(b, &L.info, n))
        return bail();
      lnx = (lnx + 1) / 2;
      lny = (lny + 1) / 2;
    }
    // first level from which the rest fits one workgroup's LDS
    S.lc = S.nlev - 1;
    for (int l = 0; l < S.nlev; l++) {
      CoarseLevels C{};
      C.n = S.nlev - l;
      if (C.n > kMaxCoarse) continue;
      for (int k = 0; k < C.n; k++) {
        C.nx[k] = S.lv[l + k].nx;
        C.ny[k] = S.lv[l + k].ny;
        C.info[k] = S.lv[l + k].info;
      }
      if (coarse_vcycle_lds_bytes(C) <= (size_t)kCoarseLdsB) {
        S.lc = l;
        S.coarse = C;
        break;
      }
    }
    if (S.coarse.n == 0) {
      b_err = "fh_batch_create: coarse levels do not fit the LDS";
      return bail();
    }
    b->coarse_lds = std::max(b->coarse_lds, coarse_vcycle_lds_bytes(S.coarse));
    b->max_lc = std::max(b->max_lc, S.lc);
    // task lists
    S.face_t0 = (int)faces.size();
    for (int by = 0; by < (S.ny + 1 + SBY - 1) / SBY; by++)
      for (int bx = 0; bx < (S.nx + 1 + SBX - 1) / SBX; bx++) faces.push_back(task_pack(s, bx, by));
    S.face_nt = (int)faces.size() - S.face_t0;
    S.cell_t0 = (int)cells.size();
    for (int by = 0; by < (S.ny + SBY - 1) / SBY; by++)
      for (int bx = 0; bx < (S.nx + SBX - 1) / SBX; bx++) cells.push_back(task_pack(s, bx, by));
    S.cell_nt = (int)cells.size() - S.cell_t0;
    for (int bx = 0; bx < (std::max(S.nx, S.ny) + 255) / 256; bx++)
      outflow.push_back(task_pack(s, bx, 0));
    S.done = 1;  // nothing to solve until a step
    b->cells_total += (long long)S.nx * S.ny;
  }
  if (upload_list(b, b->faces, faces) || upload_list(b, b->cells, cells) ||
      upload_list(b, b->outflow, outflow))
    return bail();
  for (int l = 0; l < b->max_lc; l++)
    if (build_tile_list(b, l)) return bail();
  if (balloc(b, &b->fpart, faces.size()) || balloc(b, &b->dpart, 4 * cells.size())) return bail();
  if (hipMalloc((void**)&b->d, nscenes * sizeof(BScene)) != hipSuccess ||
      hipMemcpy(b->d, b->h.data(), nscenes * sizeof(BScene), hipMemcpyHostToDevice) !=
          hipSuccess) {
    b_err = "fh_batch_create: scene table";
    return bail();
  }
  if (hipHostMalloc((void**)&b->res, nscenes * sizeof(BRes),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&b->d_res, b->res, 0) != hipSuccess) {
    b_err = "fh_batch_create: result slots";
    return bail();
  }
  memset(b->res, 0, nscenes * sizeof(BRes));
  return b;
}

void fh_batch_destroy(fh_batch* b) {
  if (!b) return;
  if (b->s) (void)hipStreamSynchronize(b->s);
  drop_graphs(b);
  for (void* p : b->allocs) (void)hipFree(p);
  for (int s = 0; s < (int)b->h.size(); s++) {
    if (b->h[s].M) (void)hipFree(b->h[s].M);
    if (b->h[s].cells) (void)hipFree(b->h[s].cells);
  }
  for (DevList* L : {&b->faces, &b->cells, &b->outflow})
    if (L->d) (void)hipFree(L->d);
  for (auto& L : b->tiles)
    if (L.d) (void)hipFree(L.d);
  if (b->d) (void)hipFree(b->d);
  if (b->res) (void)hipHostFree(b->res);
  for (hipEvent_t e : {b->ev0, b->ev1, b->evs})
    if (e) (void)hipEventDestroy(e);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

// Fields of one scene (host FP64, reference layout); labels rebuild the
// scene's multigrid hierarchy and coarsest-level operator.
int fh_batch_upload(fh_batch* b, int scene, const double* u, const double* v, const double* p,
                    const uint8_t* label, const double* dye) {
  BGUARD("HipBatch.upload()");
  if (scene < 0 || scene >= b->ns) {
    b_err = "fh_batch_upload: scene out of range";
    return FH_ERR_BAD_ARG;
  }
  BCHK(hipStreamSynchronize(b->s));
  BScene cur;
  BCHK(hipMemcpy(&cur, b->d + scene, sizeof cur, hipMemcpyDeviceToHost));
  const size_t nu = (size_t)(cur.nx + 1) * cur.ny, nv = (size_t)cur.nx * (cur.ny + 1),
               nc = (size_t)cur.nx * cur.ny;
  std::vector<float> st;
  auto up = [&](float* dst, const double* src, size_t n) -> int {
    st.resize(n);
    for (size_t k = 0; k < n; k++) st[k] = (float)src[k];
    BCHK(hipMemcpy(dst, st.data(), n * sizeof(float), hipMemcpyHostToDevice));
    return FH_OK;
  };
  int r;
  if (u && (r = up(cur.u, u, nu))) return r;
  if (v && (r = up(cur.v, v, nv))) return r;
  if (dye && (r = up(cur.dye, dye, 3 * nc))) return r;
  if (p) BCHK(hipMemcpy(cur.p, p, nc * sizeof(double), hipMemcpyHostToDevice));
  if (label) {
    BCHK(hipMemcpy(cur.lbl, label, nc, hipMemcpyHostToDevice));
    for (int l = 1; l < cur.nlev; l++)
      launch_coarsen_labels(b->s, cur.lv[l - 1].lbl, cur.lv[l - 1].nx, cur.lv[l - 1].ny,
                            cur.lv[l].lbl, cur.lv[l].nx, cur.lv[l].ny);
    for (int l = 0; l < cur.nlev; l++) {
      launch_cell_info(b->s, cur.lv[l].lbl, cur.lv[l].nx, cur.lv[l].ny, cur.lv[l].info);
      BCHK(hipMemsetAsync(cur.lv[l].x, 0, (size_t)cur.lv[l].nx * cur.lv[l].ny * 4, b->s));
      BCHK(hipMemsetAsync(cur.lv[l].xa, 0, (size_t)cur.lv[l].nx * cur.lv[l].ny * 4, b->s));
    }
    BCHK(hipGetLastError());
    const BLevel& C = cur.lv[cur.nlev - 1];
    std::vector<uint8_t> lab((size_t)C.nx * C.ny);
    BCHK(hipMemcpyAsync(lab.data(), C.lbl, lab.size(), hipMemcpyDeviceToHost, b->s));
    BCHK(hipStreamSynchronize(b->s));
    std::vector<double> M;
    std::vector<int> cl;
    build_pseudo_inverse(C.nx, C.ny, lab, M, cl);
    if (cl.size() > 1024) {
      b_err = "coarsest level too large for the direct solve";
      return FH_ERR_BAD_ARG;
    }
    BScene& H = b->h[scene];
    if (H.M) BCHK(hipFree(H.M));
    if (H.cells) BCHK(hipFree(H.cells));
    H.M = nullptr;
    H.cells = nullptr;
    H.ncoarse = (int)cl.size();
    if (!cl.empty()) {
      BCHK(hipMalloc((void**)&H.M, M.size() * sizeof(double)));
      BCHK(hipMalloc((void**)&H.cells, cl.size() * sizeof(int)));
      BCHK(hipMemcpy(H.M, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
      BCHK(hipMemcpy(H.cells, cl.data(), cl.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    cur.M = H.M;
    cur.cells = H.cells;
    cur.ncoarse = H.ncoarse;
  }
  BCHK(hipMemcpy(b->d + scene, &cur, sizeof cur, hipMemcpyHostToDevice));
  return FH_OK;
}

int fh_batch_set_forces(fh_batch* b, int scene, double buoyancy, double confinement,
                        const double* decay) {
  BGUARD("HipBatch.setForces()");
  if (scene < 0 || scene >= b->ns) {
    b_err = "fh_batch_set_forces: scene out of range";
    return FH_ERR_BAD_ARG;
  }
  BCHK(hipStreamSynchronize(b->s));
  BScene cur;
  BCHK(hipMemcpy(&cur, b->d + scene, sizeof cur, hipMemcpyDeviceToHost));
  cur.buoyancy = buoyancy;
  cur.confinement = confinement;
  for (int c = 0; c < 3; c++) cur.decay[c] = decay ? decay[c] : 0.0;
  BCHK(hipMemcpy(b->d + scene, &cur, sizeof cur, hipMemcpyHostToDevice));
  return FH_OK;
}

// One step of every scene (simulation.ts:467-587 per scene, dt from each
// scene's CFL).  graph_cycles > 0: the step up to graph_cycles V-cycles is
// one captured hipGraph (re-captured when an argument changes); 0: the same
// kernels are launched directly.  Scenes not converged after those cycles
// get further cycles, up to max_cycles, one cycle (graph) per host check.
int fh_batch_step(fh_batch* b, int scheme, double cfl, double dt_max, double rho, double tol,
                  int max_cycles, int graph_cycles) {
  BGUARD("HipBatch.step()");
  if (!kBatchKernelsEnabled) {
    b_err = kBatchOff;
    return FH_ERR_HIP;
  }
  if ((scheme != FH_SEMI_LAGRANGIAN && scheme != FH_MACCORMACK) || !(cfl > 0) || !(dt_max > 0) ||
      !(rho > 0) || max_cycles < 0 || graph_cycles < 0) {
    b_err = "fh_batch_step: bad arguments";
    return FH_ERR_BAD_ARG;
  }
  StepArgs a{scheme, cfl, dt_max, rho, tol};
  int k0 = std::min(graph_cycles > 0 ? graph_cycles : max_cycles, max_cycles);
  BCHK(hipEventRecord(b->ev0, b->s));
  if (graph_cycles > 0) {
    double key[8] = {(double)scheme, cfl, dt_max, rho, tol, (double)k0, 0, 0};
    if (!b->graph_ok || memcmp(key, b->key, sizeof key) != 0) {
      drop_graphs(b);
      int st = capture(b, &b->g_main, body_main, a, k0);
      if (st == FH_OK) st = capture(b, &b->g_cycle, body_cycle, a, 1);
      if (st == FH_OK) st = capture(b, &b->g_post, body_post, a, 0);
      if (st != FH_OK) {
        drop_graphs(b);
        return st;
      }
      memcpy(b->key, key, sizeof key);
      b->graph_ok = true;
    }
    BCHK(hipGraphLaunch(b->g_main, b->s));
  } else {
    int st = body_main(b, a, k0);
    if (st != FH_OK) return st;
  }
  int done_cycles = k0;
  for (;;) {
    BCHK(hipEventRecord(b->evs, b->s));
    BCHK(hipEventSynchronize(b->evs));
    if (all_done(b) || done_cycles >= max_cycles) break;
    if (graph_cycles > 0) {
      BCHK(hipGraphLaunch(b->g_cycle, b->s));
    } else {
      int st = enq_cycle(b, a);
      if (st != FH_OK) return st;
    }
    done_cycles++;
  }
  if (graph_cycles > 0) {
    BCHK(hipGraphLaunch(b->g_post, b->s));
  } else {
    int st = enq_post(b, a);
    if (st != FH_OK) return st;
  }
  BCHK(hipEventRecord(b->ev1, b->s));
  BCHK(hipEventSynchronize(b->ev1));
  BCHK(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  return FH_OK;
}

// Per scene of the last step: dt, V-cycles and final relative residual.
int fh_batch_results(fh_batch* b, double* dt, int* iters, double* resid) {
  BGUARD("HipBatch.results()");
  BCHK(hipStreamSynchronize(b->s));
  for (int s = 0; s < b->ns; s++) {
    if (dt) dt[s] = b->res[s].dt;
    if (iters) iters[s] = b->res[s].iters;
    if (resid) resid[s] = b->res[s].rel;
  }
  return FH_OK;
}

int fh_batch_last_step_ms(fh_batch* b, float* ms) {
  if (!b || !ms) return FH_ERR_BAD_ARG;
  *ms = b->last_ms;
  return FH_OK;
}

long long fh_batch_cells(fh_batch* b) { return b ? b->cells_total : 0; }

// Current fields of one scene (any pointer may be null).
int fh_batch_read(fh_batch* b, int scene, double* u, double* v, double* p, double* dye) {
  BGUARD("HipBatch.read()");
  if (scene < 0 || scene >= b->ns) {
    b_err = "fh_batch_read: scene out of range";
    return FH_ERR_BAD_ARG;
  }
  BCHK(hipStreamSynchronize(b->s));
  BScene cur;
  BCHK(hipMemcpy(&cur, b->d + scene, sizeof cur, hipMemcpyDeviceToHost));
  const size_t nu = (size_t)(cur.nx + 1) * cur.ny, nv = (size_t)cur.nx * (cur.ny + 1),
               nc = (size_t)cur.nx * cur.ny;
  std::vector<float> st;
  auto down = [&](double* dst, const float* src, size_t n) -> int {
    st.resize(n);
    BCHK(hipMemcpy(st.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; k++) dst[k] = st[k];
    return FH_OK;
  };
  int r;
  if (u && (r = down(u, cur.u, nu))) return r;
  if (v && (r = down(v, cur.v, nv))) return r;
  if (dye && (r = down(dye, cur.dye, 3 * nc))) return r;
  if (p) BCHK(hipMemcpy(p, cur.p, nc * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

// Tile shape of every tiled level (an index of fh_batch_tile_name).
int fh_batch_set_tile(fh_batch* b, int level, int shape) {
  BGUARD("HipBatch.setTile()");
  if (shape < 0 || shape >= kNumTileShapes || level < -1 || level >= kMaxLev) {
    b_err = "fh_batch_set_tile: bad level or shape";
    return FH_ERR_BAD_ARG;
  }
  BCHK(hipStreamSynchronize(b->s));
  for (int l = 0; l < b->max_lc; l++)
    if (level < 0 || level == l) {
      b->tile[l] = shape;
      int st = build_tile_list(b, l);
      if (st != FH_OK) return st;
    }
  drop_graphs(b);
  return FH_OK;
}

int fh_batch_get_tiles(fh_batch* b, int* shapes, int* nlevels) {
  if (!b) return FH_ERR_BAD_ARG;
  if (nlevels) *nlevels = b->max_lc;
  if (shapes)
    for (int l = 0; l < b->max_lc; l++) shapes[l] = b->tile[l];
  return FH_OK;
}

// Autotuner: for each tiled level, times `reps` pre-smooth+restrict and
// prolong+post-smooth passes of every candidate tile shape over the whole
// batch and keeps the fastest.  Touches only multigrid scratch (the level
// unknowns and coarse right-hand sides), never p or the velocity.  ms_out
// (optional, max_lc * fh_batch_num_tiles() floats) gets every timing.
int fh_batch_tune(fh_batch* b, int reps, float* ms_out) {
  BGUARD("HipBatch.tune()");
  if (!kBatchKernelsEnabled) {
    b_err = kBatchOff;
    return FH_ERR_HIP;
  }
  if (reps < 1) reps = 1;
  BCHK(hipStreamSynchronize(b->s));
  k_b_set_done<<<(b->ns + 63) / 64, 64, 0, b->s>>>(b->d, b->ns, 0);
  BCHK(hipGetLastError());
  for (int l = 0; l < b->max_lc; l++) {
    float best = 1e30f;
    int best_t = 0;
    for (int t = 0; t < kNumTileShapes; t++) {
      b->tile[l] = t;
      int st = build_tile_list(b, l);
      if (st != FH_OK) return st;
      const DevList& L = b->tiles[l];
      for (int w = 0; w < 2; w++) {  // warm-up
        launch_b_smooth_restrict(t, pgrid(b, L.n), b->s, b->d, L.d, L.n, l);
        launch_b_prolong_smooth(t, pgrid(b, L.n), b->s, b->d, L.d, L.n, l);
      }
      BCHK(hipEventRecord(b->ev0, b->s));
      for (int r = 0; r < reps; r++) {
        launch_b_smooth_restrict(t, pgrid(b, L.n), b->s, b->d, L.d, L.n, l);
        launch_b_prolong_smooth(t, pgrid(b, L.n), b->s, b->d, L.d, L.n, l);
      }
      BCHK(hipEventRecord(b->ev1, b->s));
      BCHK(hipEventSynchronize(b->ev1));
      BCHK(hipGetLastError());
      float ms = 0.f;
      BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      ms /= reps;
      if (ms_out) ms_out[l * kNumTileShapes + t] = ms;
      if (ms < best) {
        best = ms;
        best_t = t;
      }
    }
    b->tile[l] = best_t;
    int st = build_tile_list(b, l);
    if (st != FH_OK) return st;
  }
  k_b_set_done<<<(b->ns + 63) / 64, 64, 0, b->s>>>(b->d, b->ns, 1);
  BCHK(hipGetLastError());
  BCHK(hipStreamSynchronize(b->s));
  drop_graphs(b);
  return FH_OK;
}

}  // extern "C"
