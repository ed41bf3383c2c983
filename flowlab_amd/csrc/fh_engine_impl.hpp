// The state of one libflowhip engine (one MAC grid on one HIP stream),
// shared by the engine's entry points (fh_engine.hip) and the batched step
// (fh_batch.hip), which launches over many engines' buffers at once.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <vector>

#include "../../include/flowhip.h"
#include "fh_launch.hpp"

namespace fh {

struct Level {
  int nx, ny;
  uint8_t* lbl;  // level 0 aliases the engine's label array
  float* x;      // the level's unknown after a cycle
  float* b;
  float* xa;     // pre-smoothed unknown (tiled kernels read halos of x
                 // while writing interiors, so they ping-pong x <-> xa)
  uint8_t* info; // per-cell Fluid flag + open-neighbour count (cell_info)
};

// LDS the one-workgroup coarse sub-cycle may use.  64 KB: with gfx950's
// 160 KB the 128x64 level of a 2048x1024 grid would fit too, but one
// workgroup then takes longer over it (measured +31 us) than the two tiled
// launches it replaces (17 us).
constexpr size_t kCoarseLds = 65536;
constexpr int kMaxTiledLevels = 16;

}  // namespace fh

struct fh_engine {
  int nx, ny, ds, dnx, dny;
  double h;
  hipStream_t s = nullptr;
  // fields
  float *u = nullptr, *v = nullptr, *u2 = nullptr, *v2 = nullptr, *uh = nullptr, *vh = nullptr;
  float *dye = nullptr, *dye2 = nullptr, *dyeh = nullptr, *w = nullptr;
  uint8_t* lbl = nullptr;
  // MacCormack: the forward passes' stencil origins of the limiter (one byte
  // per face-grid index / dye cell), read by the correction passes
  uint8_t *sten_f = nullptr, *sten_d = nullptr;
  double *p = nullptr, *p2 = nullptr, *div = nullptr;
  // flexible CG (allocated by the first such solve): FP64 residual, two FP32
  // directions (ping-pong by cycle parity), {alpha, beta, <d, A d>}
  double* rvec = nullptr;
  float* dvec[2] = {nullptr, nullptr};
  double* fcg = nullptr;
  // multigrid: levels for coarsest side 4 (reference); the defect-correction
  // hierarchy (coarsest side 16) is a prefix of it.
  std::vector<fh::Level> lv;
  int nlev_ref = 0, nlev_dc = 0;
  int lc_ref = 0, lc_dc = 0;  // first level run by the one-workgroup sub-cycle
  int tile[fh::kMaxTiledLevels] = {};  // tile shape per tiled level (kTileShapes)
  int coarse_group = 0;            // coarse sub-cycle workgroup (kCoarseGroups)
  bool advect_lds = true;          // LDS-staged advection kernels
  bool dye_one_pass = true;        // MacCormack dye (scale 1) in one pass
  bool faces_one_pass = true;      // MacCormack velocity in one pass
  int* one_pass_fail = nullptr;    // the one-pass kernel's reach mark (a launch's epoch)
  int one_pass_epoch = 0;          // its launch count
  // a defect-correction cycle's finest prolongation + post-smoothing and its
  // update pass run as one kernel (launch_prolong_smooth_dc); its residual
  // goes to b2, then b2 and the finest level's b swap
  bool fuse_update = true;
  float* b2 = nullptr;
  // the last enqueued defect-correction cycle ran fused (a cycle the device
  // skipped has its b swap undone with its p swap)
  bool last_dc_fused = false;
  // every level from cinfo_from on with its info padded as the coarse
  // sub-cycle stages it (launch_pad_info); level l at byte cinfo_off[l]
  uint8_t* cinfo = nullptr;
  int cinfo_from = 0;
  std::vector<int> cinfo_off;
  double* M = nullptr;  // coarsest-level pseudo-inverse, n*n
  int* cells = nullptr;
  int ncoarse = 0;
  bool labels_dirty = true;
  // reductions (room for partials_n 4-slot partials and their fold)
  double* partials = nullptr;
  size_t partials_n = 0;
  unsigned* counter = nullptr;
  double* sums = nullptr;
  float* fpartials = nullptr;
  float* fmax = nullptr;
  double* h_sums = nullptr;  // pinned
  float* h_fmax = nullptr;   // pinned, mapped; d_fmax is its device address
  float* d_fmax = nullptr;
  // Largest face speed of the resident velocity: every write of u/v bumps
  // vel_epoch; a step enqueues the reduction of its final velocity (and the
  // copy to h_fmax, then fmax_ev) and tags it with fmax_epoch.
  long vel_epoch = 0, fmax_epoch = -1;
  hipEvent_t fmax_ev = nullptr;
  // defect-correction loop: device stop flag, two host result slots (one per
  // cycle in flight) that the norm kernel writes, and the events that mark
  // those kernels' completion
  int* done = nullptr;
  double* h_res = nullptr;  // mapped, coherent; d_res is its device address: two cycle
                            // slots of 8 and the initial pass's at 16
  double* d_res = nullptr;
  hipEvent_t cyc_ev[3] = {};  // cycle slots 0, 1 and the initial pass (2)
  double rho_first = -1.0;     // r_0 / r_init of the last solve (its first cycle's contraction)
  // patch / forces
  float* patch = nullptr;
  int pi0 = 0, pj0 = 0, pnx = 0, pny = 0, has_patch = 0;
  int* patch_tiles = nullptr;  // tiles of the patch with non-zero coverage
  int patch_ntiles = 0;
  double buoyancy = 0, confinement = 0;
  fh::StirBox stir{};
  // state
  std::atomic<bool> busy{false};  // calls may come from several host threads
  bool lost = false;
  std::vector<double> hist;
  bool timing = false;
  hipEvent_t ev[FH_NUM_PHASES + 1] = {};
  bool ev_valid = false;
  // pinned host staging of the FP32 <-> FP64 field conversions (uploads and
  // readbacks), allocated at the first transfer with room for the largest
  // field: DMA from pinned memory instead of the runtime's pageable path
  float* stage = nullptr;
  size_t stage_n = 0;
  // per-kernel tracing
  int trace_kid = 0;
  std::vector<hipEvent_t> trace_ev;
  int trace_n = 0;
  // learned warm start: training-sample capture and inference
  fh_unet* capture = nullptr;
  fh_unet* warm = nullptr;
  double* psave = nullptr;
  float* wstats = nullptr;
  // warm start extrapolated in time (fh_set_warm_extrapolation): its order
  // (0: the previous pressure, the reference's start), the pressures earlier
  // fh_step solves ended with (ph[0] the newest) and how many are valid, the
  // dt of the step that produced each, and the dt of the step whose solve
  // left the resident p (0: p is not such a solution, history unusable)
  int extrap = 0;
  double* ph[2] = {nullptr, nullptr};
  int ph_n = 0;
  double ph_dt[2] = {0.0, 0.0};
  double p_dt = 0.0;
  // device time of the last pressure solve
  hipEvent_t solve_ev[2] = {};
  bool solve_ev_valid = false;
};

namespace fh {
// Makes an engine ready for the batched step (fh_batch.hip): its multigrid
// hierarchy current for its labels and the flexible-CG vectors allocated.
int engine_batch_prepare(fh_engine* e);

// The start of an fh_step solve of step size dt: with the extrapolation on
// (fh_set_warm_extrapolation) and the resident p the last step's solution,
// the start is p extrapolated over the history (order: the entries there
// are, at most the set order; 0 while the history is empty), formed into p2
// by the solve's first pass; off, the start is p itself.
struct StartPlan {
  bool on = false;
  int order = 0;
  Extrapolation xp{};
};
StartPlan plan_start(fh_engine* e, double dt);
// After the start is in p2: it becomes p, the old p (the last solution)
// becomes the newest history entry, and the buffer of the entry that falls
// out of the history becomes p2.
void rotate_history(fh_engine* e);
// The resident p is no longer the last fh_step solve's result.
void drop_pressure_history(fh_engine* e);
// The coarse sub-cycle's description of levels [lc, nlev) of an engine.
CoarseLevels engine_coarse_levels(const fh_engine* e, int lc, int nlev);
}  // namespace fh
