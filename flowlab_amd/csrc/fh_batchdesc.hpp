// The batched step's view of its scenes (fh_batch.hip): one descriptor per
// scene with the device addresses and sizes of that scene's engine, per-step
// scalars, and for each batched launch a work list of (scene, block) tasks.
//
// A batched launch is a 1-D grid over the blocks of every scene taking part:
// scene s owns the flat blocks [off_s, off_s + gx_s * gy_s), off_s a
// multiple of 8, and chunk[f >> 3] names the scene of flat block f.  Within
// its scene a block takes the index it would have in that scene's own
// single-grid launch (the same XCD-aware remap over the scene's gx x gy
// grid, the same partials slot), so every scene computes exactly the bits of
// its own launch.  Flat blocks past a scene's last (the padding to a
// multiple of 8) return at once.  With off_s a multiple of 8, flat & 7 ==
// local & 7: the remap still gives each XCD a contiguous run of the scene's
// blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace fh {

constexpr int kBatchMaxScenes = 255;
constexpr int kBatchMaxLevels = 16;
constexpr int kBatchSpeedBlocks = 128;  // blocks of the max-speed reduction per scene

struct LevelDesc {
  float *x, *xa, *b;
  const uint8_t* info;
  int nx, ny;
};

// A scene's buffers.  Pairs that the single-grid step swaps once per step
// (u, v, dye) are held as [2] and indexed by the step's parity q: the step
// reads [q] and leaves its result in [q ^ 1].  The pressure buffers of a
// step are in its BStep (they rotate with the extrapolated start's history).
struct SceneDesc {
  int nx, ny;
  float *u[2], *v[2], *uh, *vh;
  float *dye[2], *dyeh, *w;
  const uint8_t* lbl;
  uint8_t *sten_f, *sten_d;
  double *div, *rvec, *fcg;
  float* dvec[2];
  double *partials, *sums, *d_res;
  int* done;
  int red_blocks;  // partials of the defect / CG passes
  int lc;          // levels [0, lc) tiled; [lc, nlev) in the one-workgroup sub-cycle
  LevelDesc lv[kBatchMaxLevels];
  CoarseLevels cl;
  const double* M;
  const int* cells;
  int ncoarse;
};

// Per-scene scalars of one step, as the single-grid step derives them from
// the scene's dt (simulation.ts:474-479).
struct BStep {
  float c;             // dt / h
  float k0, k1, k2;    // dye keep factors
  float gain, eps;     // buoyancy * dt, confinement
  float hf, dtf, inv_hf;
  double scale, grad_scale, inv_h;
  // The step's pressure buffers, as fh_step uses them: the initial pass
  // reads p0 (the last solution) and, with the extrapolated start
  // (plan_start), the history in xp, and stores the start it forms to ps;
  // the first cycle's update reads the start (ps, or p0 when ps is null) and
  // writes pr, where the solve's result stays.
  const double* p0;
  double* ps;
  double* pr;
  Extrapolation xp;
  int xo;  // order of the extrapolated start (ps non-null)
};

struct BSpan {
  int off, gx, gy, pad;
};

struct BatchGrid {
  const BSpan* span;     // per scene (gx = gy = 0: not in this launch)
  const uint8_t* chunk;  // scene of each run of 8 flat blocks
  int blocks;            // flat blocks (a multiple of 8)
};

// Scene and in-scene block of this flat block; false for padding.
__device__ __forceinline__ bool batch_block(const BatchGrid& g, int* s, int* bx, int* by, int* gy,
                                            int* bid) {
  const int flat = (int)blockIdx.x;
  const int sc = g.chunk[flat >> 3];
  const BSpan sp = g.span[sc];
  const int local = flat - sp.off;
  const int n = sp.gx * sp.gy;
  if (local >= n) return false;
  xcd_map(local, n, sp.gx, bx, by);
  *s = sc;
  *gy = sp.gy;
  *bid = local;
  return true;
}

// ---- host: each scene's grid per batched launch ------------------------------
// The single-grid launchers' grids, so a scene's blocks are the same either way.
enum BatchAdvect { kBAdvectFaces, kBMacCormackFaces, kBAdvectDye, kBMacCormackDye };
dim3 batch_grid_advect(int which, int nx, int ny);
enum BatchProject { kBVorticity, kBForces, kBBuoyancy, kBProject, kNumBatchProject };
dim3 batch_grid_project(int which, int nx, int ny);
// The defect-correction / CG passes' grid.
dim3 batch_grid_reduce(int nx, int ny);

// ---- batched launchers (one launch over every scene of the work list) --------
// q: the step's parity (SceneDesc).  mac: the MacCormack forward pass (into
// uh/vh or dyeh, recording the limiter stencils) rather than the
// semi-Lagrangian pass into [q ^ 1].
void launch_batch_advect_faces(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                               const BStep* st, int q, int mac);
void launch_batch_maccormack_faces(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                                   const BStep* st, int q);
void launch_batch_advect_dye(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                             const BStep* st, int q, int mac);
void launch_batch_maccormack_dye(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                                 const BStep* st, int q);
// Forces on the stepped velocity [q ^ 1] with the heat of dye [q].
void launch_batch_vorticity(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                            const BStep* st, int q);
void launch_batch_forces(hipStream_t s, const SceneDesc* D, const BatchGrid& g, const BStep* st,
                         int q);
void launch_batch_buoyancy(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                           const BStep* st, int q);
void launch_batch_project(hipStream_t s, const SceneDesc* D, const BatchGrid& g, const BStep* st,
                          int q);
// max |u|, |v| of velocity [q ^ 1] per scene into host_out[s] (mapped
// pinned memory); g gives every scene kBatchSpeedBlocks x 1 blocks and
// partials holds kBatchSpeedBlocks floats per scene.
void launch_batch_max_speed(hipStream_t s, const SceneDesc* D, const BatchGrid& g, int q,
                            int nscenes, float* partials, float* host_out);
// The defect pass: init 1 (the pass before the first cycle: p0 unchanged,
// no correction), init 2 (the same, forming the divergence of velocity
// [q ^ 1] in the pass and storing it, and the extrapolated start when the
// step has one) or init 0, the first cycle's update (pr = start + e).
void launch_batch_dc_update(hipStream_t s, const SceneDesc* D, const BatchGrid& g,
                            const BStep* st, int q, int init);
// One block per scene: the norm of the last defect / CG pass into sums and
// host slot (d_res + 8 slot, the initial pass's at 16): mode 0 the initial
// pass, 1 a defect-correction cycle, 2 a CG cycle.
void launch_batch_finalize(hipStream_t s, const SceneDesc* D, int nscenes, const BStep* st,
                           int mode, int slot, double tol);
// Flexible CG of cycle c (cp = c & 1, first = c == 1) on each scene's pr.
void launch_batch_fcg(hipStream_t s, const SceneDesc* D, const BatchGrid& g, const BStep* st,
                      int nscenes, int cp, int first);
// Multigrid level l of every scene with l < lc, in tile shape `shape`.
void launch_batch_smooth_restrict(hipStream_t s, int shape, const SceneDesc* D,
                                  const BatchGrid& g, int l);
void launch_batch_prolong_smooth(hipStream_t s, int shape, const SceneDesc* D, const BatchGrid& g,
                                 int l);
// The one-workgroup sub-cycle of every scene from its level lc; lds bytes of
// dynamic LDS (the largest scene's).
void launch_batch_coarse_vcycle(hipStream_t s, const SceneDesc* D, int nscenes, size_t lds);

}  // namespace fh
