/*
 * libflowhip -- C ABI of the MI355X (gfx950) projection-method step.
 *
 * This is the drop-in boundary: plain pointers and sizes, no framework
 * types.  A host binding (the N-API addon under flowlab_amd/node/, the
 * Python ctypes layer in flowlab_amd/hip.py) wraps it into the reference's
 * three engine seams:
 *
 *   GpuStep        src/core/simulation.ts:35-84     -> fh_step, fh_upload,
 *                  fh_set_dye_patch, fh_set_buoyancy, fh_set_stir,
 *                  fh_set_labels, fh_read_dye (+ fh_read_velocity, which the
 *                  fused path's readback performs, src/gpu/stepGpu.ts:546-604)
 *   PressureSolver src/core/pressureSolver.ts:21-41 -> fh_solve_pressure
 *   Advector       src/core/advector.ts:28-63       -> fh_advect_velocity,
 *                  fh_advect_dye
 *
 * Host fields are the reference's Float64Array layouts (src/core/grid.ts:
 * 84-141): p/label/div nx*ny, u (nx+1)*ny, v nx*(ny+1), dye 3 planes of
 * (nx*s)*(ny*s) for dye scale s, all row-major with x fastest.  Labels are
 * Fluid 0, Air 1, Solid 2 (grid.ts:41-45).  On the device velocities and dye
 * are FP32; the pressure, its right-hand side and every norm are FP64.
 *
 * Every entry point returns an int status (FH_OK = 0) except fh_create
 * (NULL on failure, as device.ts:22-45 returns null) and the accessors noted.
 * fh_last_error() describes the last failure on the calling thread.
 * Calls on one engine must not overlap: an overlapping call returns
 * FH_ERR_BUSY (the reference throws "<Class>.<method>() is not re-entrant",
 * src/gpu/stepGpu.ts:402).  After a device fault every call returns
 * FH_ERR_DEVICE_LOST (src/gpu/stepGpu.ts:403).
 */
#ifndef FLOWHIP_H
#define FLOWHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fh_engine fh_engine;

enum {
  FH_OK = 0,
  FH_ERR_BAD_ARG = 1,
  FH_ERR_BUSY = 2,
  FH_ERR_DEVICE_LOST = 3,
  FH_ERR_NO_DEVICE = 4,
  FH_ERR_HIP = 5
};

/* src/core/advect.ts:21 */
enum { FH_SEMI_LAGRANGIAN = 0, FH_MACCORMACK = 1 };

/* Pressure solvers.  FH_SOLVER_RBSOR: red-black SOR, FP32, sweep budget and
 * tol checked every 12 sweeps (pressureRedBlack.ts:88-113).  FH_SOLVER_MG:
 * the reference V(2,2) multigrid, FP32, coarsened to <= 4 cells a side with
 * 8 red-black pairs on the coarsest level, `iterations` = cycle budget
 * (pressureMultigrid.ts:219-320).  FH_SOLVER_MG_DC: mixed-precision defect
 * correction -- FP64 pressure, residual and norms, FP32 V(2,2) on the
 * correction, coarsened to <= 16 cells a side with an FP64 direct solve;
 * `iterations` = cycle cap, `tol` = relative residual target.
 * FH_SOLVER_MG_FCG: a defect-correction cycle, then (if the tolerance is not
 * met) flexible conjugate gradients with one stored direction, preconditioned
 * by the same FP32 V-cycle -- the Krylov wrapper docs/WEBGPU.md:462-465 names
 * for boundary-limited multigrid contraction; `iterations` = cycle cap (one
 * V-cycle each). */
enum { FH_SOLVER_RBSOR = 1, FH_SOLVER_MG = 2, FH_SOLVER_MG_DC = 3, FH_SOLVER_MG_FCG = 4 };

/* Step phases reported by fh_get_timings (milliseconds of device time). */
enum {
  FH_PHASE_ADVECT = 0,
  FH_PHASE_FORCES = 1,
  FH_PHASE_DIVERGENCE = 2,
  FH_PHASE_SOLVE = 3,
  FH_PHASE_PROJECT = 4,
  FH_PHASE_DYE = 5,
  FH_PHASE_TOTAL = 6,
  FH_NUM_PHASES = 7
};

const char* fh_last_error(void);
int fh_device_count(void);

/* Lifecycle.  h is the cell size (simulation.ts:365 uses 1/ny); dye_scale
 * is the whole-number dye refinement (simulation.ts:358-372).  Replaces the
 * GpuStepper / GpuMultigridSolver / GpuAdvector constructors and their
 * destroy() (src/gpu/stepGpu.ts:672-689). */
fh_engine* fh_create(int nx, int ny, int dye_scale, double h);
void fh_destroy(fh_engine* e);

/* GpuStep.invalidate() + the re-upload it triggers (stepGpu.ts:419-433):
 * host truth for every field.  Any pointer may be NULL to leave that field. */
int fh_upload(fh_engine* e, const double* u, const double* v, const double* p,
              const uint8_t* label, const double* dye);
/* GpuStep.setLabels (simulation.ts:74) followed by the device-side
 * commitLabels of boundaries.ts:45-64. */
int fh_set_labels(fh_engine* e, const uint8_t* label);
/* GpuStep.setDyePatch (simulation.ts:53): data = 3 value planes + coverage,
 * each pnx*pny (dye.ts:30-36); data == NULL clears the patch. */
int fh_set_dye_patch(fh_engine* e, int i0, int j0, int pnx, int pny, const float* data);
/* GpuStep.setBuoyancy (simulation.ts:60). */
int fh_set_buoyancy(fh_engine* e, double w);
/* Vorticity-confinement epsilon (no reference counterpart; 0 = off). */
int fh_set_confinement(fh_engine* e, double eps);
/* GpuStep.setStir (simulation.ts:72): consumed by the next step. */
int fh_set_stir(fh_engine* e, int active, double x, double y, double r, double dx, double dy);

/* GpuStep.step (simulation.ts:36-46): advect -> forces -> divergence ->
 * solve -> subtract + outflow -> dye advect (x keep) -> dye patch, on the
 * device, one stream.  keep[c] = exp(-dyeDecay_c dt).  Returns the solver's
 * iteration count (sweeps or V-cycles) in *out_iters. */
int fh_step(fh_engine* e, int scheme, double dt, double scale, double grad_scale,
            const double keep[3], int solver, int iterations, double omega, double tol,
            int* out_iters);
int fh_read_velocity(fh_engine* e, double* u, double* v);
/* GpuStep.readDye (simulation.ts:83). */
int fh_read_dye(fh_engine* e, double* dye);
/* The device pressure (a diagnostic the reference lacks: f.p is not
 * updated by its fused step, simulation.ts:29-33). */
int fh_read_pressure(fh_engine* e, double* p);
int fh_read_divergence(fh_engine* e, double* div);
/* Simulation.maxFaceSpeed (simulation.ts:451-459) on the device fields. */
int fh_max_face_speed(fh_engine* e, double* out);

/* PressureSolver.solve (pressureSolver.ts:21-41): p in/out (warm start),
 * div, label as host arrays.  Returns sweeps/V-cycles in *out_iters. */
int fh_solve_pressure(fh_engine* e, int solver, double* p, const double* div,
                      const uint8_t* label, double scale, int iterations, double omega,
                      double tol, int* out_iters);

/* Advector.velocity / Advector.dye (advector.ts:28-63) on host arrays.
 * fh_advect_dye takes and returns all 3 channels (3 planes). */
int fh_advect_velocity(fh_engine* e, int scheme, const double* u_in, const double* v_in,
                       double* u_out, double* v_out, const uint8_t* label, double dt);
int fh_advect_dye(fh_engine* e, int scheme, const double* u, const double* v,
                  const double* q_in, double* q_out, const uint8_t* label, double dt);

/* The fused step's phases on the resident fields, one call each, so every
 * kernel can be diffed against the FP64 oracle from identical state. */
int fh_phase_advect_velocity(fh_engine* e, int scheme, double dt);
int fh_phase_forces(fh_engine* e, double dt);
int fh_phase_divergence(fh_engine* e);
int fh_phase_solve(fh_engine* e, int solver, double scale, int iterations, double omega,
                   double tol, int* out_iters);
int fh_phase_project(fh_engine* e, double grad_scale);
int fh_phase_dye(fh_engine* e, int scheme, double dt, const double keep[3]);

/* Relative residual (rmsRemainingDivergence / fluidDivRms, pressure.ts:
 * 102-161) before and after every cycle of the last FH_SOLVER_MG_DC or
 * FH_SOLVER_MG_FCG solve;
 * returns the number of entries written (<= max). */
int fh_solve_history(fh_engine* e, double* out, int max);

/* Per-phase device milliseconds of the last fh_step when timing is on. */
int fh_set_timing(fh_engine* e, int on);
int fh_get_timings(fh_engine* e, double* ms /* FH_NUM_PHASES */);
int fh_synchronize(fh_engine* e);

/* Live per-kernel timing on the engine's stream: with a kernel selected,
 * every launch of it is bracketed by HIP events (up to 8192 launches);
 * fh_trace_stats returns the launch count and the summed device ms. */
enum {
  FH_K_NONE = 0,
  FH_K_RB_SWEEP_FINE = 1,     /* red-black half-sweep on the finest level (on a grid that
                                 fits one workgroup's LDS, one launch of several
                                 iterations) */
  FH_K_ADVECT_FACES = 2,      /* semi-Lagrangian / MacCormack forward, faces */
  FH_K_MACCORMACK_FACES = 3,  /* MacCormack reverse+correct, faces */
  FH_K_ADVECT_DYE = 4,
  FH_K_MACCORMACK_DYE = 5,
  FH_K_DC_UPDATE = 6,         /* FP64 defect-correction update + norm */
  FH_K_RESTRICT_FINE = 7,     /* residual+restrict from the finest level */
  FH_K_SMOOTH_RESTRICT_FINE = 8, /* fused pre-smooth + restrict, finest level */
  FH_K_PROLONG_SMOOTH_FINE = 9,  /* fused prolong + post-smooth, finest level */
  FH_K_COARSE_VCYCLE = 10,       /* one-workgroup sub-cycle of the coarse levels */
  FH_K_FCG_DOTS = 11,            /* flexible CG: dot products (+ coefficients) */
  FH_K_FCG_UPDATE = 12,          /* flexible CG: direction, p, r and norm update */
  FH_K_DIVERGENCE = 13,          /* FP64 divergence of the FP32 faces */
  FH_K_FORCES = 14,              /* buoyancy (+ stir, vorticity confinement) */
  FH_K_PROJECT = 15              /* gradient subtraction + outflow */
};
int fh_trace_kernel(fh_engine* e, int kernel);
int fh_trace_stats(fh_engine* e, int* count, double* total_ms);
/* Device ms of an empty event pair on the engine's stream (mean of reps):
 * the part of each fh_trace_stats launch time that is the bracket itself. */
int fh_trace_overhead(fh_engine* e, int reps, double* ms);

/* Tile shapes of the LDS-tiled multigrid smoothers.  No counterpart in the
 * reference, whose WebGPU multigrid uses fixed 8x8 workgroups
 * (pressureGpu.ts:37-44).  fh_num_tile_shapes() candidates, named by
 * fh_tile_shape_name(); level -1 sets every tiled level.  Results do not
 * depend on the shape (temporal blocking computes the global sweep). */
int fh_num_tile_shapes(void);
const char* fh_tile_shape_name(int shape);
int fh_set_tile(fh_engine* e, int level, int shape);
int fh_get_tiles(fh_engine* e, int* shapes, int* nlevels);
/* Autotuner: times every shape on each tiled level of this engine's grid
 * (reps launch pairs each, best of three interleaved rounds) and keeps the
 * fastest, then the coarse sub-cycle's workgroup size likewise; needs labels
 * set.
 * ms_out (may be NULL): nlevels x fh_num_tile_shapes() ms. */
int fh_tune_tiles(fh_engine* e, int reps, float* ms_out);
/* The coarse sub-cycle's workgroup size (index 0: 1024, 1: 512, 2: 256
 * threads); fh_tune_tiles also picks it.  Results do not depend on it. */
int fh_set_coarse_group(fh_engine* e, int group);
int fh_coarse_group_threads(fh_engine* e);
/* Advection kernels (advect.ts:72-112, 177-210, 241-270, 291-311): on != 0
 * (the default) stages each workgroup's input windows through LDS; 0 gathers
 * straight from global memory.  Results are the same bits either way. */
int fh_set_advect_staging(fh_engine* e, int on);
/* MacCormack dye at dye scale 1 with staging on (advect.ts:241-270 then
 * 291-311): on != 0 (the default) runs the forward pass and the correction
 * as one kernel, which computes the forward-pass samples a tile's
 * correction reads in LDS instead of storing and re-reading them; 0 runs
 * the two passes.  Results are the same bits either way.  No counterpart
 * in the reference. */
int fh_set_dye_one_pass(fh_engine* e, int on);
/* The same for the MacCormack velocity (advect.ts:72-112 then 177-210). */
int fh_set_faces_one_pass(fh_engine* e, int on);

/* ---- allocation-free pressure solve (no reference counterpart) ----------
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * red-black SOR solve of fh_solve_pressure (FH_SOLVER_RBSOR, the same kernels
 * and the same bits) on caller-owned DEVICE arrays -- p (FP64, warm start in,
 * answer out), div (FP64), label (u8), nx*ny each -- on the caller's stream
 * (a hipStream_t; NULL = the null stream).  Its scratch (FP32 iterate and
 * right-hand side, norm partials) is one fixed device array of ~10 MB per
 * GPU, held until the process ends; nothing is allocated or freed by the
 * call.  Limits: nx*ny <= fh_noalloc_max_cells() (2^20) and each side
 * <= 65536, else FH_ERR_BAD_ARG (fh_noalloc_last_error says why).  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * host threads, other streams) race.  The call reads the norms on the host
 * (the zero-divergence guard, and every 12 sweeps when tol > 0), so it
 * synchronises its stream and returns with p written. */
const char* fh_noalloc_last_error(void);
long long fh_noalloc_max_cells(void);
int fh_solve_pressure_rbsor_noalloc(void* stream, double* p, const double* div,
                                    const uint8_t* label, int nx, int ny, double scale,
                                    int iterations, double omega, double tol, int* out_iters);

/* ---- learned warm start (no reference counterpart) ---------------------
 * A 4-level convolutional U-Net (3x3 convs, channels 16-32-64-64, BF16
 * weights and activations, FP32 accumulation on MFMA) predicts a coarse
 * correction to the start of the solve (the previous pressure, or its
 * extrapolation in time); the defect-correction solve then runs its
 * V-cycles to the tolerance as before, so plain multigrid decides
 * convergence.  The network runs on the fine grid gathered into F x F blocks
 * (fh_unet_coarse_grid), its output is interpolated bilinearly onto the fine
 * fluid cells, and the start moves along it by the energy-optimal step
 * (never increasing the error's A-norm).
 * Parameters: per layer weights [cout][3][3][cin] then bias [cout], FP32,
 * for the 15 layers e0a e0b e1a e1b e2a e2b b3a b3b d2a d2b d1a d1b d0a d0b
 * head.  Network input: [N][H][W][8] (channels: defect block sum / s_r,
 * solid fraction, Air fraction, fluid-mean pressure / s_p, then zeros);
 * output [N][H][W].  H and W must be multiples of 8. */
typedef struct fh_unet fh_unet;
fh_unet* fh_unet_create(unsigned long long seed);
void fh_unet_destroy(fh_unet* u);
const char* fh_unet_last_error(void);
int fh_unet_num_params(void);
/* The block size F of the coarse grid the network runs on for an nx x ny
 * grid (the smallest power of two leaving both sides at most 128) and the
 * coarse sides; 0 when F does not divide the sides into multiples of 8 (the
 * learned start then leaves the start unchanged). */
int fh_unet_coarse_grid(int nx, int ny, int* cw, int* ch);
int fh_unet_get_params(fh_unet* u, float* out, int n);
int fh_unet_set_params(fh_unet* u, const float* in, int n);
/* Output scale: the coarse correction is s_r * c0 * net. */
int fh_unet_set_scale(fh_unet* u, double c0);
int fh_unet_get_scale(fh_unet* u, double* c0);
int fh_unet_forward(fh_unet* u, int N, int H, int W, const float* in, float* out);
/* Parameter gradient of sum(dout .* net(in)). */
int fh_unet_backward(fh_unet* u, int N, int H, int W, const float* in, const float* dout,
                     float* dparams);
/* Training set on the device: every FP64 multigrid solve of an engine with
 * capture on (fh_set_capture) whose coarse grid is H x W (coarse sides)
 * appends (the network input of its start, the block-mean correction
 * p* - p over s_r) until `capacity` samples are held. */
int fh_unet_dataset_reserve(fh_unet* u, int capacity, int H, int W);
int fh_unet_dataset_size(fh_unet* u);
/* Captured sample `index`: the network input x (4 floats per coarse cell:
 * defect block sum / s_r, solid fraction, Air fraction, fluid-mean pressure
 * / s_p), the target t (block-mean correction / s_r) and {s_r, s_p}; any
 * output may be NULL. */
int fh_unet_dataset_get(fh_unet* u, int index, float* x, float* t, float* scales);
/* Sets c0 to the RMS of the captured targets (correction / s_r). */
int fh_unet_calibrate(fh_unet* u, int samples);
/* `steps` Adam steps (lr; beta 0.9/0.999) on batches of `batch`; per-step
 * masked MSE in loss_out (may be NULL). */
int fh_unet_train(fh_unet* u, int steps, int batch, double lr, unsigned long long seed,
                  double* loss_out);
int fh_set_capture(fh_engine* e, fh_unet* u);
/* Device milliseconds of the last pressure solve (fh_step's or
 * fh_solve_pressure's), warm start and every cycle included. */
int fh_last_solve_ms(fh_engine* e, double* ms);
/* NULL = start from the previous pressure (the reference's warm start). */
int fh_set_warm_start(fh_engine* e, fh_unet* u);
/* Warm start of fh_step's pressure solve extrapolated in time (no
 * counterpart in the reference, which starts from the previous pressure,
 * stepGpu.ts:19-24): order 1 (linear) or 2 (quadratic) Lagrange
 * extrapolation, at the step's own dt, of the pressures the last order + 1
 * fh_step solves ended with; 0 (the default) starts from the previous
 * pressure.  The history restarts after fh_upload, fh_set_labels, any solve
 * outside fh_step, or a failed step, and the first steps after a restart use
 * what history there is.  The learned warm start (fh_set_warm_start), when
 * set, corrects the extrapolated pressure.  The solve still runs to its
 * tolerance, so only the iteration count depends on the start. */
int fh_set_warm_extrapolation(fh_engine* e, int order);
/* On (the default): each defect-correction cycle runs its finest-level
 * prolongation + post-smoothing and the FP64 update after it as one kernel
 * (the same pressure and residual bits as the two passes; the residual
 * norm's partial sums are grouped per tile).  Off: the two passes, for
 * comparison.  No counterpart in the reference. */
int fh_set_fused_update(fh_engine* e, int on);

/* ---- Batched scenes ------------------------------------------------------
 * No counterpart in the reference, which steps one grid per Simulation
 * (simulation.ts:467-587).  A batch holds up to 255 independent grids of any
 * sizes of at least 3 x 3 (dyeScale 1, fixed at creation), each with the
 * buffers of a single-grid engine.  A step launches each phase once for all
 * scenes, over a work list of (scene, block) tasks: each scene takes dt from
 * its own CFL (simulation.ts:474-479) and solves its pressure with
 * FH_SOLVER_MG_FCG to its own convergence, bit for bit as fh_step on that
 * scene alone.  Per-scene fields are uploaded and read in the reference's
 * host layout, FP64, as fh_upload.  The defect-correction grid of a scene
 * must need at most 8192 reduction blocks (about 4 M cells); a larger one
 * fails the step with FH_ERR_BAD_ARG. */
typedef struct fh_batch fh_batch;
const char* fh_batch_last_error(void);
/* NULL on failure (fh_batch_last_error says why). */
fh_batch* fh_batch_create(int nscenes, const int* nx, const int* ny, const double* h);
void fh_batch_destroy(fh_batch* b);
/* Any pointer may be NULL (field unchanged).  Labels rebuild the scene's
 * multigrid hierarchy; call with labels at least once per scene. */
int fh_batch_upload(fh_batch* b, int scene, const double* u, const double* v, const double* p,
                    const uint8_t* label, const double* dye);
/* Buoyancy weight, confinement epsilon, dye decay rates (decay may be NULL). */
int fh_batch_set_forces(fh_batch* b, int scene, double buoyancy, double confinement,
                        const double* decay);
/* One step of every scene: dt = min(dt_max, cfl h / max face speed),
 * scale = rho h^2/dt, gradScale = dt/(rho h), dye keep exp(-decay dt).
 * Pressure: cycles until the scene's relative residual <= tol, at most
 * max_cycles. */
int fh_batch_step(fh_batch* b, int scheme, double cfl, double dt_max, double rho, double tol,
                  int max_cycles);
/* Per scene of the last step (arrays of nscenes, any may be NULL). */
int fh_batch_results(fh_batch* b, double* dt, int* iters, double* resid);
/* Every scene's solve starts from its last solutions extrapolated in time,
 * as fh_set_warm_extrapolation makes fh_step's (0: the previous pressure);
 * each scene keeps its own history, restarted by its fh_batch_upload. */
int fh_batch_set_warm_extrapolation(fh_batch* b, int order);
int fh_batch_read(fh_batch* b, int scene, double* u, double* v, double* p, double* dye);
int fh_batch_last_step_ms(fh_batch* b, float* ms);
long long fh_batch_cells(fh_batch* b);
/* Multigrid tile shapes of the batched launches, one per level index (level
 * l of every scene that tiles it): fh_batch_num_tiles() candidates by name;
 * level -1 sets every level; get_tiles reports the tiled levels of the
 * deepest scene.  Results do not depend on the shape. */
int fh_batch_num_tiles(void);
const char* fh_batch_tile_name(int shape);
int fh_batch_set_tile(fh_batch* b, int level, int shape);
int fh_batch_get_tiles(fh_batch* b, int* shapes, int* nlevels);
/* Autotuner of the batched multigrid launches: per level index, the shape
 * with the shortest launch pair; ms_out gets nlevels x num_tiles ms. */
int fh_batch_tune(fh_batch* b, int reps, float* ms_out);

#ifdef __cplusplus
}
#endif
#endif
