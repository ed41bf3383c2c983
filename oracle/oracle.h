/*
 * FP64 CPU oracle for flowlab's projection-method step.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the timed CPU baseline -- never as part of the product path.
 *
 * Every function restates the reference TypeScript (AleksandrRogachev94/flowlab,
 * src/core/ TypeScript) operation for operation in double precision, with the same
 * neighbour-summation orders. File:line citations are relative to the
 * reference root. Functions marked BUILD-DEFINED have no reference
 * counterpart (vorticity confinement, the defect-correction multigrid with a
 * direct coarsest solve); their parity is pinned by invariants and closed
 * forms in tests/, not by the reference.
 *
 * Array layouts are the reference's (src/core/grid.ts:1-36, 84-141):
 *   p, label, div : nx*ny       row-major, stride nx
 *   u             : (nx+1)*ny   stride nx+1
 *   v             : nx*(ny+1)   stride nx
 *   dye           : 3 planes of dnx*dny on the dye grid
 */
#ifndef FLOWLAB_ORACLE_H
#define FLOWLAB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FO_FLUID = 0, FO_AIR = 1, FO_SOLID = 2 };

/* grid.ts:197-315 */
double fo_sample_u(int nx, int ny, double h, const double* u, double x, double y);
double fo_sample_v(int nx, int ny, double h, const double* v, double x, double y);
double fo_sample_p(int nx, int ny, double h, const double* q, double x, double y);

/* advect.ts:126-136, 166-211, 241-270, 280-312 */
void fo_advect_velocity(int nx, int ny, double h, const double* uIn, const double* vIn,
                        double* uOut, double* vOut, const uint8_t* label, double dt);
void fo_advect_velocity_maccormack(int nx, int ny, double h, const double* uIn,
                                   const double* vIn, double* uHat, double* vHat,
                                   double* uOut, double* vOut, const uint8_t* label, double dt);
void fo_advect_scalar(int nx, int ny, double h, int dnx, int dny, double dh, const double* u,
                      const double* v, const double* qIn, double* qOut, const uint8_t* label,
                      double dt);
void fo_advect_scalar_maccormack(int nx, int ny, double h, int dnx, int dny, double dh,
                                 const double* u, const double* v, const double* qIn,
                                 double* qHat, double* qOut, const uint8_t* label, double dt);

/* divergence.ts:17-25 */
void fo_divergence(int nx, int ny, double h, const double* u, const double* v, double* out);

/* pressure.ts:40-212 */
void fo_gs_sweep(int nx, int ny, double* p, const double* div, const uint8_t* label,
                 double scale, double omega, int reverseX, int reverseY);
double fo_rms_remaining_divergence(int nx, int ny, const double* p, const double* div,
                                   const uint8_t* label, double scale);
double fo_fluid_div_rms(int n, const double* div, const uint8_t* label);
int fo_solve_pressure(int nx, int ny, double* p, const double* div, const uint8_t* label,
                      double scale, int iterations, double omega, double tol);

/* pressureRedBlack.ts:32-113 */
void fo_rb_sweep(int nx, int ny, double* p, const double* div, const uint8_t* label,
                 double scale, double omega, int color);
int fo_solve_pressure_rb(int nx, int ny, double* p, const double* div, const uint8_t* label,
                         double scale, int iterations, double omega, double tol);

/* pressureMultigrid.ts:65-320.  mode 0 = the reference CpuMultigridSolver
 * (coarsen until both sides <= 4, 8 red-black pairs on the coarsest level,
 * fixed cycle budget, tol checked after every cycle).  mode 1 = BUILD-DEFINED
 * defect correction: coarsen until both sides <= 16, FP64 direct solve on the
 * coarsest level, the V-cycle applied to the correction of an FP64 iterate. */
int fo_mg_level_count(int nx, int ny, int mode);
void fo_mg_level_size(int nx, int ny, int mode, int level, int* lnx, int* lny);
void fo_coarsen_labels(int fnx, int fny, const uint8_t* fine, int cnx, int cny, uint8_t* coarse);
void fo_mg_residual(int nx, int ny, const double* x, const double* b, const uint8_t* label,
                    double* r);
void fo_mg_restrict(int fnx, int fny, const double* r, int cnx, int cny, double* bC, double* eC);
void fo_mg_prolong_add(int fnx, int fny, double* x, const uint8_t* label, int cnx, int cny,
                       const double* e);
int fo_mg_solve(int nx, int ny, double* p, const double* div, const uint8_t* label, double scale,
                int mode, int cycles, double tol, double* resid_hist);
/* One V-cycle of mode 1 applied to (x,b) on the finest level from a zero
 * correction: x += V(b - A x).  Exposed so one GPU cycle can be diffed. */
void fo_mg_dc_cycle(int nx, int ny, double* p, const double* b, const uint8_t* label);
/* Dense FP64 solve on a coarsest-level grid: e = A^+ b (min-norm on pure
 * Neumann components).  BUILD-DEFINED. */
void fo_coarse_direct_solve(int nx, int ny, const uint8_t* label, const double* b, double* e);

/* subtractGradient.ts:40-61, boundaries.ts:45-143 */
void fo_subtract_gradient(int nx, int ny, const double* p, double* u, double* v,
                          const uint8_t* label, double gradScale);
void fo_apply_outflow(int nx, int ny, double* u, double* v, const uint8_t* label);
void fo_commit_labels(int nx, int ny, double* p, double* u, double* v, const uint8_t* label,
                      double* dye, int dye_plane);

/* buoyancy.ts:70-88 */
void fo_apply_buoyancy(int nx, int ny, double h, int dnx, int dny, double dh, double* v,
                       const double* heat, const uint8_t* label, double w, double dt);
/* dye.ts:70-87 */
void fo_apply_dye_patch(int dnx, int dny, double* dye, int dye_plane, int i0, int j0, int pnx,
                        int pny, const float* data);
/* stir.ts:49-83 */
void fo_apply_stir(int nx, int ny, double h, double* u, double* v, const uint8_t* label,
                   double sx, double sy, double sr, double sdx, double sdy);
/* vorticity.ts:24-42 */
void fo_compute_vorticity(int nx, int ny, double h, const double* u, const double* v,
                          double* out);
/* BUILD-DEFINED vorticity confinement (Fedkiw, Stam & Jensen 2001). */
void fo_vorticity_confinement(int nx, int ny, double h, double* u, double* v,
                              const uint8_t* label, const double* omega_c, double eps, double dt);

/* simulation.ts:218-221 */
double fo_optimal_omega(int nx, int ny);

/* One Simulation.step() of the CPU path, simulation.ts:467-587, with the
 * build's confinement force inserted between buoyancy and the divergence. */
typedef struct {
  int scheme;          /* 0 semi-Lagrangian, 1 MacCormack */
  int solver;          /* 0 lexicographic SOR, 1 red-black SOR, 2 MG (reference), 3 MG (defect correction) */
  int pressure_iters;  /* SOR sweep cap, or the MG cycle cap */
  int residual_every;  /* 0 = never */
  double cfl_target, dt_max, tol, omega, rho;
  double keep_decay[3]; /* dyeDecay per channel (rate, 1/s) */
  double buoyancy;
  double confinement;  /* epsilon */
  int stir_active;
  double stir[5];      /* x, y, r, dx, dy */
  /* dye patch */
  int patch_i0, patch_j0, patch_nx, patch_ny;
  const float* patch_data;
  /* outputs */
  double dt_used;
  int iters;
} fo_step_params;

typedef struct {
  int nx, ny, dnx, dny;
  double h, dh;
  double *u, *v, *p, *dye; /* dye: 3 planes */
  uint8_t* label;
  double *uNext, *vNext, *uHat, *vHat, *dyeNext, *qHat, *div, *work;
  long steps;
} fo_sim;

void fo_sim_step(fo_sim* s, fo_step_params* prm);
double fo_max_face_speed(int nx, int ny, const double* u, const double* v);

/* JS's `seed = (seed * 1103515245 + 12345) & 0x7fffffff` evaluated in
 * double precision then ToInt32 (pressure.test.ts:117-122). */
void fo_lcg_fill(int n, int seed, double* out);

#ifdef __cplusplus
}
#endif
#endif
