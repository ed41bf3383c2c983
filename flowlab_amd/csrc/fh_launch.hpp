// Host-side launchers of libflowhip's kernels (one translation unit each for
// advection, projection/forces and the pressure solvers).  Every launcher
// enqueues on `s` and returns immediately; none synchronises.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace fh {

struct StirBox {
  int active;
  float x, y, r, dx, dy;  // cell units (world / h)
  int ui0, ui1, uj0, uj1;  // inclusive u-face box (stir.ts:398-405)
  int vi0, vi1, vj0, vj1;  // inclusive v-face box (stir.ts:406-413)
};

// ---- advection (fh_advect.hip) -------------------------------------------
// lds: stage the inputs through LDS windows (faces always; dye when ds == 1;
// advect_faces only when the carrier is the source).  Same bits either way.
// sten: the MacCormack forward pass (advect_faces, advect_dye with lds)
// records the limiter's stencil origins there ((nx+1) x (ny+1) bytes for
// the faces, dnx x dny for the dye), and the correction pass given the same
// buffer reads them instead of tracing forward again; null: not recorded /
// traced again.  A correction may take sten only from a forward pass of the
// same velocity, labels and c that ran with lds.
// Semi-Lagrangian pass over every u and v face: out = src at the RK2
// backtrace along (uc, vc) with c = dt/h.  advect.ts:72-112.
void launch_advect_faces(hipStream_t s, const float* uc, const float* vc, const float* us,
                         const float* vs, float* uo, float* vo, const uint8_t* lbl, int nx, int ny,
                         float c, bool lds, uint8_t* sten = nullptr);
// MacCormack reverse pass fused with the correction and limiter.
// advect.ts:177-210.  heat (lds only; may be null): buoyancy alone
// (buoyancy.ts:70-88, dye scale 1) applied at each open interior v face's
// write, v += gain * heat at the face, as launch_forces's buoyancy pass would
// apply it after this pass.
void launch_maccormack_faces(hipStream_t s, const float* un, const float* vn, const float* uh,
                             const float* vh, float* uo, float* vo, const uint8_t* lbl, int nx,
                             int ny, float c, bool lds, const uint8_t* sten = nullptr,
                             const float* heat = nullptr, float gain = 0.f);
// Three dye channels on one backtrace, scaled by keep[c] at the write.
// advect.ts:241-270 (+ the decay of simulation.ts:569-576).
// speed (lds with ds == 1 only; may be null): each workgroup stores the
// largest |u|, |v| of its staged velocity window there, advect_dye_speed_blocks
// values in all; their maximum is the max face speed (launch_finalize_max).
void launch_advect_dye(hipStream_t s, const float* u, const float* v, const float* q, float* qo,
                       const uint8_t* lbl, int nx, int ny, int ds, float c, float k0, float k1,
                       float k2, bool lds, uint8_t* sten = nullptr, float* speed = nullptr);
int advect_dye_speed_blocks(int nx, int ny);
// advect.ts:291-311 for three channels, times keep[c].
void launch_maccormack_dye(hipStream_t s, const float* u, const float* v, const float* qn,
                           const float* qh, float* qo, const uint8_t* lbl, int nx, int ny, int ds,
                           float c, float k0, float k1, float k2, bool lds,
                           const uint8_t* sten = nullptr);
// MacCormack dye in one pass (dye scale 1): each tile computes the
// forward-pass values its correction samples itself, in LDS.  A tile whose
// velocity exceeds the staged reach writes nothing and marks the launch
// (fail = epoch; epoch must differ from every earlier launch's on fail);
// the matching *_rerun launches, enqueued right after, then run the two-pass
// kernels over every tile (the forward pass into qh and the stencils into
// sten), and otherwise return at once.  speed (may be null):
// max-speed partials, maccormack_dye_fused_blocks of them.
void launch_maccormack_dye_fused(hipStream_t s, const float* u, const float* v, const float* qn,
                                 float* qo, const uint8_t* lbl, int nx, int ny, float c, float k0,
                                 float k1, float k2, float* speed, int* fail, int epoch);
void launch_maccormack_dye_rerun(hipStream_t s, const float* u, const float* v, const float* qn,
                                 float* qh, float* qo, const uint8_t* lbl, int nx, int ny, float c,
                                 float k0, float k1, float k2, uint8_t* sten, const int* fail,
                                 int epoch);
int maccormack_dye_fused_blocks(int nx, int ny);
// The same for the velocity (advect.ts:72-112 then 177-210; heat, gain:
// buoyancy alone, as launch_maccormack_faces applies it), the forward pass
// of the two-pass kernels into uh, vh.
void launch_maccormack_faces_fused(hipStream_t s, const float* un, const float* vn, float* uo,
                                   float* vo, const uint8_t* lbl, int nx, int ny, float c,
                                   const float* heat, float gain, int* fail, int epoch);
void launch_maccormack_faces_rerun(hipStream_t s, const float* un, const float* vn, float* uh,
                                   float* vh, float* uo, float* vo, const uint8_t* lbl, int nx,
                                   int ny, float c, uint8_t* sten, const float* heat, float gain,
                                   const int* fail, int epoch);

// ---- forces, divergence, projection (fh_project.hip) ----------------------
void launch_vorticity(hipStream_t s, const float* u, const float* v, float* w, int nx, int ny,
                      float inv_h);
void launch_forces(hipStream_t s, float* u, float* v, const uint8_t* lbl, const float* heat,
                   const float* w, int nx, int ny, int ds, float h, float dt, float gain,
                   float eps, StirBox stir);
void launch_divergence(hipStream_t s, const float* u, const float* v, double* div, int nx, int ny,
                       double inv_h);
void launch_subtract_gradient(hipStream_t s, float* u, float* v, const double* p,
                              const uint8_t* lbl, int nx, int ny, double grad_scale);
void launch_outflow(hipStream_t s, float* u, float* v, const uint8_t* lbl, int nx, int ny);
// launch_subtract_gradient then launch_outflow, in one pass where nx, ny >= 3.
void launch_subtract_outflow(hipStream_t s, float* u, float* v, const double* p,
                             const uint8_t* lbl, int nx, int ny, double grad_scale);
// The patch's BX x BY tiles (64 x 4, row-major over the patch) with any
// non-zero coverage, from the host copy of the patch data (4 planes).
void dye_patch_tiles(const float* data, int pnx, int pny, std::vector<int>& tiles);
// Applies the patch on those tiles only (cells of zero coverage are left
// unchanged either way).
void launch_dye_patch(hipStream_t s, float* dye, const float* patch, int dnx, int dny, int i0,
                      int j0, int pnx, int pny, const int* tiles, int ntiles);
void launch_commit_labels(hipStream_t s, double* p, float* u, float* v, float* dye,
                          const uint8_t* lbl, int nx, int ny, int dye_plane);
// max |u|, |v| into *out (simulation.ts:451-459); u and v are read as float4
// and must be 16-byte aligned (the engine passes hipMalloc bases).
void launch_max_speed(hipStream_t s, const float* u, const float* v, int nx, int ny,
                      float* partials, unsigned* counter, float* out,
                      float* host_out = nullptr);
// The largest of n per-workgroup maxima into *out and host_out (optional).
void launch_finalize_max(hipStream_t s, const float* partials, int n, float* out,
                         float* host_out);

// ---- pressure (fh_mg.hip) --------------------------------------------------
// Warm start extrapolated in time from the previous solves' pressures:
// start = w0 p + w1 h0 + w2 h1 (h0 the newest earlier solution; terms beyond
// the order unread), formed as fma(w2, h1, fma(w1, h0, w0 p)).
struct Extrapolation {
  const double* h0;
  const double* h1;
  double w0, w1, w2;
};
// Reductions over more than kFoldAbove blocks fold their partials with
// kFoldBlocks workgroups before the one-block finalize; every partials
// buffer holds 4 * kFoldBlocks doubles beyond its producers' 4 per block.
constexpr int kFoldAbove = 8192, kFoldBlocks = 128;
// One colour half-sweep: x = (1-w) x + w (sum - scale*b)/count on fluid cells
// of colour (i+j)&1 == color.  pressureRedBlack.ts:32-72.  first != 0 treats
// x as zero on entry and writes both colours (the first sweep of a V-cycle
// on a zero correction).
void launch_rb_sweep(hipStream_t s, float* x, const float* b, const uint8_t* lbl, int nx, int ny,
                     int color, float omega, float scale, int first);
// `iters` iterations (red then black, as launch_rb_sweep with first = 0) in
// one workgroup holding the whole grid in LDS; the same bits.  Only where
// rb_sor_lds_fits (9 B per cell within the device's workgroup LDS).
bool rb_sor_lds_fits(int nx, int ny);
void launch_rb_sor_lds(hipStream_t s, float* x, const float* b, const uint8_t* lbl, int nx, int ny,
                       int iters, float omega, float scale);
// The current device's workgroup LDS limit in bytes (cached per device).
int device_max_lds();
// Coarse b = sum of child residuals, coarse x = 0.  pressureMultigrid.ts:117-178.
void launch_residual_restrict(hipStream_t s, const float* x, const float* b, const uint8_t* lbl,
                              int nx, int ny, float* bc, float* xc, int cnx, int cny);
// x += bilinear(e) on fluid cells.  pressureMultigrid.ts:192-207.
void launch_prolong(hipStream_t s, float* x, const uint8_t* lbl, int nx, int ny, const float* e,
                    int cnx, int cny);
void launch_coarsen_labels(hipStream_t s, const uint8_t* f, int fnx, int fny, uint8_t* c, int cnx,
                           int cny);
// x = M b on the coarsest level (FP64 dense apply of the pseudo-inverse).
void launch_coarse_solve(hipStream_t s, const float* b, float* x, const double* M,
                         const int* cells, int n);
// out = (float)(factor * div)
void launch_rhs32(hipStream_t s, const double* div, float* out, int n, double factor);
void launch_f64_to_f32(hipStream_t s, const double* a, float* out, int n);
void launch_f32_to_f64(hipStream_t s, const float* a, double* out, int n);
// Defect-correction update (FP64): pn = po + e;  r = b - A pn with
// b = -scale*div;  r32 = (float) r;  sums[0..3] = {sum r^2 (fluid, count>0),
// n of those, sum div^2 (fluid), n fluid}, sums[4] = relative residual,
// sums[5] = 1 once it is <= tol (tol > 0), which also sets *done; with *done
// already set nothing runs.  e may be null (no update), done may be null.
// info is the level's cell_info table (launch_cell_info).
void launch_dc_update(hipStream_t s, const double* po, double* pn, const float* e,
                      const double* div, const uint8_t* info, float* r32, int nx, int ny,
                      double scale, double* partials, double* sums, double* host_sums, int* done,
                      double tol);
// The two kernels of launch_dc_update, for callers that time them apart.
void launch_dc_correct(hipStream_t s, const double* po, double* pn, const float* e,
                       const double* div, const uint8_t* info, float* r32, int nx, int ny,
                       double scale, double* partials, const int* done, double* r64 = nullptr);
// launch_dc_correct's initial pass (no e, no pn, done unread) with the
// divergence (launch_divergence's values, stored to div) formed in the same
// pass from the faces u, v.
// pn non-null: the pass starts from po extrapolated at `order` over *xp
// (order 0: po itself) and stores that start to pn (which differs from po
// and xp's history).
void launch_div_defect(hipStream_t s, const float* u, const float* v, double inv_h, double* div,
                       const double* po, const uint8_t* info, float* r32, double* r64, int nx,
                       int ny, double scale, double* partials, double* pn = nullptr,
                       const Extrapolation* xp = nullptr, int order = 0);
void launch_dc_norm(hipStream_t s, const double* partials, int nx, int ny, double* sums,
                    double* host_sums, int* done, double scale, double tol);
// launch_dc_norm over n partials of a pass with another grid (the fused
// prolongation + update: one per tile); partials must hold 4 n + 4 kFoldBlocks.
void launch_dc_norm_n(hipStream_t s, const double* partials, int n, double* sums,
                      double* host_sums, int* done, double scale, double tol);
// The norm of the pass before the first cycle: runs whatever *done holds and
// sets *done (and sums[5]) to 1 when the divergence is zero, 0 otherwise
// (pressureMultigrid.ts:294-300), so the cycles enqueued behind it skip
// themselves without the host reading the norm first.
void launch_dc_norm_init(hipStream_t s, const double* partials, int nx, int ny, double* sums,
                         double* host_sums, int* done, double scale);
// Flexible CG after the V-cycle has left z = B r in z (first: the first
// step of a solve, beta = 0 and dprev unread).  launch_fcg_dots: the four
// dot products and the coefficients, state = {alpha, beta, <d, A d>}.
// launch_fcg_update: d = z + beta dprev into dnext, p += alpha d,
// r -= alpha A d, r32 = (float) r, norm partials.  launch_fcg_norm: the
// norm as launch_dc_norm writes it, with the divergence terms taken from
// sums0 (the initial pass's sums).
void launch_fcg_dots(hipStream_t s, const float* z, const float* dprev, const double* r,
                     const uint8_t* info, int nx, int ny, int first, double* partials,
                     double* state, const int* done);
void launch_fcg_update(hipStream_t s, const float* z, const float* dprev, float* dnext, double* p,
                       double* r, const uint8_t* info, float* r32, int nx, int ny, int first,
                       const double* state, double* partials, const int* done);
void launch_fcg_norm(hipStream_t s, const double* partials, int nx, int ny, double* sums,
                     const double* sums0, double* host_sums, int* done, double scale, double tol);
// Norms of the current FP32 or FP64 pressure: sums as in launch_dc_update,
// r computed as div + (count p - sum)/scale (pressure.ts:102-142).
void launch_norms(hipStream_t s, const float* p32, const double* p64, const double* div,
                  const uint8_t* lbl, int nx, int ny, double scale, double* partials,
                  double* sums);
// out = p extrapolated at the given order over n cells (order 0: a copy of
// p).  Every array holds an even number of doubles, 16-byte aligned, and out
// differs from the others.
void launch_extrapolate(hipStream_t s, const double* p, const Extrapolation& xp, int order,
                        double* out, long n);
void launch_set_int(hipStream_t s, int* p, int v);

// ---- fused multigrid (fh_mgfused.hip) ----------------------------------------
constexpr int kMaxCoarse = 12;
// Most unknowns of the coarsest level's direct (pseudo-inverse) solve.
constexpr int kMaxDirect = 1024;
// Most cells of one colour on a level of the one-workgroup coarse sub-cycle
// (a workgroup of NC threads holds kCoarseSweepCells / NC per thread).
constexpr int kCoarseSweepCells = 4096;
// Multiplier m of q / d = mulhi(q, m) for 0 <= q, d < 2^16 (FastDiv,
// fh_mgtile.hpp); 0 for d = 1, whose quotient is q itself.
__host__ __device__ inline unsigned fastdiv_magic(int d) {
  return d > 1 ? 0xFFFFFFFFu / (unsigned)d + 1u : 0u;
}
// Bytes of a level's info array padded by a ring of zeros (pitch nx + 2),
// rounded up to 16.
inline int padded_info_bytes(int nx, int ny) { return ((nx + 2) * (ny + 2) + 15) & ~15; }
// The levels of the one-workgroup coarse sub-cycle and its LDS layout: every
// level's unknown x (padded by a ring of zeros, pitch nx + 2), then every
// level's right-hand side b (same padding), then every level's info bytes
// (padded the same way, each level rounded up to 16 bytes), then the direct
// solve's dense vector and indices.  The offsets and the division
// multipliers of nx and of the red-black pairs per row (nx + 1) / 2 come from
// the host (coarse_levels_layout), so the kernel neither divides nor searches
// levels.  image: the info region exactly as it is staged (every level's
// padded info, launch_pad_info), built once per geometry.
struct CoarseLevels {
  int n;
  int nx[kMaxCoarse];
  int ny[kMaxCoarse];
  int xo[kMaxCoarse], bo[kMaxCoarse];  // float offsets of x and b
  int io[kMaxCoarse];                  // byte offset of info
  unsigned mnx[kMaxCoarse], mpr[kMaxCoarse];
  int cells;   // padded cells of all levels (floats of the x region)
  int ibytes;  // bytes of the info region
  const uint8_t* image;
};
// Fills the layout fields of L from n, nx and ny.
void coarse_levels_layout(CoarseLevels& L);
// out = info padded by a ring of zeros (pitch nx + 2), padded_info_bytes long.
void launch_pad_info(hipStream_t s, const uint8_t* info, int nx, int ny, uint8_t* out);
// info byte per cell: bit 7 fluid, bits 0..2 the count of non-solid
// in-domain neighbours; 0 for non-fluid cells.
void launch_cell_info(hipStream_t s, const uint8_t* lbl, int nx, int ny, uint8_t* info);
// Pre-smoothing (2 red-black sweeps, omega 1) + residual + restriction of a
// level to the next one's right-hand side bc.  zero: the unknown starts at 0
// (xin unread).  xin and xout must differ.  shape: tile shape index
// (kTileShapes, fh_mgtile.hpp); every shape gives the same bits.
void launch_smooth_restrict(hipStream_t s, int shape, const float* xin, float* xout,
                            const float* b, const uint8_t* info, int nx, int ny, float* bc, int cnx,
                            int cny, int zero, const int* done);
// Prolongation of the coarse correction e and post-smoothing (2 sweeps).
void launch_prolong_smooth(hipStream_t s, int shape, const float* xin, float* xout,
                           const float* b, const uint8_t* info, int nx, int ny, const float* e,
                           int cnx, int cny, const int* done);
// The finest level's prolongation + post-smoothing fused with the
// defect-correction update after it (prolong_smooth_dc_tile): pn = po + e,
// the residual of pn to rout (FP32; must differ from b) and r64 (may be
// null), and one norm partial per tile into partials; returns the number of
// partials (launch_dc_norm_n's n), prolong_smooth_dc_tiles(shape, nx, ny).
int launch_prolong_smooth_dc(hipStream_t s, int shape, const float* xin, const float* b,
                             const uint8_t* info, int nx, int ny, const float* e, int cnx,
                             int cny, const int* done, const double* po, double* pn,
                             const double* div, float* rout, double* r64, double scale,
                             double* partials);
int prolong_smooth_dc_tiles(int shape, int nx, int ny);
// The sub-cycle of levels L[0..n) in one workgroup: b0 is the first level's
// right-hand side, x0in its warm start (NULL = zero), x0out its result.
// direct: FP64 dense solve on the last level, else 8 red-black pairs.
// group: workgroup size index (kCoarseGroups, fh_mgtile.hpp), same bits.
void launch_coarse_vcycle(hipStream_t s, int group, const CoarseLevels& L, const float* b0,
                          const float* x0in, float* x0out, const double* M, const int* cells,
                          int ncoarse, int direct, const int* done);
size_t coarse_vcycle_lds_bytes(const CoarseLevels& L);
// out[0..3] = sums of 4-slot block partials (nblocks <= 8192).
void launch_finalize_sum4(hipStream_t s, const double* partials, int nblocks, double* out);

// Host, FP64: dense pseudo-inverse M (n x n, row-major) of the coarsest
// level's operator over its unknown cells (fluid with an open neighbour),
// with Neumann components regularised and mean-projected (fh_engine.hip).
void build_pseudo_inverse(int nx, int ny, const std::vector<uint8_t>& lab, std::vector<double>& M,
                          std::vector<int>& cells);

}  // namespace fh
