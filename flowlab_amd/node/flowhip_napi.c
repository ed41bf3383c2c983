/*
 * N-API addon over libflowhip: what a TypeScript HipEngine binds.
 *
 * Mirrors the reference's engine seams (src/core/simulation.ts:35-84
 * GpuStep, src/core/pressureSolver.ts:21-41 PressureSolver,
 * src/core/advector.ts:28-63 Advector).  TypedArrays are passed zero-copy
 * (napi_get_typedarray_info) after their element type and count are checked
 * against the engine's grid (a mismatch throws TypeError before any library
 * call); the GPU work of step/solve/advect/readbacks runs on a libuv worker
 * (napi_create_async_work) and resolves a Promise, as the reference's
 * step() is async only for its device readback (simulation.ts:461-466).
 * Every async job holds references to the engine handle and to the arrays it
 * reads or writes until it completes, so neither can be collected under it.
 * A non-OK status rejects with Error(message), with the reference's messages
 * for re-entrancy and device loss.
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/flowhip.h"

#define NAPI_CALL(env, call)                                        \
  do {                                                              \
    if ((call) != napi_ok) {                                        \
      napi_throw_error((env), NULL, "N-API call failed: " #call);   \
      return NULL;                                                  \
    }                                                               \
  } while (0)

/* The engine handle with the sizes every array argument is checked against. */
typedef struct {
  fh_engine* e;
  size_t nu, nv, nc, nd; /* (nx+1) ny, nx (ny+1), nx ny, dye cells (dnx dny) */
} engine_t;

static napi_value throw_status(napi_env env, int st) {
  const char* msg = st == FH_ERR_DEVICE_LOST ? "GPU device lost" : fh_last_error();
  napi_throw_error(env, NULL, msg && *msg ? msg : "libflowhip error");
  return NULL;
}

static void finalize_engine(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  engine_t* h = (engine_t*)data;
  fh_destroy(h->e);
  free(h);
}

/* NULL (with a TypeError pending) unless v is an engine handle. */
static engine_t* get_engine(napi_env env, napi_value v) {
  void* p = NULL;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_external ||
      napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected an engine handle from create()");
    return NULL;
  }
  return (engine_t*)p;
}

static double get_num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

static const char* type_name(napi_typedarray_type t) {
  switch (t) {
    case napi_float64_array: return "Float64Array";
    case napi_float32_array: return "Float32Array";
    case napi_uint8_array: return "Uint8Array";
    default: return "TypedArray";
  }
}

/* Data pointer of a TypedArray of type `want` holding exactly `n` elements
 * (n == 0: any count).  null/undefined give *ok = 1 and NULL when optional;
 * anything else that does not match throws TypeError and sets *ok = 0. */
static void* get_ta(napi_env env, napi_value v, napi_typedarray_type want, size_t n,
                    int optional, const char* what, int* ok) {
  *ok = 0;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return NULL;
  if (t == napi_null || t == napi_undefined) {
    if (optional) {
      *ok = 1;
      return NULL;
    }
    char m[160];
    snprintf(m, sizeof m, "%s: expected a %s, got %s", what, type_name(want),
             t == napi_null ? "null" : "undefined");
    napi_throw_type_error(env, NULL, m);
    return NULL;
  }
  bool is = false;
  napi_typedarray_type type;
  size_t length = 0, off = 0;
  void* data = NULL;
  napi_value ab;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is ||
      napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &off) != napi_ok ||
      type != want) {
    char m[160];
    snprintf(m, sizeof m, "%s: expected a %s", what, type_name(want));
    napi_throw_type_error(env, NULL, m);
    return NULL;
  }
  if (n && length != n) {
    char m[200];
    snprintf(m, sizeof m, "%s: expected %zu elements, got %zu", what, n, length);
    napi_throw_type_error(env, NULL, m);
    return NULL;
  }
  *ok = 1;
  return data;
}

/* create(nx, ny, dyeScale, h) -> engine handle (null on failure, like initGpu) */
static napi_value Create(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  const int nx = (int)get_num(env, a[0]), ny = (int)get_num(env, a[1]);
  const int ds = (int)get_num(env, a[2]);
  fh_engine* e = fh_create(nx, ny, ds, get_num(env, a[3]));
  napi_value out;
  if (!e) {
    NAPI_CALL(env, napi_get_null(env, &out));
    return out;
  }
  engine_t* h = (engine_t*)calloc(1, sizeof(engine_t));
  if (!h) {
    fh_destroy(e);
    NAPI_CALL(env, napi_get_null(env, &out));
    return out;
  }
  h->e = e;
  h->nu = (size_t)(nx + 1) * ny;
  h->nv = (size_t)nx * (ny + 1);
  h->nc = (size_t)nx * ny;
  h->nd = (size_t)nx * ds * ny * ds;
  if (napi_create_external(env, h, finalize_engine, NULL, &out) != napi_ok) {
    fh_destroy(e);
    free(h);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  return out;
}

/* upload(e, u, v, p, label, dye) -- the re-upload after invalidate(); any
 * field may be null to leave it */
static napi_value Upload(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value a[6];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  int ok[5];
  void* u = get_ta(env, a[1], napi_float64_array, h->nu, 1, "upload u", &ok[0]);
  if (!ok[0]) return NULL;
  void* v = get_ta(env, a[2], napi_float64_array, h->nv, 1, "upload v", &ok[1]);
  if (!ok[1]) return NULL;
  void* p = get_ta(env, a[3], napi_float64_array, h->nc, 1, "upload p", &ok[2]);
  if (!ok[2]) return NULL;
  void* l = get_ta(env, a[4], napi_uint8_array, h->nc, 1, "upload label", &ok[3]);
  if (!ok[3]) return NULL;
  void* d = get_ta(env, a[5], napi_float64_array, 3 * h->nd, 1, "upload dye", &ok[4]);
  if (!ok[4]) return NULL;
  int st = fh_upload(h->e, u, v, p, l, d);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value SetLabels(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  int ok;
  void* l = get_ta(env, a[1], napi_uint8_array, h->nc, 0, "setLabels label", &ok);
  if (!ok) return NULL;
  int st = fh_set_labels(h->e, l);
  if (st) return throw_status(env, st);
  return NULL;
}

/* setDyePatch(e, i0, j0, nx, ny, data|null): data = 4 planes of nx*ny */
static napi_value SetDyePatch(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value a[6];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  const int pnx = (int)get_num(env, a[3]), pny = (int)get_num(env, a[4]);
  int ok;
  const size_t n = pnx > 0 && pny > 0 ? (size_t)4 * pnx * pny : 0;
  void* d = get_ta(env, a[5], napi_float32_array, n, 1, "setDyePatch data", &ok);
  if (!ok) return NULL;
  if (d && n == 0) {
    napi_throw_type_error(env, NULL, "setDyePatch: empty patch");
    return NULL;
  }
  int st = fh_set_dye_patch(h->e, (int)get_num(env, a[1]), (int)get_num(env, a[2]), pnx, pny,
                            (const float*)d);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value SetScalar(napi_env env, napi_callback_info info, int which) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  double v = get_num(env, a[1]);
  int st = which == 0 ? fh_set_buoyancy(h->e, v)
         : which == 1 ? fh_set_confinement(h->e, v)
         : which == 2 ? fh_set_timing(h->e, v != 0.0)
                      : fh_set_warm_extrapolation(h->e, (int)v);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value SetBuoyancy(napi_env env, napi_callback_info info) {
  return SetScalar(env, info, 0);
}

static napi_value SetConfinement(napi_env env, napi_callback_info info) {
  return SetScalar(env, info, 1);
}

/* setTiming(e, on): per-phase device timing of every step (fh_set_timing) */
static napi_value SetTiming(napi_env env, napi_callback_info info) {
  return SetScalar(env, info, 2);
}

/* setWarmExtrapolation(e, order): each step's solve starts from the last
 * order + 1 solutions extrapolated in time (fh_set_warm_extrapolation) */
static napi_value SetWarmExtrapolation(napi_env env, napi_callback_info info) {
  return SetScalar(env, info, 3);
}

/* getTimings(e) -> Float64Array of FH_NUM_PHASES device ms of the last step
 * (advect, forces, divergence, solve, project, dye, total; zeros with timing
 * off).  Waits for that step's end: call it once the step has completed. */
static napi_value GetTimings(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  double ms[FH_NUM_PHASES];
  int st = fh_get_timings(h->e, ms);
  if (st) return throw_status(env, st);
  void* data = NULL;
  napi_value ab, out;
  NAPI_CALL(env, napi_create_arraybuffer(env, sizeof ms, &data, &ab));
  memcpy(data, ms, sizeof ms);
  NAPI_CALL(env, napi_create_typedarray(env, napi_float64_array, FH_NUM_PHASES, ab, 0, &out));
  return out;
}

/* setStir(e, active, x, y, r, dx, dy) */
static napi_value SetStir(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value a[7];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  int st = fh_set_stir(h->e, (int)get_num(env, a[1]), get_num(env, a[2]), get_num(env, a[3]),
                       get_num(env, a[4]), get_num(env, a[5]), get_num(env, a[6]));
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value MaxFaceSpeed(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  engine_t* h = get_engine(env, a[0]);
  if (!h) return NULL;
  double m = 0;
  int st = fh_max_face_speed(h->e, &m);
  if (st) return throw_status(env, st);
  napi_value out;
  NAPI_CALL(env, napi_create_double(env, m, &out));
  return out;
}

/* ---- async work: step / solve / readbacks / advection -------------------- */

enum { OP_STEP, OP_SOLVE, OP_READ_VEL, OP_READ_DYE, OP_READ_P, OP_ADV_VEL, OP_ADV_DYE };

typedef struct {
  int op;
  engine_t* h;
  napi_deferred deferred;
  napi_async_work work;
  napi_ref refs[7]; /* the engine handle, then each array */
  int nrefs;
  void* ptr[6];
  int nptr;
  int scheme, solver, iterations, iters, status;
  double dt, scale, grad_scale, omega, tol, keep[3];
  char err[256];
} job_t;

static void execute(napi_env env, void* data) {
  (void)env;
  job_t* j = (job_t*)data;
  fh_engine* e = j->h->e;
  switch (j->op) {
    case OP_STEP:
      j->status = fh_step(e, j->scheme, j->dt, j->scale, j->grad_scale, j->keep, j->solver,
                          j->iterations, j->omega, j->tol, &j->iters);
      break;
    case OP_SOLVE:
      j->status = fh_solve_pressure(e, j->solver, (double*)j->ptr[0], (const double*)j->ptr[1],
                                    (const uint8_t*)j->ptr[2], j->scale, j->iterations, j->omega,
                                    j->tol, &j->iters);
      break;
    case OP_READ_VEL:
      j->status = fh_read_velocity(e, (double*)j->ptr[0], (double*)j->ptr[1]);
      break;
    case OP_READ_DYE:
      j->status = fh_read_dye(e, (double*)j->ptr[0]);
      break;
    case OP_READ_P:
      j->status = fh_read_pressure(e, (double*)j->ptr[0]);
      break;
    case OP_ADV_VEL:
      j->status = fh_advect_velocity(e, j->scheme, (const double*)j->ptr[0],
                                     (const double*)j->ptr[1], (double*)j->ptr[2],
                                     (double*)j->ptr[3], (const uint8_t*)j->ptr[4], j->dt);
      break;
    case OP_ADV_DYE:
      j->status = fh_advect_dye(e, j->scheme, (const double*)j->ptr[0],
                                (const double*)j->ptr[1], (const double*)j->ptr[2],
                                (double*)j->ptr[3], (const uint8_t*)j->ptr[4], j->dt);
      break;
  }
  if (j->status) {
    const char* m = j->status == FH_ERR_DEVICE_LOST ? "GPU device lost" : fh_last_error();
    strncpy(j->err, m && *m ? m : "libflowhip error", sizeof j->err - 1);
  }
}

static void release(napi_env env, job_t* j) {
  for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
  free(j);
}

static void complete(napi_env env, napi_status status, void* data) {
  job_t* j = (job_t*)data;
  napi_value v;
  if (status != napi_ok || j->status) {
    napi_value msg;
    napi_create_string_utf8(env, j->status ? j->err : "async work failed", NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &v);
    napi_reject_deferred(env, j->deferred, v);
  } else {
    if (j->op == OP_STEP || j->op == OP_SOLVE)
      napi_create_int32(env, j->iters, &v);
    else
      napi_get_undefined(env, &v);
    napi_resolve_deferred(env, j->deferred, v);
  }
  napi_delete_async_work(env, j->work);
  release(env, j);
}

/* A job on the engine a[0], holding a reference to it; NULL (exception
 * pending) when a[0] is not an engine. */
static job_t* new_job(napi_env env, napi_value handle, int op) {
  engine_t* h = get_engine(env, handle);
  if (!h) return NULL;
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  if (!j) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  j->op = op;
  j->h = h;
  if (napi_create_reference(env, handle, 1, &j->refs[j->nrefs]) != napi_ok) {
    free(j);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_reference");
    return NULL;
  }
  j->nrefs++;
  return j;
}

/* Adds array `ta` (type t, n elements) to the job, referenced until it
 * completes; 0 (TypeError pending) if it does not match. */
static int hold(napi_env env, job_t* j, napi_value ta, napi_typedarray_type t, size_t n,
                const char* what) {
  int ok;
  void* p = get_ta(env, ta, t, n, 0, what, &ok);
  if (!ok) return 0;
  if (napi_create_reference(env, ta, 1, &j->refs[j->nrefs]) != napi_ok) {
    napi_throw_error(env, NULL, "N-API call failed: napi_create_reference");
    return 0;
  }
  j->nrefs++;
  j->ptr[j->nptr++] = p;
  return 1;
}

static napi_value start(napi_env env, job_t* j, const char* name) {
  napi_value promise, rname;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &rname) != napi_ok ||
      napi_create_async_work(env, NULL, rname, execute, complete, j, &j->work) != napi_ok) {
    release(env, j);
    napi_throw_error(env, NULL, "N-API call failed: could not start async work");
    return NULL;
  }
  if (napi_queue_async_work(env, j->work) != napi_ok) {
    napi_delete_async_work(env, j->work);
    release(env, j);
    napi_throw_error(env, NULL, "N-API call failed: napi_queue_async_work");
    return NULL;
  }
  return promise;
}

static napi_value bad_args(napi_env env, job_t* j) {
  release(env, j);
  return NULL; /* the TypeError is pending */
}

/* step(e, scheme, dt, scale, gradScale, keep[3], solver, iterations, omega, tol) -> Promise<iters> */
static napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 10;
  napi_value a[10];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = new_job(env, a[0], OP_STEP);
  if (!j) return NULL;
  j->scheme = (int)get_num(env, a[1]);
  j->dt = get_num(env, a[2]);
  j->scale = get_num(env, a[3]);
  j->grad_scale = get_num(env, a[4]);
  for (uint32_t k = 0; k < 3; k++) {
    napi_value el;
    if (napi_get_element(env, a[5], k, &el) != napi_ok) {
      napi_throw_type_error(env, NULL, "step: keep must be an array of 3 numbers");
      return bad_args(env, j);
    }
    j->keep[k] = get_num(env, el);
  }
  j->solver = (int)get_num(env, a[6]);
  j->iterations = (int)get_num(env, a[7]);
  j->omega = get_num(env, a[8]);
  j->tol = get_num(env, a[9]);
  return start(env, j, "flowhip.step");
}

/* solvePressure(e, solver, p, div, label, scale, iterations, omega, tol) -> Promise<iters> */
static napi_value SolvePressure(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value a[9];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = new_job(env, a[0], OP_SOLVE);
  if (!j) return NULL;
  const size_t nc = j->h->nc;
  j->solver = (int)get_num(env, a[1]);
  if (!hold(env, j, a[2], napi_float64_array, nc, "solvePressure p") ||
      !hold(env, j, a[3], napi_float64_array, nc, "solvePressure div") ||
      !hold(env, j, a[4], napi_uint8_array, nc, "solvePressure label"))
    return bad_args(env, j);
  j->scale = get_num(env, a[5]);
  j->iterations = (int)get_num(env, a[6]);
  j->omega = get_num(env, a[7]);
  j->tol = get_num(env, a[8]);
  return start(env, j, "flowhip.solvePressure");
}

static napi_value ReadVelocity(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = new_job(env, a[0], OP_READ_VEL);
  if (!j) return NULL;
  if (!hold(env, j, a[1], napi_float64_array, j->h->nu, "readVelocity u") ||
      !hold(env, j, a[2], napi_float64_array, j->h->nv, "readVelocity v"))
    return bad_args(env, j);
  return start(env, j, "flowhip.readVelocity");
}

static napi_value ReadOne(napi_env env, napi_callback_info info, int op, const char* name) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = new_job(env, a[0], op);
  if (!j) return NULL;
  const size_t n = op == OP_READ_DYE ? 3 * j->h->nd : j->h->nc;
  if (!hold(env, j, a[1], napi_float64_array, n, name)) return bad_args(env, j);
  return start(env, j, name);
}

static napi_value ReadDye(napi_env env, napi_callback_info info) {
  return ReadOne(env, info, OP_READ_DYE, "flowhip.readDye");
}

static napi_value ReadPressure(napi_env env, napi_callback_info info) {
  return ReadOne(env, info, OP_READ_P, "flowhip.readPressure");
}

/* advectVelocity(e, scheme, uIn, vIn, uOut, vOut, label, dt) -> Promise */
static napi_value AdvectVelocity(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value a[8];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = new_job(env, a[0], OP_ADV_VEL);
  if (!j) return NULL;
  const engine_t* h = j->h;
  j->scheme = (int)get_num(env, a[1]);
  if (!hold(env, j, a[2], napi_float64_array, h->nu, "advectVelocity uIn") ||
      !hold(env, j, a[3], napi_float64_array, h->nv, "advectVelocity vIn") ||
      !hold(env, j, a[4], napi_float64_array, h->nu, "advectVelocity uOut") ||
      !hold(env, j, a[5], napi_float64_array, h->nv, "advectVelocity vOut") ||
      !hold(env, j, a[6], napi_uint8_array, h->nc, "advectVelocity label"))
    return bad_args(env, j);
  j->dt = get_num(env, a[7]);
  return start(env, j, "flowhip.advectVelocity");
}

/* advectDye(e, scheme, u, v, qIn(3 planes), qOut(3 planes), label, dt) -> Promise */
static napi_value AdvectDye(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value a[8];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = new_job(env, a[0], OP_ADV_DYE);
  if (!j) return NULL;
  const engine_t* h = j->h;
  j->scheme = (int)get_num(env, a[1]);
  if (!hold(env, j, a[2], napi_float64_array, h->nu, "advectDye u") ||
      !hold(env, j, a[3], napi_float64_array, h->nv, "advectDye v") ||
      !hold(env, j, a[4], napi_float64_array, 3 * h->nd, "advectDye qIn (3 planes)") ||
      !hold(env, j, a[5], napi_float64_array, 3 * h->nd, "advectDye qOut (3 planes)") ||
      !hold(env, j, a[6], napi_uint8_array, h->nc, "advectDye label"))
    return bad_args(env, j);
  j->dt = get_num(env, a[7]);
  return start(env, j, "flowhip.advectDye");
}

static napi_value DeviceCount(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value out;
  NAPI_CALL(env, napi_create_int32(env, fh_device_count(), &out));
  return out;
}

#define EXPORT(name, fn)                                                          \
  do {                                                                            \
    napi_value f;                                                                 \
    if (napi_create_function(env, name, NAPI_AUTO_LENGTH, fn, NULL, &f) != napi_ok || \
        napi_set_named_property(env, exports, name, f) != napi_ok)                \
      return NULL;                                                                \
  } while (0)

static napi_value Init(napi_env env, napi_value exports) {
  EXPORT("deviceCount", DeviceCount);
  EXPORT("create", Create);
  EXPORT("upload", Upload);
  EXPORT("setLabels", SetLabels);
  EXPORT("setDyePatch", SetDyePatch);
  EXPORT("setBuoyancy", SetBuoyancy);
  EXPORT("setConfinement", SetConfinement);
  EXPORT("setStir", SetStir);
  EXPORT("setTiming", SetTiming);
  EXPORT("setWarmExtrapolation", SetWarmExtrapolation);
  EXPORT("getTimings", GetTimings);
  EXPORT("maxFaceSpeed", MaxFaceSpeed);
  EXPORT("step", Step);
  EXPORT("solvePressure", SolvePressure);
  EXPORT("readVelocity", ReadVelocity);
  EXPORT("readDye", ReadDye);
  EXPORT("readPressure", ReadPressure);
  EXPORT("advectVelocity", AdvectVelocity);
  EXPORT("advectDye", AdvectDye);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
