/*
 * N-API addon over libflowhip: what a TypeScript HipEngine binds.
 *
 * Mirrors the reference's engine seams (src/core/simulation.ts:35-84
 * GpuStep, src/core/pressureSolver.ts:21-41 PressureSolver,
 * src/core/advector.ts:28-63 Advector).  TypedArrays are passed zero-copy
 * (napi_get_typedarray_info); the GPU work of step/solve/advect runs on a
 * libuv worker (napi_create_async_work) and resolves a Promise, as the
 * reference's step() is async only for its device readback
 * (simulation.ts:461-466).  A non-OK status rejects with Error(message),
 * with the reference's messages for re-entrancy and device loss.
 */
#include <node_api.h>
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

static napi_value throw_status(napi_env env, int st) {
  const char* msg = st == FH_ERR_DEVICE_LOST ? "GPU device lost" : fh_last_error();
  napi_throw_error(env, NULL, msg && *msg ? msg : "libflowhip error");
  return NULL;
}

static void finalize_engine(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  fh_destroy((fh_engine*)data);
}

static fh_engine* get_engine(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (fh_engine*)p;
}

static double get_num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

/* data pointer of a TypedArray of the expected type (or NULL for null/undefined) */
static void* get_ta(napi_env env, napi_value v, napi_typedarray_type want, size_t* len) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_null || t == napi_undefined) return NULL;
  bool is;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return NULL;
  napi_typedarray_type type;
  size_t length;
  void* data;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &off) != napi_ok) return NULL;
  if (type != want) return NULL;
  if (len) *len = length;
  return data;
}

/* create(nx, ny, dyeScale, h) -> engine handle (null on failure, like initGpu) */
static napi_value Create(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  fh_engine* e = fh_create((int)get_num(env, a[0]), (int)get_num(env, a[1]),
                           (int)get_num(env, a[2]), get_num(env, a[3]));
  napi_value out;
  if (!e) {
    NAPI_CALL(env, napi_get_null(env, &out));
    return out;
  }
  NAPI_CALL(env, napi_create_external(env, e, finalize_engine, NULL, &out));
  return out;
}

/* upload(e, u, v, p, label, dye) -- the re-upload after invalidate() */
static napi_value Upload(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value a[6];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  fh_engine* e = get_engine(env, a[0]);
  int st = fh_upload(e, get_ta(env, a[1], napi_float64_array, NULL),
                     get_ta(env, a[2], napi_float64_array, NULL),
                     get_ta(env, a[3], napi_float64_array, NULL),
                     get_ta(env, a[4], napi_uint8_array, NULL),
                     get_ta(env, a[5], napi_float64_array, NULL));
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value SetLabels(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  int st = fh_set_labels(get_engine(env, a[0]), get_ta(env, a[1], napi_uint8_array, NULL));
  if (st) return throw_status(env, st);
  return NULL;
}

/* setDyePatch(e, i0, j0, nx, ny, data|null) */
static napi_value SetDyePatch(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value a[6];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  int st = fh_set_dye_patch(get_engine(env, a[0]), (int)get_num(env, a[1]),
                            (int)get_num(env, a[2]), (int)get_num(env, a[3]),
                            (int)get_num(env, a[4]), get_ta(env, a[5], napi_float32_array, NULL));
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value SetScalar(napi_env env, napi_callback_info info, int which) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  fh_engine* e = get_engine(env, a[0]);
  double v = get_num(env, a[1]);
  int st = which == 0 ? fh_set_buoyancy(e, v) : fh_set_confinement(e, v);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value SetBuoyancy(napi_env env, napi_callback_info info) {
  return SetScalar(env, info, 0);
}

static napi_value SetConfinement(napi_env env, napi_callback_info info) {
  return SetScalar(env, info, 1);
}

/* setStir(e, active, x, y, r, dx, dy) */
static napi_value SetStir(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value a[7];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  int st = fh_set_stir(get_engine(env, a[0]), (int)get_num(env, a[1]), get_num(env, a[2]),
                       get_num(env, a[3]), get_num(env, a[4]), get_num(env, a[5]),
                       get_num(env, a[6]));
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value MaxFaceSpeed(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  double m = 0;
  int st = fh_max_face_speed(get_engine(env, a[0]), &m);
  if (st) return throw_status(env, st);
  napi_value out;
  NAPI_CALL(env, napi_create_double(env, m, &out));
  return out;
}

/* ---- async work: step / solve / readbacks / advection -------------------- */

enum { OP_STEP, OP_SOLVE, OP_READ_VEL, OP_READ_DYE, OP_READ_P, OP_ADV_VEL, OP_ADV_DYE };

typedef struct {
  int op;
  fh_engine* e;
  napi_deferred deferred;
  napi_async_work work;
  napi_ref refs[6];
  int nrefs;
  void* ptr[6];
  int scheme, solver, iterations, iters, status;
  double dt, scale, grad_scale, omega, tol, keep[3];
  char err[256];
} job_t;

static void execute(napi_env env, void* data) {
  (void)env;
  job_t* j = (job_t*)data;
  switch (j->op) {
    case OP_STEP:
      j->status = fh_step(j->e, j->scheme, j->dt, j->scale, j->grad_scale, j->keep, j->solver,
                          j->iterations, j->omega, j->tol, &j->iters);
      break;
    case OP_SOLVE:
      j->status = fh_solve_pressure(j->e, j->solver, (double*)j->ptr[0], (const double*)j->ptr[1],
                                    (const uint8_t*)j->ptr[2], j->scale, j->iterations, j->omega,
                                    j->tol, &j->iters);
      break;
    case OP_READ_VEL:
      j->status = fh_read_velocity(j->e, (double*)j->ptr[0], (double*)j->ptr[1]);
      break;
    case OP_READ_DYE:
      j->status = fh_read_dye(j->e, (double*)j->ptr[0]);
      break;
    case OP_READ_P:
      j->status = fh_read_pressure(j->e, (double*)j->ptr[0]);
      break;
    case OP_ADV_VEL:
      j->status = fh_advect_velocity(j->e, j->scheme, (const double*)j->ptr[0],
                                     (const double*)j->ptr[1], (double*)j->ptr[2],
                                     (double*)j->ptr[3], (const uint8_t*)j->ptr[4], j->dt);
      break;
    case OP_ADV_DYE:
      j->status = fh_advect_dye(j->e, j->scheme, (const double*)j->ptr[0],
                                (const double*)j->ptr[1], (const double*)j->ptr[2],
                                (double*)j->ptr[3], (const uint8_t*)j->ptr[4], j->dt);
      break;
  }
  if (j->status) {
    const char* m = j->status == FH_ERR_DEVICE_LOST ? "GPU device lost" : fh_last_error();
    strncpy(j->err, m ? m : "libflowhip error", sizeof j->err - 1);
  }
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
  for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
  napi_delete_async_work(env, j->work);
  free(j);
}

/* keeps each TypedArray alive until the job completes */
static int hold(napi_env env, job_t* j, napi_value ta, napi_typedarray_type t) {
  void* p = get_ta(env, ta, t, NULL);
  if (!p) return 0;
  napi_create_reference(env, ta, 1, &j->refs[j->nrefs]);
  j->ptr[j->nrefs++] = p;
  return 1;
}

static napi_value start(napi_env env, job_t* j, const char* name) {
  napi_value promise, rname;
  NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
  NAPI_CALL(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &rname));
  NAPI_CALL(env, napi_create_async_work(env, NULL, rname, execute, complete, j, &j->work));
  NAPI_CALL(env, napi_queue_async_work(env, j->work));
  return promise;
}

static napi_value bad_args(napi_env env, job_t* j, const char* what) {
  for (int k = 0; k < j->nrefs; k++) napi_delete_reference(env, j->refs[k]);
  free(j);
  napi_throw_type_error(env, NULL, what);
  return NULL;
}

/* step(e, scheme, dt, scale, gradScale, keep[3], solver, iterations, omega, tol) -> Promise<iters> */
static napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 10;
  napi_value a[10];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  j->op = OP_STEP;
  j->e = get_engine(env, a[0]);
  j->scheme = (int)get_num(env, a[1]);
  j->dt = get_num(env, a[2]);
  j->scale = get_num(env, a[3]);
  j->grad_scale = get_num(env, a[4]);
  for (uint32_t k = 0; k < 3; k++) {
    napi_value el;
    napi_get_element(env, a[5], k, &el);
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
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  j->op = OP_SOLVE;
  j->e = get_engine(env, a[0]);
  j->solver = (int)get_num(env, a[1]);
  if (!hold(env, j, a[2], napi_float64_array) || !hold(env, j, a[3], napi_float64_array) ||
      !hold(env, j, a[4], napi_uint8_array))
    return bad_args(env, j, "solvePressure(e, solver, Float64Array p, Float64Array div, Uint8Array label, ...)");
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
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  j->op = OP_READ_VEL;
  j->e = get_engine(env, a[0]);
  if (!hold(env, j, a[1], napi_float64_array) || !hold(env, j, a[2], napi_float64_array))
    return bad_args(env, j, "readVelocity(e, Float64Array u, Float64Array v)");
  return start(env, j, "flowhip.readVelocity");
}

static napi_value ReadOne(napi_env env, napi_callback_info info, int op, const char* name) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  j->op = op;
  j->e = get_engine(env, a[0]);
  if (!hold(env, j, a[1], napi_float64_array)) return bad_args(env, j, "expected a Float64Array");
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
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  j->op = OP_ADV_VEL;
  j->e = get_engine(env, a[0]);
  j->scheme = (int)get_num(env, a[1]);
  if (!hold(env, j, a[2], napi_float64_array) || !hold(env, j, a[3], napi_float64_array) ||
      !hold(env, j, a[4], napi_float64_array) || !hold(env, j, a[5], napi_float64_array) ||
      !hold(env, j, a[6], napi_uint8_array))
    return bad_args(env, j, "advectVelocity(e, scheme, uIn, vIn, uOut, vOut, label, dt)");
  j->dt = get_num(env, a[7]);
  return start(env, j, "flowhip.advectVelocity");
}

/* advectDye(e, scheme, u, v, qIn(3 planes), qOut(3 planes), label, dt) -> Promise */
static napi_value AdvectDye(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value a[8];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  job_t* j = (job_t*)calloc(1, sizeof(job_t));
  j->op = OP_ADV_DYE;
  j->e = get_engine(env, a[0]);
  j->scheme = (int)get_num(env, a[1]);
  if (!hold(env, j, a[2], napi_float64_array) || !hold(env, j, a[3], napi_float64_array) ||
      !hold(env, j, a[4], napi_float64_array) || !hold(env, j, a[5], napi_float64_array) ||
      !hold(env, j, a[6], napi_uint8_array))
    return bad_args(env, j, "advectDye(e, scheme, u, v, qIn, qOut, label, dt)");
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
