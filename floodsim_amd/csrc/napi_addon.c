/*
 * N-API addon: the C ABI of libfloodsim_hip.so (include/floodsim_hip.h) as Node functions.
 *
 * The JavaScript class floodsim_amd/js/solver-hip.js builds the reference GPUSolver interface
 * (src/solver-gpu.js:13-693) on top of these functions. Calls that wait on the device (advance,
 * frame, readState, readInlets) run on libuv's thread pool through napi_async_work and return
 * Promises, so the event loop stays free as it does with the reference's mapAsync readbacks; the
 * rest are synchronous. A handle must not be driven by two calls at once: the JS class serialises
 * them, and the addon rejects a call that finds the handle busy.
 *
 * Arrays cross as typed arrays: Float64Array for layers and state, Uint8Array for the mask,
 * Int32Array for inlet indices, Float32Array for the packed frame.
 */
#ifndef NODE_GYP_MODULE_NAME
#define NODE_GYP_MODULE_NAME floodsim_napi
#endif
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "floodsim_hip.h"

#define NAPI_OK(env, call)                                                   \
  do {                                                                       \
    if ((call) != napi_ok) {                                                 \
      napi_throw_error((env), NULL, "floodsim napi: " #call " failed");      \
      return NULL;                                                           \
    }                                                                        \
  } while (0)

typedef struct Solver {
  fs_handle h;
  int busy;       /* an async call is in flight */
  int32_t nx, ny, nInlets;
} Solver;

/* ---------- argument helpers ---------- */

static napi_value undefined_value(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

static int is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return 1;
  return t == napi_null || t == napi_undefined;
}

/* typed array of the given type and (when n >= 0) length; NULL data for null/undefined */
static int typed(napi_env env, napi_value v, napi_typedarray_type want, long n, void **data, size_t *len,
                 const char *what) {
  *data = NULL;
  if (len) *len = 0;
  if (is_nullish(env, v)) return 0;
  bool ok = false;
  napi_is_typedarray(env, v, &ok);
  if (!ok) {
    char m[160];
    snprintf(m, sizeof m, "%s: expected a typed array", what);
    napi_throw_type_error(env, NULL, m);
    return -1;
  }
  napi_typedarray_type t;
  size_t l, off;
  napi_value ab;
  napi_get_typedarray_info(env, v, &t, &l, data, &ab, &off);
  if (t != want || (n >= 0 && (long)l != n)) {
    char m[160];
    snprintf(m, sizeof m, "%s: wrong typed array type or length (%zu, want %ld)", what, l, n);
    napi_throw_type_error(env, NULL, m);
    return -1;
  }
  if (len) *len = l;
  return 0;
}

static double num_prop(napi_env env, napi_value o, const char *k, double dflt) {
  bool has = false;
  napi_has_named_property(env, o, k, &has);
  if (!has) return dflt;
  napi_value v;
  napi_get_named_property(env, o, k, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b ? 1.0 : 0.0;
  }
  if (t != napi_number) return dflt;
  double d;
  napi_get_value_double(env, v, &d);
  return d;
}

static Solver *get_solver(napi_env env, napi_value v) {
  Solver *s = NULL;
  if (napi_get_value_external(env, v, (void **)&s) != napi_ok || !s) {
    napi_throw_type_error(env, NULL, "floodsim: not a solver handle");
    return NULL;
  }
  if (!s->h) {
    napi_throw_error(env, NULL, "floodsim: solver was destroyed");
    return NULL;
  }
  if (s->busy) {
    napi_throw_error(env, NULL, "floodsim: solver is busy (calls on one solver must not overlap)");
    return NULL;
  }
  return s;
}

static napi_value fail(napi_env env, fs_handle h, int rc) {
  char m[600];
  const char *e = fs_last_error(h);
  snprintf(m, sizeof m, "floodsim_hip error %d: %s", rc, e ? e : "");
  napi_throw_error(env, NULL, m);
  return NULL;
}

static napi_value make_f64(napi_env env, const double *src, size_t n) {
  void *d;
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, n * 8, &d, &ab) != napi_ok) return NULL;
  if (n) memcpy(d, src, n * 8);
  napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta);
  return ta;
}

static napi_value make_f32(napi_env env, const float *src, size_t n) {
  void *d;
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, n * 4, &d, &ab) != napi_ok) return NULL;
  if (n) memcpy(d, src, n * 4);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta);
  return ta;
}

static void set_num(napi_env env, napi_value o, const char *k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, k, x);
}

static napi_value clock_object(napi_env env, const fs_clock *c) {
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "t", c->t);
  set_num(env, o, "lastDt", c->lastDt);
  set_num(env, o, "dtAvg", c->dtAvg);
  set_num(env, o, "steps", (double)c->steps);
  set_num(env, o, "cellUpdates", (double)c->cellUpdates);
  return o;
}

/* ---------- create / destroy ---------- */

static void finalize_solver(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  Solver *s = (Solver *)data;
  if (s->h && !s->busy) fs_destroy(s->h);
  free(s);
}

/* create(params, z, manning|null, manningScalar, mask|null, infil|null, rainField|null, inlets|null) */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value a[8];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  if (argc < 8) {
    napi_throw_type_error(env, NULL, "create: 8 arguments expected");
    return NULL;
  }
  fs_params p;
  memset(&p, 0, sizeof p);
  p.nx = (int32_t)num_prop(env, a[0], "nx", 0);
  p.ny = (int32_t)num_prop(env, a[0], "ny", 0);
  p.dx = num_prop(env, a[0], "dx", 1.0);
  p.g = num_prop(env, a[0], "g", 9.81);
  p.secondOrder = (int32_t)num_prop(env, a[0], "secondOrder", 1);
  p.cfl = num_prop(env, a[0], "cfl", p.secondOrder ? 0.48 : 0.9);
  p.hdry = num_prop(env, a[0], "hdry", 1e-5);
  p.hfront = num_prop(env, a[0], "hfront", 1e-3);
  p.dtDry = num_prop(env, a[0], "dtDry", 1.0);
  p.limiter = (int32_t)num_prop(env, a[0], "limiter", 0);
  p.bc[0] = (int32_t)num_prop(env, a[0], "bcWest", 0);
  p.bc[1] = (int32_t)num_prop(env, a[0], "bcEast", 0);
  p.bc[2] = (int32_t)num_prop(env, a[0], "bcSouth", 0);
  p.bc[3] = (int32_t)num_prop(env, a[0], "bcNorth", 0);
  p.lts = (int32_t)num_prop(env, a[0], "lts", 0);
  p.rows = (int32_t)num_prop(env, a[0], "rows", 0);
  p.untiled = (int32_t)num_prop(env, a[0], "untiled", 0);
  p.hazard = (int32_t)num_prop(env, a[0], "hazard", 0);
  p.implicitFriction = (int32_t)num_prop(env, a[0], "implicitFriction", 0);
  p.stepFlux = (int32_t)num_prop(env, a[0], "stepFlux", 0);
  if (p.nx <= 0 || p.ny <= 0) {
    napi_throw_range_error(env, NULL, "create: nx and ny must be positive");
    return NULL;
  }
  const long n = (long)p.nx * p.ny;
  void *z, *man, *mask, *infil, *rf, *inl;
  size_t nin = 0;
  double ms = 0;
  if (is_nullish(env, a[1])) {
    napi_throw_type_error(env, NULL, "create: z is required");
    return NULL;
  }
  if (typed(env, a[1], napi_float64_array, n, &z, NULL, "z")) return NULL;
  if (typed(env, a[2], napi_float64_array, n, &man, NULL, "manning")) return NULL;
  napi_get_value_double(env, a[3], &ms);
  if (typed(env, a[4], napi_uint8_array, n, &mask, NULL, "mask")) return NULL;
  if (typed(env, a[5], napi_float64_array, n, &infil, NULL, "infil")) return NULL;
  if (typed(env, a[6], napi_float64_array, n, &rf, NULL, "rainField")) return NULL;
  if (typed(env, a[7], napi_int32_array, -1, &inl, &nin, "inlets")) return NULL;
  fs_handle h = NULL;
  int rc = fs_create(&p, (const double *)z, (const double *)man, ms, (const uint8_t *)mask,
                     (const double *)infil, (const double *)rf, (const int32_t *)inl, (int32_t)nin, &h);
  if (rc) return fail(env, NULL, rc);
  Solver *s = (Solver *)calloc(1, sizeof(Solver));
  s->h = h;
  s->nx = p.nx;
  s->ny = p.ny;
  s->nInlets = (int32_t)nin;
  napi_value ext;
  if (napi_create_external(env, s, finalize_solver, NULL, &ext) != napi_ok) {
    fs_destroy(h);
    free(s);
    napi_throw_error(env, NULL, "create: napi_create_external failed");
    return NULL;
  }
  return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = NULL;
  if (napi_get_value_external(env, a[0], (void **)&s) != napi_ok || !s) {
    napi_throw_type_error(env, NULL, "destroy: not a solver handle");
    return NULL;
  }
  if (s->busy) {
    napi_throw_error(env, NULL, "destroy: solver is busy");
    return NULL;
  }
  if (s->h) fs_destroy(s->h);
  s->h = NULL;
  return undefined_value(env);
}

/* ---------- synchronous calls ---------- */

/* setState(h, depth, hu, hv, bookInitial) */
static napi_value js_set_state(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value a[5];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  const long n = (long)s->nx * s->ny;
  void *h, *hu, *hv;
  if (typed(env, a[1], napi_float64_array, n, &h, NULL, "h")) return NULL;
  if (typed(env, a[2], napi_float64_array, n, &hu, NULL, "hu")) return NULL;
  if (typed(env, a[3], napi_float64_array, n, &hv, NULL, "hv")) return NULL;
  bool book = true;
  if (argc > 4) napi_get_value_bool(env, a[4], &book);
  int rc = fs_set_state(s->h, (const double *)h, (const double *)hu, (const double *)hv, book ? 1 : 0);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_set_rain(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  double r = 0;
  napi_get_value_double(env, a[1], &r);
  int rc = fs_set_rain(s->h, r);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_set_rain_field(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *f;
  if (typed(env, a[1], napi_float64_array, (long)s->nx * s->ny, &f, NULL, "rainField")) return NULL;
  int rc = fs_set_rain_field(s->h, (const double *)f);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_set_inlet_rates(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *q;
  size_t n = 0;
  if (typed(env, a[1], napi_float64_array, -1, &q, &n, "rates")) return NULL;
  int rc = fs_set_inlet_rates(s->h, (const double *)q, (int32_t)n);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

/* setStageCells(h, Int32Array cells) -> fs_set_stage_cells; setStage(h, z) -> fs_set_stage */
static napi_value js_set_stage_cells(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *c;
  size_t n = 0;
  if (typed(env, a[1], napi_int32_array, -1, &c, &n, "stageCells")) return NULL;
  int rc = fs_set_stage_cells(s->h, (const int32_t *)c, (int32_t)n);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_set_stage(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  double z = 0;
  napi_get_value_double(env, a[1], &z);
  int rc = fs_set_stage(s->h, z);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

/* setCrests(h, Int32Array i, j, axis, Float64Array z, C, ground) -> fs_set_crests;
 * setCrestLevels(h, Float64Array z) -> fs_set_crest_levels; readBarriers(h, n) -> Float64Array(9 n) */
static napi_value js_set_crests(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value a[7];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *fi, *fj, *ax, *z, *c, *g;
  size_t n = 0, n2 = 0;
  if (typed(env, a[1], napi_int32_array, -1, &fi, &n, "crest i")) return NULL;
  if (typed(env, a[2], napi_int32_array, (long)n, &fj, &n2, "crest j")) return NULL;
  if (typed(env, a[3], napi_int32_array, (long)n, &ax, &n2, "crest axis")) return NULL;
  if (typed(env, a[4], napi_float64_array, (long)n, &z, &n2, "crest z")) return NULL;
  if (typed(env, a[5], napi_float64_array, (long)n, &c, &n2, "crest C")) return NULL;
  if (typed(env, a[6], napi_float64_array, (long)n, &g, &n2, "crest ground")) return NULL;
  int rc = fs_set_crests(s->h, (const int32_t *)fi, (const int32_t *)fj, (const int32_t *)ax, (const double *)z,
                         (const double *)c, (const double *)g, (int32_t)n);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_set_crest_levels(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *z;
  size_t n = 0;
  if (typed(env, a[1], napi_float64_array, -1, &z, &n, "crest levels")) return NULL;
  int rc = fs_set_crest_levels(s->h, (const double *)z);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_read_barriers(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  int32_t n = 0;
  napi_get_value_int32(env, a[1], &n);
  if (n < 0) n = 0;
  double *out = (double *)calloc((size_t)(n ? 9 * n : 1), sizeof(double));
  if (!out) { napi_throw_error(env, NULL, "readBarriers: out of memory"); return NULL; }
  int rc = fs_read_barriers(s->h, out);
  napi_value r = rc ? NULL : make_f64(env, out, (size_t)9 * n);
  free(out);
  return rc ? fail(env, s->h, rc) : r;
}

/* readInfiltration(h) -> Float32Array(nx*ny): fs_read_infiltration */
static napi_value js_read_infiltration(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  const size_t n = (size_t)s->nx * s->ny;
  float *F = (float *)calloc(n ? n : 1, sizeof(float));
  if (!F) { napi_throw_error(env, NULL, "readInfiltration: out of memory"); return NULL; }
  int rc = fs_read_infiltration(s->h, F);
  napi_value r = rc ? NULL : make_f32(env, F, n);
  free(F);
  return rc ? fail(env, s->h, rc) : r;
}

/* setGa(h, Float64Array psiD, fmax, perv) -> fs_set_ga; setGi(h, Float64Array rate, cap, drain) -> fs_set_gi */
static napi_value js_set_layers3(napi_env env, napi_callback_info info, int gi) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *p[3];
  size_t n0 = 0, n = 0;
  if (typed(env, a[1], napi_float64_array, -1, &p[0], &n0, gi ? "gi.rate" : "ga.psiD")) return NULL;
  if (typed(env, a[2], napi_float64_array, (long)n0, &p[1], &n, gi ? "gi.cap" : "ga.fmax")) return NULL;
  if (typed(env, a[3], napi_float64_array, (long)n0, &p[2], &n, gi ? "gi.drain" : "ga.perv")) return NULL;
  int rc = gi ? fs_set_gi(s->h, (const double *)p[0], (const double *)p[1], (const double *)p[2])
              : fs_set_ga(s->h, (const double *)p[0], (const double *)p[1], (const double *)p[2]);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}
static napi_value js_set_ga(napi_env env, napi_callback_info info) { return js_set_layers3(env, info, 0); }
static napi_value js_set_gi(napi_env env, napi_callback_info info) { return js_set_layers3(env, info, 1); }

static napi_value js_clock(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  fs_clock c;
  int rc = fs_clock_get(s->h, &c);
  return rc ? fail(env, s->h, rc) : clock_object(env, &c);
}

static napi_value js_tile_stats(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  fs_tiles t;
  int rc = fs_tile_stats(s->h, &t);
  if (rc) return fail(env, s->h, rc);
  napi_value o, b;
  napi_create_object(env, &o);
  napi_get_boolean(env, t.tiled != 0, &b);
  napi_set_named_property(env, o, "tiled", b);
  set_num(env, o, "tiles", (double)t.tiles);
  set_num(env, o, "update", (double)t.update);
  set_num(env, o, "flux", (double)t.flux);
  set_num(env, o, "active", (double)t.cellsActive);
  set_num(env, o, "cells", (double)t.cellsTotal);
  return o;
}

/* debugNoSteps(h, on) -> fs_pool_debug(FS_DEBUG_NO_STEPS): the stall path's test switch */
static napi_value js_debug_no_steps(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  bool on = false;
  napi_get_value_bool(env, a[1], &on);
  int rc = fs_pool_debug(fs_solver_pool(s->h), FS_DEBUG_NO_STEPS, on ? 1 : 0);
  return rc ? fail(env, s->h, rc) : undefined_value(env);
}

static napi_value js_levels(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  int32_t tx = 0, ty = 0;
  int rc = fs_lts_levels(s->h, NULL, &tx, &ty);
  if (rc) return fail(env, s->h, rc);
  void *d;
  napi_value ab, ta;
  NAPI_OK(env, napi_create_arraybuffer(env, (size_t)tx * ty, &d, &ab));
  rc = fs_lts_levels(s->h, (uint8_t *)d, NULL, NULL);
  if (rc) return fail(env, s->h, rc);
  NAPI_OK(env, napi_create_typedarray(env, napi_uint8_array, (size_t)tx * ty, ab, 0, &ta));
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "levels", ta);
  set_num(env, o, "tx", tx);
  set_num(env, o, "ty", ty);
  return o;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_string_utf8(env, fs_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

/* ---------- asynchronous calls (libuv thread pool) ---------- */

enum { W_ADVANCE, W_FRAME, W_READ_STATE, W_READ_INLETS, W_SNAPSHOT, W_RESTORE, W_HAZARD };

typedef struct Work {
  napi_async_work work;
  napi_deferred deferred;
  Solver *s;
  int kind;
  double t;
  int32_t K;
  int rc;
  char err[512];
  fs_clock clk;
  fs_mass mass;
  float *out;            /* frame: 2*nx*ny; readInlets: app, 4 per inlet */
  double *f[4];          /* readState: h, hu, hv, hmax; readInlets: vol, h */
  float *hz[6];          /* readHazard: dv, force, duration, arrival, vmax, tWet */
  uint8_t *cls[3];       /* readHazard: arr, defra, defraUrban */
  void *blob;            /* snapshot (owned) / restore (a copy of the caller's bytes) */
  size_t blobBytes;
} Work;

static void free_blob(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  free(data);
}

static void work_execute(napi_env env, void *data) {
  (void)env;
  Work *w = (Work *)data;
  fs_handle h = w->s->h;
  const size_t n = (size_t)w->s->nx * w->s->ny;
  switch (w->kind) {
    case W_ADVANCE:
      w->rc = fs_advance(h, w->t, w->K, &w->clk);
      break;
    case W_FRAME:
      w->rc = fs_frame(h, w->t, w->K, w->out, &w->clk);
      break;
    case W_READ_STATE:
      w->rc = fs_read_state(h, w->f[0], w->f[1], w->f[2], w->f[3], &w->mass);
      if (!w->rc) w->rc = fs_clock_get(h, &w->clk);
      break;
    case W_READ_INLETS:
      w->rc = fs_read_inlets(h, w->f[0], w->f[1], w->s->nInlets);
      if (!w->rc) w->rc = fs_read_inlet_approach(h, w->out, w->s->nInlets);
      break;
    case W_SNAPSHOT:
      w->rc = fs_snapshot(h, w->blob, w->blobBytes);
      break;
    case W_HAZARD:
      w->rc = fs_read_hazard(h, w->hz[0], w->hz[1], w->hz[2], w->hz[3], w->hz[4], w->hz[5], w->cls[0], w->cls[1],
                             w->cls[2]);
      break;
    case W_RESTORE:
      w->rc = fs_restore(h, w->blob, w->blobBytes);
      break;
  }
  (void)n;
  if (w->rc) {
    const char *e = fs_last_error(h);
    snprintf(w->err, sizeof w->err, "%s", e ? e : "");
  }
}

static void work_complete(napi_env env, napi_status status, void *data) {
  Work *w = (Work *)data;
  w->s->busy = 0;
  const size_t n = (size_t)w->s->nx * w->s->ny;
  napi_value result = NULL;
  if (status != napi_ok || w->rc) {
    char m[640];
    if (w->rc == FS_E_STALL)
      snprintf(m, sizeof m, "GPU solver made no progress (device error?)");  /* solver-gpu.js:601 */
    else
      snprintf(m, sizeof m, "floodsim_hip error %d: %s", w->rc, w->err);
    napi_value msg, err;
    napi_create_string_utf8(env, m, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_value code;
    napi_create_int32(env, w->rc, &code);
    napi_set_named_property(env, err, "code", code);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    switch (w->kind) {
      case W_ADVANCE:
        result = clock_object(env, &w->clk);
        break;
      case W_FRAME: {
        napi_create_object(env, &result);
        napi_set_named_property(env, result, "clock", clock_object(env, &w->clk));
        napi_value out;
        if (w->out)
          out = make_f32(env, w->out, 2 * n);
        else
          napi_get_null(env, &out);
        napi_set_named_property(env, result, "out", out);
        break;
      }
      case W_READ_STATE: {
        napi_create_object(env, &result);
        static const char *names[4] = {"h", "hu", "hv", "hmax"};
        for (int k = 0; k < 4; ++k) napi_set_named_property(env, result, names[k], make_f64(env, w->f[k], n));
        napi_value m;
        napi_create_object(env, &m);
        set_num(env, m, "volume", w->mass.volume);
        set_num(env, m, "rain", w->mass.rain);
        set_num(env, m, "infil", w->mass.infil);
        set_num(env, m, "giStored", w->mass.giStored);
        set_num(env, m, "outflow", w->mass.outflow);
        set_num(env, m, "exchanged", w->mass.exchanged);
        set_num(env, m, "clipped", w->mass.clipped);
        set_num(env, m, "expected", w->mass.expected);
        set_num(env, m, "rel", w->mass.rel);
        set_num(env, m, "relExClip", w->mass.relExClip);
        set_num(env, m, "initial", w->mass.initial);
        napi_set_named_property(env, result, "mass", m);
        napi_set_named_property(env, result, "clock", clock_object(env, &w->clk));
        break;
      }
      case W_READ_INLETS: {
        napi_create_object(env, &result);
        napi_set_named_property(env, result, "vol", make_f64(env, w->f[0], (size_t)w->s->nInlets));
        napi_set_named_property(env, result, "h", make_f64(env, w->f[1], (size_t)w->s->nInlets));
        napi_set_named_property(env, result, "app", make_f32(env, w->out, 4 * (size_t)w->s->nInlets));
        break;
      }
      case W_SNAPSHOT: {
        /* the snapshot buffer becomes the ArrayBuffer's memory (freed by its finalizer) */
        napi_value ab;
        if (napi_create_external_arraybuffer(env, w->blob, w->blobBytes, free_blob, NULL, &ab) == napi_ok) {
          napi_create_typedarray(env, napi_uint8_array, w->blobBytes, ab, 0, &result);
          w->blob = NULL;
        } else {
          napi_get_null(env, &result);
        }
        break;
      }
      case W_RESTORE:
        napi_get_undefined(env, &result);
        break;
      case W_HAZARD: {
        static const char *fn[6] = {"dv", "force", "duration", "arrival", "vmax", "tWet"};
        static const char *cn[3] = {"arr", "defra", "defraUrban"};
        napi_create_object(env, &result);
        for (int k = 0; k < 6; ++k) napi_set_named_property(env, result, fn[k], make_f32(env, w->hz[k], n));
        napi_value c;
        napi_create_object(env, &c);
        for (int k = 0; k < 3; ++k) {
          void *d;
          napi_value ab, ta;
          napi_create_arraybuffer(env, n, &d, &ab);
          if (n) memcpy(d, w->cls[k], n);
          napi_create_typedarray(env, napi_uint8_array, n, ab, 0, &ta);
          napi_set_named_property(env, c, cn[k], ta);
        }
        napi_set_named_property(env, result, "cls", c);
        break;
      }
    }
    napi_resolve_deferred(env, w->deferred, result);
  }
  napi_delete_async_work(env, w->work);
  free(w->out);
  for (int k = 0; k < 4; ++k) free(w->f[k]);
  for (int k = 0; k < 6; ++k) free(w->hz[k]);
  for (int k = 0; k < 3; ++k) free(w->cls[k]);
  free(w->blob);
  free(w);
}

static napi_value queue_work_blob(napi_env env, Solver *s, int kind, void *blob, size_t bytes);

static napi_value queue_work(napi_env env, Solver *s, int kind, double t, int32_t K, int readOut) {
  Work *w = (Work *)calloc(1, sizeof(Work));
  if (!w) {
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  w->s = s;
  w->kind = kind;
  w->t = t;
  w->K = K;
  const size_t n = (size_t)s->nx * s->ny;
  int oom = 0;
  if (kind == W_FRAME && readOut) oom |= !(w->out = (float *)malloc(2 * n * sizeof(float)));
  if (kind == W_READ_STATE)
    for (int k = 0; k < 4; ++k) oom |= !(w->f[k] = (double *)malloc(n * sizeof(double)));
  if (kind == W_READ_INLETS) {
    for (int k = 0; k < 2; ++k) oom |= !(w->f[k] = (double *)calloc((size_t)s->nInlets + 1, sizeof(double)));
    oom |= !(w->out = (float *)calloc(4 * (size_t)s->nInlets + 1, sizeof(float)));
  }
  if (kind == W_HAZARD) {
    for (int k = 0; k < 6; ++k) oom |= !(w->hz[k] = (float *)malloc(n * sizeof(float) + 4));
    for (int k = 0; k < 3; ++k) oom |= !(w->cls[k] = (uint8_t *)malloc(n + 1));
  }
  if (oom) {
    free(w->out);
    for (int k = 0; k < 4; ++k) free(w->f[k]);
    for (int k = 0; k < 6; ++k) free(w->hz[k]);
    for (int k = 0; k < 3; ++k) free(w->cls[k]);
    free(w);
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_string_utf8(env, "floodsim", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, work_execute, work_complete, w, &w->work);
  s->busy = 1;
  if (napi_queue_async_work(env, w->work) != napi_ok) {
    s->busy = 0;
    napi_delete_async_work(env, w->work);
    free(w->out);
    for (int k = 0; k < 4; ++k) free(w->f[k]);
    free(w);
    napi_throw_error(env, NULL, "floodsim: napi_queue_async_work failed");
    return NULL;
  }
  return promise;
}

static napi_value queue_work_blob(napi_env env, Solver *s, int kind, void *blob, size_t bytes) {
  Work *w = (Work *)calloc(1, sizeof(Work));
  if (!w) {
    free(blob);
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  w->s = s;
  w->kind = kind;
  w->blob = blob;
  w->blobBytes = bytes;
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_string_utf8(env, "floodsim", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, work_execute, work_complete, w, &w->work);
  s->busy = 1;
  if (napi_queue_async_work(env, w->work) != napi_ok) {
    s->busy = 0;
    napi_delete_async_work(env, w->work);
    free(w->blob);
    free(w);
    napi_throw_error(env, NULL, "floodsim: napi_queue_async_work failed");
    return NULL;
  }
  return promise;
}

/* snapshot(h) -> Promise<Uint8Array> */
static napi_value js_snapshot(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  const size_t n = fs_snapshot_bytes(s->h);
  void *blob = malloc(n ? n : 1);
  if (!blob) {
    napi_throw_error(env, NULL, "floodsim: out of host memory for the snapshot");
    return NULL;
  }
  return queue_work_blob(env, s, W_SNAPSHOT, blob, n);
}

/* restore(h, Uint8Array) -> Promise<undefined> (the bytes are copied first) */
static napi_value js_restore(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  void *data;
  size_t n = 0;
  if (typed(env, a[1], napi_uint8_array, -1, &data, &n, "snapshot")) return NULL;
  void *blob = malloc(n ? n : 1);
  if (!blob) {
    napi_throw_error(env, NULL, "floodsim: out of host memory for the snapshot");
    return NULL;
  }
  if (n) memcpy(blob, data, n);
  return queue_work_blob(env, s, W_RESTORE, blob, n);
}

/* advance(h, tTarget, K) -> Promise<clock> */
static napi_value js_advance(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  double t = 0;
  int32_t K = 64;
  napi_get_value_double(env, a[1], &t);
  if (argc > 2 && !is_nullish(env, a[2])) napi_get_value_int32(env, a[2], &K);
  return queue_work(env, s, W_ADVANCE, t, K, 0);
}

/* frame(h, tEnd, K, readOut) -> Promise<{clock, out: Float32Array|null}> */
static napi_value js_frame(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  double t = 0;
  int32_t K = 64;
  bool ro = true;
  napi_get_value_double(env, a[1], &t);
  if (argc > 2 && !is_nullish(env, a[2])) napi_get_value_int32(env, a[2], &K);
  if (argc > 3 && !is_nullish(env, a[3])) napi_get_value_bool(env, a[3], &ro);
  return queue_work(env, s, W_FRAME, t, K, ro ? 1 : 0);
}

static napi_value js_read_state(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  return queue_work(env, s, W_READ_STATE, 0, 0, 0);
}

static napi_value js_read_hazard(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  return queue_work(env, s, W_HAZARD, 0, 0, 0);
}

static napi_value js_read_inlets(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  return queue_work(env, s, W_READ_INLETS, 0, 0, 0);
}

/* ---------- module ---------- */

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"version", NULL, js_version, NULL, NULL, NULL, napi_enumerable, NULL},
      {"create", NULL, js_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setState", NULL, js_set_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setRain", NULL, js_set_rain, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setRainField", NULL, js_set_rain_field, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setInletRates", NULL, js_set_inlet_rates, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setStageCells", NULL, js_set_stage_cells, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setStage", NULL, js_set_stage, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setCrests", NULL, js_set_crests, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setGa", NULL, js_set_ga, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setGi", NULL, js_set_gi, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setCrestLevels", NULL, js_set_crest_levels, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readBarriers", NULL, js_read_barriers, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readInfiltration", NULL, js_read_infiltration, NULL, NULL, NULL, napi_enumerable, NULL},
      {"clock", NULL, js_clock, NULL, NULL, NULL, napi_enumerable, NULL},
      {"tileStats", NULL, js_tile_stats, NULL, NULL, NULL, napi_enumerable, NULL},
      {"debugNoSteps", NULL, js_debug_no_steps, NULL, NULL, NULL, napi_enumerable, NULL},
      {"levels", NULL, js_levels, NULL, NULL, NULL, napi_enumerable, NULL},
      {"advance", NULL, js_advance, NULL, NULL, NULL, napi_enumerable, NULL},
      {"frame", NULL, js_frame, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readState", NULL, js_read_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readInlets", NULL, js_read_inlets, NULL, NULL, NULL, napi_enumerable, NULL},
      {"snapshot", NULL, js_snapshot, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readHazard", NULL, js_read_hazard, NULL, NULL, NULL, napi_enumerable, NULL},
      {"restore", NULL, js_restore, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
