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

static napi_value mass_object(napi_env env, const fs_mass *ms) {
  napi_value m;
  napi_create_object(env, &m);
  set_num(env, m, "volume", ms->volume);
  set_num(env, m, "rain", ms->rain);
  set_num(env, m, "infil", ms->infil);
  set_num(env, m, "giStored", ms->giStored);
  set_num(env, m, "outflow", ms->outflow);
  set_num(env, m, "exchanged", ms->exchanged);
  set_num(env, m, "clipped", ms->clipped);
  set_num(env, m, "expected", ms->expected);
  set_num(env, m, "rel", ms->rel);
  set_num(env, m, "relExClip", ms->relExClip);
  set_num(env, m, "initial", ms->initial);
  return m;
}

/* ---------- create / destroy ---------- */

static void finalize_solver(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  Solver *s = (Solver *)data;
  if (s->h && !s->busy) fs_destroy(s->h);
  free(s);
}

/* fs_params from the options object the JS classes build (defaults of solver-gpu.js:39-47) */
static int parse_params(napi_env env, napi_value o, fs_params *p) {
  memset(p, 0, sizeof *p);
  p->nx = (int32_t)num_prop(env, o, "nx", 0);
  p->ny = (int32_t)num_prop(env, o, "ny", 0);
  p->dx = num_prop(env, o, "dx", 1.0);
  p->g = num_prop(env, o, "g", 9.81);
  p->secondOrder = (int32_t)num_prop(env, o, "secondOrder", 1);
  p->cfl = num_prop(env, o, "cfl", p->secondOrder ? 0.48 : 0.9);
  p->hdry = num_prop(env, o, "hdry", 1e-5);
  p->hfront = num_prop(env, o, "hfront", 1e-3);
  p->dtDry = num_prop(env, o, "dtDry", 1.0);
  p->limiter = (int32_t)num_prop(env, o, "limiter", 0);
  p->bc[0] = (int32_t)num_prop(env, o, "bcWest", 0);
  p->bc[1] = (int32_t)num_prop(env, o, "bcEast", 0);
  p->bc[2] = (int32_t)num_prop(env, o, "bcSouth", 0);
  p->bc[3] = (int32_t)num_prop(env, o, "bcNorth", 0);
  p->lts = (int32_t)num_prop(env, o, "lts", 0);
  p->rows = (int32_t)num_prop(env, o, "rows", 0);
  p->untiled = (int32_t)num_prop(env, o, "untiled", 0);
  p->hazard = (int32_t)num_prop(env, o, "hazard", 0);
  p->implicitFriction = (int32_t)num_prop(env, o, "implicitFriction", 0);
  p->stepFlux = (int32_t)num_prop(env, o, "stepFlux", 0);
  if (p->nx <= 0 || p->ny <= 0) {
    napi_throw_range_error(env, NULL, "create: nx and ny must be positive");
    return -1;
  }
  return 0;
}

/* the model layers of fs_create / fs_pool_add: a[0..6] = z, manning|null, manningScalar, mask|null,
 * infil|null, rainField|null, inlets|null */
typedef struct Layers {
  void *z, *man, *mask, *infil, *rf, *inl;
  double ms;
  size_t nin;
} Layers;

static int parse_layers(napi_env env, const napi_value *a, const fs_params *p, Layers *L) {
  const long n = (long)p->nx * p->ny;
  memset(L, 0, sizeof *L);
  if (is_nullish(env, a[0])) {
    napi_throw_type_error(env, NULL, "create: z is required");
    return -1;
  }
  if (typed(env, a[0], napi_float64_array, n, &L->z, NULL, "z")) return -1;
  if (typed(env, a[1], napi_float64_array, n, &L->man, NULL, "manning")) return -1;
  napi_get_value_double(env, a[2], &L->ms);
  if (typed(env, a[3], napi_uint8_array, n, &L->mask, NULL, "mask")) return -1;
  if (typed(env, a[4], napi_float64_array, n, &L->infil, NULL, "infil")) return -1;
  if (typed(env, a[5], napi_float64_array, n, &L->rf, NULL, "rainField")) return -1;
  if (typed(env, a[6], napi_int32_array, -1, &L->inl, &L->nin, "inlets")) return -1;
  return 0;
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
  Layers L;
  if (parse_params(env, a[0], &p) || parse_layers(env, a + 1, &p, &L)) return NULL;
  fs_handle h = NULL;
  int rc = fs_create(&p, (const double *)L.z, (const double *)L.man, L.ms, (const uint8_t *)L.mask,
                     (const double *)L.infil, (const double *)L.rf, (const int32_t *)L.inl, (int32_t)L.nin, &h);
  if (rc) return fail(env, NULL, rc);
  Solver *s = (Solver *)calloc(1, sizeof(Solver));
  if (!s) {
    fs_destroy(h);
    napi_throw_error(env, NULL, "create: out of host memory");
    return NULL;
  }
  s->h = h;
  s->nx = p.nx;
  s->ny = p.ny;
  s->nInlets = (int32_t)L.nin;
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
        napi_set_named_property(env, result, "mass", mass_object(env, &w->mass));
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

/* snapshotLayout(h) -> [record, planes, barriers, soil, rainField] bytes (fs_snapshot_layout) */
static napi_value js_snapshot_layout(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  Solver *s = get_solver(env, a[0]);
  if (!s) return NULL;
  uint64_t b[FS_SNAP_PARTS];
  int rc = fs_snapshot_layout(s->h, b);
  if (rc) return fail(env, s->h, rc);
  napi_value arr;
  napi_create_array_with_length(env, FS_SNAP_PARTS, &arr);
  for (uint32_t k = 0; k < FS_SNAP_PARTS; ++k) {
    napi_value x;
    napi_create_double(env, (double)b[k], &x);
    napi_set_element(env, arr, k, x);
  }
  return arr;
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

/* ---------- pools: batched / served instances (fs_pool_*) ----------
 *
 * A pool handle holds many instances (slots) advanced together by one launch sequence
 * (floodsim_amd/js/pool-hip.js builds Pool and Server on these). No counterpart in the reference,
 * which runs one solve at a time (src/batch.js:1-4). As with a solver handle, advance, step and
 * readState run on libuv's thread pool and return Promises, and a pool must not be driven by two calls
 * at once: the JS class serialises them and the addon rejects a call that finds the pool busy.
 */

typedef struct PoolH {
  fs_pool_t p;
  int busy;
  int32_t cap;       /* slots described by dims */
  int32_t *dims;     /* per slot: nx, ny (0, 0: no instance) */
} PoolH;

static void finalize_pool(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  PoolH *q = (PoolH *)data;
  if (q->p && !q->busy) fs_pool_destroy(q->p);
  free(q->dims);
  free(q);
}

static PoolH *get_pool(napi_env env, napi_value v) {
  PoolH *q = NULL;
  if (napi_get_value_external(env, v, (void **)&q) != napi_ok || !q) {
    napi_throw_type_error(env, NULL, "floodsim: not a pool handle");
    return NULL;
  }
  if (!q->p) {
    napi_throw_error(env, NULL, "floodsim: pool was destroyed");
    return NULL;
  }
  if (q->busy) {
    napi_throw_error(env, NULL, "floodsim: pool is busy (calls on one pool must not overlap)");
    return NULL;
  }
  return q;
}

/* the slot argument, checked against the instances the pool holds */
static int slot_arg(napi_env env, const PoolH *q, napi_value v, int32_t *slot) {
  int32_t k = -1;
  napi_get_value_int32(env, v, &k);
  if (k < 0 || k >= q->cap || q->dims[2 * k] == 0) {
    napi_throw_range_error(env, NULL, "floodsim: no instance in that pool slot");
    return -1;
  }
  *slot = k;
  return 0;
}

static napi_value pool_fail(napi_env env, const PoolH *q, int rc) {
  char m[600];
  const char *e = fs_pool_last_error(q ? q->p : NULL);
  snprintf(m, sizeof m, "floodsim_hip pool error %d: %s", rc, e ? e : "");
  napi_throw_error(env, NULL, m);
  return NULL;
}

#define POOL_ARGS(N)                                                     \
  size_t argc = (N);                                                     \
  napi_value a[(N) > 0 ? (N) : 1];                                       \
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));       \
  PoolH *q = get_pool(env, a[0]);                                        \
  if (!q) return NULL;

#define POOL_SLOT_ARGS(N)                                                \
  POOL_ARGS(N)                                                           \
  int32_t slot = -1;                                                     \
  if (slot_arg(env, q, a[1], &slot)) return NULL;

#define POOL_RESULT(rc) ((rc) ? pool_fail(env, q, (rc)) : undefined_value(env))

/* poolCreate(hbmBudgetBytes) -> pool handle (0 = the device's free memory) */
static napi_value js_pool_create(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  double budget = 0;
  if (argc > 0 && !is_nullish(env, a[0])) napi_get_value_double(env, a[0], &budget);
  if (!(budget >= 0)) {
    napi_throw_range_error(env, NULL, "poolCreate: the budget must be >= 0 bytes");
    return NULL;
  }
  fs_pool_t p = NULL;
  int rc = fs_pool_create((size_t)budget, &p);
  if (rc) {
    char m[600];
    const char *e = fs_pool_last_error(NULL);
    snprintf(m, sizeof m, "floodsim_hip pool error %d: %s", rc, e ? e : "");
    napi_throw_error(env, NULL, m);
    return NULL;
  }
  PoolH *q = (PoolH *)calloc(1, sizeof(PoolH));
  if (!q) {
    fs_pool_destroy(p);
    napi_throw_error(env, NULL, "poolCreate: out of host memory");
    return NULL;
  }
  q->p = p;
  napi_value ext;
  if (napi_create_external(env, q, finalize_pool, NULL, &ext) != napi_ok) {
    fs_pool_destroy(p);
    free(q);
    napi_throw_error(env, NULL, "poolCreate: napi_create_external failed");
    return NULL;
  }
  return ext;
}

static napi_value js_pool_destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  PoolH *q = NULL;
  if (napi_get_value_external(env, a[0], (void **)&q) != napi_ok || !q) {
    napi_throw_type_error(env, NULL, "poolDestroy: not a pool handle");
    return NULL;
  }
  if (q->busy) {
    napi_throw_error(env, NULL, "poolDestroy: pool is busy");
    return NULL;
  }
  if (q->p) fs_pool_destroy(q->p);
  q->p = NULL;
  return undefined_value(env);
}

/* poolBytesNeeded(params, withInfil, withRainField) -> bytes one instance takes */
static napi_value js_pool_bytes_needed(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(env, napi_get_cb_info(env, info, &argc, a, NULL, NULL));
  fs_params p;
  if (parse_params(env, a[0], &p)) return NULL;
  bool infil = false, rf = false;
  if (argc > 1) napi_get_value_bool(env, a[1], &infil);
  if (argc > 2) napi_get_value_bool(env, a[2], &rf);
  napi_value v;
  napi_create_double(env, (double)fs_pool_bytes_needed(&p, infil ? 1 : 0, rf ? 1 : 0), &v);
  return v;
}

static napi_value js_pool_bytes_used(napi_env env, napi_callback_info info) {
  POOL_ARGS(1)
  napi_value v;
  napi_create_double(env, (double)fs_pool_bytes_used(q->p), &v);
  return v;
}

/* poolAdd(p, params, z, manning|null, manningScalar, mask|null, infil|null, rainField|null, inlets|null)
 * -> slot, or -1 when the pool's HBM budget has no room (FS_E_BUDGET) */
static napi_value js_pool_add(napi_env env, napi_callback_info info) {
  POOL_ARGS(9)
  if (argc < 9) {
    napi_throw_type_error(env, NULL, "poolAdd: 9 arguments expected");
    return NULL;
  }
  fs_params p;
  Layers L;
  if (parse_params(env, a[1], &p) || parse_layers(env, a + 2, &p, &L)) return NULL;
  int32_t slot = -1;
  int rc = fs_pool_add(q->p, &p, (const double *)L.z, (const double *)L.man, L.ms, (const uint8_t *)L.mask,
                       (const double *)L.infil, (const double *)L.rf, (const int32_t *)L.inl, (int32_t)L.nin, &slot);
  napi_value v;
  if (rc == FS_E_BUDGET) {
    napi_create_int32(env, -1, &v);
    return v;
  }
  if (rc) return pool_fail(env, q, rc);
  if (slot >= q->cap) {
    const int32_t cap = slot + 1 > 2 * q->cap ? slot + 1 : 2 * q->cap;
    int32_t *d = (int32_t *)realloc(q->dims, (size_t)cap * 2 * sizeof(int32_t));
    if (!d) {
      fs_pool_remove(q->p, slot);
      napi_throw_error(env, NULL, "poolAdd: out of host memory");
      return NULL;
    }
    memset(d + 2 * q->cap, 0, (size_t)(cap - q->cap) * 2 * sizeof(int32_t));
    q->dims = d;
    q->cap = cap;
  }
  q->dims[2 * slot] = p.nx;
  q->dims[2 * slot + 1] = p.ny;
  napi_create_int32(env, slot, &v);
  return v;
}

static napi_value js_pool_remove(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(2)
  int rc = fs_pool_remove(q->p, slot);
  if (!rc) q->dims[2 * slot] = q->dims[2 * slot + 1] = 0;
  return POOL_RESULT(rc);
}

static napi_value js_pool_set_rows(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  int32_t rows = 0;
  napi_get_value_int32(env, a[2], &rows);
  return POOL_RESULT(fs_pool_set_rows(q->p, slot, rows));
}

/* poolSetState(p, slot, h, hu|null, hv|null, bookInitial) */
static napi_value js_pool_set_state(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(6)
  const long n = (long)q->dims[2 * slot] * q->dims[2 * slot + 1];
  void *h, *hu, *hv;
  if (typed(env, a[2], napi_float64_array, n, &h, NULL, "h")) return NULL;
  if (typed(env, a[3], napi_float64_array, n, &hu, NULL, "hu")) return NULL;
  if (typed(env, a[4], napi_float64_array, n, &hv, NULL, "hv")) return NULL;
  bool book = true;
  if (argc > 5) napi_get_value_bool(env, a[5], &book);
  return POOL_RESULT(fs_pool_set_state(q->p, slot, (const double *)h, (const double *)hu, (const double *)hv, book ? 1 : 0));
}

static napi_value js_pool_set_rain(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  double r = 0;
  napi_get_value_double(env, a[2], &r);
  return POOL_RESULT(fs_pool_set_rain(q->p, slot, r));
}

static napi_value js_pool_set_inlet_rates(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  void *r;
  size_t n = 0;
  if (typed(env, a[2], napi_float64_array, -1, &r, &n, "rates")) return NULL;
  return POOL_RESULT(fs_pool_set_inlet_rates(q->p, slot, (const double *)r, (int32_t)n));
}

static napi_value js_pool_set_stage_cells(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  void *c;
  size_t n = 0;
  if (typed(env, a[2], napi_int32_array, -1, &c, &n, "stageCells")) return NULL;
  return POOL_RESULT(fs_pool_set_stage_cells(q->p, slot, (const int32_t *)c, (int32_t)n));
}

static napi_value js_pool_set_stage(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  double z = 0;
  napi_get_value_double(env, a[2], &z);
  return POOL_RESULT(fs_pool_set_stage(q->p, slot, z));
}

/* poolSetCrests(p, slot, Int32Array i, j, axis, Float64Array z, C, ground) */
static napi_value js_pool_set_crests(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(8)
  void *fi, *fj, *ax, *z, *c, *g;
  size_t n = 0, n2 = 0;
  if (typed(env, a[2], napi_int32_array, -1, &fi, &n, "crest i")) return NULL;
  if (typed(env, a[3], napi_int32_array, (long)n, &fj, &n2, "crest j")) return NULL;
  if (typed(env, a[4], napi_int32_array, (long)n, &ax, &n2, "crest axis")) return NULL;
  if (typed(env, a[5], napi_float64_array, (long)n, &z, &n2, "crest z")) return NULL;
  if (typed(env, a[6], napi_float64_array, (long)n, &c, &n2, "crest C")) return NULL;
  if (typed(env, a[7], napi_float64_array, (long)n, &g, &n2, "crest ground")) return NULL;
  return POOL_RESULT(fs_pool_set_crests(q->p, slot, (const int32_t *)fi, (const int32_t *)fj, (const int32_t *)ax,
                                        (const double *)z, (const double *)c, (const double *)g, (int32_t)n));
}

static napi_value js_pool_set_crest_levels(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  void *z;
  size_t n = 0;
  if (typed(env, a[2], napi_float64_array, -1, &z, &n, "crest levels")) return NULL;
  return POOL_RESULT(fs_pool_set_crest_levels(q->p, slot, (const double *)z));
}

/* poolReadBarriers(p, slot, nBarriers) -> Float64Array(9 n) */
static napi_value js_pool_read_barriers(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(3)
  int32_t n = 0;
  napi_get_value_int32(env, a[2], &n);
  if (n < 0) n = 0;
  double *out = (double *)calloc((size_t)(n ? 9 * n : 1), sizeof(double));
  if (!out) {
    napi_throw_error(env, NULL, "poolReadBarriers: out of memory");
    return NULL;
  }
  int rc = fs_pool_read_barriers(q->p, slot, out);
  napi_value r = rc ? NULL : make_f64(env, out, (size_t)9 * n);
  free(out);
  return rc ? pool_fail(env, q, rc) : r;
}

/* poolSetGa(p, slot, psiD, fmax, perv) / poolSetGi(p, slot, rate, cap, drain): Float64Array(nx*ny) each */
static napi_value pool_set_layers3(napi_env env, napi_callback_info info, int gi) {
  POOL_SLOT_ARGS(5)
  const long n = (long)q->dims[2 * slot] * q->dims[2 * slot + 1];
  void *v[3];
  static const char *nm[2][3] = {{"ga.psiD", "ga.fmax", "ga.perv"}, {"gi.rate", "gi.cap", "gi.drain"}};
  for (int k = 0; k < 3; ++k)
    if (typed(env, a[2 + k], napi_float64_array, n, &v[k], NULL, nm[gi][k])) return NULL;
  int rc = gi ? fs_pool_set_gi(q->p, slot, (const double *)v[0], (const double *)v[1], (const double *)v[2])
              : fs_pool_set_ga(q->p, slot, (const double *)v[0], (const double *)v[1], (const double *)v[2]);
  return POOL_RESULT(rc);
}
static napi_value js_pool_set_ga(napi_env env, napi_callback_info info) { return pool_set_layers3(env, info, 0); }
static napi_value js_pool_set_gi(napi_env env, napi_callback_info info) { return pool_set_layers3(env, info, 1); }

static napi_value js_pool_clock(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(2)
  fs_clock c;
  int rc = fs_pool_clock(q->p, slot, &c);
  return rc ? pool_fail(env, q, rc) : clock_object(env, &c);
}

static napi_value js_pool_set_timing(napi_env env, napi_callback_info info) {
  POOL_ARGS(2)
  bool on = false;
  napi_get_value_bool(env, a[1], &on);
  return POOL_RESULT(fs_pool_set_timing(q->p, on ? 1 : 0));
}

/* poolKernelMs(p) -> {stage1, stage2, launches} (timing mode; -1 when not measured) */
static napi_value js_pool_kernel_ms(napi_env env, napi_callback_info info) {
  POOL_ARGS(1)
  double s1 = -1, s2 = -1;
  int64_t n = 0;
  int rc = fs_pool_kernel_ms(q->p, &s1, &s2, &n);
  if (rc) return pool_fail(env, q, rc);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "stage1", s1);
  set_num(env, o, "stage2", s2);
  set_num(env, o, "launches", (double)n);
  return o;
}

enum { PW_ADVANCE, PW_STEP, PW_READ_STATE };

typedef struct PoolWork {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref keep;          /* the pool's handle object, held until the call completes */
  PoolH *q;
  int kind;
  int32_t n, K, slot;
  int32_t *slots;
  double *targets;
  fs_clock *clks;
  double *f[4];           /* readState: h, hu, hv, hmax */
  fs_mass mass;
  fs_clock clk;
  size_t cells;
  int rc;
  char err[512];
} PoolWork;

static void pool_work_free(PoolWork *w) {
  free(w->slots);
  free(w->targets);
  free(w->clks);
  for (int k = 0; k < 4; ++k) free(w->f[k]);
  free(w);
}

static void pool_work_execute(napi_env env, void *data) {
  (void)env;
  PoolWork *w = (PoolWork *)data;
  switch (w->kind) {
    case PW_ADVANCE:
      w->rc = fs_pool_advance(w->q->p, w->slots, w->targets, w->n, w->K, w->clks);
      break;
    case PW_STEP:
      w->rc = fs_pool_step(w->q->p, w->slots, w->targets, w->n, w->K);
      break;
    case PW_READ_STATE:
      w->rc = fs_pool_read_state(w->q->p, w->slot, w->f[0], w->f[1], w->f[2], w->f[3], &w->mass);
      if (!w->rc) w->rc = fs_pool_clock(w->q->p, w->slot, &w->clk);
      break;
  }
  if (w->rc) {
    const char *e = fs_pool_last_error(w->q->p);
    snprintf(w->err, sizeof w->err, "%s", e ? e : "");
  }
}

static void pool_work_complete(napi_env env, napi_status status, void *data) {
  PoolWork *w = (PoolWork *)data;
  w->q->busy = 0;
  if (status != napi_ok || w->rc) {
    char m[640];
    if (w->rc == FS_E_STALL)
      snprintf(m, sizeof m, "GPU solver made no progress (device error?)");  /* solver-gpu.js:601 */
    else
      snprintf(m, sizeof m, "floodsim_hip pool error %d: %s", w->rc, w->err);
    napi_value msg, err, code;
    napi_create_string_utf8(env, m, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_create_int32(env, w->rc, &code);
    napi_set_named_property(env, err, "code", code);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_value result = NULL;
    switch (w->kind) {
      case PW_ADVANCE:
        napi_create_array_with_length(env, (size_t)w->n, &result);
        for (int32_t k = 0; k < w->n; ++k) napi_set_element(env, result, (uint32_t)k, clock_object(env, &w->clks[k]));
        break;
      case PW_STEP:
        napi_get_undefined(env, &result);
        break;
      case PW_READ_STATE: {
        static const char *names[4] = {"h", "hu", "hv", "hmax"};
        napi_create_object(env, &result);
        for (int k = 0; k < 4; ++k) napi_set_named_property(env, result, names[k], make_f64(env, w->f[k], w->cells));
        napi_set_named_property(env, result, "mass", mass_object(env, &w->mass));
        napi_set_named_property(env, result, "clock", clock_object(env, &w->clk));
        break;
      }
    }
    napi_resolve_deferred(env, w->deferred, result);
  }
  napi_delete_async_work(env, w->work);
  napi_delete_reference(env, w->keep);
  pool_work_free(w);
}

static napi_value pool_queue(napi_env env, napi_value handle, PoolWork *w) {
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_reference(env, handle, 1, &w->keep);
  napi_create_string_utf8(env, "floodsim-pool", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, pool_work_execute, pool_work_complete, w, &w->work);
  w->q->busy = 1;
  if (napi_queue_async_work(env, w->work) != napi_ok) {
    w->q->busy = 0;
    napi_delete_async_work(env, w->work);
    napi_delete_reference(env, w->keep);
    pool_work_free(w);
    napi_throw_error(env, NULL, "floodsim: napi_queue_async_work failed");
    return NULL;
  }
  return promise;
}

/* poolAdvance(p, Int32Array slots, Float64Array targets, K) -> Promise<clock[]>;
 * poolStep(p, slots, targets, K) -> Promise<undefined> (exactly K steps, stopping at each target) */
static napi_value pool_advance_or_step(napi_env env, napi_callback_info info, int kind) {
  POOL_ARGS(4)
  void *sl, *tg;
  size_t n = 0, n2 = 0;
  if (typed(env, a[1], napi_int32_array, -1, &sl, &n, "slots")) return NULL;
  if (typed(env, a[2], napi_float64_array, (long)n, &tg, &n2, "targets")) return NULL;
  for (size_t k = 0; k < n; ++k) {
    const int32_t s = ((const int32_t *)sl)[k];
    if (s < 0 || s >= q->cap || q->dims[2 * s] == 0) {
      napi_throw_range_error(env, NULL, "floodsim: no instance in that pool slot");
      return NULL;
    }
  }
  int32_t K = 64;
  if (argc > 3 && !is_nullish(env, a[3])) napi_get_value_int32(env, a[3], &K);
  PoolWork *w = (PoolWork *)calloc(1, sizeof(PoolWork));
  if (!w) {
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  w->q = q;
  w->kind = kind;
  w->n = (int32_t)n;
  w->K = K;
  w->slots = (int32_t *)malloc((n ? n : 1) * sizeof(int32_t));
  w->targets = (double *)malloc((n ? n : 1) * sizeof(double));
  w->clks = (fs_clock *)calloc(n ? n : 1, sizeof(fs_clock));
  if (!w->slots || !w->targets || !w->clks) {
    pool_work_free(w);
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  if (n) {
    memcpy(w->slots, sl, n * sizeof(int32_t));
    memcpy(w->targets, tg, n * sizeof(double));
  }
  return pool_queue(env, a[0], w);
}
static napi_value js_pool_advance(napi_env env, napi_callback_info info) { return pool_advance_or_step(env, info, PW_ADVANCE); }
static napi_value js_pool_step(napi_env env, napi_callback_info info) { return pool_advance_or_step(env, info, PW_STEP); }

/* poolReadState(p, slot) -> Promise<{h, hu, hv, hmax, mass, clock}> */
static napi_value js_pool_read_state(napi_env env, napi_callback_info info) {
  POOL_SLOT_ARGS(2)
  PoolWork *w = (PoolWork *)calloc(1, sizeof(PoolWork));
  if (!w) {
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  w->q = q;
  w->kind = PW_READ_STATE;
  w->slot = slot;
  w->cells = (size_t)q->dims[2 * slot] * q->dims[2 * slot + 1];
  int oom = 0;
  for (int k = 0; k < 4; ++k) oom |= !(w->f[k] = (double *)malloc(w->cells * sizeof(double)));
  if (oom) {
    pool_work_free(w);
    napi_throw_error(env, NULL, "floodsim: out of host memory");
    return NULL;
  }
  return pool_queue(env, a[0], w);
}

#define POOL_FN(name, fn) {name, NULL, fn, NULL, NULL, NULL, napi_enumerable, NULL},
#define POOL_FUNCTIONS                                        \
  POOL_FN("poolCreate", js_pool_create)                       \
  POOL_FN("poolDestroy", js_pool_destroy)                     \
  POOL_FN("poolBytesNeeded", js_pool_bytes_needed)            \
  POOL_FN("poolBytesUsed", js_pool_bytes_used)                \
  POOL_FN("poolAdd", js_pool_add)                             \
  POOL_FN("poolRemove", js_pool_remove)                       \
  POOL_FN("poolSetRows", js_pool_set_rows)                    \
  POOL_FN("poolSetState", js_pool_set_state)                  \
  POOL_FN("poolSetRain", js_pool_set_rain)                    \
  POOL_FN("poolSetInletRates", js_pool_set_inlet_rates)       \
  POOL_FN("poolSetStageCells", js_pool_set_stage_cells)       \
  POOL_FN("poolSetStage", js_pool_set_stage)                  \
  POOL_FN("poolSetCrests", js_pool_set_crests)                \
  POOL_FN("poolSetCrestLevels", js_pool_set_crest_levels)     \
  POOL_FN("poolReadBarriers", js_pool_read_barriers)          \
  POOL_FN("poolSetGa", js_pool_set_ga)                        \
  POOL_FN("poolSetGi", js_pool_set_gi)                        \
  POOL_FN("poolClock", js_pool_clock)                         \
  POOL_FN("poolSetTiming", js_pool_set_timing)                \
  POOL_FN("poolKernelMs", js_pool_kernel_ms)                  \
  POOL_FN("poolAdvance", js_pool_advance)                     \
  POOL_FN("poolStep", js_pool_step)                           \
  POOL_FN("poolReadState", js_pool_read_state)

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
      {"snapshotLayout", NULL, js_snapshot_layout, NULL, NULL, NULL, napi_enumerable, NULL},
      POOL_FUNCTIONS
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
