/*
 * bd_node.c -- N-API addon over libbd.so: the Node side of the C-ABI (include/bd.h).
 *
 * Each function takes the solver as an external handle and JS typed arrays for the data; typed
 * arrays are read in place for the duration of the call (the library copies what it keeps, as
 * queue.writeBuffer does in the reference), readbacks fill caller arrays. A failing call throws a JS
 * Error with bd_last_error's message, as the reference's GpuSolver constructor throws for its limits
 * (src/simulation/gpu/solver.ts:298-363). gpu_solver.js wraps these into the GpuSolver-shaped class.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bd.h"

#define CHECK_NAPI(call)                                  \
  do {                                                    \
    if ((call) != napi_ok) {                              \
      napi_throw_error(env, NULL, "N-API call failed: " #call); \
      return NULL;                                        \
    }                                                     \
  } while (0)

static napi_value throw_bd(napi_env env, const bd_solver* s, int rc) {
  char msg[512];
  snprintf(msg, sizeof msg, "%s (code %d)", bd_last_error(s), rc);
  napi_throw_error(env, NULL, msg);
  return NULL;
}

static napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

/* property helpers: a missing property leaves the default */
static int has(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool h = false;
  if (napi_has_named_property(env, obj, name, &h) != napi_ok || !h) return 0;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return 0;
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

static double num(napi_env env, napi_value obj, const char* name, double def) {
  napi_value v;
  double d = def;
  if (has(env, obj, name, &v)) {
    bool b;
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_boolean && napi_get_value_bool(env, v, &b) == napi_ok) return b ? 1.0 : 0.0;
    napi_get_value_double(env, v, &d);
  }
  return d;
}

/* a typed array's data (NULL when absent); *count receives its length */
static void* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* count) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return NULL;
  napi_typedarray_type t;
  size_t n, off;
  void* data;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off) != napi_ok || t != want) return NULL;
  if (count) *count = n;
  return data;
}

static void* prop_typed(napi_env env, napi_value obj, const char* name, napi_typedarray_type want, size_t* count) {
  napi_value v;
  if (!has(env, obj, name, &v)) return NULL;
  return typed(env, v, want, count);
}

static bd_solver* handle(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (bd_solver*)p;
}

static void finalize(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  bd_destroy((bd_solver*)data);
}

static void read_yarn(napi_env env, napi_value o, bd_yarn* y) {
  y->linear_density = num(env, o, "linearDensity", 0);
  y->ea_low = num(env, o, "eaLow", 0);
  y->ea_high = num(env, o, "eaHigh", 0);
  y->crimp_strain = num(env, o, "crimpStrain", 0);
  y->ea_compress = num(env, o, "eaCompress", 0);
  y->axial_viscosity = num(env, o, "axialViscosity", 0);
  y->bend_rigidity = num(env, o, "bendRigidity", 0);
  y->bend_viscosity = num(env, o, "bendViscosity", 0);
  y->strength = num(env, o, "strength", 0);
  y->softening_opening = num(env, o, "softeningOpening", 0);
  y->yarns_per_span = (int32_t)num(env, o, "yarnsPerSpan", 1);
}

static int read_layer(napi_env env, napi_value o, bd_layer* L) {
  memset(L, 0, sizeof *L);
  L->nx = (int32_t)num(env, o, "nx", 0);
  L->ny = (int32_t)num(env, o, "ny", 0);
  L->node_offset = (int32_t)num(env, o, "nodeOffset", 0);
  L->bundle_u = (int32_t)num(env, o, "bundleU", 1);
  L->bundle_v = (int32_t)num(env, o, "bundleV", 1);
  L->field_offset = (int32_t)num(env, o, "fieldOffset", 0);
  L->field_k = (int32_t)num(env, o, "fieldK", 0);
  L->field_m = (int32_t)num(env, o, "fieldM", 0);
  L->weave_pitch_u = num(env, o, "weavePitchU", 0);
  L->weave_pitch_v = num(env, o, "weavePitchV", 0);
  napi_value m, ys, y;
  if (!has(env, o, "material", &m)) return 0;
  L->material.pitch_u = num(env, m, "pitchU", 0);
  L->material.pitch_v = num(env, m, "pitchV", 0);
  L->material.shear_stiffness = num(env, m, "shearStiffness", 0);
  L->material.crossing_friction = num(env, m, "crossingFriction", 0);
  L->material.crossing_normal0 = num(env, m, "crossingNormal0", 0);
  L->material.crossing_wrap = num(env, m, "crossingWrap", 0);
  L->material.air_drag = num(env, m, "airDrag", 0);
  L->material.thickness = num(env, m, "thickness", 0);
  if (!has(env, m, "yarns", &ys)) return 0;
  for (uint32_t f = 0; f < 2; f++) {
    if (napi_get_element(env, ys, f, &y) != napi_ok) return 0;
    read_yarn(env, y, &L->material.yarns[f]);
  }
  return 1;
}

/* GPU_DEFAULTS (gpu/solver.ts:61-72) where a property is absent */
static void read_params(napi_env env, napi_value o, bd_params* p) {
  memset(p, 0, sizeof *p);
  p->h = num(env, o, "h", 1.0 / 120.0);
  napi_value g, e;
  p->gravity[2] = -9.81;
  if (has(env, o, "gravity", &g))
    for (uint32_t q = 0; q < 3; q++)
      if (napi_get_element(env, g, q, &e) == napi_ok) napi_get_value_double(env, e, &p->gravity[q]);
  p->newton_iterations = (int32_t)num(env, o, "newtonIterations", 2);
  p->newton_max = (int32_t)num(env, o, "newtonMax", p->newton_iterations);
  p->newton_tolerance = num(env, o, "newtonTolerance", 1e-5);
  p->pcg_iterations = (int32_t)num(env, o, "pcgIterations", 2);
  p->pcg_max_iterations = (int32_t)num(env, o, "pcgMaxIterations", 200);
  p->pcg_rel_tol = num(env, o, "pcgRelTol", 0);
  p->coarse_sweeps = (int32_t)num(env, o, "coarseSweeps", 2);
  p->max_coarsest_nodes = (int32_t)num(env, o, "maxCoarsestNodes", 400);
  p->lag_coarse = (int32_t)num(env, o, "lagCoarse", 1);
  p->line_search = (int32_t)num(env, o, "lineSearch", 1);
  p->history = (int32_t)num(env, o, "history", 1);
  p->precision = (int32_t)num(env, o, "precision", BD_PRECISION_FP32);
  p->self_contact = (int32_t)num(env, o, "selfContact", 0);
  p->coarse_solve = (int32_t)num(env, o, "coarseSolve", BD_COARSE_SWEEPS);
  p->smooth_sweeps = (int32_t)num(env, o, "smoothSweeps", 1);
  p->net_level = (int32_t)num(env, o, "netLevel", 0);
}

/* create(scene, params) -> handle (new GpuSolver(device, scene, params), gpu/solver.ts:275) */
static napi_value create(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2], layers;
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) {
    napi_throw_type_error(env, NULL, "create(scene, params)");
    return NULL;
  }
  bd_scene sc;
  memset(&sc, 0, sizeof sc);
  sc.node_count = (int32_t)num(env, argv[0], "nodeCount", 0);
  sc.crossing_count = (int32_t)num(env, argv[0], "crossingCount", 0);
  uint32_t nl = 0;
  if (!has(env, argv[0], "layers", &layers) || napi_get_array_length(env, layers, &nl) != napi_ok || nl == 0 || nl > 8) {
    napi_throw_type_error(env, NULL, "scene.layers: 1..8 layer objects");
    return NULL;
  }
  bd_layer L[8];
  for (uint32_t l = 0; l < nl; l++) {
    napi_value lo;
    CHECK_NAPI(napi_get_element(env, layers, l, &lo));
    if (!read_layer(env, lo, &L[l])) {
      napi_throw_type_error(env, NULL, "scene.layers[]: nx, ny, nodeOffset and a material with 2 yarns");
      return NULL;
    }
  }
  sc.layer_count = (int32_t)nl;
  sc.layers = L;
  const size_t N = sc.node_count > 0 ? (size_t)sc.node_count : 0, Cn = sc.crossing_count > 0 ? (size_t)sc.crossing_count : 0;
  struct {
    const char* name;
    napi_typedarray_type type;
    size_t len;
    const void** dst;
  } arrays[] = {
      {"position", napi_float64_array, 3 * N, (const void**)&sc.position},
      {"mass", napi_float64_array, N, (const void**)&sc.mass},
      {"pinned", napi_uint8_array, N, (const void**)&sc.pinned},
      {"restLength", napi_float64_array, 2 * N, (const void**)&sc.rest_length},
      {"softening", napi_float64_array, 2 * N, (const void**)&sc.softening},
      {"segFlags", napi_uint16_array, 2 * N, (const void**)&sc.seg_flags},
      {"intact", napi_float64_array, 2 * N, (const void**)&sc.intact},
      {"cutMask", napi_uint32_array, 2 * N, (const void**)&sc.cut_mask},
      {"failSite", napi_float64_array, 2 * N, (const void**)&sc.fail_site},
      {"crossingNormal0", napi_float64_array, N, (const void**)&sc.crossing_normal0},
      {"tails", napi_float64_array, 4 * N, (const void**)&sc.tails},
      {"fieldStrength", napi_float32_array, 2 * Cn, (const void**)&sc.field_strength},
      {"fieldDamage", napi_uint8_array, 2 * Cn, (const void**)&sc.field_damage},
      {"fieldArea", napi_uint8_array, 2 * Cn, (const void**)&sc.field_area},
  };
  for (size_t a = 0; a < sizeof arrays / sizeof arrays[0]; a++) {
    size_t n = 0;
    *arrays[a].dst = prop_typed(env, argv[0], arrays[a].name, arrays[a].type, &n);
    if (!*arrays[a].dst || n < arrays[a].len) {
      char msg[160];
      snprintf(msg, sizeof msg, "scene.%s: a typed array of the right type with at least %zu entries", arrays[a].name,
               arrays[a].len);
      napi_throw_type_error(env, NULL, msg);
      return NULL;
    }
  }
  bd_params p;
  read_params(env, argv[1], &p);
  bd_solver* s = NULL;
  const int rc = bd_create(&sc, &p, &s);
  if (rc != BD_OK) return throw_bd(env, NULL, rc);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, s, finalize, NULL, &ext));
  return ext;
}

static napi_value args(napi_env env, napi_callback_info info, size_t want, napi_value* argv, bd_solver** s) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok || argc < 1 || !(*s = handle(env, argv[0]))) {
    napi_throw_type_error(env, NULL, "the first argument is a solver handle");
    return NULL;
  }
  for (size_t i = argc; i < want; i++) napi_get_undefined(env, &argv[i]);
  return argv[0];
}

static int32_t int_arg(napi_env env, napi_value v, int32_t def) {
  int32_t x = def;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_int32(env, v, &x);
  return x;
}

static double dbl_arg(napi_env env, napi_value v, double def) {
  double x = def;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_double(env, v, &x);
  return x;
}

/* substep(h, min, max): encodeSubstep (gpu/solver.ts:737-760) */
static napi_value substep(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  bd_solver* s;
  if (!args(env, info, 3, argv, &s)) return NULL;
  const int rc = bd_substep(s, int_arg(env, argv[1], 0), int_arg(env, argv[2], 0));
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

/* setState(h, x Float64Array(3N), v Float64Array(3N) | null) (gpu/solver.ts:831-836) */
static napi_value set_state(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  bd_solver* s;
  if (!args(env, info, 3, argv, &s)) return NULL;
  const double* x = typed(env, argv[1], napi_float64_array, NULL);
  const double* v = typed(env, argv[2], napi_float64_array, NULL);
  if (!x) {
    napi_throw_type_error(env, NULL, "setState: x is a Float64Array of 3N");
    return NULL;
  }
  const int rc = bd_set_state(s, x, v);
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

/* setGrips(h, {node, grip, weight, offset, goal, stiffness}) (ReferenceGrip, solver.ts:25-37) */
static napi_value set_grips(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  bd_solver* s;
  if (!args(env, info, 2, argv, &s)) return NULL;
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  if (t != napi_object) {
    const int rc = bd_set_grips(s, NULL);
    return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
  }
  size_t ne = 0, ng = 0;
  bd_grips g;
  g.node = prop_typed(env, argv[1], "node", napi_int32_array, &ne);
  g.grip = prop_typed(env, argv[1], "grip", napi_int32_array, NULL);
  g.weight = prop_typed(env, argv[1], "weight", napi_float64_array, NULL);
  g.offset = prop_typed(env, argv[1], "offset", napi_float64_array, NULL);
  g.goal = prop_typed(env, argv[1], "goal", napi_float64_array, &ng);
  g.stiffness = prop_typed(env, argv[1], "stiffness", napi_float64_array, NULL);
  if (!g.node || !g.grip || !g.weight || !g.offset || !g.goal || !g.stiffness) {
    napi_throw_type_error(env, NULL, "setGrips: node, grip (Int32Array), weight, offset, goal, stiffness (Float64Array)");
    return NULL;
  }
  g.entry_count = (int32_t)ne;
  g.grip_count = (int32_t)(ng / 3);
  const int rc = bd_set_grips(s, &g);
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

/* setGripGoals(h, goals Float64Array(3 per grip)) */
static napi_value set_grip_goals(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  bd_solver* s;
  if (!args(env, info, 2, argv, &s)) return NULL;
  size_t n = 0;
  const double* goals = typed(env, argv[1], napi_float64_array, &n);
  if (!goals) {
    napi_throw_type_error(env, NULL, "setGripGoals: a Float64Array of 3 per grip");
    return NULL;
  }
  const int rc = bd_set_grip_goals(s, (int)(n / 3), goals);
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

/* read(h, buffer, level, typedArray): bd_read into the array's bytes */
static napi_value read_buf(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  bd_solver* s;
  if (!args(env, info, 4, argv, &s)) return NULL;
  bool is = false;
  napi_is_typedarray(env, argv[3], &is);
  if (!is) {
    napi_throw_type_error(env, NULL, "read: the destination is a typed array");
    return NULL;
  }
  napi_typedarray_type t;
  size_t n, off;
  void* data;
  napi_value ab;
  CHECK_NAPI(napi_get_typedarray_info(env, argv[3], &t, &n, &data, &ab, &off));
  size_t esize = (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
                 : (t == napi_float32_array || t == napi_int32_array || t == napi_uint32_array) ? 4
                 : (t == napi_int16_array || t == napi_uint16_array) ? 2 : 1;
  const int rc = bd_read(s, (bd_buffer)int_arg(env, argv[1], 0), int_arg(env, argv[2], 0), data, n * esize);
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

static void set_num(napi_env env, napi_value o, const char* name, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, name, x);
}

static void set_arr(napi_env env, napi_value o, const char* name, const double* v, int n) {
  napi_value a, x;
  napi_create_array_with_length(env, (size_t)n, &a);
  for (int i = 0; i < n; i++) {
    napi_create_double(env, v[i], &x);
    napi_set_element(env, a, (uint32_t)i, x);
  }
  napi_set_named_property(env, o, name, a);
}

/* stats(h) -> {newtonIterations, pcgIterations[], relResidual[], alpha[], ruptures, ...} */
static napi_value stats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  bd_solver* s;
  if (!args(env, info, 1, argv, &s)) return NULL;
  bd_stats st;
  const int rc = bd_get_stats(s, &st);
  if (rc != BD_OK) return throw_bd(env, s, rc);
  napi_value o;
  CHECK_NAPI(napi_create_object(env, &o));
  const int n = st.newton_iterations > 0 ? st.newton_iterations : 1;
  double it[BD_MAX_NEWTON];
  for (int k = 0; k < BD_MAX_NEWTON; k++) it[k] = st.pcg_iterations[k];
  set_num(env, o, "newtonIterations", st.newton_iterations);
  set_arr(env, o, "pcgIterations", it, n);
  set_arr(env, o, "relResidual", st.rel_residual, n);
  set_arr(env, o, "alpha", st.alpha, n);
  set_arr(env, o, "warmRelResidual", st.warm_relres, n);
  set_num(env, o, "rejections", st.rejections);
  set_num(env, o, "ruptures", st.ruptures);
  set_num(env, o, "pulledOut", st.pulled_out);
  set_num(env, o, "learnedFallbacks", st.learned_fallbacks);
  set_num(env, o, "maxSpeed", st.max_speed);
  double lv[BD_MAX_LEVELS];
  for (int k = 0; k < BD_MAX_LEVELS; k++) lv[k] = st.level_nodes[k];
  set_arr(env, o, "levelNodes", lv, st.level_count);
  return o;
}

/* setLearned(h, on), learnRecord(h, capacity), learnTrain(h, epochs, lr, batch, targets) -> Float64Array */
static napi_value set_learned(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  bd_solver* s;
  if (!args(env, info, 2, argv, &s)) return NULL;
  const int rc = bd_set_learned(s, int_arg(env, argv[1], 0));
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

static napi_value learn_record(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  bd_solver* s;
  if (!args(env, info, 2, argv, &s)) return NULL;
  const int rc = bd_learn_record(s, int_arg(env, argv[1], 0));
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

static napi_value learn_train(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  bd_solver* s;
  if (!args(env, info, 5, argv, &s)) return NULL;
  const int epochs = int_arg(env, argv[1], 1);
  napi_value ab, arr;
  void* data;
  CHECK_NAPI(napi_create_arraybuffer(env, sizeof(double) * (size_t)(epochs > 0 ? epochs : 1), &data, &ab));
  CHECK_NAPI(napi_create_typedarray(env, napi_float64_array, (size_t)(epochs > 0 ? epochs : 1), ab, 0, &arr));
  const int rc = bd_learn_train(s, epochs, dbl_arg(env, argv[2], 1e-3), int_arg(env, argv[3], 1), int_arg(env, argv[4], 3),
                                (double*)data);
  return rc == BD_OK ? arr : throw_bd(env, s, rc);
}

static napi_value sync_solver(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  bd_solver* s;
  if (!args(env, info, 1, argv, &s)) return NULL;
  const int rc = bd_sync(s);
  return rc == BD_OK ? undefined(env) : throw_bd(env, s, rc);
}

static napi_value init(napi_env env, napi_value exports) {
  const struct {
    const char* name;
    napi_callback fn;
  } fns[] = {{"create", create},         {"substep", substep},       {"setState", set_state},
             {"setGrips", set_grips},     {"setGripGoals", set_grip_goals}, {"read", read_buf},
             {"stats", stats},            {"setLearned", set_learned}, {"learnRecord", learn_record},
             {"learnTrain", learn_train}, {"sync", sync_solver}};
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
    napi_value f;
    napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f);
    napi_set_named_property(env, exports, fns[i].name, f);
  }
  napi_value v;
  napi_create_int32(env, BD_MAX_LINE, &v);
  napi_set_named_property(env, exports, "BD_MAX_LINE", v);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
