/* Thin N-API addon over the C ABI of include/sunaengine.h.
 *
 * Each function converts JS numbers / typed arrays to the C ABI's plain
 * pointers and sizes, calls one sn_* entry point and turns a non-zero return
 * into a thrown JS Error carrying sn_last_error() — the same "[aether] ..."
 * text the reference engine throws (src/engine.js:2174).  No state lives
 * here besides the engine handle wrapped in a JS external.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sunaengine.h"

#define CALL(env, x)                                 \
  do {                                               \
    if ((x) != napi_ok) {                            \
      napi_throw_error((env), NULL, "napi failure"); \
      return NULL;                                   \
    }                                                \
  } while (0)

static napi_value throw_sn(napi_env env, sn_engine* e, int rc) {
  const char* m = e ? sn_last_error(e) : "sn call failed";
  char buf[512];
  snprintf(buf, sizeof buf, "%s (code %d)", (m && *m) ? m : "sn call failed", rc);
  napi_throw_error(env, NULL, buf);
  return NULL;
}

static int get_args(napi_env env, napi_callback_info info, size_t n, napi_value* argv) {
  size_t argc = n;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return -1;
  return (int)argc;
}
static uint32_t u32(napi_env env, napi_value v) {
  uint32_t x = 0;
  napi_get_value_uint32(env, v, &x);
  return x;
}
static int32_t i32(napi_env env, napi_value v) {
  int32_t x = 0;
  napi_get_value_int32(env, v, &x);
  return x;
}
static sn_engine* handle(napi_env env, napi_value v) {
  void* p = NULL;
  napi_get_value_external(env, v, &p);
  return (sn_engine*)p;
}
static void* typed_data(napi_env env, napi_value v, size_t* len) {
  napi_typedarray_type t;
  size_t n = 0, off = 0;
  void* d = NULL;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &n, &d, &ab, &off) != napi_ok) return NULL;
  if (len) *len = n;
  return d;
}
static napi_value new_typed(napi_env env, napi_typedarray_type t, size_t n, size_t elt, void** data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, n * elt, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, n, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}
static napi_value num(napi_env env, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  return v;
}

/* create(maxParticles, maxCells, maxScenes, iters, streamCap, device, luts:Int32Array, mats:Int32Array
 *        [, debug = 1 [, materials = 0]]) */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value a[10];
  const int argc = get_args(env, info, 10, a);
  if (argc < 8) {
    napi_throw_error(env, NULL, "create: 8 arguments expected");
    return NULL;
  }
  sn_opts o;
  memset(&o, 0, sizeof(o));
  o.no_debug = (argc > 8 && u32(env, a[8]) == 0) ? 1u : 0u;
  o.materials = (argc > 9 && u32(env, a[9]) != 0) ? 1u : 0u;
  o.max_particles = u32(env, a[0]);
  o.max_cells = u32(env, a[1]);
  o.max_scenes = u32(env, a[2]);
  o.iters = u32(env, a[3]);
  o.stream_cap = u32(env, a[4]);
  o.device = u32(env, a[5]);
  const int32_t* luts = (const int32_t*)typed_data(env, a[6], NULL);
  const int32_t* mats = (const int32_t*)typed_data(env, a[7], NULL);
  sn_engine* e = NULL;
  int rc = sn_create(&o, luts, mats, &e);
  if (rc) {
    napi_value r = throw_sn(env, e, rc);
    if (e) sn_destroy(e);
    return r;
  }
  napi_value ext;
  CALL(env, napi_create_external(env, e, NULL, NULL, &ext));
  return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  get_args(env, info, 1, a);
  sn_destroy(handle(env, a[0]));
  return NULL;
}

/* loadScenes(h, [{n, nFluid, state:Int32Array, domW, domH, cellShift, stAcc?:Int32Array}]) */
static napi_value js_load(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t count = 0;
  napi_get_array_length(env, a[1], &count);
  sn_scene* sc = (sn_scene*)calloc(count ? count : 1, sizeof(sn_scene));
  for (uint32_t k = 0; k < count; ++k) {
    napi_value o, v;
    napi_get_element(env, a[1], k, &o);
    napi_get_named_property(env, o, "n", &v);
    sc[k].n = u32(env, v);
    napi_get_named_property(env, o, "nFluid", &v);
    sc[k].n_fluid = u32(env, v);
    napi_get_named_property(env, o, "state", &v);
    sc[k].state8 = (const int32_t*)typed_data(env, v, NULL);
    napi_get_named_property(env, o, "domW", &v);
    sc[k].dom_w = i32(env, v);
    napi_get_named_property(env, o, "domH", &v);
    sc[k].dom_h = i32(env, v);
    napi_get_named_property(env, o, "cellShift", &v);
    sc[k].cell_shift = u32(env, v);
    bool has = false;
    napi_has_named_property(env, o, "stAcc", &has);
    sc[k].st_acc = NULL;
    if (has) {
      napi_get_named_property(env, o, "stAcc", &v);
      napi_valuetype t;
      napi_typeof(env, v, &t);
      if (t == napi_object) sc[k].st_acc = (const int32_t*)typed_data(env, v, NULL);
    }
  }
  int rc = sn_load_scenes(e, sc, count);
  free(sc);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  int rc = sn_step(e, u32(env, a[1]));
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value a[1];
  get_args(env, info, 1, a);
  sn_engine* e = handle(env, a[0]);
  int rc = sn_sync(e);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

typedef int (*read8_fn)(sn_engine*, uint32_t, uint32_t*);

static napi_value read_words(napi_env env, napi_callback_info info, int which) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t scene = u32(env, a[1]);
  size_t n = which == 2 ? 16 : 8;
  void* d = NULL;
  napi_value arr = new_typed(env, napi_uint32_array, n, 4, &d);
  int rc = which == 0 ? sn_read_hash(e, scene, (uint32_t*)d)
           : which == 1 ? sn_read_chain(e, scene, (uint32_t*)d)
                        : sn_debug_counters(e, scene, (uint32_t*)d);
  if (rc) return throw_sn(env, e, rc);
  return arr;
}
static napi_value js_read_hash(napi_env env, napi_callback_info info) { return read_words(env, info, 0); }
static napi_value js_read_chain(napi_env env, napi_callback_info info) { return read_words(env, info, 1); }
static napi_value js_debug(napi_env env, napi_callback_info info) { return read_words(env, info, 2); }

static napi_value js_substep_count(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t v = 0;
  int rc = sn_substep_count(e, u32(env, a[1]), &v);
  if (rc) return throw_sn(env, e, rc);
  return num(env, v);
}

/* readHashStream(h, scene, max) -> Uint32Array(k*8) */
static napi_value js_read_stream(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t scene = u32(env, a[1]), maxs = u32(env, a[2]);
  uint32_t* tmp = (uint32_t*)malloc((size_t)(maxs ? maxs : 1) * 32);
  size_t n = 0;
  int rc = sn_read_hash_stream(e, scene, tmp, maxs, &n);
  if (rc) {
    free(tmp);
    return throw_sn(env, e, rc);
  }
  void* d = NULL;
  napi_value arr = new_typed(env, napi_uint32_array, n * 8, 4, &d);
  memcpy(d, tmp, n * 32);
  free(tmp);
  return arr;
}

static napi_value js_reset_stream(napi_env env, napi_callback_info info) {
  napi_value a[1];
  get_args(env, info, 1, a);
  sn_engine* e = handle(env, a[0]);
  int rc = sn_reset_stream(e);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

/* readState(h, scene, n) / readDerived(h, scene, n) */
static napi_value read_slots(napi_env env, napi_callback_info info, int derived) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t scene = u32(env, a[1]), n = u32(env, a[2]);
  size_t words = (size_t)n * (derived ? 16 : 8);
  void* d = NULL;
  napi_value arr = new_typed(env, napi_int32_array, words ? words : 0, 4, &d);
  if (!words) return arr;
  int rc = derived ? sn_read_derived(e, scene, (int32_t*)d, words) : sn_read_state(e, scene, (int32_t*)d, words);
  if (rc) return throw_sn(env, e, rc);
  return arr;
}
/* writeStateDerived(h, scene, state:Int32Array(n*8), derived:Int32Array(n*16) | null) */
static napi_value js_write_state_derived(napi_env env, napi_callback_info info) {
  napi_value a[4];
  const int argc = get_args(env, info, 4, a);
  sn_engine* e = handle(env, a[0]);
  size_t sw = 0, dw = 0;
  const int32_t* st = (const int32_t*)typed_data(env, a[2], &sw);
  const int32_t* dv = NULL;
  if (argc > 3) {
    napi_valuetype t;
    napi_typeof(env, a[3], &t);
    if (t == napi_object) dv = (const int32_t*)typed_data(env, a[3], &dw);
  }
  int rc = sn_write_state_derived(e, u32(env, a[1]), st, sw, dv, dw);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

/* readBond(h, scene, n) -> Uint32Array(n * 48) */
static napi_value js_read_bond(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t scene = u32(env, a[1]), n = u32(env, a[2]);
  size_t words = (size_t)n * 48;
  void* d = NULL;
  napi_value arr = new_typed(env, napi_uint32_array, words, 4, &d);
  if (!words) return arr;
  int rc = sn_read_bond(e, scene, (uint32_t*)d, words);
  if (rc) return throw_sn(env, e, rc);
  return arr;
}
static napi_value js_read_state(napi_env env, napi_callback_info info) { return read_slots(env, info, 0); }
static napi_value js_read_derived(napi_env env, napi_callback_info info) { return read_slots(env, info, 1); }

/* applyImpulse(h, scene, cx, cy, r, ix, iy, limit) -> hits */
static napi_value js_impulse(napi_env env, napi_callback_info info) {
  napi_value a[8];
  get_args(env, info, 8, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t hits = 0;
  int rc = sn_apply_impulse(e, u32(env, a[1]), i32(env, a[2]), i32(env, a[3]), i32(env, a[4]), i32(env, a[5]),
                            i32(env, a[6]), u32(env, a[7]), &hits);
  if (rc) return throw_sn(env, e, rc);
  return num(env, hits);
}

/* pushHits(h, scene) -> particles moved by stir events and applyImpulse since load */
static napi_value js_push_hits(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  uint32_t v = 0;
  int rc = sn_read_push_hits(e, u32(env, a[1]), &v);
  if (rc) return throw_sn(env, e, rc);
  return num(env, v);
}

/* spawnFluid(h, fluid8:Int32Array, flags:number) -> new n
 * (flags: 1 = {noWeld: true}, 2 = {ownGroup: true}) */
static napi_value js_spawn(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  size_t nf = 0;
  const int32_t* f = (const int32_t*)typed_data(env, a[1], &nf);
  uint32_t flags = 0;
  napi_get_value_uint32(env, a[2], &flags);
  uint32_t n = 0;
  int rc = sn_spawn_fluid(e, f, (uint32_t)(nf / 8), flags, &n);
  if (rc) return throw_sn(env, e, rc);
  return num(env, n);
}

/* setEvents(h, events:Int32Array of 8-word records) */
static napi_value js_set_events(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  size_t n = 0;
  const int32_t* w = (const int32_t*)typed_data(env, a[1], &n);
  int rc = sn_set_events(e, (const sn_event*)w, (uint32_t)(n / 8));
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

static napi_value js_set_active(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn_engine* e = handle(env, a[0]);
  size_t n = 0;
  const uint8_t* m = (const uint8_t*)typed_data(env, a[1], &n);
  int rc = sn_set_active(e, m, (uint32_t)n);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

/* ---- slotted worlds (continuous batching) and the non-blocking query ---- */

/* reserveSlots(h, capParticles:Uint32Array, capCells:Uint32Array) */
static napi_value js_reserve_slots(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  size_t n = 0, m = 0;
  const uint32_t* cp = (const uint32_t*)typed_data(env, a[1], &n);
  const uint32_t* cc = (const uint32_t*)typed_data(env, a[2], &m);
  if (n != m) {
    napi_throw_error(env, NULL, "reserveSlots: capacity arrays differ in length");
    return NULL;
  }
  int rc = sn_reserve_slots(e, (uint32_t)n, cp, cc);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

/* loadSlot(h, slot, {n, nFluid, state:Int32Array, domW, domH, cellShift}) */
static napi_value js_load_slot(napi_env env, napi_callback_info info) {
  napi_value a[3], v;
  get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  sn_scene sc;
  memset(&sc, 0, sizeof(sc));
  napi_get_named_property(env, a[2], "n", &v);
  sc.n = u32(env, v);
  napi_get_named_property(env, a[2], "nFluid", &v);
  sc.n_fluid = u32(env, v);
  napi_get_named_property(env, a[2], "state", &v);
  sc.state8 = (const int32_t*)typed_data(env, v, NULL);
  napi_get_named_property(env, a[2], "domW", &v);
  sc.dom_w = i32(env, v);
  napi_get_named_property(env, a[2], "domH", &v);
  sc.dom_h = i32(env, v);
  napi_get_named_property(env, a[2], "cellShift", &v);
  sc.cell_shift = u32(env, v);
  int rc = sn_load_slot(e, u32(env, a[1]), &sc);
  if (rc) return throw_sn(env, e, rc);
  return NULL;
}

/* query(h) -> true while enqueued work is still running */
static napi_value js_query(napi_env env, napi_callback_info info) {
  napi_value a[1], r;
  get_args(env, info, 1, a);
  sn_engine* e = handle(env, a[0]);
  int rc = sn_query(e);
  if (rc < 0) return throw_sn(env, e, rc);
  napi_get_boolean(env, rc == 1, &r);
  return r;
}

/* autotune(h, trialSubsteps, lookahead) -> {chosen: {tileCells, graphSubsteps, msPerSubstep}, trials: [...]} */
static napi_value tune_obj(napi_env env, const sn_tune_trial* t) {
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "tileCells", num(env, t->tile_cells));
  napi_set_named_property(env, o, "graphSubsteps", num(env, t->graph_substeps));
  napi_set_named_property(env, o, "msPerSubstep", num(env, t->ms_per_substep));
  return o;
}
static napi_value js_autotune(napi_env env, napi_callback_info info) {
  napi_value a[3], r, arr;
  const int argc = get_args(env, info, 3, a);
  sn_engine* e = handle(env, a[0]);
  sn_tune_trial trials[16], best;
  uint32_t n = 0;
  int rc = sn_autotune(e, u32(env, a[1]), argc > 2 ? u32(env, a[2]) : 0u, trials, 16, &n, &best);
  if (rc) return throw_sn(env, e, rc);
  napi_create_object(env, &r);
  napi_create_array_with_length(env, n < 16 ? n : 16, &arr);
  for (uint32_t i = 0; i < n && i < 16; ++i) napi_set_element(env, arr, i, tune_obj(env, &trials[i]));
  napi_set_named_property(env, r, "trials", arr);
  napi_set_named_property(env, r, "chosen", tune_obj(env, &best));
  return r;
}

/* ---- the 3D solver ---- */
static sn3_engine* handle3(napi_env env, napi_value v) {
  void* p = NULL;
  napi_get_value_external(env, v, &p);
  return (sn3_engine*)p;
}
static napi_value throw_sn3(napi_env env, sn3_engine* e, int rc) {
  const char* m = e ? sn3_last_error(e) : "sn3 call failed";
  char buf[512];
  snprintf(buf, sizeof buf, "%s (code %d)", (m && *m) ? m : "sn3 call failed", rc);
  napi_throw_error(env, NULL, buf);
  return NULL;
}
static int scene3_of(napi_env env, napi_value o, sn3_scene* sc) {
  napi_value v;
  napi_get_named_property(env, o, "n", &v);
  sc->n = u32(env, v);
  napi_get_named_property(env, o, "parts", &v);
  sc->parts16 = (const int32_t*)typed_data(env, v, NULL);
  napi_get_named_property(env, o, "gx", &v);
  sc->gx = u32(env, v);
  napi_get_named_property(env, o, "gy", &v);
  sc->gy = u32(env, v);
  napi_get_named_property(env, o, "gz", &v);
  sc->gz = u32(env, v);
  return 0;
}

/* create3(maxParticles, maxCells, maxScenes, iters, streamCap, device, luts3:Int32Array [, debug = 1]) */
static napi_value js_create3(napi_env env, napi_callback_info info) {
  napi_value a[8];
  const int argc = get_args(env, info, 8, a);
  if (argc < 7) {
    napi_throw_error(env, NULL, "create3: 7 arguments expected");
    return NULL;
  }
  sn_opts o;
  memset(&o, 0, sizeof(o));
  o.max_particles = u32(env, a[0]);
  o.max_cells = u32(env, a[1]);
  o.max_scenes = u32(env, a[2]);
  o.iters = u32(env, a[3]);
  o.stream_cap = u32(env, a[4]);
  o.device = u32(env, a[5]);
  o.no_debug = (argc > 7 && u32(env, a[7]) == 0) ? 1u : 0u;
  size_t words = 0;
  const int32_t* luts = (const int32_t*)typed_data(env, a[6], &words);
  sn3_engine* e = NULL;
  int rc = sn3_create(&o, luts, (uint32_t)words, &e);
  if (rc) {
    napi_value r = throw_sn3(env, e, rc);
    if (e) sn3_destroy(e);
    return r;
  }
  napi_value ext;
  CALL(env, napi_create_external(env, e, NULL, NULL, &ext));
  return ext;
}
static napi_value js_destroy3(napi_env env, napi_callback_info info) {
  napi_value a[1];
  get_args(env, info, 1, a);
  sn3_destroy(handle3(env, a[0]));
  return NULL;
}
/* load3(h, [{n, parts:Int32Array (n*16), gx, gy, gz}]) */
static napi_value js_load3(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn3_engine* e = handle3(env, a[0]);
  uint32_t count = 0;
  napi_get_array_length(env, a[1], &count);
  sn3_scene* sc = (sn3_scene*)calloc(count ? count : 1, sizeof(sn3_scene));
  for (uint32_t k = 0; k < count; ++k) {
    napi_value o;
    napi_get_element(env, a[1], k, &o);
    scene3_of(env, o, &sc[k]);
  }
  int rc = sn3_load(e, sc, count);
  free(sc);
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
/* reserveSlots3(h, capParticles:Uint32Array, capCells:Uint32Array) */
static napi_value js_reserve_slots3(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn3_engine* e = handle3(env, a[0]);
  size_t n = 0, m = 0;
  const uint32_t* cp = (const uint32_t*)typed_data(env, a[1], &n);
  const uint32_t* cc = (const uint32_t*)typed_data(env, a[2], &m);
  if (n != m) {
    napi_throw_error(env, NULL, "reserveSlots3: capacity arrays differ in length");
    return NULL;
  }
  int rc = sn3_reserve_slots(e, (uint32_t)n, cp, cc);
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
/* loadSlot3(h, slot, {n, parts, gx, gy, gz}) */
static napi_value js_load_slot3(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn3_engine* e = handle3(env, a[0]);
  sn3_scene sc;
  scene3_of(env, a[2], &sc);
  int rc = sn3_load_slot(e, u32(env, a[1]), &sc);
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
/* setEvents3(h, Int32Array of 10-word events sorted by (scene, substep)) */
static napi_value js_set_events3(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn3_engine* e = handle3(env, a[0]);
  size_t n = 0;
  const int32_t* w = (const int32_t*)typed_data(env, a[1], &n);
  int rc = sn3_set_events(e, (const sn3_event*)w, (uint32_t)(n / 10));
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
static napi_value js_set_active3(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn3_engine* e = handle3(env, a[0]);
  size_t n = 0;
  const uint8_t* m = (const uint8_t*)typed_data(env, a[1], &n);
  int rc = sn3_set_active(e, m, (uint32_t)n);
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
static napi_value js_step3(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn3_engine* e = handle3(env, a[0]);
  int rc = sn3_step(e, u32(env, a[1]));
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
static napi_value js_sync3(napi_env env, napi_callback_info info) {
  napi_value a[1];
  get_args(env, info, 1, a);
  sn3_engine* e = handle3(env, a[0]);
  int rc = sn3_sync(e);
  if (rc) return throw_sn3(env, e, rc);
  return NULL;
}
static napi_value js_query3(napi_env env, napi_callback_info info) {
  napi_value a[1], r;
  get_args(env, info, 1, a);
  sn3_engine* e = handle3(env, a[0]);
  int rc = sn3_query(e);
  if (rc < 0) return throw_sn3(env, e, rc);
  napi_get_boolean(env, rc == 1, &r);
  return r;
}
/* readHashStream3(h, scene, maxSubsteps) -> Uint32Array (8 words per substep) */
static napi_value js_read_stream3(napi_env env, napi_callback_info info) {
  napi_value a[3];
  get_args(env, info, 3, a);
  sn3_engine* e = handle3(env, a[0]);
  uint32_t scene = u32(env, a[1]), maxs = u32(env, a[2]);
  uint32_t* tmp = (uint32_t*)malloc((size_t)(maxs ? maxs : 1) * 32);
  size_t n = 0;
  int rc = sn3_read_stream(e, scene, tmp, maxs, &n);
  if (rc) {
    free(tmp);
    return throw_sn3(env, e, rc);
  }
  void* d = NULL;
  napi_value arr = new_typed(env, napi_uint32_array, n * 8, 4, &d);
  memcpy(d, tmp, n * 32);
  free(tmp);
  return arr;
}
/* readHash3(h, scene) -> Uint32Array(8) */
static napi_value js_read_hash3(napi_env env, napi_callback_info info) {
  napi_value a[2];
  get_args(env, info, 2, a);
  sn3_engine* e = handle3(env, a[0]);
  void* d = NULL;
  napi_value arr = new_typed(env, napi_uint32_array, 8, 4, &d);
  int rc = sn3_read_hash(e, u32(env, a[1]), (uint32_t*)d);
  if (rc) return throw_sn3(env, e, rc);
  return arr;
}

static napi_value init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"create", js_create},          {"destroy", js_destroy},         {"loadScenes", js_load},
      {"step", js_step},              {"sync", js_sync},               {"readHash", js_read_hash},
      {"readChain", js_read_chain},   {"debugCounters", js_debug},     {"substepCount", js_substep_count},
      {"readHashStream", js_read_stream}, {"resetStream", js_reset_stream}, {"readState", js_read_state},
      {"readDerived", js_read_derived}, {"readBond", js_read_bond}, {"autotune", js_autotune}, {"writeStateDerived", js_write_state_derived}, {"applyImpulse", js_impulse},  {"spawnFluid", js_spawn}, {"pushHits", js_push_hits},
      {"setEvents", js_set_events},   {"setActive", js_set_active},
      {"reserveSlots", js_reserve_slots}, {"loadSlot", js_load_slot}, {"query", js_query},
      {"create3", js_create3},        {"destroy3", js_destroy3},       {"load3", js_load3},
      {"reserveSlots3", js_reserve_slots3}, {"loadSlot3", js_load_slot3}, {"setEvents3", js_set_events3},
      {"setActive3", js_set_active3}, {"step3", js_step3},             {"sync3", js_sync3},
      {"query3", js_query3},          {"readHashStream3", js_read_stream3}, {"readHash3", js_read_hash3},
  };
  for (size_t k = 0; k < sizeof fns / sizeof fns[0]; ++k) {
    napi_value f;
    napi_create_function(env, fns[k].name, NAPI_AUTO_LENGTH, fns[k].cb, NULL, &f);
    napi_set_named_property(env, exports, fns[k].name, f);
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
