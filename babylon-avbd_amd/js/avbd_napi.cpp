// Node N-API addon over the C ABI of libavbd_hip.so (include/avbd.h): the binding the reference's JavaScript world
// API (src/native/pipeline.js, src/native/solver3D.js) calls for its "hip" backend. Each export is one C entry point;
// a non-zero status becomes a thrown Error carrying avbd_last_error(), as the reference throws on invalid use.
// Typed arrays cross without copies on the way in; reads return fresh typed arrays (the reference returns fresh
// arrays from its readbacks, solver.ts:1421-1475).
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/avbd.h"

namespace {

#define NAPI_OK(call)                                       \
  do {                                                      \
    if ((call) != napi_ok) {                                \
      napi_throw_error(env, nullptr, "N-API call failed");  \
      return nullptr;                                       \
    }                                                       \
  } while (0)

napi_value throwStatus(napi_env env, int status) {
  const char *msg = avbd_last_error();
  std::string text = std::string("avbd: ") + (msg && *msg ? msg : "error") + " (status " + std::to_string(status) + ")";
  napi_throw_error(env, nullptr, text.c_str());
  return nullptr;
}

#define AVBD_OK_OR_THROW(call)                  \
  do {                                          \
    int st_ = (call);                           \
    if (st_ != AVBD_OK) return throwStatus(env, st_); \
  } while (0)

bool args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  if (argc < want) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

avbd_world *worldOf(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a world handle");
    return nullptr;
  }
  return static_cast<avbd_world *>(p);
}

bool u32(napi_env env, napi_value v, uint32_t *out) {
  if (napi_get_value_uint32(env, v, out) != napi_ok) {
    napi_throw_type_error(env, nullptr, "expected an unsigned integer");
    return false;
  }
  return true;
}

/** Data pointer and element count of a typed array of the given type (or null/undefined when allowed). */
bool typed(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *length, bool optional) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (optional && (t == napi_null || t == napi_undefined)) {
    *data = nullptr;
    *length = 0;
    return true;
  }
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, nullptr, "expected a typed array");
    return false;
  }
  napi_typedarray_type type;
  napi_value buffer;
  size_t offset;
  if (napi_get_typedarray_info(env, v, &type, length, data, &buffer, &offset) != napi_ok || type != want) {
    napi_throw_type_error(env, nullptr, "typed array of the wrong element type");
    return false;
  }
  return true;
}

napi_value newTyped(napi_env env, napi_typedarray_type type, size_t elements, size_t elementBytes, void **data) {
  napi_value buffer, array;
  if (napi_create_arraybuffer(env, elements * elementBytes, data, &buffer) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, type, elements, buffer, 0, &array) != napi_ok) return nullptr;
  return array;
}

double numberProp(napi_env env, napi_value obj, const char *name, double fallback) {
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return fallback;
  napi_value v;
  double d = fallback;
  napi_get_named_property(env, obj, name, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_number) napi_get_value_double(env, v, &d);
  return d;
}

napi_value prop(napi_env env, napi_value obj, const char *name) {
  napi_value v = nullptr;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

// --- exports ------------------------------------------------------------------------------------------------------

napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  NAPI_OK(napi_create_string_utf8(env, avbd_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value DeviceCount(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_OK(napi_create_int32(env, avbd_device_count(), &v));
  return v;
}

/**
 * createWorld(scenes, options): scenes = [{bodies: Float32Array, joints: Float32Array, jointInfo: Int32Array,
 * ignore: Uint32Array, filters: Uint32Array | null, dt, gravity, up: [3], alpha, gamma, betaLin, betaAng, iterations,
 * flags}]; options = {colorCap, colorRounds, pairCapacity, graph, device, adaptive, broadphase (0 auto, 1 grid,
 * 2 hploc), bvhRebuildInterval, solver (0 auto, 1 colour launches, 2 dataflow, 3 workgroup per world), coloringRounds
 * (0 auto, 1 launches, 2 listed), worldBlock (threads per world's workgroup: 128, 256, 512, 1024; 0: 256), dataflowShare (the dataflow grid's share
 * of the device: 1/dataflowShare; 0: all)}. Returns
 * the world handle (avbd_batch_create).
 */
napi_value CreateWorld(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  uint32_t n = 0;
  NAPI_OK(napi_get_array_length(env, argv[0], &n));
  std::vector<AvbdSceneDesc> descs(n);
  for (uint32_t k = 0; k < n; k++) {
    napi_value s;
    NAPI_OK(napi_get_element(env, argv[0], k, &s));
    AvbdSceneDesc &d = descs[k];
    std::memset(&d, 0, sizeof d);
    void *p;
    size_t len;
    if (!typed(env, prop(env, s, "bodies"), napi_float32_array, &p, &len, false)) return nullptr;
    if (len % AVBD_BODY_FLOATS) return throwStatus(env, AVBD_ERR_INVALID);
    d.body_count = uint32_t(len / AVBD_BODY_FLOATS);
    d.bodies = static_cast<const float *>(p);
    if (!typed(env, prop(env, s, "joints"), napi_float32_array, &p, &len, true)) return nullptr;
    d.joint_count = uint32_t(len / AVBD_JOINT_FLOATS);
    d.joints = static_cast<const float *>(p);
    if (!typed(env, prop(env, s, "jointInfo"), napi_int32_array, &p, &len, true)) return nullptr;
    if (len != 4ull * d.joint_count) {
      napi_throw_range_error(env, nullptr, "jointInfo needs 4 integers per joint");
      return nullptr;
    }
    d.joint_info = static_cast<const int32_t *>(p);
    if (!typed(env, prop(env, s, "ignore"), napi_uint32_array, &p, &len, true)) return nullptr;
    d.ignore_count = uint32_t(len / 2);
    d.ignore_pairs = static_cast<const uint32_t *>(p);
    if (!typed(env, prop(env, s, "filters"), napi_uint32_array, &p, &len, true)) return nullptr;
    d.filters = static_cast<const uint32_t *>(p);
    d.dt = float(numberProp(env, s, "dt", 1.0 / 60));
    d.gravity = float(numberProp(env, s, "gravity", -10));
    napi_value up = prop(env, s, "up");
    for (uint32_t c = 0; c < 3; c++) {
      napi_value e;
      double v = c == 1 ? 1 : 0;
      if (napi_get_element(env, up, c, &e) == napi_ok) napi_get_value_double(env, e, &v);
      d.up[c] = float(v);
    }
    d.alpha = float(numberProp(env, s, "alpha", 0.99));
    d.gamma = float(numberProp(env, s, "gamma", 0.999));
    d.beta_lin = float(numberProp(env, s, "betaLin", 1e4));
    d.beta_ang = float(numberProp(env, s, "betaAng", 100));
    d.iterations = uint32_t(numberProp(env, s, "iterations", 10));
    d.flags = uint32_t(numberProp(env, s, "flags", AVBD_FLAGS_DEFAULT));
    if (!typed(env, prop(env, s, "hulls"), napi_uint32_array, &p, &len, true)) return nullptr;
    if (len % 4) {
      napi_throw_range_error(env, nullptr, "hulls needs 4 words per record word");
      return nullptr;
    }
    d.hull_words = uint32_t(len / 4);
    d.hulls = static_cast<const uint32_t *>(p);
  }
  AvbdOptions o;
  std::memset(&o, 0, sizeof o);
  o.color_cap = uint32_t(numberProp(env, argv[1], "colorCap", 0));
  o.color_rounds = uint32_t(numberProp(env, argv[1], "colorRounds", 0));
  o.pair_capacity = uint32_t(numberProp(env, argv[1], "pairCapacity", 0));
  o.use_graph = uint32_t(numberProp(env, argv[1], "graph", 0));
  o.device = int32_t(numberProp(env, argv[1], "device", 0));
  o.adaptive = uint32_t(numberProp(env, argv[1], "adaptive", 0));
  o.broadphase = uint32_t(numberProp(env, argv[1], "broadphase", 0));
  o.bvh_rebuild_interval = uint32_t(numberProp(env, argv[1], "bvhRebuildInterval", 0));
  o.solver = uint32_t(numberProp(env, argv[1], "solver", 0));
  o.coloring_rounds = uint32_t(numberProp(env, argv[1], "coloringRounds", 0));
  o.world_block = uint32_t(numberProp(env, argv[1], "worldBlock", 0));
  o.dataflow_share = uint32_t(numberProp(env, argv[1], "dataflowShare", 0));
  avbd_world *w = nullptr;
  AVBD_OK_OR_THROW(avbd_batch_create(descs.data(), n, &o, &w));
  napi_value handle;
  NAPI_OK(napi_create_external(env, w, nullptr, nullptr, &handle));
  return handle;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_destroy(w));
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t n;
  if (!w || !u32(env, argv[1], &n)) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_step(w, n));
  return nullptr;
}

napi_value StepTimed(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t n;
  if (!w || !u32(env, argv[1], &n)) return nullptr;
  void *data;
  napi_value out = newTyped(env, napi_float32_array, n, 4, &data);
  if (!out) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_step_timed(w, n, static_cast<float *>(data)));
  return out;
}

napi_value Synchronize(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_synchronize(w));
  return nullptr;
}

napi_value Counts(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  uint32_t bodies = 0, joints = 0;
  AVBD_OK_OR_THROW(avbd_world_body_count(w, &bodies));
  AVBD_OK_OR_THROW(avbd_world_joint_count(w, &joints));
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  NAPI_OK(napi_create_uint32(env, bodies, &v));
  NAPI_OK(napi_set_named_property(env, o, "bodies", v));
  NAPI_OK(napi_create_uint32(env, joints, &v));
  NAPI_OK(napi_set_named_property(env, o, "joints", v));
  return o;
}

template <int FLOATS, int (*READ)(avbd_world *, uint32_t, uint32_t, float *)>
napi_value ReadRecords(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t first, count;
  if (!w || !u32(env, argv[1], &first) || !u32(env, argv[2], &count)) return nullptr;
  void *data;
  napi_value out = newTyped(env, napi_float32_array, size_t(count) * FLOATS, 4, &data);
  if (!out) return nullptr;
  if (count) AVBD_OK_OR_THROW(READ(w, first, count, static_cast<float *>(data)));
  return out;
}

template <int FLOATS, int (*WRITE)(avbd_world *, uint32_t, uint32_t, const float *)>
napi_value WriteRecords(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t first;
  void *data;
  size_t len;
  if (!w || !u32(env, argv[1], &first) || !typed(env, argv[2], napi_float32_array, &data, &len, false)) return nullptr;
  if (len % FLOATS) return throwStatus(env, AVBD_ERR_INVALID);
  AVBD_OK_OR_THROW(WRITE(w, first, uint32_t(len / FLOATS), static_cast<const float *>(data)));
  return nullptr;
}

napi_value ReadSelected(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx;
  size_t n;
  bool posesOnly = false;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false)) return nullptr;
  napi_get_value_bool(env, argv[2], &posesOnly);
  void *data;
  napi_value out = newTyped(env, napi_float32_array, n * (posesOnly ? 8 : AVBD_BODY_FLOATS), 4, &data);
  if (!out) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_read_selected(w, static_cast<const uint32_t *>(idx), uint32_t(n), posesOnly ? 1 : 0,
                                            static_cast<float *>(data)));
  return out;
}

napi_value ReadCounters(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  AvbdCounters c;
  AVBD_OK_OR_THROW(avbd_world_read_counters(w, &c));
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  const struct {
    const char *name;
    uint32_t value;
  } fields[] = {{"pairs", c.pairs},       {"contacts", c.contacts}, {"manifolds", c.manifolds},
                {"joints", c.joints},     {"overflow", c.overflow}, {"clashes", c.clashes},
                {"colors", c.colors},     {"sleeping", 0},          {"pivotFallbacks", c.pivot_fallbacks},
                {"solveAborts", c.solve_aborts}};
  for (const auto &f : fields) {
    NAPI_OK(napi_create_uint32(env, f.value, &v));
    NAPI_OK(napi_set_named_property(env, o, f.name, v));
  }
  return o;
}

napi_value ReadPairs(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  uint32_t count = 0;
  AVBD_OK_OR_THROW(avbd_world_read_pairs(w, nullptr, 0, &count));
  void *data;
  napi_value out = newTyped(env, napi_uint32_array, 2ull * count, 4, &data);
  if (!out) return nullptr;
  uint32_t got = 0;
  if (count) AVBD_OK_OR_THROW(avbd_world_read_pairs(w, static_cast<uint32_t *>(data), count, &got));
  return out;
}

napi_value ReadContacts(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  uint32_t m = 0, c = 0;
  AVBD_OK_OR_THROW(avbd_world_read_contacts(w, nullptr, 0, &m, nullptr, 0, &c));
  void *md, *cd;
  napi_value mo = newTyped(env, napi_uint32_array, size_t(m) * AVBD_MANIFOLD_WORDS, 4, &md);
  napi_value co = newTyped(env, napi_uint32_array, size_t(c) * AVBD_CONTACT_WORDS, 4, &cd);
  if (!mo || !co) return nullptr;
  uint32_t gm = 0, gc = 0;
  AVBD_OK_OR_THROW(avbd_world_read_contacts(w, static_cast<uint32_t *>(md), m, &gm, static_cast<uint32_t *>(cd), c, &gc));
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  NAPI_OK(napi_set_named_property(env, o, "manifolds", mo));
  NAPI_OK(napi_set_named_property(env, o, "contacts", co));
  return o;
}

napi_value Adapt(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  int changed = 0;
  AVBD_OK_OR_THROW(avbd_world_adapt(w, &changed));
  napi_value v;
  NAPI_OK(napi_get_boolean(env, changed != 0, &v));
  return v;
}

napi_value SetColoring(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t cap, rounds;
  if (!w || !u32(env, argv[1], &cap) || !u32(env, argv[2], &rounds)) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_set_coloring(w, cap, rounds));
  return nullptr;
}

napi_value DeviceBytes(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  uint64_t bytes = 0;
  AVBD_OK_OR_THROW(avbd_world_device_bytes(w, &bytes));
  napi_value v;
  NAPI_OK(napi_create_double(env, double(bytes), &v));
  return v;
}

napi_value EnqueueCommands(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *data;
  size_t length;
  if (!w || !typed(env, argv[1], napi_uint32_array, &data, &length, false)) return nullptr;
  if (length % AVBD_COMMAND_WORDS) {
    napi_throw_range_error(env, nullptr, "commands are 12 words each");
    return nullptr;
  }
  AVBD_OK_OR_THROW(avbd_world_enqueue_commands(w, static_cast<const uint32_t *>(data),
                                               uint32_t(length / AVBD_COMMAND_WORDS)));
  return nullptr;
}

napi_value SetFilters(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!args(env, info, 4, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx, *groups, *masks;
  size_t n, ng, nm;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false) ||
      !typed(env, argv[2], napi_uint32_array, &groups, &ng, false) ||
      !typed(env, argv[3], napi_uint32_array, &masks, &nm, false))
    return nullptr;
  if (ng != n || nm != n) {
    napi_throw_range_error(env, nullptr, "Each body index needs a group and collision mask");
    return nullptr;
  }
  AVBD_OK_OR_THROW(avbd_world_set_filters(w, static_cast<const uint32_t *>(idx), static_cast<const uint32_t *>(groups),
                                          static_cast<const uint32_t *>(masks), uint32_t(n)));
  return nullptr;
}

napi_value SetRestitution(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx, *values;
  size_t n, nv;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false) ||
      !typed(env, argv[2], napi_float32_array, &values, &nv, false))
    return nullptr;
  if (nv != n) {
    napi_throw_range_error(env, nullptr, "Each body index needs a restitution value");
    return nullptr;
  }
  AVBD_OK_OR_THROW(avbd_world_set_restitution(w, static_cast<const uint32_t *>(idx), static_cast<const float *>(values),
                                              uint32_t(n)));
  return nullptr;
}

napi_value GetRestitution(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t n;
  if (!w || !u32(env, argv[1], &n)) return nullptr;
  void *data;
  napi_value out = newTyped(env, napi_float32_array, n, 4, &data);
  if (!out) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_get_restitution(w, static_cast<float *>(data)));
  return out;
}

napi_value EnableSleeping(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!args(env, info, 5, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  double speed, delay;
  int32_t islands;
  bool start;
  if (!w) return nullptr;
  if (napi_get_value_double(env, argv[1], &speed) != napi_ok || napi_get_value_double(env, argv[2], &delay) != napi_ok ||
      napi_get_value_int32(env, argv[3], &islands) != napi_ok || napi_get_value_bool(env, argv[4], &start) != napi_ok) {
    napi_throw_type_error(env, nullptr, "expected speed, delay, islands and startAsleep");
    return nullptr;
  }
  AVBD_OK_OR_THROW(avbd_world_enable_sleeping(w, float(speed), float(delay), islands, start ? 1 : 0));
  return nullptr;
}

template <int (*F)(avbd_world *)>
napi_value WorldCall(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  AVBD_OK_OR_THROW(F(w));
  return nullptr;
}

/** {enabled, islands} of the world's sleeping. */
napi_value SleepMode(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  if (!w) return nullptr;
  int islands = 0;
  const int on = avbd_world_sleeping(w, &islands);
  napi_value out, a, b;
  napi_create_object(env, &out);
  napi_get_boolean(env, on == 1, &a);
  napi_get_boolean(env, on == 1 && islands == 1, &b);
  napi_set_named_property(env, out, "enabled", a);
  napi_set_named_property(env, out, "islands", b);
  return out;
}

napi_value WakeBodies(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx;
  size_t n;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false)) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_wake_bodies(w, static_cast<const uint32_t *>(idx), uint32_t(n)));
  return nullptr;
}

napi_value SetSleepEnabled(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx, *flags;
  size_t n, nf;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false) ||
      !typed(env, argv[2], napi_uint8_array, &flags, &nf, false))
    return nullptr;
  if (nf != n) {
    napi_throw_range_error(env, nullptr, "Each body index needs a sleep flag");
    return nullptr;
  }
  AVBD_OK_OR_THROW(avbd_world_set_sleep_enabled(w, static_cast<const uint32_t *>(idx),
                                                static_cast<const uint8_t *>(flags), uint32_t(n)));
  return nullptr;
}

/** {sleeping, wakeRequested} of one world of the batch after the last step. */
napi_value SleepStats(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t k;
  if (!w || !u32(env, argv[1], &k)) return nullptr;
  uint32_t s[2];
  AVBD_OK_OR_THROW(avbd_world_sleep_stats(w, k, s));
  napi_value out, a, b;
  napi_create_object(env, &out);
  napi_create_uint32(env, s[0], &a);
  napi_get_boolean(env, s[1] != 0, &b);
  napi_set_named_property(env, out, "sleeping", a);
  napi_set_named_property(env, out, "wakeRequested", b);
  return out;
}

/** cast(world, queries: Float32Array (8 per query), ignore: Uint32Array, collidesWith, membership (-1: none)) ->
 * Float32Array hits (8 per query). */
napi_value Cast(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!args(env, info, 5, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *q, *ign;
  size_t nq, ni;
  uint32_t mask;
  int64_t membership;
  if (!w || !typed(env, argv[1], napi_float32_array, &q, &nq, false) ||
      !typed(env, argv[2], napi_uint32_array, &ign, &ni, false) || !u32(env, argv[3], &mask))
    return nullptr;
  if (napi_get_value_int64(env, argv[4], &membership) != napi_ok) {
    napi_throw_type_error(env, nullptr, "expected a membership mask or -1");
    return nullptr;
  }
  const uint32_t n = uint32_t(nq / 8);
  void *data;
  napi_value out = newTyped(env, napi_float32_array, size_t(n) * 8, 4, &data);
  if (!out) return nullptr;
  AvbdQueryOptions o{static_cast<const uint32_t *>(ign), uint32_t(ni), mask,
                     membership < 0 ? 0u : uint32_t(membership), membership < 0 ? 0 : 1};
  AVBD_OK_OR_THROW(avbd_world_cast(w, static_cast<const float *>(q), n, &o, static_cast<float *>(data)));
  return out;
}

napi_value WatchContacts(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t pairs, events;
  if (!w || !u32(env, argv[1], &pairs) || !u32(env, argv[2], &events)) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_watch_contacts(w, pairs, events));
  return nullptr;
}

napi_value SetWatched(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx, *flags;
  size_t n, nf;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false) ||
      !typed(env, argv[2], napi_uint8_array, &flags, &nf, false))
    return nullptr;
  if (nf != n) {
    napi_throw_range_error(env, nullptr, "Each watched index needs a flag");
    return nullptr;
  }
  AVBD_OK_OR_THROW(avbd_world_set_watched(w, static_cast<const uint32_t *>(idx), static_cast<const uint8_t *>(flags),
                                          uint32_t(n)));
  return nullptr;
}

napi_value ResetContactHistory(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  void *idx;
  size_t n;
  if (!w || !typed(env, argv[1], napi_uint32_array, &idx, &n, false)) return nullptr;
  AVBD_OK_OR_THROW(avbd_world_reset_contact_history(w, static_cast<const uint32_t *>(idx), uint32_t(n)));
  return nullptr;
}

/** readContactEvents(world, capacity) -> {words: Uint32Array (12 per event), written, full}. */
napi_value ReadContactEvents(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t cap;
  if (!w || !u32(env, argv[1], &cap)) return nullptr;
  std::vector<uint32_t> buf(size_t(cap) * 12);
  uint32_t count = 0, written = 0, full = 0;
  AVBD_OK_OR_THROW(avbd_world_read_contact_events(w, buf.data(), cap, &count, &written, &full));
  void *data;
  napi_value words = newTyped(env, napi_uint32_array, size_t(count) * 12, 4, &data);
  if (!words) return nullptr;
  std::memcpy(data, buf.data(), size_t(count) * 48);
  napi_value out, a, b;
  napi_create_object(env, &out);
  napi_create_uint32(env, written, &a);
  napi_get_boolean(env, full != 0, &b);
  napi_set_named_property(env, out, "words", words);
  napi_set_named_property(env, out, "written", a);
  napi_set_named_property(env, out, "full", b);
  return out;
}

napi_value GetParams(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t k;
  if (!w || !u32(env, argv[1], &k)) return nullptr;
  AvbdParams p{};
  AVBD_OK_OR_THROW(avbd_world_get_params(w, k, &p));
  napi_value o, v, up;
  NAPI_OK(napi_create_object(env, &o));
  const struct {
    const char *name;
    double value;
  } fields[] = {{"dt", p.dt},       {"gravity", p.gravity},   {"alpha", p.alpha},
                {"gamma", p.gamma}, {"betaLin", p.beta_lin},  {"betaAng", p.beta_ang},
                {"iterations", double(p.iterations)}, {"flags", double(p.flags)}};
  for (const auto &f : fields) {
    NAPI_OK(napi_create_double(env, f.value, &v));
    NAPI_OK(napi_set_named_property(env, o, f.name, v));
  }
  NAPI_OK(napi_create_array_with_length(env, 3, &up));
  for (uint32_t i = 0; i < 3; i++) {
    NAPI_OK(napi_create_double(env, p.up[i], &v));
    NAPI_OK(napi_set_element(env, up, i, v));
  }
  NAPI_OK(napi_set_named_property(env, o, "up", up));
  return o;
}

/** setParams(handle, world, {dt, gravity, up, alpha, gamma, betaLin, betaAng, iterations, flags}); missing fields keep
 * their current values. */
napi_value SetParams(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  avbd_world *w = worldOf(env, argv[0]);
  uint32_t k;
  if (!w || !u32(env, argv[1], &k)) return nullptr;
  AvbdParams p{};
  AVBD_OK_OR_THROW(avbd_world_get_params(w, k, &p));
  p.dt = float(numberProp(env, argv[2], "dt", p.dt));
  p.gravity = float(numberProp(env, argv[2], "gravity", p.gravity));
  p.alpha = float(numberProp(env, argv[2], "alpha", p.alpha));
  p.gamma = float(numberProp(env, argv[2], "gamma", p.gamma));
  p.beta_lin = float(numberProp(env, argv[2], "betaLin", p.beta_lin));
  p.beta_ang = float(numberProp(env, argv[2], "betaAng", p.beta_ang));
  p.iterations = uint32_t(numberProp(env, argv[2], "iterations", p.iterations));
  p.flags = uint32_t(numberProp(env, argv[2], "flags", p.flags));
  napi_value up = prop(env, argv[2], "up");
  bool isArray = false;
  if (up) napi_is_array(env, up, &isArray);
  if (isArray) {
    for (uint32_t i = 0; i < 3; i++) {
      napi_value e;
      double d = p.up[i];
      if (napi_get_element(env, up, i, &e) == napi_ok) napi_get_value_double(env, e, &d);
      p.up[i] = float(d);
    }
  }
  AVBD_OK_OR_THROW(avbd_world_set_params(w, k, &p));
  return nullptr;
}

napi_value setNumber(napi_env env, napi_value obj, const char *name, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, name, x);
  return obj;
}
napi_value setNumbers(napi_env env, napi_value obj, const char *name, const double *v, uint32_t n) {
  napi_value a;
  napi_create_array_with_length(env, n, &a);
  for (uint32_t k = 0; k < n; k++) {
    napi_value x;
    napi_create_double(env, v[k], &x);
    napi_set_element(env, a, k, x);
  }
  napi_set_named_property(env, obj, name, a);
  return obj;
}

/** A cooked hull as {record: Uint32Array (header at word 0), volume, moments, size, radius, center, rotation}, or null
 * when the input makes no hull (the reference's convexHull / hullFromTriangles return null). */
napi_value hullResult(napi_env env, int code, const std::vector<uint32_t> &record, const AvbdHullInfo &info) {
  napi_value out;
  if (code != AVBD_OK) {
    napi_get_null(env, &out);
    return out;
  }
  void *data;
  napi_value rec = newTyped(env, napi_uint32_array, 4ull * info.words, 4, &data);
  std::memcpy(data, record.data(), 16ull * info.words);
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "record", rec);
  setNumber(env, out, "volume", info.volume);
  setNumbers(env, out, "moments", info.moments, 3);
  setNumbers(env, out, "size", info.size, 3);
  setNumber(env, out, "radius", info.radius);
  setNumbers(env, out, "center", info.center, 3);
  setNumbers(env, out, "rotation", info.rotation, 4);
  return out;
}

/** hullFromPoints(Float64Array xyz) (avbd_hull_from_points). */
napi_value HullFromPoints(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  void *p;
  size_t len;
  if (!typed(env, argv[0], napi_float64_array, &p, &len, false)) return nullptr;
  std::vector<uint32_t> record(4ull * AVBD_HULL_MAX_WORDS);
  AvbdHullInfo hi;
  const int code = avbd_hull_from_points(static_cast<const double *>(p), uint32_t(len / 3), 0, record.data(),
                                         AVBD_HULL_MAX_WORDS, &hi);
  return hullResult(env, code, record, hi);
}

/** hullFromTriangles(Float64Array xyz, Uint32Array indices) (avbd_hull_from_triangles). */
napi_value HullFromTriangles(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  void *p, *q;
  size_t len, ni;
  if (!typed(env, argv[0], napi_float64_array, &p, &len, false)) return nullptr;
  if (!typed(env, argv[1], napi_uint32_array, &q, &ni, false)) return nullptr;
  std::vector<uint32_t> record(4ull * AVBD_HULL_MAX_WORDS);
  AvbdHullInfo hi;
  const int code = avbd_hull_from_triangles(static_cast<const double *>(p), uint32_t(len / 3),
                                            static_cast<const uint32_t *>(q), uint32_t(ni), 0, record.data(),
                                            AVBD_HULL_MAX_WORDS, &hi);
  return hullResult(env, code, record, hi);
}

napi_value Init(napi_env env, napi_value exports) {
  const napi_property_descriptor props[] = {
      {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"deviceCount", nullptr, DeviceCount, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"createWorld", nullptr, CreateWorld, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"step", nullptr, Step, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"stepTimed", nullptr, StepTimed, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"synchronize", nullptr, Synchronize, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"counts", nullptr, Counts, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readBodies", nullptr, ReadRecords<AVBD_BODY_FLOATS, avbd_world_read_bodies>, nullptr, nullptr, nullptr,
       napi_enumerable, nullptr},
      {"writeBodies", nullptr, WriteRecords<AVBD_BODY_FLOATS, avbd_world_write_bodies>, nullptr, nullptr, nullptr,
       napi_enumerable, nullptr},
      {"readJoints", nullptr, ReadRecords<AVBD_JOINT_FLOATS, avbd_world_read_joints>, nullptr, nullptr, nullptr,
       napi_enumerable, nullptr},
      {"writeJoints", nullptr, WriteRecords<AVBD_JOINT_FLOATS, avbd_world_write_joints>, nullptr, nullptr, nullptr,
       napi_enumerable, nullptr},
      {"readSelected", nullptr, ReadSelected, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readCounters", nullptr, ReadCounters, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readPairs", nullptr, ReadPairs, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readContacts", nullptr, ReadContacts, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"adapt", nullptr, Adapt, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setColoring", nullptr, SetColoring, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"deviceBytes", nullptr, DeviceBytes, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"enqueueCommands", nullptr, EnqueueCommands, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setFilters", nullptr, SetFilters, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setRestitution", nullptr, SetRestitution, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"getRestitution", nullptr, GetRestitution, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"enableSleeping", nullptr, EnableSleeping, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"disableSleeping", nullptr, WorldCall<avbd_world_disable_sleeping>, nullptr, nullptr, nullptr, napi_enumerable,
       nullptr},
      {"wakeAll", nullptr, WorldCall<avbd_world_wake_all>, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"sleepMode", nullptr, SleepMode, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"wakeBodies", nullptr, WakeBodies, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setSleepEnabled", nullptr, SetSleepEnabled, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"sleepStats", nullptr, SleepStats, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cast", nullptr, Cast, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"watchContacts", nullptr, WatchContacts, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setWatched", nullptr, SetWatched, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"resetContactHistory", nullptr, ResetContactHistory, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readContactEvents", nullptr, ReadContactEvents, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"getParams", nullptr, GetParams, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setParams", nullptr, SetParams, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"hullFromPoints", nullptr, HullFromPoints, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"hullFromTriangles", nullptr, HullFromTriangles, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
  };
  if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return nullptr;
  return exports;
}

}  // namespace

NAPI_MODULE(avbd_napi, Init)
