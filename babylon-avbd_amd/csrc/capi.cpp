// extern "C" entry points of libavbd_hip.so (declared in include/avbd.h) over the host engine.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/avbd.h"
#include "engine.h"
#include "hull_cook.h"
#include "server.h"

using namespace avbd;

struct avbd_world {
  std::unique_ptr<Engine> engine;
};

struct avbd_server {
  std::unique_ptr<Server> server;
};

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

template <class F> int guarded(F &&f) {
  try {
    g_error.clear();
    return f();
  } catch (const std::invalid_argument &e) {
    return fail(AVBD_ERR_INVALID, e.what());
  } catch (const std::out_of_range &e) {
    return fail(AVBD_ERR_INVALID, e.what());
  } catch (const std::length_error &e) {
    return fail(AVBD_ERR_CAPACITY, e.what());
  } catch (const std::bad_alloc &) {
    return fail(AVBD_ERR_DEVICE, "out of memory");
  } catch (const std::exception &e) {
    return fail(AVBD_ERR_DEVICE, e.what());
  }
}

HostWorld toHost(const AvbdSceneDesc &s) {
  if (s.body_count && !s.bodies) throw std::invalid_argument("bodies is NULL");
  if (s.joint_count && (!s.joints || !s.joint_info)) throw std::invalid_argument("joints or joint_info is NULL");
  if (s.ignore_count && !s.ignore_pairs) throw std::invalid_argument("ignore_pairs is NULL");
  if (!(s.dt > 0.f)) throw std::invalid_argument("timeStep must be positive");
  if (s.iterations == 0) throw std::invalid_argument("iterations must be a positive integer");
  float upLen = std::sqrt(s.up[0] * s.up[0] + s.up[1] * s.up[1] + s.up[2] * s.up[2]);
  if (std::fabs(upLen - 1.f) > 1e-5f) throw std::invalid_argument("up must be a unit vector");
  HostWorld w;
  w.bodies.resize(s.body_count);
  if (s.body_count) std::memcpy(w.bodies.data(), s.bodies, size_t(s.body_count) * sizeof(Body));
  w.joints.resize(s.joint_count);
  w.info.resize(s.joint_count);
  if (s.joint_count) {
    std::memcpy(w.joints.data(), s.joints, size_t(s.joint_count) * sizeof(Joint));
    std::memcpy(w.info.data(), s.joint_info, size_t(s.joint_count) * sizeof(int4));
  }
  for (uint32_t k = 0; k < s.ignore_count; k++) w.ignorePairs.emplace_back(s.ignore_pairs[2 * k], s.ignore_pairs[2 * k + 1]);
  if (s.hull_words && !s.hulls) throw std::invalid_argument("hulls is NULL");
  w.hulls.resize(s.hull_words);
  if (s.hull_words) std::memcpy(w.hulls.data(), s.hulls, size_t(s.hull_words) * sizeof(uint4));
  if (s.filters) {
    w.filters.resize(s.body_count);
    for (uint32_t i = 0; i < s.body_count; i++) w.filters[i] = make_uint2(s.filters[2 * i], s.filters[2 * i + 1]);
  }
  WorldParams p{};
  p.dt = s.dt;
  p.gravity = s.gravity;
  p.betaLin = s.beta_lin;
  p.betaAng = s.beta_ang;
  p.gamma = s.gamma;
  p.alpha = s.alpha;
  p.up[0] = s.up[0];
  p.up[1] = s.up[1];
  p.up[2] = s.up[2];
  p.flags = s.flags;
  p.step = 0;
  w.params = p;
  w.iterations = s.iterations;
  return w;
}

EngineOptions toOptions(const AvbdOptions *o) {
  EngineOptions e;
  if (o) {
    if (o->color_cap) e.colorCap = o->color_cap;
    if (o->color_rounds) e.colorRounds = o->color_rounds;
    e.pairCapacity = o->pair_capacity;
    e.useGraph = o->use_graph != 0;
    e.adaptive = o->adaptive != 0;
    if (o->broadphase > 2) throw std::invalid_argument("broadphase must be 0 (auto), 1 (grid) or 2 (hploc)");
    e.broadphase = o->broadphase;
    if (o->bvh_rebuild_interval) e.rebuildInterval = o->bvh_rebuild_interval;
    if (o->solver > 3)
      throw std::invalid_argument("solver must be 0 (auto), 1 (colour launches), 2 (dataflow) or 3 (workgroup per world)");
    e.solver = o->solver;
    if (o->world_block != 0 && o->world_block != 128 && o->world_block != 256 && o->world_block != 512 &&
        o->world_block != 1024)
      throw std::invalid_argument("world_block must be 0, 128, 256, 512 or 1024");
    e.worldBlock = o->world_block;
    if (o->dataflow_share > 1024) throw std::invalid_argument("dataflow_share must be at most 1024");
    e.dataflowShare = o->dataflow_share;
    if (o->coloring_rounds > 2)
      throw std::invalid_argument("coloring_rounds must be 0 (auto), 1 (one launch per round) or 2 (one block)");
    e.coloringRounds = o->coloring_rounds;
    e.device = o->device;
  }
  if (e.colorCap > 64) throw std::invalid_argument("color_cap must be at most 64");
  return e;
}

#define CHECK_WORLD(w)                                                              \
  if (!(w) || !(w)->engine) return fail(AVBD_ERR_DESTROYED, "the world has been destroyed")

}  // namespace

extern "C" {

const char *avbd_last_error(void) { return g_error.c_str(); }
const char *avbd_version(void) { return "avbd-mi355x 0.1.0 (gfx950)"; }

int avbd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int avbd_batch_create(const AvbdSceneDesc *scenes, uint32_t n, const AvbdOptions *options, avbd_world **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    *out = nullptr;
    if (!scenes || n == 0) throw std::invalid_argument("a batch needs at least one scene");
    int devices = avbd_device_count();
    if (devices <= 0) return fail(AVBD_ERR_DEVICE, "no HIP device is available");
    std::vector<HostWorld> worlds;
    worlds.reserve(n);
    for (uint32_t k = 0; k < n; k++) worlds.push_back(toHost(scenes[k]));
    auto w = std::make_unique<avbd_world>();
    w->engine = std::make_unique<Engine>(worlds, toOptions(options));
    *out = w.release();
    return AVBD_OK;
  });
}

int avbd_world_create(const AvbdSceneDesc *scene, const AvbdOptions *options, avbd_world **out) {
  return avbd_batch_create(scene, 1, options, out);
}

int avbd_world_destroy(avbd_world *world) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine.reset();
    delete world;
    return AVBD_OK;
  });
}

int avbd_world_step(avbd_world *world, uint32_t n) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->step(n);
    return AVBD_OK;
  });
}

int avbd_world_step_timed(avbd_world *world, uint32_t n, float *ms) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (n && !ms) throw std::invalid_argument("ms is NULL");
    if (n) world->engine->stepTimed(n, ms);
    return AVBD_OK;
  });
}

int avbd_world_step_timed_begin(avbd_world *world, uint32_t n) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (n > 4096) throw std::invalid_argument("at most 4096 timed steps at once");
    if (world->engine->timedPending()) throw std::invalid_argument("timed steps already pending: end them first");
    world->engine->stepTimedBegin(n);
    return AVBD_OK;
  });
}

int avbd_world_step_timed_end(avbd_world *world, float *ms, uint32_t capacity, uint32_t *count) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!count) throw std::invalid_argument("count is NULL");
    const uint32_t n = world->engine->timedPending();
    if (n > capacity) throw std::length_error("ms capacity below the number of timed steps");
    if (n && !ms) throw std::invalid_argument("ms is NULL");
    *count = world->engine->stepTimedEnd(ms);
    return AVBD_OK;
  });
}

int avbd_world_synchronize(avbd_world *world) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->synchronize();
    return AVBD_OK;
  });
}

int avbd_world_body_count(const avbd_world *world, uint32_t *out) {
  if (!world || !world->engine) return fail(AVBD_ERR_DESTROYED, "the world has been destroyed");
  if (!out) return fail(AVBD_ERR_INVALID, "out is NULL");
  *out = world->engine->bodyCount();
  return AVBD_OK;
}

int avbd_world_joint_count(const avbd_world *world, uint32_t *out) {
  if (!world || !world->engine) return fail(AVBD_ERR_DESTROYED, "the world has been destroyed");
  if (!out) return fail(AVBD_ERR_INVALID, "out is NULL");
  *out = world->engine->jointCount();
  return AVBD_OK;
}

int avbd_world_read_bodies(avbd_world *world, uint32_t first, uint32_t count, float *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out && count) throw std::invalid_argument("out is NULL");
    world->engine->readBodies(first, count, reinterpret_cast<Body *>(out));
    return AVBD_OK;
  });
}

int avbd_world_write_bodies(avbd_world *world, uint32_t first, uint32_t count, const float *in) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!in && count) throw std::invalid_argument("in is NULL");
    world->engine->writeBodies(first, count, reinterpret_cast<const Body *>(in));
    return AVBD_OK;
  });
}

int avbd_world_read_selected(avbd_world *world, const uint32_t *indices, uint32_t n, int poses_only, float *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (n && (!indices || !out)) throw std::invalid_argument("indices or out is NULL");
    Engine &e = *world->engine;
    for (uint32_t k = 0; k < n; k++)
      if (indices[k] >= e.bodyCount()) throw std::out_of_range("body index out of range");
    std::vector<Body> all(e.bodyCount());
    e.readBodies(0, e.bodyCount(), all.data());
    for (uint32_t k = 0; k < n; k++) {
      const Body &b = all[indices[k]];
      if (poses_only) {
        std::memcpy(out + 8 * size_t(k), &b.pos, 16);
        std::memcpy(out + 8 * size_t(k) + 4, &b.rot, 16);
      } else {
        std::memcpy(out + 40 * size_t(k), &b, sizeof(Body));
      }
    }
    return AVBD_OK;
  });
}

int avbd_world_read_joints(avbd_world *world, uint32_t first, uint32_t count, float *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out && count) throw std::invalid_argument("out is NULL");
    world->engine->readJoints(first, count, reinterpret_cast<Joint *>(out));
    return AVBD_OK;
  });
}

int avbd_world_write_joints(avbd_world *world, uint32_t first, uint32_t count, const float *in) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!in && count) throw std::invalid_argument("in is NULL");
    world->engine->writeJoints(first, count, reinterpret_cast<const Joint *>(in));
    return AVBD_OK;
  });
}

int avbd_world_read_counters(avbd_world *world, AvbdCounters *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out) throw std::invalid_argument("out is NULL");
    Counters c = world->engine->readCounters();
    out->pairs = c.pairs;
    out->contacts = c.contacts;
    out->manifolds = c.manifolds;
    out->overflow = ((c.overflow & 1u) ? 1u : 0u) | ((c.overflow & 2u) ? 2u : 0u) | ((c.overflow & 4u) ? 4u : 0u);
    out->clashes = c.clashes;
    out->colors = c.colors;
    out->joints = world->engine->jointCount();
    out->bodies = world->engine->bodyCount();
    out->pivot_fallbacks = c.pivot64;
    out->solve_aborts = c.solveAborts;
    return AVBD_OK;
  });
}

int avbd_world_read_pairs(avbd_world *world, uint32_t *out, uint32_t capacity, uint32_t *count) {
  return guarded([&] {
    CHECK_WORLD(world);
    std::vector<uint2> pairs = world->engine->readPairs();
    if (count) *count = uint32_t(pairs.size());
    if (out) {
      uint32_t n = std::min<uint32_t>(capacity, uint32_t(pairs.size()));
      std::memcpy(out, pairs.data(), size_t(n) * 8);
    }
    return AVBD_OK;
  });
}

int avbd_world_read_contacts(avbd_world *world, uint32_t *manifolds, uint32_t manifold_capacity, uint32_t *manifold_count,
                             uint32_t *contacts, uint32_t contact_capacity, uint32_t *contact_count) {
  return guarded([&] {
    CHECK_WORLD(world);
    std::vector<Manifold> m;
    std::vector<Contact> c;
    world->engine->readContacts(m, c);
    if (manifold_count) *manifold_count = uint32_t(m.size());
    if (contact_count) *contact_count = uint32_t(c.size());
    if (manifolds) std::memcpy(manifolds, m.data(), std::min<size_t>(manifold_capacity, m.size()) * sizeof(Manifold));
    if (contacts) std::memcpy(contacts, c.data(), std::min<size_t>(contact_capacity, c.size()) * sizeof(Contact));
    return AVBD_OK;
  });
}

int avbd_world_read_colors(avbd_world *world, uint32_t *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out) throw std::invalid_argument("out is NULL");
    std::vector<uint32_t> c = world->engine->readColors();
    std::memcpy(out, c.data(), c.size() * 4);
    return AVBD_OK;
  });
}

int avbd_world_read_adjacency(avbd_world *world, uint32_t *start, uint32_t *list, uint32_t capacity, uint32_t *count) {
  return guarded([&] {
    CHECK_WORLD(world);
    std::vector<uint32_t> s, l;
    world->engine->readAdjacency(s, l);
    if (count) *count = uint32_t(l.size());
    if (start) std::memcpy(start, s.data(), s.size() * 4);
    if (list) std::memcpy(list, l.data(), std::min<size_t>(capacity, l.size()) * 4);
    return AVBD_OK;
  });
}

int avbd_world_read_refs(avbd_world *world, float *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out) throw std::invalid_argument("out is NULL");
    std::vector<float4> r = world->engine->readRefs();
    std::memcpy(out, r.data(), r.size() * sizeof(float4));
    return AVBD_OK;
  });
}

int avbd_world_read_steps(avbd_world *world, uint32_t *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out) throw std::invalid_argument("out is NULL");
    std::vector<WorldParams> p = world->engine->readWorldParams();
    for (size_t k = 0; k < p.size(); k++) out[k] = p[k].step;
    return AVBD_OK;
  });
}

int avbd_world_write_state(avbd_world *world, const float *refs, const uint32_t *colors, uint32_t manifold_count,
                           const uint32_t *manifolds, uint32_t contact_count, const uint32_t *contacts,
                           const uint32_t *steps) {
  return guarded([&] {
    CHECK_WORLD(world);
    if ((manifold_count && !manifolds) || (contact_count && !contacts))
      throw std::invalid_argument("manifolds or contacts is NULL");
    std::vector<Manifold> m(manifold_count);
    std::vector<Contact> c(contact_count);
    if (manifold_count) std::memcpy(m.data(), manifolds, size_t(manifold_count) * sizeof(Manifold));
    if (contact_count) std::memcpy(c.data(), contacts, size_t(contact_count) * sizeof(Contact));
    uint32_t n = world->engine->bodyCount();
    for (const Manifold &mf : m)
      if (mf.ids.x >= n || mf.ids.y >= n || mf.ids.x <= mf.ids.y || (mf.ids.w & 15u) > 8u)
        throw std::invalid_argument("manifold records must name bodies a > b with at most 8 contacts");
    world->engine->writeState(reinterpret_cast<const float4 *>(refs), colors, m, c, steps);
    return AVBD_OK;
  });
}

int avbd_world_set_coloring(avbd_world *world, uint32_t color_cap, uint32_t color_rounds) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (color_cap == 0 || color_cap > 64) throw std::invalid_argument("color_cap must be in 1..64");
    world->engine->setColorCap(color_cap);
    if (color_rounds) world->engine->setColorRounds(color_rounds);
    return AVBD_OK;
  });
}

int avbd_world_get_coloring(avbd_world *world, uint32_t *color_cap, uint32_t *color_rounds, uint32_t *pair_capacity) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (color_cap) *color_cap = world->engine->colorCap();
    if (color_rounds) *color_rounds = world->engine->colorRounds();
    if (pair_capacity) *pair_capacity = world->engine->pairCapacity();
    return AVBD_OK;
  });
}

int avbd_world_release_joints(avbd_world *world, const uint32_t *slots, uint32_t n) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->releaseJoints(slots, n);
    return AVBD_OK;
  });
}

int avbd_world_disable_constraint(avbd_world *world, uint32_t slot) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->disableConstraint(slot);
    return AVBD_OK;
  });
}

int avbd_world_set_world_anchor(avbd_world *world, uint32_t slot, const float *point) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->setWorldAnchor(slot, point);
    return AVBD_OK;
  });
}

int avbd_world_set_motor(avbd_world *world, uint32_t slot, const float *speed, const float *max_torque) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->setMotor(slot, speed, max_torque);
    return AVBD_OK;
  });
}

int avbd_world_set_spring_material(avbd_world *world, uint32_t slot, float break_strain, float break_force) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->setSpringMaterial(slot, break_strain, break_force);
    return AVBD_OK;
  });
}

int avbd_world_set_angular_constraint(avbd_world *world, uint32_t slot, int32_t type, const float *axis_a,
                                      const float *axis_b, float lo, float hi) {
  return guarded([&] {
    CHECK_WORLD(world);
    world->engine->setAngularConstraint(slot, type, axis_a, axis_b, lo, hi);
    return AVBD_OK;
  });
}

int avbd_world_joint_info(avbd_world *world, uint32_t slot, int32_t *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out) throw std::invalid_argument("out is NULL");
    const int4 f = world->engine->jointInfo(slot);
    out[0] = f.x;
    out[1] = f.y;
    out[2] = f.z;
    out[3] = f.w;
    return AVBD_OK;
  });
}

int avbd_world_adapt(avbd_world *world, int *changed) {
  return guarded([&] {
    CHECK_WORLD(world);
    bool c = world->engine->adapt();
    if (changed) *changed = c ? 1 : 0;
    return AVBD_OK;
  });
}

int avbd_world_profile_step(avbd_world *world, float *out) {
  return guarded([&] {
    CHECK_WORLD(world);
    if (!out) throw std::invalid_argument("out is NULL");
    PhaseTimes p = world->engine->profileStep();
    const float v[13] = {p.total, p.collision, p.adjacency, p.coloring, p.warmStart, p.primal, p.dual, p.velocities,
                         float(p.primalLaunches), float(p.dualLaunches), p.broadphase, p.narrowphase, p.pairSearch};
    std::memcpy(out, v, sizeof v);
    return AVBD_OK;
  });
}

int avbd_world_device_bytes(const avbd_world *world, uint64_t *out) {
  if (!world || !world->engine) return fail(AVBD_ERR_DESTROYED, "the world has been destroyed");
  if (!out) return fail(AVBD_ERR_INVALID, "out is NULL");
  *out = world->engine->deviceBytes();
  return AVBD_OK;
}

int avbd_world_solver_kind(const avbd_world *world, uint32_t *out) {
  if (!world || !world->engine) return fail(AVBD_ERR_DESTROYED, "the world has been destroyed");
  if (!out) return fail(AVBD_ERR_INVALID, "out is NULL");
  *out = world->engine->solverKind();
  return AVBD_OK;
}

int avbd_world_enqueue_commands(avbd_world *world, const uint32_t *commands, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->applyCommands(commands, n);
    return AVBD_OK;
  });
}

int avbd_world_get_params(avbd_world *world, uint32_t world_index, AvbdParams *out) {
  CHECK_WORLD(world);
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (world_index >= world->engine->worldCount()) throw std::invalid_argument("world index out of range");
    const WorldParams p = world->engine->readWorldParams()[world_index];
    out->dt = p.dt;
    out->gravity = p.gravity;
    out->up[0] = p.up[0];
    out->up[1] = p.up[1];
    out->up[2] = p.up[2];
    out->alpha = p.alpha;
    out->gamma = p.gamma;
    out->beta_lin = p.betaLin;
    out->beta_ang = p.betaAng;
    out->iterations = world->engine->iterations();
    out->flags = p.flags;
    return AVBD_OK;
  });
}

int avbd_world_set_params(avbd_world *world, uint32_t world_index, const AvbdParams *params) {
  CHECK_WORLD(world);
  return guarded([&] {
    if (!params) throw std::invalid_argument("params is NULL");
    ParamEdit e{};
    e.dt = params->dt;
    e.gravity = params->gravity;
    e.betaLin = params->beta_lin;
    e.betaAng = params->beta_ang;
    e.gamma = params->gamma;
    e.alpha = params->alpha;
    e.up[0] = params->up[0];
    e.up[1] = params->up[1];
    e.up[2] = params->up[2];
    e.flags = params->flags;
    if (params->iterations == 0) throw std::invalid_argument("iterations must be a positive integer");
    world->engine->setParams(world_index, e);
    world->engine->setIterations(params->iterations);
    return AVBD_OK;
  });
}

int avbd_world_set_filters(avbd_world *world, const uint32_t *indices, const uint32_t *groups,
                           const uint32_t *collides_with, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->setFilters(indices, groups, collides_with, n);
    return AVBD_OK;
  });
}

int avbd_world_set_restitution(avbd_world *world, const uint32_t *indices, const float *values, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->setRestitution(indices, values, n);
    return AVBD_OK;
  });
}

int avbd_world_get_restitution(avbd_world *world, float *out) {
  CHECK_WORLD(world);
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    world->engine->getRestitution(out);
    return AVBD_OK;
  });
}

int avbd_world_enable_sleeping(avbd_world *world, float speed_threshold, float time_threshold, int islands,
                               int start_asleep) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->enableSleeping(speed_threshold, time_threshold, islands, start_asleep != 0);
    return AVBD_OK;
  });
}

int avbd_world_disable_sleeping(avbd_world *world) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->disableSleeping();
    return AVBD_OK;
  });
}

int avbd_world_sleeping(avbd_world *world, int *islands) {
  CHECK_WORLD(world);
  if (islands) *islands = world->engine->sleepIslands() ? 1 : 0;
  return world->engine->sleeping() ? 1 : 0;
}

int avbd_world_wake_all(avbd_world *world) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->wakeAll();
    return AVBD_OK;
  });
}

int avbd_world_wake_bodies(avbd_world *world, const uint32_t *indices, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->wakeBodies(indices, n);
    return AVBD_OK;
  });
}

int avbd_world_set_sleep_enabled(avbd_world *world, const uint32_t *indices, const uint8_t *flags, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->setSleepEnabled(indices, flags, n);
    return AVBD_OK;
  });
}

int avbd_world_sleep_stats(avbd_world *world, uint32_t world_index, uint32_t out[2]) {
  CHECK_WORLD(world);
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    world->engine->sleepStats(world_index, out);
    return AVBD_OK;
  });
}

int avbd_world_read_sleep(avbd_world *world, uint32_t *states, uint32_t *globals) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->readSleep(states, globals);
    return AVBD_OK;
  });
}

int avbd_world_cast(avbd_world *world, const float *queries, uint32_t n, const AvbdQueryOptions *options,
                    float *hits) {
  CHECK_WORLD(world);
  return guarded([&] {
    avbd::QueryFilter f{};
    f.mask = 0xffffffffu;
    if (options) {
      if (options->ignore_count > 16 || (options->ignore_count && !options->ignore))
        throw std::invalid_argument("Queries can ignore up to 16 live body indices");
      for (uint32_t k = 0; k < options->ignore_count; k++) f.ignore[k] = options->ignore[k];
      f.ignoreCount = options->ignore_count;
      f.mask = options->collides_with;
      f.membership = options->membership;
      f.useMembership = options->use_membership ? 1u : 0u;
    }
    world->engine->cast(queries, n, f, hits);
    return AVBD_OK;
  });
}

int avbd_world_cast_noalloc(avbd_world *world, const float *queries, uint32_t n, const AvbdQueryOptions *options,
                            float *hits) {
  CHECK_WORLD(world);
  return guarded([&] {
    avbd::QueryFilter f{};
    f.mask = 0xffffffffu;
    if (options) {
      if (options->ignore_count > 16 || (options->ignore_count && !options->ignore))
        throw std::invalid_argument("Queries can ignore up to 16 live body indices");
      for (uint32_t k = 0; k < options->ignore_count; k++) f.ignore[k] = options->ignore[k];
      f.ignoreCount = options->ignore_count;
      f.mask = options->collides_with;
      f.membership = options->membership;
      f.useMembership = options->use_membership ? 1u : 0u;
    }
    world->engine->castNoAlloc(queries, n, f, hits);
    return AVBD_OK;
  });
}

int avbd_world_watch_contacts(avbd_world *world, uint32_t max_pairs, uint32_t max_events) {
  CHECK_WORLD(world);
  return guarded([&] {
    world->engine->watchContacts(max_pairs, max_events);
    return AVBD_OK;
  });
}

int avbd_world_set_watched(avbd_world *world, const uint32_t *indices, const uint8_t *flags, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    if (n && (!indices || !flags)) throw std::invalid_argument("indices or flags is NULL");
    world->engine->setWatched(indices, flags, n);
    return AVBD_OK;
  });
}

int avbd_world_reset_contact_history(avbd_world *world, const uint32_t *indices, uint32_t n) {
  CHECK_WORLD(world);
  return guarded([&] {
    if (n && !indices) throw std::invalid_argument("indices is NULL");
    world->engine->resetContactHistory(indices, n);
    return AVBD_OK;
  });
}

int avbd_world_read_contact_events(avbd_world *world, uint32_t *out, uint32_t cap, uint32_t *count, uint32_t *written,
                                   uint32_t *full) {
  CHECK_WORLD(world);
  return guarded([&] {
    const uint32_t n = world->engine->readContactEvents(out, cap, written, full);
    if (count) *count = n;
    return AVBD_OK;
  });
}

// --- serving loop ----------------------------------------------------------------------------------------------------

#define CHECK_SERVER(s) \
  if (!(s) || !(s)->server) return fail(AVBD_ERR_DESTROYED, "the server has been destroyed")

// --- convex hull cooking (host only) -----------------------------------------------------------------------------------

static int cookedHull(bool ok, const std::string &err, const avbd::hullcook::Shape &shape, uint32_t at, uint32_t *record,
                      uint32_t capacity, AvbdHullInfo *info) {
  if (!ok) throw std::invalid_argument(err);
  const size_t words = avbd::hullcook::recordWords(shape);
  if (uint64_t(at) + words >= (1ull << 24)) throw std::invalid_argument("hull buffer beyond 2^24 words");
  if (record) {
    if (capacity < words) throw std::invalid_argument("record capacity too small for the hull");
    avbd::hullcook::pack(shape, at, record);
  }
  if (info) {
    *info = AvbdHullInfo{};
    info->vertex_count = uint32_t(shape.vertices.size() / 3);
    info->face_count = uint32_t(shape.faces.size());
    info->edge_count = uint32_t(shape.edges.size());
    info->words = uint32_t(words);
    info->volume = shape.volume;
    info->radius = shape.radius;
    for (int k = 0; k < 3; k++) {
      info->moments[k] = shape.moments[k];
      info->size[k] = shape.size[k];
      info->center[k] = shape.center[k];
    }
    for (int k = 0; k < 4; k++) info->rotation[k] = shape.rotation[k];
  }
  return AVBD_OK;
}

int avbd_hull_from_points(const double *xyz, uint32_t n_points, uint32_t at, uint32_t *record, uint32_t capacity_words,
                          AvbdHullInfo *info) {
  return guarded([&] {
    if (n_points && !xyz) throw std::invalid_argument("xyz is NULL");
    for (uint64_t k = 0; k < 3ull * n_points; k++)
      if (!std::isfinite(xyz[k])) throw std::invalid_argument("Hull points must contain finite xyz triples");
    avbd::hullcook::Shape shape;
    std::string err;
    const bool ok = avbd::hullcook::fromPoints(xyz, n_points, shape, err);
    return cookedHull(ok, ok ? err : "A convex hull requires finite points spanning a volume", shape, at, record,
                      capacity_words, info);
  });
}

int avbd_hull_from_triangles(const double *xyz, uint32_t n_points, const uint32_t *indices, uint32_t n_indices,
                             uint32_t at, uint32_t *record, uint32_t capacity_words, AvbdHullInfo *info) {
  return guarded([&] {
    if ((n_points && !xyz) || (n_indices && !indices)) throw std::invalid_argument("xyz or indices is NULL");
    for (uint64_t k = 0; k < 3ull * n_points; k++)
      if (!std::isfinite(xyz[k])) throw std::invalid_argument("positions must be finite xyz triples");
    avbd::hullcook::Shape shape;
    std::string err;
    const bool ok = avbd::hullcook::fromTriangles(xyz, n_points, indices, n_indices, shape, err);
    return cookedHull(ok, err, shape, at, record, capacity_words, info);
  });
}

int avbd_server_create(const AvbdServerOptions *options, avbd_server **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    *out = nullptr;
    if (avbd_device_count() <= 0) return fail(AVBD_ERR_DEVICE, "no HIP device is available");
    ServerOptions o;
    o.engine.useGraph = true;
    if (options) {
      if (options->hbm_budget_bytes) o.budgetBytes = size_t(options->hbm_budget_bytes);
      if (options->max_active) o.maxActive = options->max_active;
      if (options->lookahead) o.lookahead = options->lookahead;
      o.device = options->device;
      if (options->color_cap) o.engine.colorCap = options->color_cap;
      if (options->color_rounds) o.engine.colorRounds = options->color_rounds;
      if (options->broadphase > 2) throw std::invalid_argument("broadphase must be 0 (auto), 1 (grid) or 2 (hploc)");
      o.engine.broadphase = options->broadphase;
      o.smallWorldBodies = options->small_world_bodies;
      const uint32_t b = options->small_world_block;
      if (b != 0 && b != 128 && b != 256 && b != 512 && b != 1024)
        throw std::invalid_argument("small_world_block must be 0, 128, 256, 512 or 1024");
      if (b) o.smallWorldBlock = b;
      o.largeWorldBodies = options->large_world_bodies;
      if (options->large_world_share > 1024) throw std::invalid_argument("large_world_share must be at most 1024");
      if (options->large_world_share) o.largeWorldShare = options->large_world_share;
    }
    if (o.engine.colorCap > 64) throw std::invalid_argument("color_cap must be at most 64");
    auto s = std::make_unique<avbd_server>();
    s->server = std::make_unique<Server>(o);
    *out = s.release();
    return AVBD_OK;
  });
}

int avbd_server_destroy(avbd_server *server) {
  return guarded([&] {
    CHECK_SERVER(server);
    server->server.reset();
    delete server;
    return AVBD_OK;
  });
}

int avbd_server_add_world(avbd_server *server, const AvbdSceneDesc *scene, uint32_t *world_id) {
  return guarded([&] {
    CHECK_SERVER(server);
    if (!scene || !world_id) throw std::invalid_argument("scene or world_id is NULL");
    *world_id = server->server->addWorld(toHost(*scene));
    return AVBD_OK;
  });
}

int avbd_server_world_body_count(avbd_server *server, uint32_t world_id, uint32_t *out) {
  return guarded([&] {
    CHECK_SERVER(server);
    if (!out) throw std::invalid_argument("out is NULL");
    *out = server->server->bodyCount(world_id);
    return AVBD_OK;
  });
}

int avbd_server_submit(avbd_server *server, uint32_t world_id, uint32_t steps, uint64_t *request_id) {
  return guarded([&] {
    CHECK_SERVER(server);
    if (!request_id) throw std::invalid_argument("request_id is NULL");
    *request_id = server->server->submit(world_id, steps);
    return AVBD_OK;
  });
}

int avbd_server_wait(avbd_server *server, uint64_t request_id, double *latency_ms) {
  return guarded([&] {
    CHECK_SERVER(server);
    double ms = server->server->wait(request_id);
    if (latency_ms) *latency_ms = ms;
    return AVBD_OK;
  });
}

int avbd_server_poll(avbd_server *server, uint64_t request_id, int *done) {
  return guarded([&] {
    CHECK_SERVER(server);
    if (!done) throw std::invalid_argument("done is NULL");
    *done = server->server->done(request_id) ? 1 : 0;
    return AVBD_OK;
  });
}

int avbd_server_read_bodies(avbd_server *server, uint32_t world_id, float *out) {
  return guarded([&] {
    CHECK_SERVER(server);
    if (!out) throw std::invalid_argument("out is NULL");
    std::vector<Body> b;
    server->server->readBodies(world_id, b);
    std::memcpy(out, b.data(), b.size() * sizeof(Body));
    return AVBD_OK;
  });
}

int avbd_server_stats(avbd_server *server, AvbdServerStats *out) {
  return guarded([&] {
    CHECK_SERVER(server);
    if (!out) throw std::invalid_argument("out is NULL");
    ServerStats s = server->server->stats();
    out->submitted = s.submitted;
    out->completed = s.completed;
    out->admissions = s.admissions;
    out->evictions = s.evictions;
    out->restores = s.restores;
    out->steps = s.steps;
    out->body_steps = s.bodySteps;
    out->wall_seconds = s.wallSeconds;
    out->step_ms_p50 = s.stepMsP50;
    out->step_ms_p99 = s.stepMsP99;
    out->step_ms_max = s.stepMsMax;
    out->request_ms_p50 = s.requestMsP50;
    out->request_ms_p99 = s.requestMsP99;
    out->resident_bytes = s.residentBytes;
    out->peak_resident_bytes = s.peakResidentBytes;
    out->resident_worlds = s.residentWorlds;
    out->worlds = s.worlds;
    out->solve_aborts = s.solveAborts;
    return AVBD_OK;
  });
}

int avbd_server_reset_stats(avbd_server *server) {
  return guarded([&] {
    CHECK_SERVER(server);
    server->server->resetStats();
    return AVBD_OK;
  });
}

}  // extern "C"
