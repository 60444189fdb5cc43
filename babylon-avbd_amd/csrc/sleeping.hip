// Sleeping (src/gpu/worldSleeping.js, worldSleeping.wgsl, islandSleeping.wgsl, sleepContact.wgsl). A sleeping body is a
// fixed support: its mass (size.w) is 0 and its velocities 0; its awake mass stays in its SleepState and comes back
// when it wakes. Whole-world mode: a body sleeps after `delay` seconds of being quiet (linear speed + angular speed ×
// bounding radius at most `speed`) and supported (gravity 0, or a contact below it with a static, sleeping or quiet
// partner); an impact on a sleeper wakes its world. Island mode: contacts and joints join bodies into islands of a
// union-find forest (links among sleepers are kept while they sleep, since their pairs are not collided); an island
// sleeps only when no member is busy, and wakes as a whole. The forest is flattened after every step, so each body's
// parent is its island's root (the island's smallest index) and the forest does not depend on thread timing.
//
// Per step: before (wake requests, forest reset) ahead of the step; after the narrowphase, impacts on sleepers from
// the contacts and joints just found, the hit sleepers woken, and, if any woke, the pairs and narrowphase run again
// (gated kernels) so woken neighbours keep their mutual contacts in this very step; after the solve, islands, supports
// and impacts, busy islands, and the rest pass. Every flag is set by idempotent stores or ors, and the counts are sums,
// so the outcome does not depend on scheduling.
#include "kernels.h"

namespace avbd {

namespace {

constexpr uint32_t SF_READY = 1, SF_HELD = 2, SF_START_QUIET = 4, SF_GATHERED = 8;  // SleepState::flags
constexpr uint32_t SW_IMPACT = 1, SW_GROUNDED = 2, SW_BUSY = 4, SW_REQUEST = 8;     // SleepState::wake (at a root)
constexpr uint32_t NO_ROOT = 0xffffffffu;
constexpr uint32_t FOREST_STEPS = 1024;  // bound on a root walk or a union's retries; past it the world wakes

AVBD_D uint32_t ldRelaxed(const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
AVBD_D void stRelaxed(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

AVBD_D uint32_t *globalsOf(const DeviceWorld &w, uint32_t i) { return w.sleepGlobals + 4ull * w.bodyWorld[i]; }
AVBD_D const WorldParams &paramsOf(const DeviceWorld &w, uint32_t i) { return w.worlds[w.bodyWorld[i]]; }
AVBD_D f3 upOf(const WorldParams &p) { return mk3(p.up[0], p.up[1], p.up[2]); }
AVBD_D uint32_t manifoldCount(const DeviceWorld &w) { return umin(w.counters[C_MANIFOLDS], w.pairCapacity); }
AVBD_D bool liveManifold(const Manifold &m) { return (m.ids.w & 15u) != 0u && (m.ids.w & 0x80000000u) == 0u; }

/** Linear speed + angular speed × bounding radius (worldSleeping.wgsl speed()). */
AVBD_D float motion(const Body &b) { return length(xyz(b.vel)) + length(xyz(b.angVel)) * b.moment.w; }

/** wake(): the awake mass back; quiet time and flags reset. Only body i's thread calls it for body i. */
AVBD_D void wakeOne(const DeviceWorld &w, uint32_t i) {
  SleepState &s = w.sleep[i];
  if (s.mass > 0.f) {
    w.bodies[i].size.w = s.mass;
    s.asleep = 0u;
    s.quiet = 0.f;
    s.flags = 0u;
  }
}

/** The body's island root, halving the path on the way (parents only move to ancestors: parent ≤ index always). A
 * walk past the bound, or a parent above its child, is never expected: it asks the body's world to wake. */
AVBD_D uint32_t rootOf(const DeviceWorld &w, uint32_t i) {
  const uint32_t start = i;
  for (uint32_t k = 0; k < FOREST_STEPS && i < w.nBodies; k++) {
    const uint32_t p = ldRelaxed(&w.sleep[i].parent);
    if (p == i) return i;
    if (p > i) break;
    const uint32_t gp = ldRelaxed(&w.sleep[p].parent);
    if (gp != p) atomicMin(&w.sleep[i].parent, gp);  // p is not a root: halve (a no-op write is skipped)
    i = gp;
  }
  stRelaxed(globalsOf(w, start), 1u);
  return NO_ROOT;
}

/** Link the higher of two island roots below the lower (islandSleeping.wgsl connect()). */
AVBD_D void connect(const DeviceWorld &w, uint32_t a, uint32_t b) {
  if (a >= w.nBodies || b >= w.nBodies || w.sleep[a].mass <= 0.f || w.sleep[b].mass <= 0.f) return;
  for (uint32_t k = 0; k < FOREST_STEPS; k++) {
    const uint32_t ra = rootOf(w, a), rb = rootOf(w, b);
    if (ra == rb || ra == NO_ROOT || rb == NO_ROOT) return;
    const uint32_t hi = umax(ra, rb), lo = umin(ra, rb);
    if (atomicCAS(&w.sleep[hi].parent, hi, lo) == hi) return;
  }
  stRelaxed(globalsOf(w, a), 1u);
}

/** Or bits into an island root's wake word; skipped when they are already set (one island's members would otherwise
 * all hit the same word). */
AVBD_D void orWake(const DeviceWorld &w, uint32_t r, uint32_t bits) {
  if ((ldRelaxed(&w.sleep[r].wake) & bits) != bits) atomicOr(&w.sleep[r].wake, bits);
}

/** orWake for the lanes with `want` set, called by every lane of the wave that entered: the lanes sharing the first
 * wanting lane's root are covered by its one or. */
AVBD_D void orWakeWave(const DeviceWorld &w, uint32_t r, uint32_t bits, bool want) {
  const uint64_t lanes = __ballot(want);
  if (!lanes) return;
  const int leader = __ffsll(static_cast<unsigned long long>(lanes)) - 1;
  const uint32_t r0 = __shfl(r, leader);
  if (want && (r != r0 || int(__lane_id()) == leader)) orWake(w, r, bits);
}

AVBD_D void signalWake(const DeviceWorld &w, const SleepParams &p, uint32_t i) {
  uint32_t *g = globalsOf(w, i);
  if (p.islands) {
    if (!ldRelaxed(g + 3)) stRelaxed(g + 3, 1u);
    const uint32_t r = rootOf(w, i);
    if (r != NO_ROOT) orWake(w, r, SW_IMPACT);
  } else if (!ldRelaxed(g)) {
    stRelaxed(g, 1u);
  }
}

/** A joint or free motion: a moving body (past twice the threshold) joined to a sleeper wakes it. */
AVBD_D void impactBodies(const DeviceWorld &w, const SleepParams &p, uint32_t a, uint32_t b) {
  if (w.sleep[a].asleep && !w.sleep[b].asleep && motion(w.bodies[b]) > 2.f * p.speed) signalWake(w, p, a);
}

/** sleepContact.wgsl contactVelocity(): a point's velocity now, or over the step just solved (from the step-start and
 * inertial poses, less this step's gravity for a dynamic body); 0 for a sleeper. */
AVBD_D f3 pointVelocity(const DeviceWorld &w, uint32_t i, f3 r, bool previous) {
  if (w.sleep[i].asleep) return mk3(0.f, 0.f, 0.f);
  const Body &B = w.bodies[i];
  f3 v = xyz(B.vel), om = xyz(B.angVel);
  float4 q = B.rot;
  if (previous) {
    const WorldParams &p = paramsOf(w, i);
    const float dt = p.dt;
    v = divs(sub(xyz(B.inertialPos), xyz(B.ipos)), dt);
    if (B.size.w > 0.f) v = sub(v, scale(upOf(p), p.gravity * dt));
    const float4 a = B.inertialRot, b = B.irot;
    const f3 av = xyz(a), bv = xyz(b);
    const f3 t = scale(sub(add(scale(bv, -a.w), scale(av, b.w)), cross(av, bv)), 2.f);
    om = divs(t, qlengthSq(b) * dt);
    q = b;
  }
  return add(v, cross(om, qrotate(q, r)));
}

AVBD_D void contactImpact(const DeviceWorld &w, const SleepParams &p, uint32_t a, uint32_t b, f3 n, f3 rA, f3 rB,
                          bool beforeSolve) {
  if (!w.sleep[a].asleep || w.sleep[b].asleep) return;
  float incoming = 0.f;
  if (!beforeSolve) incoming = dot(sub(pointVelocity(w, b, rB, true), pointVelocity(w, a, rA, true)), n);
  const float outgoing = dot(sub(pointVelocity(w, b, rB, false), pointVelocity(w, a, rA, false)), n);
  if (rmax(incoming, outgoing) > 2.f * p.speed) signalWake(w, p, a);
}

/** Every point of a manifold between a sleeper and a waker, both ways (a rotating face can approach at one corner
 * while another recedes). */
AVBD_D void manifoldImpact(const DeviceWorld &w, const SleepParams &p, const Manifold &m, bool beforeSolve) {
  const uint32_t a = m.ids.x, b = m.ids.y;
  if (w.sleep[a].asleep == w.sleep[b].asleep) return;
  const f3 n = xyz(m.geo);
  for (uint32_t k = 0; k < (m.ids.w & 15u); k++) {
    const Contact c = w.contacts[m.ids.z + k];
    contactImpact(w, p, a, b, n, xyz(c.rA), xyz(c.rB), beforeSolve);
    contactImpact(w, p, b, a, neg(n), xyz(c.rB), xyz(c.rA), beforeSolve);
  }
}

/** b supports a: b is static or asleep, or was quiet and supported at the last rest pass and is still quiet. */
AVBD_D void support(const DeviceWorld &w, const SleepParams &p, uint32_t a, uint32_t b) {
  const SleepState &s = w.sleep[b];
  if (s.mass <= 0.f || s.asleep || ((s.flags & SF_READY) && motion(w.bodies[b]) <= p.speed))
    stRelaxed(&w.sleep[a].support, 1u);
}

/** Add `hit` lanes to word `slot` of their world's globals: one atomic per wave for the lanes of the first active
 * lane's world (normally all of them), one per lane for the others. Every lane of the wave that entered calls it. */
AVBD_D void countPerWorld(const DeviceWorld &w, uint32_t world, bool hit, uint32_t slot) {
  const uint64_t active = __ballot(1);
  const int leader = __ffsll(static_cast<unsigned long long>(active)) - 1;
  const uint32_t w0 = __shfl(world, leader);
  const uint64_t hits = __ballot(hit && world == w0);
  if (int(__lane_id()) == leader && hits) atomicAdd(&w.sleepGlobals[4ull * w0 + slot], uint32_t(__popcll(hits)));
  if (hit && world != w0) atomicAdd(&w.sleepGlobals[4ull * world + slot], 1u);
}

}  // namespace

/** register(): a range of bodies' states from their current masses; eligibility from `eligible` (one byte per body
 * of the range), or kept when it is null. */
__global__ void k_sleep_init(DeviceWorld w, const uint8_t *eligible, uint32_t first, uint32_t count) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint32_t i = first + k;
  SleepState &s = w.sleep[i];
  const uint32_t e = eligible ? uint32_t(eligible[k]) : s.eligible;
  s = SleepState{w.bodies[i].size.w, 0.f, 0u, 0u, e, i, 0u, 0u};
}

/** Eligibility edits: (index, flag) pairs. */
__global__ void k_sleep_eligible(DeviceWorld w, const uint32_t *words, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) w.sleep[words[2 * k]].eligible = words[2 * k + 1];
}

/** wakeBody(): the bodies are held awake (their flags become exactly SF_HELD). */
__global__ void k_sleep_hold(DeviceWorld w, const uint32_t *indices, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) w.sleep[indices[k]].flags = SF_HELD;
}

/** requestWake: a held sleeper asks its island to wake (before the forest is reset). */
__global__ void k_sleep_request(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies || !w.sleep[i].asleep || !(w.sleep[i].flags & SF_HELD)) return;
  const uint32_t r = rootOf(w, i);
  if (r != NO_ROOT) orWake(w, r, SW_REQUEST);
}

/** gatherWake: the members of a requesting island are marked to wake. */
__global__ void k_sleep_gather(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  const uint32_t r = rootOf(w, i);
  if (r != NO_ROOT && (ldRelaxed(&w.sleep[r].wake) & SW_REQUEST)) w.sleep[i].flags |= SF_GATHERED;
}

/** before(): wake-all requests (forced or asked for by the last step), held and gathered bodies wake; awake bodies
 * leave their islands (the links among sleepers stay); startAsleep marks the still eligible bodies. */
__global__ void k_sleep_before(DeviceWorld w, SleepParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  SleepState &s = w.sleep[i];
  const bool all = (p.force & 1u) || ldRelaxed(globalsOf(w, i)) != 0u;
  if (p.islands) {
    if (all) {
      wakeOne(w, i);
      s.parent = i;
    } else if (s.flags & (SF_HELD | SF_GATHERED)) {
      const uint32_t held = s.flags & SF_HELD;
      wakeOne(w, i);
      s.flags |= held;
    }
    if (!s.asleep) s.parent = i;
    if ((p.force & 2u) && motion(w.bodies[i]) <= p.speed && s.eligible) s.flags |= SF_START_QUIET;
    s.wake = 0u;
  } else if (all || (s.flags & SF_HELD)) {
    wakeOne(w, i);
  }
  s.support = 0u;
}

/** clear(): every world's globals, and the refresh gate. */
__global__ void k_sleep_clear(DeviceWorld w) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= 4u * w.nWorlds) w.sleepGlobals[k] = 0u;
}

/** Impacts on sleepers from this step's contacts (before the solve: velocities now) and joints. Items: manifold slots,
 * then joints. */
__global__ void k_sleep_impact(DeviceWorld w, SleepParams p) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < w.pairCapacity) {
    if (t >= manifoldCount(w)) return;
    const Manifold m = w.manifolds[t];
    if (liveManifold(m)) manifoldImpact(w, p, m, true);
    return;
  }
  const uint32_t j = t - w.pairCapacity;
  if (j >= w.nJoints) return;
  const int4 f = w.info[j];
  if (f.x == 0 || f.y < 0) return;
  impactBodies(w, p, uint32_t(f.y), uint32_t(f.z));
  impactBodies(w, p, uint32_t(f.z), uint32_t(f.y));
}

/** applyImpact: the sleepers of a hit island (of a hit world, in whole-world mode) wake; their count opens the gate of
 * the collision refresh. */
__global__ void k_sleep_apply_impact(DeviceWorld w, SleepParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  const uint32_t *g = globalsOf(w, i);
  bool woke = false;
  if (w.sleep[i].asleep && (ldRelaxed(g + 3) != 0u || ldRelaxed(g) != 0u)) {
    bool hit = ldRelaxed(g) != 0u;
    if (!hit && p.islands) {
      const uint32_t r = rootOf(w, i);
      hit = r == NO_ROOT || (ldRelaxed(&w.sleep[r].wake) & SW_IMPACT) != 0u;
    }
    if (hit) {
      wakeOne(w, i);
      woke = true;
    }
  }
  countPerWorld(w, w.bodyWorld[i], woke, 2);
  const uint64_t any = __ballot(woke);
  if (any && int(__lane_id()) == __ffsll(static_cast<unsigned long long>(any)) - 1)
    atomicAdd(&w.sleepGlobals[4ull * w.nWorlds], uint32_t(__popcll(any)));
}

/** Before the gated pairs and narrowphase run again: their counters and the adjacency degrees they accumulate. */
__global__ void k_sleep_refresh_prepare(DeviceWorld w) {
  if (gatedOff(w)) return;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= w.nBodies; i += stride) {
    if (i < w.nBodies) w.adjFill[i] = 0u;
    w.adjStart[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.counters[C_PAIRS] = 0u;
    w.counters[C_CONTACTS] = 0u;
    w.counters[C_MANIFOLDS] = 0u;
    w.counters[C_CANDIDATES] = 0u;
  }
}

/** Before the unions: each awake dynamic body (a root after the reset) hooks below its smallest dynamic partner in its
 * adjacency list (contacts and active joints), one store to its own parent; with the pointer jumping after it most of
 * the unions then find their two roots equal. Links go to lower indices only, so the islands and their roots are those
 * the unions alone would give. */
__global__ void k_sleep_hook(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies || w.sleep[i].mass <= 0.f || w.sleep[i].parent != i) return;
  uint32_t best = i;
  const uint32_t end = umin(w.adjStart[i + 1], w.adjCapacity);
  for (uint32_t e = w.adjStart[i]; e < end; e++) {
    const uint32_t key = w.adjList[e];
    uint32_t j;
    if (key < w.nJoints) {
      const int4 f = w.info[key];
      if (f.x == 0 || f.y < 0 || f.z < 0) continue;
      j = uint32_t(f.y) == i ? uint32_t(f.z) : uint32_t(f.y);
    } else {
      const Manifold m = w.manifolds[key - w.nJoints];
      if (!liveManifold(m)) continue;
      j = m.ids.x == i ? m.ids.y : m.ids.x;
    }
    if (j < best && w.sleep[j].mass > 0.f) best = j;
  }
  if (best != i) stRelaxed(&w.sleep[i].parent, best);
}

/** A linking pass before the unions: for each contact or joint whose ends have different roots, the higher root's
 * parent drops to the lower root (an atomic min, no retry: a link it overrides is made again by the unions, which see
 * every contact and joint). */
__global__ void k_sleep_link(DeviceWorld w) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a, b;
  if (t < w.pairCapacity) {
    if (t >= manifoldCount(w)) return;
    const Manifold m = w.manifolds[t];
    if (!liveManifold(m)) return;
    a = m.ids.x;
    b = m.ids.y;
  } else {
    const uint32_t j = t - w.pairCapacity;
    if (j >= w.nJoints) return;
    const int4 f = w.info[j];
    if (f.x == 0 || f.y < 0 || f.z < 0) return;
    a = uint32_t(f.y);
    b = uint32_t(f.z);
  }
  if (a >= w.nBodies || b >= w.nBodies || w.sleep[a].mass <= 0.f || w.sleep[b].mass <= 0.f) return;
  const uint32_t ra = rootOf(w, a), rb = rootOf(w, b);
  if (ra == rb || ra == NO_ROOT || rb == NO_ROOT) return;
  atomicMin(&w.sleep[umax(ra, rb)].parent, umin(ra, rb));
}

/** Island links of this step's contacts and joints (island mode). Items: manifold slots, then joints. */
__global__ void k_sleep_connect(DeviceWorld w) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < w.pairCapacity) {
    if (t >= manifoldCount(w)) return;
    const Manifold m = w.manifolds[t];
    if (liveManifold(m)) connect(w, m.ids.x, m.ids.y);
    return;
  }
  const uint32_t j = t - w.pairCapacity;
  if (j >= w.nJoints) return;
  const int4 f = w.info[j];
  if (f.x != 0 && f.y >= 0 && f.z >= 0) connect(w, uint32_t(f.y), uint32_t(f.z));
}

/** Every body's parent becomes its island root. */
__global__ void k_sleep_flatten(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  const uint32_t r = rootOf(w, i);
  if (r != NO_ROOT) stRelaxed(&w.sleep[i].parent, r);
}

/** After the solve: impacts (incoming and outgoing velocities), supports from the contacts below each body, and for
 * startAsleep the islands resting on still static support. Items: manifold slots, then joints. */
__global__ void k_sleep_contacts(DeviceWorld w, SleepParams p) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < w.pairCapacity) {
    if (t >= manifoldCount(w)) return;
    const Manifold m = w.manifolds[t];
    if (!liveManifold(m)) return;
    const uint32_t a = m.ids.x, b = m.ids.y;
    manifoldImpact(w, p, m, false);
    const float up = dot(xyz(m.geo), upOf(paramsOf(w, a)));
    if (up > 0.3f) support(w, p, a, b);
    if (up < -0.3f) support(w, p, b, a);
    if (p.islands && (p.force & 2u)) {
      const SleepState &sa = w.sleep[a], &sb = w.sleep[b];
      if (sb.mass <= 0.f && motion(w.bodies[b]) <= p.speed && up > 0.3f && sa.mass > 0.f) {
        const uint32_t r = rootOf(w, a);
        if (r != NO_ROOT) orWake(w, r, SW_GROUNDED);
      }
      if (sa.mass <= 0.f && motion(w.bodies[a]) <= p.speed && up < -0.3f && sb.mass > 0.f) {
        const uint32_t r = rootOf(w, b);
        if (r != NO_ROOT) orWake(w, r, SW_GROUNDED);
      }
    }
    return;
  }
  const uint32_t j = t - w.pairCapacity;
  if (j >= w.nJoints) return;
  const int4 f = w.info[j];
  if (f.x == 0 || f.y < 0) return;
  impactBodies(w, p, uint32_t(f.y), uint32_t(f.z));
  impactBodies(w, p, uint32_t(f.z), uint32_t(f.y));
}

/** activity (island mode): a moving member, or an eligible one not yet quiet and supported long enough, keeps its
 * island awake (at startAsleep: any member that was not still). */
__global__ void k_sleep_activity(DeviceWorld w, SleepParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  const SleepState &s = w.sleep[i];
  if (s.mass <= 0.f) return;
  const WorldParams &wp = paramsOf(w, i);
  bool busy = false;
  if (p.force & 2u) {
    busy = (s.flags & SF_START_QUIET) == 0u;
  } else if (!s.asleep) {
    busy = motion(w.bodies[i]) > p.speed;
    if (s.eligible && !(s.flags & SF_HELD))
      busy = busy || s.quiet + wp.dt < p.delay || (wp.gravity != 0.f && !ldRelaxed(&s.support));
  }
  const uint32_t r = busy ? rootOf(w, i) : NO_ROOT;
  orWakeWave(w, r, SW_BUSY, r != NO_ROOT);
}

/** rest: the quiet timer, falling asleep (mass and velocities to 0), waking on impacts, overflow or clashes, and the
 * per-world sleeping count. */
__global__ void k_sleep_rest(DeviceWorld w, SleepParams p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  SleepState &s = w.sleep[i];
  bool counted = false;
  if (s.mass > 0.f) {
    const uint32_t *g = globalsOf(w, i);
    uint32_t islandWake = 0u;
    bool wakeNow = !s.eligible;
    if (!wakeNow) {
      if (p.islands) {
        const uint32_t r = rootOf(w, i);
        islandWake = r == NO_ROOT ? SW_IMPACT : ldRelaxed(&w.sleep[r].wake);
      }
      wakeNow = (islandWake & SW_IMPACT) || ldRelaxed(g) != 0u || w.counters[C_OVERFLOW] != 0u ||
                w.counters[C_CLASHES] != 0u;
    }
    if (wakeNow) {
      wakeOne(w, i);
    } else {
      const WorldParams &wp = paramsOf(w, i);
      const bool startQuiet = (s.flags & SF_START_QUIET) != 0u, held = (s.flags & SF_HELD) != 0u;
      const bool ready = motion(w.bodies[i]) <= p.speed && (wp.gravity == 0.f || ldRelaxed(&s.support) != 0u);
      s.flags = ready ? SF_READY : 0u;
      if (!s.asleep) {
        s.quiet = ready ? s.quiet + wp.dt : 0.f;
        const bool now = p.islands ? !held && ((s.quiet >= p.delay && !(islandWake & SW_BUSY)) ||
                                               ((p.force & 2u) && startQuiet && (islandWake & 6u) == SW_GROUNDED))
                                   : s.quiet >= p.delay;
        if (now) {
          s.asleep = 1u;
          Body &B = w.bodies[i];
          B.size.w = 0.f;
          B.vel = make_float4(0.f, 0.f, 0.f, 0.f);
          B.angVel = make_float4(0.f, 0.f, 0.f, B.angVel.w);
        }
      }
      counted = s.asleep != 0u;
    }
  }
  countPerWorld(w, w.bodyWorld[i], counted, 1);
}

static inline uint32_t blocks(uint64_t n, uint32_t bs) {
  const uint64_t g = (n + bs - 1) / bs;
  return uint32_t(g < 1 ? 1 : g);
}

void launch_sleep_init(const DeviceWorld &w, const uint8_t *eligible, uint32_t first, uint32_t count, hipStream_t s) {
  if (count) k_sleep_init<<<blocks(count, 256), 256, 0, s>>>(w, eligible, first, count);
}
void launch_sleep_eligible(const DeviceWorld &w, const uint32_t *words, uint32_t n, hipStream_t s) {
  if (n) k_sleep_eligible<<<blocks(n, 256), 256, 0, s>>>(w, words, n);
}
void launch_sleep_hold(const DeviceWorld &w, const uint32_t *indices, uint32_t n, bool islands, hipStream_t s) {
  if (!n) return;
  k_sleep_hold<<<blocks(n, 256), 256, 0, s>>>(w, indices, n);
  if (!islands) return;
  k_sleep_request<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
  k_sleep_gather<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
}
void launch_sleep_before(const DeviceWorld &w, const SleepParams &p, hipStream_t s) {
  k_sleep_before<<<blocks(w.nBodies, 256), 256, 0, s>>>(w, p);
  k_sleep_clear<<<blocks(4ull * w.nWorlds + 1, 256), 256, 0, s>>>(w);
}
void launch_sleep_impact(const DeviceWorld &w, const SleepParams &p, hipStream_t s) {
  k_sleep_impact<<<blocks(uint64_t(w.pairCapacity) + w.nJoints, 256), 256, 0, s>>>(w, p);
  k_sleep_apply_impact<<<blocks(w.nBodies, 256), 256, 0, s>>>(w, p);
}
void launch_sleep_refresh_prepare(const DeviceWorld &w, hipStream_t s) {
  k_sleep_refresh_prepare<<<blocks(uint64_t(w.nBodies) + 1, 256), 256, 0, s>>>(w);
}
void launch_sleep_after(const DeviceWorld &w, const SleepParams &p, hipStream_t s) {
  const uint32_t items = blocks(uint64_t(w.pairCapacity) + w.nJoints, 256);
  if (p.islands) {
    k_sleep_hook<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
    k_sleep_flatten<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
    k_sleep_link<<<items, 256, 0, s>>>(w);
    k_sleep_flatten<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
    k_sleep_connect<<<items, 256, 0, s>>>(w);
    k_sleep_flatten<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
  }
  k_sleep_contacts<<<items, 256, 0, s>>>(w, p);
  if (p.islands) k_sleep_activity<<<blocks(w.nBodies, 256), 256, 0, s>>>(w, p);
  k_sleep_rest<<<blocks(w.nBodies, 256), 256, 0, s>>>(w, p);
}

}  // namespace avbd
