// The device-side view of a batch of worlds (a single world is a batch of one): pointers into HBM and sizes, passed by
// value to every kernel. Worlds occupy contiguous body and joint ranges; per-step products (pairs, manifolds, contacts,
// adjacency, colour buckets) are shared arrays over the whole batch, so one launch per phase steps every world.
#pragma once
#include "avbd_device.h"

namespace avbd {

/**
 * Work of one step's dataflow solve (written on the device before it): unit u = it × unitsPerIter + r is wave
 * r - waveStart[c] of colour c's primal (its heavy waves first) in iteration it, with the duals its bodies finish.
 * Stamps of the step count from `base` (each step's base is past the last one's stamps).
 */
struct DataflowPlan {
  uint32_t base, iterations, unitsPerIter, total;
  uint32_t waveStart[65];
  uint32_t heavyWaves[64];
};

/** One body's sleeping state (32 B, the reference's per-body sleep record). */
struct SleepState {
  float mass;         // the body's mass while awake (0: static)
  float quiet;        // seconds it has been quiet and supported without a break
  uint32_t asleep;
  uint32_t support;   // this step: a contact holds it from below
  uint32_t eligible;  // it may sleep
  uint32_t parent;    // island forest (island mode): parent <= index; each body's island root after every step
  uint32_t wake;      // at an island root, this step: 1 hit, 2 rests on still static support, 4 busy, 8 wake requested
  uint32_t flags;     // 1 quiet and supported at the last rest pass, 2 held awake, 4 still at start, 8 island woken
};
static_assert(sizeof(SleepState) == 32, "SleepState is 32 bytes");

/** Sleeping parameters of a launch: thresholds, mode, and force bits (1 wake all, 2 start asleep). */
struct SleepParams {
  float speed, delay;
  uint32_t islands, force;
};

struct DeviceWorld {
  // Scene
  Body *bodies;
  uint32_t *bodyWorld;   // world of each body
  WorldParams *worlds;
  float4 *refs;          // 2 per body: reference pos, reference rot (contact reuse)
  float4 *irotInv;       // per body: the inverse of its step-start rotation, formed by the warm start
  uint2 *filters;        // groups, collides-with
  Joint *joints;
  int4 *info;            // type, a (-1 = world), b, 0
  uint32_t *noCollide;   // (hi, lo, joint slot or ~0) triples, sorted by (hi, lo)
  uint32_t *ncStart;     // nBodies + 1: body hi's triples are noCollide[ncStart[hi] .. ncStart[hi + 1])
  uint32_t nBodies, nJoints, nWorlds, noCollideCount;
  uint32_t roundedShapes;  // any sphere or capsule in the batch (else the narrowphase runs its box-only variant)
  uint32_t hullShapes;     // any hull in the batch: pairs with a hull go to the hull narrowphase
  const uint4 *hulls;      // hull records (4 u32 per word; a hull body's shape code is 3 + its header word)
  uint32_t *hullPairs;     // pair slots with a hull, listed by the narrowphase (C_HULL_PAIRS of them)
  uint32_t *counters;    // COUNTER_WORDS
  uint32_t *scanScratch; // block sums of the multi-level prefix scan

  // Uniform-grid broadphase
  uint32_t *gridCount;   // tableSize + 1 (counts, then bucket starts after the scan)
  uint32_t *gridCursor;  // tableSize
  uint32_t *gridSorted;  // nBodies
  int4 *cellOf;          // per body: cell xyz, world (w = -1: large body)
  float4 *sortedPR;      // cell-sorted copies of (pos.xyz, bounding radius)
  int4 *sortedCell;      // cell-sorted copies of cellOf
  uint32_t tableMask;
  float cellSize;
  uint32_t *largeStart;  // nWorlds + 1
  uint32_t *largeList;
  uint32_t largeCount;

  // Pairs: per-body counts (scanned into offsets, nBodies + 1), the pair list (a > b)
  uint32_t *pairCount;
  uint32_t *pairStash;   // PAIR_STASH partners per body from the counting pass
  uint2 *pairs;
  uint32_t pairCapacity;

  // Manifolds: per pair slot (narrowphase output), compacted list (current), last step's list, scan scratch
  Manifold *mfSlots;
  Manifold *manifolds;
  Manifold *prevManifolds;
  Contact *contacts;      // 8 per pair slot
  uint32_t *pointMap;     // the current contact points, densely: manifold index << 3 | point (C_CONTACTS of them)
  float *restitution;     // per body: restitution coefficient in [0, 1] (0: the body does not bounce)
  float4 *bounceVel;      // per body, restitution passes: incoming linear, incoming angular, current linear, angular
  float4 *bounce;         // per manifold slot, restitution passes: (impulse, delta, rebound, mass), lever A, lever B
  uint32_t *bodyDone;     // per body: stamp of its last primal update (dataflow solve)
  uint32_t *dualDone;     // per body: stamp of its last dual of the pairs it finishes (dataflow solve)
  DataflowPlan *dfPlan;
  uint32_t dfSpin;        // polls of a dataflow wait before it gives up (0: give up at the first unmet poll)
  Contact *prevContacts;
  uint32_t *table;        // last step's manifolds by pair hash (value = index + 1)
  uint32_t hashMask;

  // Adjacency: degree/start (nBodies + 1), fill cursor (nBodies), list
  uint32_t *adjStart;
  uint32_t *adjFill;
  uint32_t *adjList;
  uint4 *adjEntry;        // 2 per adjacency entry: what the colouring and the primal read of it (ENTRY_* layout)
  uint32_t adjCapacity;

  // Colouring: state A, state B, rank, histogram (64 × groups + 1), starts (65), bucketed bodies
  uint32_t *colorA;
  uint32_t *colorB;
  uint32_t *colorFlags;   // [0] the colouring state's buffer (0 colorA, 1 colorB), [1 + r] round r has work
  uint32_t *colorPending; // the bodies the mark pass left pending (C_PENDING of them, in no particular order)
  uint32_t *colorRank;
  uint32_t *colorHist;
  uint32_t *colorStart;
  uint32_t *colorBodies;
  uint4 *colorSlot;       // per bucketed body: body, first and end adjacency entry, world
  uint32_t colorGroups;
  uint32_t colorCap;
  // Sleeping
  SleepState *sleep;        // per body
  uint32_t *sleepGlobals;   // per world: wake all, sleeping, woken by an impact, island hit; then bodies woken (gate)
  const uint32_t *gate;     // null, or: the collision refresh's kernels return while *gate == 0
};

/** A shape query (32 B): origin xyz, maximum distance; unit direction xyz, sweep radius (0: a ray). */
struct QueryRec {
  float4 origin;
  float4 direction;
};

/** Which bodies a batch of queries sees: body group & mask != 0; with membership, the body's mask & membership != 0;
 * never the ignored ones. */
struct QueryFilter {
  uint32_t ignore[16];
  uint32_t ignoreCount, mask, membership, useMembership;
};

/** Contact event storage: head words (events written, a set overflowed, 2 spare), then two hash sets of mask + 1 slots
 * (2 words: a + 1, b); events of 3 float4 (kind, a, b, step bits; point, impulse; normal); watch words per body (bit
 * 0: watched, above it: lifetime epoch) now and at the last step. */
constexpr uint32_t CW_HEAD = 4, CW_BEGIN = 1, CW_END = 2;
struct ContactWatchDev {
  uint32_t *table;
  float4 *events;
  uint32_t *watch;
  uint32_t *previousWatch;
  uint32_t mask, capacity;
};

/** A gated launch (the collision refresh after sleepers woke) has nothing to do. */
__device__ __forceinline__ bool gatedOff(const DeviceWorld &w) { return w.gate != nullptr && *w.gate == 0u; }

/** The writable parameters of a world (its step counter, body range and large-body radius are kept). */
struct ParamEdit {
  float dt, gravity, betaLin, betaAng, gamma, alpha;
  float up[3];
  uint32_t flags;
};

constexpr int COLOR_WG = 256;

/**
 * Primal lanes: a body with at most LANES adjacency entries takes LANES lanes (one entry per lane), a heavier one
 * HEAVY_LANES. Each colour is bucketed as (heavy bodies, light bodies): bucket key 2 × colour + (light ? 1 : 0), so
 * colorStart[2c], colorStart[2c + 1], colorStart[2c + 2] bound colour c's heavy and light bodies.
 */
constexpr uint32_t LANES = 8;
constexpr uint32_t HEAVY_LANES = 32;
constexpr uint32_t COLOR_KEYS = 2 * MAX_COLORS;
constexpr uint32_t COLOR_FLAG_WORDS = 64;  // colorFlags: rounds past 61 always run

/**
 * Adjacency entry record, written once per step after the per-body sort (2 × uint4 = 32 B):
 *   [0].x partner body (ENTRY_NONE for a world anchor), [0].y manifold: first contact / joint: joint id,
 *   [0].z flags: contact count (bits 0-3), ENTRY_IS_A, ENTRY_JOINT, ENTRY_DYNAMIC (partner is dynamic),
 *   [0].w manifold friction (f32 bits); [1].xyz manifold normal (f32 bits), [1].w the normal's tangentScale (f32
 *   bits, formed once per step: every contact row of the pair takes its basis from it).
 * Reading it replaces the chain adjacency id → manifold/joint record → partner body in the colouring and the primal.
 */
constexpr uint32_t ENTRY_NONE = 0xffffffffu;
constexpr uint32_t ENTRY_IS_A = 1u << 4, ENTRY_JOINT = 1u << 5, ENTRY_DYNAMIC = 1u << 6;
/** [0].z bits 8-15: a dynamic partner's colour, written after the colouring for the dataflow solve. */
constexpr uint32_t ENTRY_COLOR_SHIFT = 8;

}  // namespace avbd
