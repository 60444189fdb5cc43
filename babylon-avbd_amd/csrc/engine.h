// Host engine: owns the device memory of a batch of worlds (one world is a batch of one) and orders the step's kernels on
// one HIP stream, optionally replaying them from a captured hipGraph (one graph per contact-buffer parity).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "engine_types.h"
#include "hploc.h"

namespace avbd {

/** One world of a batch as the host builds it: reference records plus parameters. */
struct HostWorld {
  std::vector<Body> bodies;
  std::vector<Joint> joints;
  std::vector<int4> info;                       // type, a, b (world-local; a = -1: world)
  std::vector<std::pair<uint32_t, uint32_t>> ignorePairs;  // world-local body pairs that never collide
  std::vector<uint2> filters;                   // empty: collide with everything
  std::vector<uint4> hulls;                     // hull records (headers at the words body shape codes name)
  WorldParams params;
  uint32_t iterations = 10;
};

struct EngineOptions {
  uint32_t colorCap = 16;       // colours dispatched per iteration (and the colouring's cap)
  uint32_t colorRounds = 16;    // Jones–Plassmann rounds per step (even)
  uint32_t pairCapacity = 0;    // 0: chosen from the body count
  bool useGraph = false;
  bool adaptive = false;        // adapt colour cap/rounds and storage from counters read back asynchronously
  int device = 0;
  uint32_t broadphase = 0;      // 0 auto (H-PLOC for a world or batch of >= 10,000 bodies, else grid), 1 grid, 2 H-PLOC
  uint32_t rebuildInterval = 64;  // H-PLOC: rebuild every this many steps, refit in between
  uint32_t solver = 0;          // 0 auto (batches of small worlds: one workgroup per world; else colour launches),
                                // 1 one launch per colour, 2 dataflow (one launch for all iterations; its grid must
                                // fit the device beside the other dataflow solves running: dataflowShare), 3 one
                                // workgroup per world
  uint32_t worldBlock = 0;      // threads per world's workgroup with solver 3 (128, 256, 512, 1024; 0: 256)
  uint32_t dataflowShare = 0;   // solver 2: the dataflow grid takes 1/share of the device's resident workgroups (0, 1:
                                // all), so `share` engines' solves can run at once without waiting on one another
  uint32_t coloringRounds = 0;  // 0 auto (listed while few bodies are left pending, with adaptive), 1 one launch per
                                // round, 2 always listed: one block over the pending bodies (same colours)
};

/** Milliseconds of one profiled step by phase, measured with HIP events on the engine's stream. */
struct PhaseTimes {
  float total, collision, adjacency, coloring, warmStart, primal, dual, velocities;
  uint32_t primalLaunches, dualLaunches;
  float broadphase, narrowphase;  // parts of collision: tree or grid and pair search; contacts and compaction
  float pairSearch;               // part of broadphase: the pair search alone
};

/** Everything the next step of a batch depends on, in host memory (a world evicted from HBM keeps this). */
struct SavedState {
  std::vector<Body> bodies;
  std::vector<Joint> joints;
  std::vector<float4> refs;
  std::vector<uint32_t> colors;
  std::vector<Manifold> manifolds;
  std::vector<Contact> contacts;
  std::vector<uint32_t> steps;
};

struct Counters {
  uint32_t pairs;  // pairs the broadphase found (above the capacity when it overflowed), as the reference's counter
  uint32_t contacts, manifolds, overflow, clashes, colors, prevManifolds, pivot64;
  uint32_t solveAborts;  // dataflow solve waits given up in the last step (0 unless something is wrong)
  uint32_t pending = 0;  // bodies the colouring's mark pass left pending
};

class Engine {
 public:
  Engine(const std::vector<HostWorld> &worlds, const EngineOptions &opt);
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  void step(uint32_t n);
  /** n steps with HIP events around each on the engine's stream; ms[k] = step k's device time. Waits. */
  void stepTimed(uint32_t n, float *ms);
  /** stepTimed in two halves: enqueue n event-bracketed steps; later wait and read their times (returns n). */
  void stepTimedBegin(uint32_t n);
  uint32_t stepTimedEnd(float *ms);
  uint32_t timedPending() const { return timedSteps_; }
  void synchronize();
  /** One step (not from a graph) with HIP events around each phase and each primal/dual launch. */
  PhaseTimes profileStep();

  uint32_t bodyCount() const { return nBodies_; }
  uint32_t jointCount() const { return nJoints_; }
  uint32_t worldCount() const { return nWorlds_; }
  uint32_t iterations() const { return iterations_; }
  /** Which solve the step runs: 1 colour launches, 2 dataflow, 3 one workgroup per world. */
  uint32_t solverKind() const { return useWorlds_ ? 3u : useDataflow_ ? 2u : 1u; }
  hipStream_t stream() const { return stream_; }
  const DeviceWorld &view() const { return views_[parity_]; }
  const std::vector<uint32_t> &worldFirstBody() const { return firstBody_; }
  const std::vector<uint32_t> &worldFirstJoint() const { return firstJoint_; }

  // Reads (synchronous; copy into caller buffers)
  void readBodies(uint32_t first, uint32_t count, Body *out);
  void writeBodies(uint32_t first, uint32_t count, const Body *in);
  void readJoints(uint32_t first, uint32_t count, Joint *out);
  void writeJoints(uint32_t first, uint32_t count, const Joint *in);
  Counters readCounters();
  /** The last step's pair list (a > b) in emission order. */
  std::vector<uint2> readPairs();
  /** The last step's manifolds (first = index into readContacts) and contact slots. */
  void readContacts(std::vector<Manifold> &manifolds, std::vector<Contact> &contacts);
  std::vector<uint32_t> readColors();
  std::vector<float4> readRefs();
  std::vector<WorldParams> readWorldParams();
  /** Adjacency of the last step: start (bodies + 1), list. */
  void readAdjacency(std::vector<uint32_t> &start, std::vector<uint32_t> &list);

  /** Seed the persistent state: reference poses, colour state, last step's manifolds/contacts, step counters. */
  void writeState(const float4 *refs, const uint32_t *colors, const std::vector<Manifold> &manifolds,
                  const std::vector<Contact> &contacts, const uint32_t *steps);

  /** Copy the persistent state to the host; restore() on an engine built from the same worlds continues the run
   * exactly where it stopped. */
  SavedState save();
  void restore(const SavedState &s);

  /**
   * Body commands, 12 words each (body, kind, has point, 0 | value xyzw | point xyzw; bodies index the whole batch),
   * checked on the host, then applied on the stream in submission order: after the steps enqueued before, before the
   * steps enqueued after. A teleport also drops the warm start of the last step's manifolds on the teleported body.
   */
  void applyCommands(const uint32_t *words, uint32_t n);
  /** Collision filters (groups, collides-with) of n bodies, from the next step on. */
  void setFilters(const uint32_t *indices, const uint32_t *groups, const uint32_t *masks, uint32_t n);
  /** Parameters of world k of the batch, from the next step on. */
  void setParams(uint32_t world, const ParamEdit &e);
  /** Restitution coefficients in [0, 1] of n bodies, from the next step on (the bounce passes run while any body's
   * coefficient is above 0). */
  void setRestitution(const uint32_t *indices, const float *values, uint32_t n);
  void getRestitution(float *out) const;
  /**
   * Sleeping (off until enabled; enabling again keeps the first settings): speed threshold (linear speed + angular
   * speed × bounding radius) and quiet time before sleeping, island mode (-1: islands when worlds average 4096 bodies
   * or more), and start asleep (island mode: the next step puts still, grounded groups to sleep). Eligibility comes from
   * setSleepEnabled (every body until then).
   */
  void enableSleeping(float speed, float delay, int islands, bool startAsleep);
  /** Wake every body (masses back) and stop sleeping. */
  void disableSleeping();
  bool sleeping() const { return sleeping_; }
  bool sleepIslands() const { return sleepIslands_; }
  /** Wake every body now (on the stream, before the next step). */
  void wakeAll();
  /** Hold n bodies awake through the next step; in island mode their islands wake at its start. */
  void wakeBodies(const uint32_t *indices, uint32_t n);
  /** Sleep eligibility of n bodies (0 or 1); a body allowed to sleep enables sleeping; every body wakes. */
  void setSleepEnabled(const uint32_t *indices, const uint8_t *flags, uint32_t n);
  /** World k after the last step: out[0] bodies asleep, out[1] a wake-all request pending. */
  void sleepStats(uint32_t world, uint32_t *out);
  /** Every body's sleep state, 8 words each (mass and quiet as f32 bits, asleep, support, eligible, parent, wake,
   * flags), and the 4 words per world of sleep globals. */
  void readSleep(uint32_t *states, uint32_t *globals);
  /**
   * Shape queries against the current poses (queued edits first, in stream order): n queries of 8 floats (origin xyz,
   * maximum distance, direction xyz (normalised here), sweep radius); hits of 8 floats (normal xyz, distance or -1,
   * body index bits (~0: none), 0, 0, 0).
   */
  void cast(const float *queries, uint32_t n, const QueryFilter &filter, float *hits);
  /** cast() without allocating: the queries' records, best words and hits staging live in a fixed device array shared
   * by every call on the device (at most NOALLOC_MAX_QUERIES queries per call; overlapping calls on one device race). */
  void castNoAlloc(const float *queries, uint32_t n, const QueryFilter &filter, float *hits);
  /** Contact events (ContactWatch): storage for maxPairs touching pairs per step and maxEvents unread events, made on
   * the first call (every body watched); later calls keep it. */
  void watchContacts(uint32_t maxPairs, uint32_t maxEvents);
  bool watchingContacts() const { return cw_.table != nullptr; }
  /** Which bodies report their contacts (flags 0 or 1). */
  void setWatched(const uint32_t *indices, const uint8_t *flags, uint32_t n);
  /** New lifetimes for n bodies (sorted, unique): their pairs begin again and do not end. */
  void resetContactHistory(const uint32_t *indices, uint32_t n);
  /** Unread events (12 words each: kind, a, b, step, point xyz, impulse, normal xyz, 0) into out (room for cap);
   * returns how many were copied; *written = all events since the last read, *full = a set overflowed. Clears the log. */
  uint32_t readContactEvents(uint32_t *out, uint32_t cap, uint32_t *written, uint32_t *full);
  /** Constraint edits (applied on the stream after the steps enqueued before): release slots (zeroed, type none: the
   * bodies collide again; released or unknown slots are ignored), disable one (penalties, stiffnesses and multipliers
   * zeroed), move a world joint's anchor, change a motor's speed and/or torque bound, set a spring's break strain and
   * force (Infinity: unbreakable), make a slot a motor (type 3: lo speed, hi torque bound) or limit (type 4: lo, hi
   * angles) about body-local axes, its frames captured from the current poses. Sleeping bodies involved wake. */
  void releaseJoints(const uint32_t *slots, uint32_t n);
  void disableConstraint(uint32_t slot);
  void setWorldAnchor(uint32_t slot, const float *point);
  void setMotor(uint32_t slot, const float *speed, const float *maxTorque);
  void setSpringMaterial(uint32_t slot, float breakStrain, float breakForce);
  void setAngularConstraint(uint32_t slot, int32_t type, const float *axisA, const float *axisB, float lo, float hi);
  /** Type, body A (-1: world), body B of a constraint slot (batch indices). */
  int4 jointInfo(uint32_t slot) const;
  /** Solver iterations per step (all worlds of the batch). */
  void setIterations(uint32_t n);

  void setColorCap(uint32_t cap);
  void setColorRounds(uint32_t rounds);
  uint32_t colorCap() const { return colorCap_; }
  uint32_t colorRounds() const { return colorRounds_; }
  uint32_t pairCapacity() const { return pairCap_; }
  bool usesBvh() const { return useBvh_; }
  /** Grow contact storage / colour cap from the last counters (the reference's adapt); returns true if it changed. */
  bool adapt();
  /** The reference's adapt policy applied to given counters (solver.ts:1484-1519). */
  bool adaptFrom(const Counters &c);

  /** Device bytes held by this engine. */
  size_t deviceBytes() const { return deviceBytes_; }

 private:
  void allocate();
  void captureFrames(const std::vector<uint32_t> &slots);
  /** (slot, word, value bits) triples, at most one per (slot, word); words 32..34 set the info word. */
  void editJointWords(const std::vector<uint32_t> &triples);
  void wakeConstraint(uint32_t slot);
  void allocateContacts(uint32_t pairCap);
  void encodeStep(const DeviceWorld &w, bool rebuild, uint32_t sleepForce = 0);
  void uploadWatch();
  SleepParams sleepParams(uint32_t force) const { return SleepParams{sleepSpeed_, sleepDelay_, sleepIslands_ ? 1u : 0u, force}; }
  /** Impacts on sleepers after the narrowphase, then the pairs and narrowphase again, gated on a body having woken. */
  void encodeSleepImpact(const DeviceWorld &w, uint32_t force);
  void encodeCollision(const DeviceWorld &w, bool rebuild);
  void allocateBvh(const std::vector<uint32_t> &counts);
  void *dalloc(size_t bytes);
  template <class T> T *dallocT(size_t count) { return static_cast<T *>(dalloc(count * sizeof(T))); }
  void resetGraphs();
  void pollCounters();
  /** After a step whose dataflow solve gave up a wait: colour launches from the next step on. */
  void noteAborts(const Counters &c);
  /** Copy words to a device buffer on the stream (through one of two pinned staging buffers); returns the buffer. */
  uint32_t *uploadEdit(const uint32_t *words, size_t n);

  int device_;
  hipStream_t stream_ = nullptr;
  uint32_t nBodies_ = 0, nJoints_ = 0, nWorlds_ = 0, noCollideCount_ = 0;
  uint32_t iterations_ = 10, colorCap_ = 16, colorRounds_ = 16;
  uint32_t pairCap_ = 0, tableSize_ = 0, hashSize_ = 0, colorGroups_ = 0, maxDynamicBodies_ = 0;
  bool useGraph_ = false;
  int parity_ = 0;
  std::vector<uint32_t> firstBody_, firstJoint_;
  std::vector<void *> allocations_;
  std::vector<void *> contactAllocations_;
  size_t deviceBytes_ = 0, contactBytes_ = 0;

  DeviceWorld base_{};
  DeviceWorld views_[2]{};
  Manifold *manifoldBuf_[2] = {nullptr, nullptr};
  Contact *contactBuf_[2] = {nullptr, nullptr};
  hipGraphExec_t graphs_[4] = {nullptr, nullptr, nullptr, nullptr};  // parity × (refit, rebuild)
  std::vector<int4> infoHost_;      // host copy of the constraint info words (type, a, b, 0)
  std::vector<uint4> hullsHost_;    // every world's hull records, concatenated (headers rebased)
  std::vector<char> hullHeader_;    // per word of hullsHost_: a checked record's header starts there
  std::vector<float> restitution_;  // host copy of the per-body coefficients
  uint32_t bouncing_ = 0;           // bodies with a coefficient above 0
  bool sleeping_ = false, sleepIslands_ = false;
  float sleepSpeed_ = 0.03f, sleepDelay_ = 0.5f;
  uint32_t sleepForce_ = 0;           // force bits of the next step (2: start asleep)
  std::vector<uint8_t> sleepEligible_;  // host copy of the per-body eligibility
  QueryRec *queryDev_ = nullptr;        // shape query batches (allocated on first use, grown as needed)
  uint32_t *queryBest_ = nullptr;
  float4 *queryHits_ = nullptr;
  uint32_t queryCap_ = 0;
  ContactWatchDev cw_{};
  std::vector<uint32_t> watchWords_;
  bool listedRounds_ = false;  // colouring rounds over the pending list in one block (few pending bodies lately)
  uint32_t coloringRounds_ = 0;
  uint32_t listedVotes_ = 0;
  bool useBvh_ = false;
  bool useWorlds_ = false;     // one workgroup per world solves the batch (solver 3)
  uint32_t worldBlock_ = 256;
  uint32_t *worldOrder_ = nullptr, *worldKeyStart_ = nullptr, *worldJointStart_ = nullptr;
  bool useDataflow_ = false;
  bool wantDataflow_ = false;  // asked for and possible; used while the arrays fit 32-bit offsets
  uint32_t dfGrid_ = 0;
  uint32_t rebuildInterval_ = 64;
  uint64_t stepIndex_ = 0;
  Bvh bvh_{};
  BvhPlan plan_;
  // Asynchronous counter readback (pinned host copies, one per in-flight slot)
  bool adaptive_ = false;
  uint32_t *hostCounters_ = nullptr;
  hipEvent_t counterEvents_[4] = {nullptr, nullptr, nullptr, nullptr};
  bool counterPending_[4] = {false, false, false, false};
  uint32_t counterSlot_ = 0;
  uint32_t shrinkVotes_ = 0;
  std::vector<hipEvent_t> timingEvents_;
  uint32_t timedSteps_ = 0;
  // Edits: device upload buffer, pinned staging pair with the events of their last copies, teleport marks per body
  uint32_t *editDev_ = nullptr;
  size_t editDevWords_ = 0;
  uint32_t *editHost_[2] = {nullptr, nullptr};
  size_t editHostWords_[2] = {0, 0};
  hipEvent_t editEvents_[2] = {nullptr, nullptr};
  bool editPending_[2] = {false, false};
  int editSlot_ = 0;
  uint32_t *marks_ = nullptr;
};

/**
 * Auto solver: batches whose worlds have at most this many bodies each are solved one workgroup per world. A world's
 * solve then runs on one compute unit, whose issue rate bounds its latency (about 2 ms for 2,048 boxes at 10
 * iterations); up to a few hundred bodies that beats the colour launches' fixed cost.
 */
constexpr uint32_t WORLDS_AUTO_MAX = 256;

/** Largest bounding radius still sized into grid cells (reference/three-avbd/src/avbd3d/gpu/radii.ts:47-55). */
double largestSmall(std::vector<double> radii);

}  // namespace avbd
