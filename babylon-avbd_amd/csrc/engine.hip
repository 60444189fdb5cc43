// Host engine of the AVBD 3D step: batch assembly, HBM allocation, the per-step launch order, hipGraph replay, reads,
// state seeding and capacity adaptation. The launch order is the reference's GpuSolver3D.step
// (reference/three-avbd/src/avbd3d/gpu/solver.ts:1281-1417): collision → adjacency → colouring → warm starts →
// iterations × (primal per colour, dual) → velocities; its capacity policy is GpuSolver3D.adapt (solver.ts:1484-1519).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <tuple>

#include "engine.h"
#include "kernels.h"
#include "pairtest.h"

namespace avbd {

#define HIP_CHECK(x)                                                                                         \
  do {                                                                                                       \
    hipError_t err_ = (x);                                                                                   \
    if (err_ != hipSuccess)                                                                                  \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err_) + " at " #x);             \
  } while (0)

/** The header word a hull shape code names, checked to be a whole number inside a buffer of `words` words. */
static uint32_t hullHeaderOf(float code, uint32_t words) {
  const float h = code - SHAPE_HULL;
  if (!(h >= 0.f) || h != std::floor(h) || !(h + 1.f < float(words)))
    throw std::invalid_argument("a hull body's shape code names no hull record of its world");
  return uint32_t(h);
}

/**
 * A hull record's every offset and index inside its buffer, with the narrowphase's limits (shapes.ts:59-61): at most
 * 32 vertices, 255 faces and 255 edges, faces of 3 to 32 vertices.
 */
static void checkHullRecord(const std::vector<uint4> &buf, uint32_t h) {
  const uint64_t words = buf.size();
  auto bad = [](const char *what) { throw std::invalid_argument(std::string("hull record: ") + what); };
  if (uint64_t(h) + 2 > words) bad("header outside the buffer");
  const uint4 h0 = buf[h], h1 = buf[h + 1];
  const uint32_t vs = h0.x, V = h0.y, fs = h0.z, F = h0.w, es = h1.x, E = h1.y, is = h1.z;
  if (V < 4 || V > 32 || F < 4 || F > 255 || E < 6 || E > 255) bad("vertex, face or edge count out of range");
  if (uint64_t(vs) + V > words || uint64_t(fs) + 2ull * F > words || uint64_t(es) + E > words || is >= words)
    bad("a section lies outside the buffer");
  for (uint32_t f = 0; f < F; f++) {
    const uint4 face = buf[fs + 2ull * f + 1];
    if (face.y < 3 || face.y > 32) bad("a face has fewer than 3 or more than 32 vertices");
    const uint64_t last = uint64_t(is) * 4 + face.x + face.y;
    if (last > words * 4) bad("a face's vertex indices lie outside the buffer");
    for (uint32_t j = 0; j < face.y; j++) {
      const uint64_t k = uint64_t(face.x) + j;
      const uint4 u = buf[is + k / 4];
      const uint32_t v = k % 4 == 0 ? u.x : (k % 4 == 1 ? u.y : (k % 4 == 2 ? u.z : u.w));
      if (v >= V) bad("a face names a vertex the hull does not have");
    }
  }
  for (uint32_t e = 0; e < E; e++) {
    const uint4 edge = buf[es + e];
    if (edge.x >= V || edge.y >= V || edge.z >= F || edge.w >= F) bad("an edge names a vertex or face out of range");
  }
}

static uint32_t pow2AtLeast(uint64_t n) {
  uint64_t p = 2;
  while (p < n) p <<= 1;
  return uint32_t(p);
}

double largestSmall(std::vector<double> radii) {
  const size_t n = radii.size();
  const double LARGE_FACTOR = 2;
  const size_t MAX_LARGE = 64;
  double median = 1;
  if (n > 0) {
    std::nth_element(radii.begin(), radii.begin() + (n >> 1), radii.end());
    median = radii[n >> 1];
  }
  double maxSmall = 0;
  for (double r : radii)
    if (r <= LARGE_FACTOR * median) maxSmall = std::max(maxSmall, r);
  if (n > MAX_LARGE) {
    std::nth_element(radii.begin(), radii.begin() + (n - MAX_LARGE - 1), radii.end());
    maxSmall = std::max(maxSmall, radii[n - MAX_LARGE - 1]);
  }
  return maxSmall;
}

void *Engine::dalloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  HIP_CHECK(hipMalloc(&p, bytes));
  HIP_CHECK(hipMemsetAsync(p, 0, bytes, stream_));
  allocations_.push_back(p);
  deviceBytes_ += bytes;
  return p;
}

Engine::Engine(const std::vector<HostWorld> &worlds, const EngineOptions &opt) {
  if (worlds.empty()) throw std::invalid_argument("a batch needs at least one world");
  device_ = opt.device;
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  useGraph_ = opt.useGraph;
  adaptive_ = opt.adaptive;
  if (adaptive_) {
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&hostCounters_), 4 * COUNTER_WORDS * sizeof(uint32_t),
                            hipHostMallocDefault));
    for (auto &ev : counterEvents_) HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  colorCap_ = std::min<uint32_t>(std::max<uint32_t>(opt.colorCap, 1), MAX_COLORS);
  colorRounds_ = opt.colorRounds + (opt.colorRounds & 1u);
  nWorlds_ = uint32_t(worlds.size());
  iterations_ = worlds[0].iterations;

  std::vector<Body> bodies;
  std::vector<uint32_t> bodyWorld;
  std::vector<Joint> joints;
  std::vector<int4> info;
  std::vector<uint2> filters;
  std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> noCollide;
  std::vector<WorldParams> params(nWorlds_);
  std::vector<uint32_t> largeStart(nWorlds_ + 1, 0), largeList;
  double cellSize = 1e-3;
  for (uint32_t k = 0; k < nWorlds_; k++) {
    const HostWorld &hw = worlds[k];
    const uint32_t hullBase = uint32_t(hullsHost_.size());
    if (uint64_t(hullBase) + hw.hulls.size() >= (1ull << 24) - 4)
      throw std::invalid_argument("hull buffers beyond 2^24 words (shape codes are exact f32 integers)");
    if (hw.iterations != iterations_) throw std::invalid_argument("every world of a batch takes the same iterations");
    const uint32_t base = uint32_t(bodies.size());
    const uint32_t jbase = uint32_t(joints.size());
    const uint32_t n = uint32_t(hw.bodies.size());
    firstBody_.push_back(base);
    firstJoint_.push_back(jbase);
    bodies.insert(bodies.end(), hw.bodies.begin(), hw.bodies.end());
    bodyWorld.insert(bodyWorld.end(), n, k);
    // Hull records: each one a body names is checked, then the world's buffer is appended with its offsets rebased
    std::vector<char> header(hw.hulls.size(), 0);
    for (uint32_t i = 0; i < n; i++) {
      const float code = hw.bodies[i].angVel.w;
      if (!(code >= SHAPE_HULL)) continue;
      const uint32_t h = hullHeaderOf(code, uint32_t(hw.hulls.size()));
      if (!header[h]) checkHullRecord(hw.hulls, h);
      header[h] = 1;
      bodies[base + i].angVel.w = SHAPE_HULL + float(hullBase + h);
    }
    for (size_t q = 0; q + 1 < hw.hulls.size(); q++)
      if (header[q] && header[q + 1]) throw std::invalid_argument("hull record: two headers overlap");
    for (size_t q = 0; q < hw.hulls.size(); q++) {
      uint4 u = hw.hulls[q];
      if (header[q]) {
        u.x += hullBase;
        u.z += hullBase;
      }
      hullsHost_.push_back(u);
      hullHeader_.push_back(header[q]);
    }
    for (size_t q = 0; q < hw.hulls.size(); q++)
      if (header[q]) {
        uint4 &h1 = hullsHost_[hullBase + q + 1];
        h1.x += hullBase;
        h1.z += hullBase;
      }
    if (!hw.filters.empty() && hw.filters.size() != n) throw std::invalid_argument("filters: one pair per body");
    for (uint32_t i = 0; i < n; i++) filters.push_back(hw.filters.empty() ? make_uint2(~0u, ~0u) : hw.filters[i]);
    for (size_t j = 0; j < hw.joints.size(); j++) {
      int4 f = hw.info[j];
      if (f.z < 0 || uint32_t(f.z) >= n || f.y >= int(n) || f.y == f.z)
        throw std::invalid_argument("joint endpoints must be distinct bodies of the world");
      if (f.x < T_NONE || f.x > T_LIMIT) throw std::invalid_argument("unknown joint type");
      if (f.x == T_SPRING && f.y < 0) throw std::invalid_argument("A spring requires two bodies");
      int a = f.y < 0 ? -1 : int(base) + f.y;
      int b = int(base) + f.z;
      joints.push_back(hw.joints[j]);
      info.push_back(make_int4(f.x, a, b, 0));
      if (a >= 0) noCollide.emplace_back(uint32_t(std::max(a, b)), uint32_t(std::min(a, b)), jbase + uint32_t(j));
    }
    for (auto &pr : hw.ignorePairs) {
      if (pr.first >= n || pr.second >= n || pr.first == pr.second)
        throw std::invalid_argument("ignored pairs must be distinct bodies of the world");
      uint32_t a = base + pr.first, b = base + pr.second;
      noCollide.emplace_back(std::max(a, b), std::min(a, b), 0xffffffffu);
    }
    // Large/small classification and grid cell size (solver.ts:860-881, radii.ts), per world
    std::vector<double> radii(n);
    for (uint32_t i = 0; i < n; i++) radii[i] = hw.bodies[i].moment.w;
    double maxSmall = largestSmall(radii);
    double minLarge = INFINITY;
    for (double r : radii)
      if (r > maxSmall) minLarge = std::min(minLarge, r);
    double threshold = std::isinf(minLarge) ? maxSmall * 1.5 : (maxSmall + minLarge) / 2;
    cellSize = std::max(cellSize, std::max(2 * maxSmall * (1 + 1e-4), 1e-3));
    WorldParams p = hw.params;
    p.firstBody = base;
    p.bodyCount = n;
    p.maxSmall = float(threshold);
    params[k] = p;
    largeStart[k] = uint32_t(largeList.size());
    for (uint32_t i = 0; i < n; i++)
      if (double(hw.bodies[i].moment.w) > double(p.maxSmall)) largeList.push_back(base + i);
  }
  largeStart[nWorlds_] = uint32_t(largeList.size());
  std::sort(noCollide.begin(), noCollide.end(),
            [](auto &x, auto &y) { return std::tie(std::get<0>(x), std::get<1>(x)) < std::tie(std::get<0>(y), std::get<1>(y)); });
  nBodies_ = uint32_t(bodies.size());
  nJoints_ = uint32_t(joints.size());
  noCollideCount_ = uint32_t(noCollide.size());
  for (auto &b : bodies)
    if (b.size.w > 0.f) maxDynamicBodies_++;

  pairCap_ = opt.pairCapacity ? opt.pairCapacity : std::max<uint32_t>(4096, 8 * nBodies_);
  tableSize_ = pow2AtLeast(2ull * std::max<uint32_t>(nBodies_, 1));
  colorGroups_ = (std::max<uint32_t>(nBodies_, 1) + COLOR_WG - 1) / COLOR_WG;
  rebuildInterval_ = std::max<uint32_t>(opt.rebuildInterval, 1);
  coloringRounds_ = opt.coloringRounds;
  listedRounds_ = coloringRounds_ == 2;
  if (opt.broadphase == 2) useBvh_ = true;
  else if (opt.broadphase == 1) useBvh_ = false;
  else useBvh_ = nBodies_ >= 10000;  // one world or a batch (each world its own tree) of 10,000 bodies or more
  if (opt.solver > 3)
    throw std::invalid_argument("solver must be 0 (auto), 1 (colour launches), 2 (dataflow) or 3 (workgroup per world)");
  if (opt.worldBlock != 0 && opt.worldBlock != 128 && opt.worldBlock != 256 && opt.worldBlock != 512 &&
      opt.worldBlock != 1024)
    throw std::invalid_argument("world_block must be 0, 128, 256, 512 or 1024");
  if (opt.solver == 2) {
    dfGrid_ = dataflow_grid();
    // A share of the device: engines whose solves run together (each on its own stream) must not hold more
    // workgroups than fit at once, or each would wait for its own workgroups that cannot start
    if (opt.dataflowShare > 1) dfGrid_ = std::max<uint32_t>(1, dfGrid_ / opt.dataflowShare);
    useDataflow_ = dfGrid_ > 0;
  }
  wantDataflow_ = useDataflow_;
  // Auto: a batch of worlds of at most WORLDS_AUTO_MAX bodies each is solved one workgroup per world
  uint32_t largestWorld = 0;
  for (const HostWorld &hw : worlds) largestWorld = std::max<uint32_t>(largestWorld, uint32_t(hw.bodies.size()));
  useWorlds_ = opt.solver == 3 || (opt.solver == 0 && nWorlds_ > 1 && largestWorld <= WORLDS_AUTO_MAX);
  worldBlock_ = opt.worldBlock ? opt.worldBlock : 256;

  DeviceWorld &w = base_;
  w.nBodies = nBodies_;
  w.nJoints = nJoints_;
  w.nWorlds = nWorlds_;
  w.noCollideCount = noCollideCount_;
  w.roundedShapes = 0;
  w.hullShapes = 0;
  for (const Body &b : bodies) {
    if (b.angVel.w == SHAPE_SPHERE || b.angVel.w == SHAPE_CAPSULE) w.roundedShapes = 1;
    if (b.angVel.w >= SHAPE_HULL) w.hullShapes = 1;
  }
  w.hulls = dallocT<uint4>(std::max<size_t>(hullsHost_.size(), 1));
  w.hullPairs = nullptr;  // with the pair storage
  w.bodies = dallocT<Body>(nBodies_);
  w.bodyWorld = dallocT<uint32_t>(nBodies_);
  w.worlds = dallocT<WorldParams>(nWorlds_);
  w.refs = dallocT<float4>(2ull * nBodies_);
  w.irotInv = dallocT<float4>(nBodies_);
  w.restitution = dallocT<float>(nBodies_);
  w.bounceVel = dallocT<float4>(4ull * nBodies_);
  restitution_.assign(nBodies_, 0.f);
  w.sleep = dallocT<SleepState>(nBodies_);
  w.sleepGlobals = dallocT<uint32_t>(4ull * nWorlds_ + 1);
  w.gate = nullptr;
  sleepEligible_.assign(nBodies_, 1u);
  w.bodyDone = dallocT<uint32_t>(nBodies_);
  w.dualDone = dallocT<uint32_t>(nBodies_);
  w.dfPlan = dallocT<DataflowPlan>(1);
  // Polls of a dataflow wait before it gives up (about 65 ms); AVBD_DATAFLOW_SPIN lowers it to force the give-up path
  w.dfSpin = 1u << 16;
  if (const char *spin = std::getenv("AVBD_DATAFLOW_SPIN")) w.dfSpin = uint32_t(std::strtoul(spin, nullptr, 10));
  w.filters = dallocT<uint2>(nBodies_);
  w.joints = dallocT<Joint>(nJoints_);
  w.info = dallocT<int4>(nJoints_);
  w.noCollide = dallocT<uint32_t>(3ull * noCollideCount_);
  w.ncStart = dallocT<uint32_t>(uint64_t(nBodies_) + 1);
  w.counters = dallocT<uint32_t>(COUNTER_WORDS);
  w.gridCount = dallocT<uint32_t>(uint64_t(tableSize_) + 2);
  w.gridCursor = dallocT<uint32_t>(tableSize_);
  w.gridSorted = dallocT<uint32_t>(nBodies_);
  w.cellOf = dallocT<int4>(nBodies_);
  w.sortedPR = dallocT<float4>(nBodies_);
  w.sortedCell = dallocT<int4>(nBodies_);
  w.tableMask = tableSize_ - 1;
  w.cellSize = float(cellSize);
  w.largeStart = dallocT<uint32_t>(nWorlds_ + 1);
  w.largeList = dallocT<uint32_t>(largeList.size());
  w.pairCount = dallocT<uint32_t>(uint64_t(nBodies_) + 1);
  w.pairStash = dallocT<uint32_t>(uint64_t(nBodies_) * PAIR_STASH);
  w.largeCount = uint32_t(largeList.size());
  w.adjStart = dallocT<uint32_t>(uint64_t(nBodies_) + 1);
  w.adjFill = dallocT<uint32_t>(nBodies_);
  w.colorA = dallocT<uint32_t>(nBodies_);
  w.colorB = dallocT<uint32_t>(nBodies_);
  w.colorFlags = dallocT<uint32_t>(COLOR_FLAG_WORDS);
  w.colorPending = dallocT<uint32_t>(nBodies_);
  w.colorRank = dallocT<uint32_t>(nBodies_);
  w.colorHist = dallocT<uint32_t>(uint64_t(COLOR_KEYS) * colorGroups_ + 1);
  w.colorStart = dallocT<uint32_t>(COLOR_KEYS + 1);
  w.colorBodies = dallocT<uint32_t>(nBodies_);
  w.colorSlot = dallocT<uint4>(nBodies_);
  w.colorGroups = colorGroups_;
  w.colorCap = colorCap_;
  if (useWorlds_) {
    // Worlds in launch order: most bodies first (ties by index), so the longest solves start first
    std::vector<uint32_t> order(nWorlds_);
    for (uint32_t k = 0; k < nWorlds_; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return worlds[a].bodies.size() > worlds[b].bodies.size(); });
    worldOrder_ = dallocT<uint32_t>(nWorlds_);
    worldKeyStart_ = dallocT<uint32_t>((uint64_t(nWorlds_) + 1) * COLOR_KEYS);
    HIP_CHECK(hipMemcpyAsync(worldOrder_, order.data(), order.size() * 4, hipMemcpyHostToDevice, stream_));
    if (nJoints_) {
      std::vector<uint32_t> js(firstJoint_);
      js.push_back(nJoints_);
      worldJointStart_ = dallocT<uint32_t>(js.size());
      HIP_CHECK(hipMemcpyAsync(worldJointStart_, js.data(), js.size() * 4, hipMemcpyHostToDevice, stream_));
    }
  }

  auto up = [&](void *dst, const void *src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
  };
  up(w.bodies, bodies.data(), bodies.size() * sizeof(Body));
  up(const_cast<uint4 *>(w.hulls), hullsHost_.data(), hullsHost_.size() * sizeof(uint4));
  up(w.bodyWorld, bodyWorld.data(), bodyWorld.size() * 4);
  up(w.worlds, params.data(), params.size() * sizeof(WorldParams));
  up(w.filters, filters.data(), filters.size() * sizeof(uint2));
  up(w.joints, joints.data(), joints.size() * sizeof(Joint));
  up(w.info, info.data(), info.size() * sizeof(int4));
  infoHost_ = info;
  // Frames of the constraints from the starting poses: rest orientations of angle-locked joints, reference tangents
  // of motors and limits (scenes.js:666-673, appGpuSolver3D.js:500-526)
  std::vector<uint32_t> frameSlots;
  for (uint32_t j = 0; j < nJoints_; j++)
    if ((info[j].x == T_JOINT && joints[j].penAng.w > 0.f) || info[j].x == T_MOTOR || info[j].x == T_LIMIT)
      frameSlots.push_back(j);
  captureFrames(frameSlots);
  std::vector<uint32_t> nc;
  for (auto &t : noCollide) {
    nc.push_back(std::get<0>(t));
    nc.push_back(std::get<1>(t));
    nc.push_back(std::get<2>(t));
  }
  up(w.noCollide, nc.data(), nc.size() * 4);
  std::vector<uint32_t> ncStart(uint64_t(nBodies_) + 1, 0);
  for (auto &t : noCollide) ncStart[std::get<0>(t) + 1]++;
  for (uint32_t i = 0; i < nBodies_; i++) ncStart[i + 1] += ncStart[i];
  up(w.ncStart, ncStart.data(), ncStart.size() * 4);
  up(w.largeStart, largeStart.data(), largeStart.size() * 4);
  up(w.largeList, largeList.data(), largeList.size() * 4);
  std::vector<uint32_t> noColor(nBodies_, NO_COLOR);
  up(w.colorA, noColor.data(), noColor.size() * 4);
  if (useBvh_) {
    std::vector<uint32_t> counts(nWorlds_);
    for (uint32_t k = 0; k < nWorlds_; k++) counts[k] = uint32_t(worlds[k].bodies.size());
    allocateBvh(counts);
  }
  allocateContacts(pairCap_);
  HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::allocateBvh(const std::vector<uint32_t> &counts) {
  for (uint32_t c : counts)
    if (c == 0) throw std::invalid_argument("H-PLOC needs at least one body per world");
  plan_ = makeBvhPlan(counts);
  const uint32_t n = nBodies_;
  const uint32_t tiles = (n + 255) / 256;
  bvh_.nodeCount = plan_.nodeCount;
  bvh_.nodes = dallocT<BvhNode>(plan_.nodeCount);
  bvh_.pairs = dallocT<BvhPair>(plan_.nodeCount);
  bvh_.quads = dallocT<BvhQuad>(plan_.nodeCount - n);
  if (plan_.height > 160) throw std::length_error("BVH too deep for the traversal stack");
  // A walk descends two levels per quad and keeps at most three entries of each quad on its stack
  bvh_.stackDepth = 3 * ((plan_.height + 1) / 2) + 1;
  bvh_.order = dallocT<uint32_t>(n);
  bvh_.rank = dallocT<uint32_t>(n);
  bvh_.maxPos = dallocT<uint32_t>(plan_.nodeCount);
  bvh_.candStash = dallocT<uint32_t>(uint64_t(n) * BVH_CAND);
  bvh_.candCount = dallocT<uint32_t>(n);
  bvh_.candMask = dallocT<uint32_t>(uint64_t(n) * BVH_CAND_WORDS);
  bvh_.parent = dallocT<uint32_t>(plan_.nodeCount);
  bvh_.keys = dallocT<uint64_t>(n);
  bvh_.keysAlt = dallocT<uint64_t>(n);
  bvh_.vals = dallocT<uint32_t>(n);
  bvh_.valsAlt = dallocT<uint32_t>(n);
  bvh_.clustersA = dallocT<uint32_t>(n);
  bvh_.clustersB = dallocT<uint32_t>(n);
  bvh_.roots = dallocT<uint32_t>(nWorlds_);
  bvh_.bounds = dallocT<uint32_t>(6ull * nWorlds_);
  bvh_.sortHist = dallocT<uint32_t>(256ull * tiles + 1);
  bvh_.groups = dallocT<BvhGroup>(std::max<size_t>(plan_.groups.size(), 1));
  if (!plan_.groups.empty())
    HIP_CHECK(hipMemcpyAsync(bvh_.groups, plan_.groups.data(), plan_.groups.size() * sizeof(BvhGroup),
                             hipMemcpyHostToDevice, stream_));
}

void Engine::allocateContacts(uint32_t pairCap) {
  // The dataflow solve addresses bodies, joints and contact slots with 32-bit buffer offsets
  const uint64_t span = std::max<uint64_t>(std::max<uint64_t>(uint64_t(nBodies_) * sizeof(Body),
                                                              uint64_t(nJoints_) * sizeof(Joint)),
                                           uint64_t(pairCap) * MAX_CONTACTS * sizeof(Contact));
  useDataflow_ = wantDataflow_ && span < (1ull << 32);
  // Keep last step's manifolds and contacts (the next step's warm-start source) across a resize
  std::vector<Manifold> keepM;
  std::vector<Contact> keepC;
  uint32_t counters[COUNTER_WORDS] = {0};
  bool had = manifoldBuf_[0] != nullptr;
  if (had) {
    readContacts(keepM, keepC);
    HIP_CHECK(hipMemcpy(counters, base_.counters, sizeof counters, hipMemcpyDeviceToHost));
    for (void *p : contactAllocations_) {
      HIP_CHECK(hipFree(p));
    }
    contactAllocations_.clear();
    deviceBytes_ -= contactBytes_;
  }
  pairCap_ = pairCap;
  hashSize_ = pow2AtLeast(2ull * pairCap_);
  const size_t bytesBefore = deviceBytes_;
  auto alloc = [&](size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 16;
    HIP_CHECK(hipMalloc(&p, bytes));
    HIP_CHECK(hipMemsetAsync(p, 0, bytes, stream_));
    contactAllocations_.push_back(p);
    deviceBytes_ += bytes;
    return p;
  };
  DeviceWorld &w = base_;
  w.pairCapacity = pairCap_;
  w.pairs = static_cast<uint2 *>(alloc(size_t(pairCap_) * sizeof(uint2)));
  w.mfSlots = static_cast<Manifold *>(alloc(size_t(pairCap_) * sizeof(Manifold)));
  w.hullPairs = static_cast<uint32_t *>(alloc(w.hullShapes ? size_t(pairCap_) * sizeof(uint32_t) : 16));
  w.pointMap = static_cast<uint32_t *>(alloc(size_t(pairCap_) * MAX_CONTACTS * sizeof(uint32_t)));
  for (int k = 0; k < 2; k++) {
    manifoldBuf_[k] = static_cast<Manifold *>(alloc(size_t(pairCap_) * sizeof(Manifold)));
    contactBuf_[k] = static_cast<Contact *>(alloc(size_t(pairCap_) * MAX_CONTACTS * sizeof(Contact)));
  }
  w.table = static_cast<uint32_t *>(alloc(size_t(hashSize_) * 4));
  w.hashMask = hashSize_ - 1;
  w.bounce = static_cast<float4 *>(alloc(size_t(pairCap_) * 3 * sizeof(float4)));
  w.adjCapacity = 2 * (nJoints_ + pairCap_);
  w.adjList = static_cast<uint32_t *>(alloc(size_t(w.adjCapacity) * 4));
  w.adjEntry = static_cast<uint4 *>(alloc(size_t(w.adjCapacity) * 2 * sizeof(uint4)));
  uint32_t scanNeed = scan_scratch_words(std::max<uint32_t>({tableSize_ + 2, nBodies_ + 1, pairCap_ + 1,
                                                              COLOR_KEYS * colorGroups_ + 1,
                                                              256 * ((nBodies_ + 255) / 256) + 1}));
  w.scanScratch = static_cast<uint32_t *>(alloc(size_t(scanNeed) * 4));
  contactBytes_ = deviceBytes_ - bytesBefore;
  for (int p = 0; p < 2; p++) {
    views_[p] = base_;
    views_[p].manifolds = manifoldBuf_[p];
    views_[p].prevManifolds = manifoldBuf_[1 - p];
    views_[p].contacts = contactBuf_[p];
    views_[p].prevContacts = contactBuf_[1 - p];
  }
  resetGraphs();
  if (had) {
    // The next step reads them as "previous" from the buffers of views_[parity_]
    HIP_CHECK(hipMemcpyAsync(views_[parity_].prevManifolds, keepM.data(), keepM.size() * sizeof(Manifold),
                             hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipMemcpyAsync(views_[parity_].prevContacts, keepC.data(), keepC.size() * sizeof(Contact),
                             hipMemcpyHostToDevice, stream_));
    counters[C_MANIFOLDS] = uint32_t(keepM.size());
    counters[C_CONTACTS] = uint32_t(keepC.size());
    HIP_CHECK(hipMemcpyAsync(base_.counters, counters, sizeof counters, hipMemcpyHostToDevice, stream_));
  }
  HIP_CHECK(hipStreamSynchronize(stream_));
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  resetGraphs();
  if (queryDev_) (void)hipFree(queryDev_);
  if (queryBest_) (void)hipFree(queryBest_);
  if (queryHits_) (void)hipFree(queryHits_);
  for (void *p : {static_cast<void *>(cw_.table), static_cast<void *>(cw_.events), static_cast<void *>(cw_.watch),
                  static_cast<void *>(cw_.previousWatch)})
    if (p) (void)hipFree(p);
  for (auto &ev : counterEvents_)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : timingEvents_) (void)hipEventDestroy(ev);
  if (hostCounters_) (void)hipHostFree(hostCounters_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (int k = 0; k < 2; k++) {
    if (editEvents_[k]) (void)hipEventDestroy(editEvents_[k]);
    if (editHost_[k]) (void)hipHostFree(editHost_[k]);
  }
  if (editDev_) (void)hipFree(editDev_);
  for (void *p : contactAllocations_) (void)hipFree(p);
  for (void *p : allocations_) (void)hipFree(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::resetGraphs() {
  for (auto &g : graphs_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void Engine::encodeCollision(const DeviceWorld &w, bool rebuild) {
  hipStream_t s = stream_;
  launch_begin_frame(w, s);
  launch_refs_and_hash(w, pairCap_, s);
  if (useBvh_) {
    if (rebuild) launch_bvh_build(w, bvh_, plan_, s);
    else launch_bvh_refit(w, bvh_, plan_, s);
    launch_bvh_pairs(w, bvh_, s);  // writes the pair count
  } else {
    launch_grid(w, s);
    launch_pairs(w, s);
  }
  launch_narrowphase(w, s);
}

void Engine::encodeSleepImpact(const DeviceWorld &w, uint32_t force) {
  hipStream_t s = stream_;
  launch_sleep_impact(w, sleepParams(force), s);
  DeviceWorld g = w;
  g.gate = w.sleepGlobals + 4ull * nWorlds_;
  launch_sleep_refresh_prepare(g, s);
  if (useBvh_) {
    launch_bvh_refit(g, bvh_, plan_, s);  // the woken leaves are dynamic again for the pair search's pruning
    launch_bvh_pairs(g, bvh_, s);
  } else {
    launch_pairs(g, s);
  }
  launch_narrowphase(g, s);
}

void Engine::encodeStep(const DeviceWorld &w, bool rebuild, uint32_t sleepForce) {
  hipStream_t s = stream_;
  if (sleeping_) launch_sleep_before(w, sleepParams(sleepForce), s);
  if (bouncing_) launch_restitution_before(w, s);
  encodeCollision(w, rebuild);
  if (sleeping_) encodeSleepImpact(w, sleepForce);
  launch_adjacency(w, pairCap_, s);
  launch_coloring(w, colorRounds_, useDataflow_ ? iterations_ : 0u, listedRounds_, s);
  launch_warm_start(w, s);
  if (useWorlds_) {
    launch_solve_worlds(w, worldOrder_, worldKeyStart_, worldJointStart_, iterations_, worldBlock_, s);  // + velocities
  } else if (useDataflow_) {
    launch_dataflow(w, dfGrid_, iterations_, s);
    launch_velocities(w, s);
  } else {
    for (uint32_t it = 0; it < iterations_; it++) {
      for (uint32_t c = 0; c < colorCap_; c++) launch_primal(w, c, maxDynamicBodies_, s);
      launch_dual(w, pairCap_, s);
    }
    launch_velocities(w, s);
  }
  if (bouncing_) launch_restitution_after(w, s);
  if (sleeping_) launch_sleep_after(w, sleepParams(sleepForce), s);
  if (cw_.table) launch_contact_events(w, cw_, &w == &views_[1] ? 1u : 0u, s);
}

void Engine::pollCounters() {
  // The oldest in-flight readback that has landed drives the policy (one or more steps late, as the reference's)
  for (uint32_t k = 0; k < 4; k++) {
    uint32_t slot = (counterSlot_ + k) % 4;
    if (!counterPending_[slot]) continue;
    if (hipEventQuery(counterEvents_[slot]) != hipSuccess) continue;
    counterPending_[slot] = false;
    const uint32_t *c = hostCounters_ + slot * COUNTER_WORDS;
    Counters cc{c[C_PAIR_DEMAND], c[C_CONTACTS], c[C_MANIFOLDS], c[C_OVERFLOW], c[C_CLASHES], c[C_NUM_COLORS],
                c[C_PREV_MANIFOLDS], c[C_PIVOT64], c[C_DF_ABORT], c[C_PENDING]};
    adaptFrom(cc);
  }
}

void Engine::step(uint32_t n) {
  HIP_CHECK(hipSetDevice(device_));
  for (uint32_t k = 0; k < n; k++) {
    if (adaptive_) pollCounters();
    DeviceWorld &w = views_[parity_];
    w.colorCap = colorCap_;
    const bool rebuild = useBvh_ && stepIndex_ % rebuildInterval_ == 0;
    if (sleepForce_) {  // the start-asleep step runs once, outside the graphs
      encodeStep(w, rebuild, sleepForce_);
      sleepForce_ = 0;
    } else if (useGraph_) {
      hipGraphExec_t &exec = graphs_[parity_ * 2 + (rebuild ? 1 : 0)];
      if (!exec) {
        hipGraph_t g;
        HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        encodeStep(w, rebuild);
        HIP_CHECK(hipStreamEndCapture(stream_, &g));
        HIP_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        HIP_CHECK(hipGraphDestroy(g));
      }
      HIP_CHECK(hipGraphLaunch(exec, stream_));
    } else {
      encodeStep(w, rebuild);
    }
    stepIndex_++;
    parity_ ^= 1;
    if (adaptive_ && !counterPending_[counterSlot_]) {
      HIP_CHECK(hipMemcpyAsync(hostCounters_ + counterSlot_ * COUNTER_WORDS, base_.counters,
                               COUNTER_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipEventRecord(counterEvents_[counterSlot_], stream_));
      counterPending_[counterSlot_] = true;
      counterSlot_ = (counterSlot_ + 1) % 4;
    }
  }
  HIP_CHECK(hipGetLastError());
}

void Engine::stepTimed(uint32_t n, float *ms) {
  stepTimedBegin(n);
  stepTimedEnd(ms);
}

void Engine::stepTimedBegin(uint32_t n) {
  HIP_CHECK(hipSetDevice(device_));
  while (timingEvents_.size() < 2ull * n) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    timingEvents_.push_back(e);
  }
  for (uint32_t k = 0; k < n; k++) {
    HIP_CHECK(hipEventRecord(timingEvents_[2 * k], stream_));
    step(1);
    HIP_CHECK(hipEventRecord(timingEvents_[2 * k + 1], stream_));
  }
  timedSteps_ = n;
}

uint32_t Engine::stepTimedEnd(float *ms) {
  const uint32_t n = timedSteps_;
  if (n == 0) return 0;
  HIP_CHECK(hipEventSynchronize(timingEvents_[2 * n - 1]));
  for (uint32_t k = 0; k < n; k++)
    HIP_CHECK(hipEventElapsedTime(&ms[k], timingEvents_[2 * k], timingEvents_[2 * k + 1]));
  timedSteps_ = 0;
  return n;
}

void Engine::synchronize() { HIP_CHECK(hipStreamSynchronize(stream_)); }

PhaseTimes Engine::profileStep() {
  HIP_CHECK(hipSetDevice(device_));
  const uint32_t launches = iterations_ * (colorCap_ + 1);
  std::vector<hipEvent_t> ev(9 + 2 * launches);
  for (auto &x : ev) HIP_CHECK(hipEventCreate(&x));
  hipStream_t s = stream_;
  DeviceWorld &w = views_[parity_];
  w.colorCap = colorCap_;
  const bool rebuild = useBvh_ && stepIndex_ % rebuildInterval_ == 0;
  HIP_CHECK(hipEventRecord(ev[0], s));
  const uint32_t force = sleepForce_;
  sleepForce_ = 0;
  if (sleeping_) launch_sleep_before(w, sleepParams(force), s);
  if (bouncing_) launch_restitution_before(w, s);
  // encodeCollision, with events after the frame set-up and after the broadphase
  launch_begin_frame(w, s);
  launch_refs_and_hash(w, pairCap_, s);
  HIP_CHECK(hipEventRecord(ev[2 + 2 * launches + 4], s));
  if (useBvh_) {
    if (rebuild) launch_bvh_build(w, bvh_, plan_, s);
    else launch_bvh_refit(w, bvh_, plan_, s);
    HIP_CHECK(hipEventRecord(ev[2 + 2 * launches + 6], s));
    launch_bvh_pairs(w, bvh_, s);  // writes the pair count
  } else {
    launch_grid(w, s);
    HIP_CHECK(hipEventRecord(ev[2 + 2 * launches + 6], s));
    launch_pairs(w, s);
  }
  HIP_CHECK(hipEventRecord(ev[2 + 2 * launches + 5], s));
  launch_narrowphase(w, s);
  if (sleeping_) encodeSleepImpact(w, force);
  HIP_CHECK(hipEventRecord(ev[1], s));
  launch_adjacency(w, pairCap_, s);
  HIP_CHECK(hipEventRecord(ev[2], s));
  launch_coloring(w, colorRounds_, useDataflow_ ? iterations_ : 0u, listedRounds_, s);
  HIP_CHECK(hipEventRecord(ev[3], s));
  launch_warm_start(w, s);
  HIP_CHECK(hipEventRecord(ev[4], s));
  uint32_t k = 6;
  if (useWorlds_) {
    // One launch for every iteration and the velocities: its time is reported as the primal's
    HIP_CHECK(hipEventRecord(ev[k++], s));
    launch_solve_worlds(w, worldOrder_, worldKeyStart_, worldJointStart_, iterations_, worldBlock_, s);
    HIP_CHECK(hipEventRecord(ev[k++], s));
  } else if (useDataflow_) {
    // One launch for every iteration: its time is reported as the primal's, the dual's is in it
    HIP_CHECK(hipEventRecord(ev[k++], s));
    launch_dataflow(w, dfGrid_, iterations_, s);
    HIP_CHECK(hipEventRecord(ev[k++], s));
  } else {
    for (uint32_t it = 0; it < iterations_; it++) {
      for (uint32_t c = 0; c < colorCap_; c++) {
        HIP_CHECK(hipEventRecord(ev[k++], s));
        launch_primal(w, c, maxDynamicBodies_, s);
        HIP_CHECK(hipEventRecord(ev[k++], s));
      }
      HIP_CHECK(hipEventRecord(ev[k++], s));
      launch_dual(w, pairCap_, s);
      HIP_CHECK(hipEventRecord(ev[k++], s));
    }
  }
  if (!useWorlds_) launch_velocities(w, s);
  if (bouncing_) launch_restitution_after(w, s);  // timed with the velocities
  if (sleeping_) launch_sleep_after(w, sleepParams(force), s);
  if (cw_.table) launch_contact_events(w, cw_, &w == &views_[1] ? 1u : 0u, s);
  HIP_CHECK(hipEventRecord(ev[5], s));
  HIP_CHECK(hipEventSynchronize(ev[5]));
  stepIndex_++;
  parity_ ^= 1;
  auto ms = [&](int a, int b) {
    float t = 0;
    HIP_CHECK(hipEventElapsedTime(&t, ev[a], ev[b]));
    return t;
  };
  PhaseTimes p{};
  p.total = ms(0, 5);
  p.collision = ms(0, 1);
  p.broadphase = ms(2 + 2 * launches + 4, 2 + 2 * launches + 5);
  p.pairSearch = ms(2 + 2 * launches + 6, 2 + 2 * launches + 5);
  p.narrowphase = ms(2 + 2 * launches + 5, 1);
  p.adjacency = ms(1, 2);
  p.coloring = ms(2, 3);
  p.warmStart = ms(3, 4);
  k = 6;
  if (useDataflow_ || useWorlds_) {
    p.primal = ms(6, 7);
    p.velocities = ms(7, 5);
    p.primalLaunches = 1;
    p.dualLaunches = 0;
  } else {
    for (uint32_t it = 0; it < iterations_; it++) {
      for (uint32_t c = 0; c < colorCap_; c++, k += 2) p.primal += ms(k, k + 1);
      p.dual += ms(k, k + 1);
      k += 2;
    }
    p.velocities = ms(k - 1, 5);
    p.primalLaunches = iterations_ * colorCap_;
    p.dualLaunches = iterations_;
  }
  for (auto &x : ev) HIP_CHECK(hipEventDestroy(x));
  return p;
}

void Engine::readBodies(uint32_t first, uint32_t count, Body *out) {
  if (size_t(first) + count > nBodies_) throw std::out_of_range("body range");
  HIP_CHECK(hipMemcpyAsync(out, base_.bodies + first, size_t(count) * sizeof(Body), hipMemcpyDeviceToHost, stream_));
  synchronize();
}
void Engine::writeBodies(uint32_t first, uint32_t count, const Body *in) {
  if (size_t(first) + count > nBodies_) throw std::out_of_range("body range");
  // A hull body's shape code must name a record this engine holds (the narrowphase reads it unchecked)
  for (uint32_t k = 0; k < count; k++) {
    const float code = in[k].angVel.w;
    if (!(code >= SHAPE_HULL)) continue;
    const uint32_t h = hullHeaderOf(code, uint32_t(hullsHost_.size()));
    if (!hullHeader_[h]) throw std::invalid_argument("a hull body's shape code names no hull record of this world");
  }
  if (!base_.roundedShapes)
    for (uint32_t k = 0; k < count; k++)
      if (in[k].angVel.w == SHAPE_SPHERE || in[k].angVel.w == SHAPE_CAPSULE) {
        // A sphere or capsule appears: the narrowphase switches to its full variant from the next step on
        base_.roundedShapes = views_[0].roundedShapes = views_[1].roundedShapes = 1;
        resetGraphs();
        break;
      }
  HIP_CHECK(hipMemcpyAsync(base_.bodies + first, in, size_t(count) * sizeof(Body), hipMemcpyHostToDevice, stream_));
  if (sleeping_) {  // the written records' masses are the awake masses (rewriteBodies re-registers them)
    launch_sleep_init(base_, nullptr, first, count, stream_);
    wakeAll();
  }
  synchronize();
}
void Engine::readJoints(uint32_t first, uint32_t count, Joint *out) {
  if (size_t(first) + count > nJoints_) throw std::out_of_range("joint range");
  if (count)
    HIP_CHECK(hipMemcpyAsync(out, base_.joints + first, size_t(count) * sizeof(Joint), hipMemcpyDeviceToHost, stream_));
  synchronize();
}
void Engine::writeJoints(uint32_t first, uint32_t count, const Joint *in) {
  if (size_t(first) + count > nJoints_) throw std::out_of_range("joint range");
  if (count)
    HIP_CHECK(hipMemcpyAsync(base_.joints + first, in, size_t(count) * sizeof(Joint), hipMemcpyHostToDevice, stream_));
  synchronize();
}

Counters Engine::readCounters() {
  uint32_t c[COUNTER_WORDS];
  HIP_CHECK(hipMemcpyAsync(c, base_.counters, sizeof c, hipMemcpyDeviceToHost, stream_));
  synchronize();
  Counters out{c[C_PAIR_DEMAND], c[C_CONTACTS], c[C_MANIFOLDS], c[C_OVERFLOW], c[C_CLASHES], c[C_NUM_COLORS],
               c[C_PREV_MANIFOLDS], c[C_PIVOT64], c[C_DF_ABORT], c[C_PENDING]};
  noteAborts(out);
  return out;
}

void Engine::noteAborts(const Counters &c) {
  // A dataflow wait gave up (its workgroups were not all resident for 65 ms: another kernel held the device): that
  // step's result is not valid (the counter says so) and every later step runs the colour launches instead
  if (c.solveAborts && useDataflow_) {
    useDataflow_ = false;
    wantDataflow_ = false;
    resetGraphs();
  }
}

std::vector<uint2> Engine::readPairs() {
  Counters c = readCounters();
  std::vector<uint2> out(c.pairs);
  if (c.pairs)
    HIP_CHECK(hipMemcpyAsync(out.data(), base_.pairs, size_t(c.pairs) * sizeof(uint2), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return out;
}

__global__ void k_contact_counts(const Manifold *m, uint32_t count, uint32_t *out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    out[i] = m[i].ids.w & 15u;
}

/** Dense copy of each manifold's contacts at its scanned offset; the copied manifold points there. */
__global__ void k_gather_contacts(const Manifold *m, uint32_t count, const uint32_t *offsets, const Contact *src,
                                  Contact *dst, Manifold *mout) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    Manifold mf = m[i];
    uint32_t o = offsets[i], n = mf.ids.w & 15u;
    for (uint32_t k = 0; k < n; k++) dst[o + k] = src[mf.ids.z + k];
    mf.ids.z = o;
    mout[i] = mf;
  }
}

void Engine::readContacts(std::vector<Manifold> &manifolds, std::vector<Contact> &contacts) {
  uint32_t c[COUNTER_WORDS];
  HIP_CHECK(hipMemcpyAsync(c, base_.counters, sizeof c, hipMemcpyDeviceToHost, stream_));
  synchronize();
  const uint32_t m = c[C_MANIFOLDS];
  manifolds.resize(m);
  contacts.clear();
  if (m == 0) return;
  // The last step wrote the buffers the next step reads as "previous"
  const Manifold *dm = views_[parity_].prevManifolds;
  const Contact *dc = views_[parity_].prevContacts;
  uint32_t *offsets = nullptr, *scratch = nullptr;
  Contact *dense = nullptr;
  Manifold *mout = nullptr;
  const uint32_t words = scan_scratch_words(m + 1);
  HIP_CHECK(hipMalloc(&offsets, (size_t(m) + 1) * 4));
  HIP_CHECK(hipMalloc(&scratch, size_t(words) * 4));
  HIP_CHECK(hipMalloc(&dense, size_t(m) * MAX_CONTACTS * sizeof(Contact)));
  HIP_CHECK(hipMalloc(&mout, size_t(m) * sizeof(Manifold)));
  HIP_CHECK(hipMemsetAsync(offsets, 0, (size_t(m) + 1) * 4, stream_));
  HIP_CHECK(hipMemsetAsync(scratch, 0, size_t(words) * 4, stream_));  // the scan's tile statuses start at zero
  uint32_t g = std::min<uint32_t>((m + 255) / 256, 4096);
  k_contact_counts<<<g, 256, 0, stream_>>>(dm, m, offsets);
  launch_scan_exclusive(offsets, m + 1, scratch, stream_);
  k_gather_contacts<<<g, 256, 0, stream_>>>(dm, m, offsets, dc, dense, mout);
  uint32_t total = 0, gaveUp = 0;
  HIP_CHECK(hipMemcpyAsync(&total, offsets + m, 4, hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipMemcpyAsync(&gaveUp, scratch + 2, 4, hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipMemcpyAsync(manifolds.data(), mout, size_t(m) * sizeof(Manifold), hipMemcpyDeviceToHost, stream_));
  synchronize();
  if (gaveUp) total = 0;
  contacts.resize(total);
  if (total)
    HIP_CHECK(hipMemcpyAsync(contacts.data(), dense, size_t(total) * sizeof(Contact), hipMemcpyDeviceToHost, stream_));
  synchronize();
  HIP_CHECK(hipFree(offsets));
  HIP_CHECK(hipFree(scratch));
  HIP_CHECK(hipFree(dense));
  HIP_CHECK(hipFree(mout));
  if (gaveUp) throw std::runtime_error("reading contacts: the offset scan gave up a wait");
}

std::vector<uint32_t> Engine::readColors() {
  std::vector<uint32_t> out(nBodies_);
  HIP_CHECK(hipMemcpyAsync(out.data(), base_.colorA, size_t(nBodies_) * 4, hipMemcpyDeviceToHost, stream_));
  synchronize();
  return out;
}
std::vector<float4> Engine::readRefs() {
  std::vector<float4> out(2ull * nBodies_);
  HIP_CHECK(hipMemcpyAsync(out.data(), base_.refs, out.size() * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return out;
}
std::vector<WorldParams> Engine::readWorldParams() {
  std::vector<WorldParams> out(nWorlds_);
  HIP_CHECK(hipMemcpyAsync(out.data(), base_.worlds, out.size() * sizeof(WorldParams), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return out;
}
void Engine::readAdjacency(std::vector<uint32_t> &start, std::vector<uint32_t> &list) {
  start.resize(uint64_t(nBodies_) + 1);
  HIP_CHECK(hipMemcpyAsync(start.data(), base_.adjStart, start.size() * 4, hipMemcpyDeviceToHost, stream_));
  synchronize();
  list.resize(start.back());
  if (!list.empty())
    HIP_CHECK(hipMemcpyAsync(list.data(), base_.adjList, list.size() * 4, hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void Engine::writeState(const float4 *refs, const uint32_t *colors, const std::vector<Manifold> &manifolds,
                        const std::vector<Contact> &contacts, const uint32_t *steps) {
  if (manifolds.size() > pairCap_ || contacts.size() > size_t(pairCap_) * MAX_CONTACTS)
    throw std::length_error("state exceeds the contact capacity");
  for (const Manifold &m : manifolds)
    if (size_t(m.ids.z) + (m.ids.w & 15u) > contacts.size()) throw std::out_of_range("manifold contact range");
  if (refs)
    HIP_CHECK(hipMemcpyAsync(base_.refs, refs, 2ull * nBodies_ * sizeof(float4), hipMemcpyHostToDevice, stream_));
  if (colors) HIP_CHECK(hipMemcpyAsync(base_.colorA, colors, size_t(nBodies_) * 4, hipMemcpyHostToDevice, stream_));
  if (!manifolds.empty())
    HIP_CHECK(hipMemcpyAsync(views_[parity_].prevManifolds, manifolds.data(), manifolds.size() * sizeof(Manifold),
                             hipMemcpyHostToDevice, stream_));
  if (!contacts.empty())
    HIP_CHECK(hipMemcpyAsync(views_[parity_].prevContacts, contacts.data(), contacts.size() * sizeof(Contact),
                             hipMemcpyHostToDevice, stream_));
  uint32_t c[COUNTER_WORDS] = {0};
  c[C_MANIFOLDS] = uint32_t(manifolds.size());
  c[C_CONTACTS] = uint32_t(contacts.size());
  HIP_CHECK(hipMemcpyAsync(base_.counters, c, sizeof c, hipMemcpyHostToDevice, stream_));
  if (steps) {
    std::vector<WorldParams> p = readWorldParams();
    for (uint32_t k = 0; k < nWorlds_; k++) p[k].step = steps[k];
    HIP_CHECK(hipMemcpyAsync(base_.worlds, p.data(), p.size() * sizeof(WorldParams), hipMemcpyHostToDevice, stream_));
  }
  synchronize();
}

SavedState Engine::save() {
  SavedState s;
  s.bodies.resize(nBodies_);
  readBodies(0, nBodies_, s.bodies.data());
  s.joints.resize(nJoints_);
  readJoints(0, nJoints_, s.joints.data());
  s.refs = readRefs();
  s.colors = readColors();
  readContacts(s.manifolds, s.contacts);
  std::vector<WorldParams> p = readWorldParams();
  for (const WorldParams &q : p) s.steps.push_back(q.step);
  return s;
}

void Engine::restore(const SavedState &s) {
  if (s.bodies.size() != nBodies_ || s.joints.size() != nJoints_ || s.refs.size() != 2ull * nBodies_ ||
      s.colors.size() != nBodies_ || s.steps.size() != nWorlds_)
    throw std::invalid_argument("saved state does not match this batch");
  writeBodies(0, nBodies_, s.bodies.data());
  writeJoints(0, nJoints_, s.joints.data());
  writeState(s.refs.data(), s.colors.data(), s.manifolds, s.contacts, s.steps.data());
}

// --- edits between steps --------------------------------------------------------------------------------------------

uint32_t *Engine::uploadEdit(const uint32_t *words, size_t n) {
  const int s = editSlot_;
  editSlot_ ^= 1;
  if (!editEvents_[s]) HIP_CHECK(hipEventCreateWithFlags(&editEvents_[s], hipEventDisableTiming));
  if (editPending_[s]) {
    HIP_CHECK(hipEventSynchronize(editEvents_[s]));  // the staging buffer's last copy has been read
    editPending_[s] = false;
  }
  if (editHostWords_[s] < n) {
    if (editHost_[s]) HIP_CHECK(hipHostFree(editHost_[s]));
    editHost_[s] = nullptr;
    const size_t cap = std::max<size_t>(2 * n, 1024);
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&editHost_[s]), cap * 4, hipHostMallocDefault));
    editHostWords_[s] = cap;
  }
  std::memcpy(editHost_[s], words, n * 4);
  if (editDevWords_ < n) {
    HIP_CHECK(hipStreamSynchronize(stream_));  // edits enqueued earlier may still read the old buffer
    if (editDev_) {
      HIP_CHECK(hipFree(editDev_));
      deviceBytes_ -= editDevWords_ * 4;
    }
    editDev_ = nullptr;
    const size_t cap = std::max<size_t>(2 * n, 1024);
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&editDev_), cap * 4));
    editDevWords_ = cap;
    deviceBytes_ += cap * 4;
  }
  HIP_CHECK(hipMemcpyAsync(editDev_, editHost_[s], n * 4, hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipEventRecord(editEvents_[s], stream_));
  editPending_[s] = true;
  return editDev_;
}

namespace {
bool finite(const float *v, int n) {
  for (int k = 0; k < n; k++)
    if (!std::isfinite(v[k])) return false;
  return true;
}
const char *const kCommandNames[7] = {"linearVelocity", "angularVelocity", "impulse", "force",
                                      "position",       "angularImpulse",  "torque"};
}  // namespace

void Engine::applyCommands(const uint32_t *words, uint32_t n) {
  if (n == 0) return;
  if (!words) throw std::invalid_argument("commands is NULL");
  HIP_CHECK(hipSetDevice(device_));
  bool teleports = false;
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t *c = words + 12ull * k;
    float f[8];
    std::memcpy(f, c + 4, sizeof f);
    if (c[0] >= nBodies_) throw std::invalid_argument("A live 3D GPU body index is required");
    if (c[1] > 6) throw std::invalid_argument("command kind must be 0 to 6");
    if (c[2] > 1) throw std::invalid_argument("has_point must be 0 or 1");
    const uint32_t kind = c[1];
    const bool hasPoint = c[2] != 0;
    const std::string name = kCommandNames[kind];
    if (!finite(f, 3)) throw std::invalid_argument(name + " requires three finite components");
    if ((kind == 3 || kind == 6) && !finite(f + 3, 1)) throw std::invalid_argument(name + " requires a finite time step");
    if ((kind == 2 || kind == 3) && hasPoint && !finite(f + 4, 3))
      throw std::invalid_argument("point requires three finite components");
    if (kind == 4 && hasPoint) {
      if (!finite(f + 4, 4)) throw std::invalid_argument("rotation requires four finite components");
      if (f[4] == 0.f && f[5] == 0.f && f[6] == 0.f && f[7] == 0.f)
        throw std::invalid_argument("rotation must be nonzero");
    }
    teleports = teleports || kind == 4;
  }
  // Group by body, each body's commands in submission order (commands of different bodies commute)
  std::vector<uint32_t> order(n);
  for (uint32_t k = 0; k < n; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return words[12ull * x] < words[12ull * y]; });
  std::vector<uint32_t> buf(12ull * n);
  std::vector<uint32_t> ranges;
  for (uint32_t k = 0; k < n; k++) {
    std::memcpy(&buf[12ull * k], words + 12ull * order[k], 48);
    if (k == 0 || words[12ull * order[k]] != words[12ull * order[k - 1]]) {
      ranges.push_back(k);
      ranges.push_back(0);
    }
    ranges.back()++;
  }
  const uint32_t nRanges = uint32_t(ranges.size() / 2);
  buf.insert(buf.end(), ranges.begin(), ranges.end());
  if (teleports && !marks_) marks_ = dallocT<uint32_t>(nBodies_);
  uint32_t *dev = uploadEdit(buf.data(), buf.size());
  const uint2 *devRanges = reinterpret_cast<const uint2 *>(dev + 12ull * n);
  wakeAll();  // body commands act on awake masses (every body wakes first)
  launch_body_commands(base_.bodies, dev, devRanges, nRanges, marks_, stream_);
  if (teleports) {
    launch_invalidate_manifolds(views_[parity_].prevManifolds, base_.counters, pairCap_, marks_, stream_);
    launch_clear_marks(dev, devRanges, nRanges, marks_, stream_);
  }
  HIP_CHECK(hipGetLastError());
}

void Engine::setFilters(const uint32_t *indices, const uint32_t *groups, const uint32_t *masks, uint32_t n) {
  if (n == 0) return;
  if (!indices || !groups || !masks) throw std::invalid_argument("indices, groups or collides_with is NULL");
  HIP_CHECK(hipSetDevice(device_));
  std::vector<uint32_t> buf(3ull * n);
  for (uint32_t k = 0; k < n; k++) {
    if (indices[k] >= nBodies_) throw std::invalid_argument("Collision filters require live body indices");
    buf[3 * k] = indices[k];
    buf[3 * k + 1] = groups[k];
    buf[3 * k + 2] = masks[k];
  }
  wakeAll();
  launch_set_filters(base_.filters, uploadEdit(buf.data(), buf.size()), n, stream_);
  HIP_CHECK(hipGetLastError());
}

void Engine::setRestitution(const uint32_t *indices, const float *values, uint32_t n) {
  if (n == 0) return;
  if (!indices || !values) throw std::invalid_argument("indices or values is NULL");
  for (uint32_t k = 0; k < n; k++) {
    if (indices[k] >= nBodies_) throw std::invalid_argument("Restitution requires live body indices");
    if (!(values[k] >= 0.f && values[k] <= 1.f)) throw std::invalid_argument("restitution must be between 0 and 1");
  }
  HIP_CHECK(hipSetDevice(device_));
  std::vector<uint32_t> buf(2ull * n);
  const uint32_t before = bouncing_;
  for (uint32_t k = 0; k < n; k++) {
    float &r = restitution_[indices[k]];
    bouncing_ = bouncing_ - (r > 0.f ? 1u : 0u) + (values[k] > 0.f ? 1u : 0u);
    r = values[k];
    buf[2 * k] = indices[k];
    std::memcpy(&buf[2 * k + 1], &values[k], 4);
  }
  wakeAll();
  launch_set_restitution(base_.restitution, uploadEdit(buf.data(), buf.size()), n, stream_);
  HIP_CHECK(hipGetLastError());
  if ((before == 0) != (bouncing_ == 0)) resetGraphs();  // the step gains or loses its bounce passes
}

void Engine::getRestitution(float *out) const { std::copy(restitution_.begin(), restitution_.end(), out); }

void Engine::enableSleeping(float speed, float delay, int islands, bool startAsleep) {
  if (!(std::isfinite(speed) && speed > 0.f && std::isfinite(delay) && delay > 0.f))
    throw std::invalid_argument("Sleep speedThreshold and timeThreshold must be positive and finite");
  if (islands < -1 || islands > 1) throw std::invalid_argument("islands must be -1 (auto), 0 or 1");
  if (sleeping_) return;
  HIP_CHECK(hipSetDevice(device_));
  sleepSpeed_ = speed;
  sleepDelay_ = delay;
  sleepIslands_ = islands < 0 ? nBodies_ / std::max(1u, nWorlds_) >= 4096u : islands == 1;
  std::vector<uint32_t> packed((sleepEligible_.size() + 3) / 4, 0u);
  std::memcpy(packed.data(), sleepEligible_.data(), sleepEligible_.size());
  const uint8_t *eligible = reinterpret_cast<const uint8_t *>(uploadEdit(packed.data(), packed.size()));
  launch_sleep_init(base_, eligible, 0, nBodies_, stream_);
  launch_fill_u32(base_.sleepGlobals, 0u, 4ull * nWorlds_ + 1, stream_);
  HIP_CHECK(hipGetLastError());
  sleeping_ = true;
  sleepForce_ = startAsleep && sleepIslands_ ? 2u : 0u;
  resetGraphs();
}

void Engine::disableSleeping() {
  if (!sleeping_) return;
  wakeAll();
  sleeping_ = false;
  sleepForce_ = 0;
  resetGraphs();
}

void Engine::wakeAll() {
  if (!sleeping_) return;
  HIP_CHECK(hipSetDevice(device_));
  launch_sleep_before(base_, sleepParams(1u), stream_);
  HIP_CHECK(hipGetLastError());
}

void Engine::wakeBodies(const uint32_t *indices, uint32_t n) {
  if (n == 0) return;
  if (!indices) throw std::invalid_argument("indices is NULL");
  for (uint32_t k = 0; k < n; k++)
    if (indices[k] >= nBodies_) throw std::invalid_argument("Sleeping wake requires a live GPU body index");
  if (!sleeping_) return;
  HIP_CHECK(hipSetDevice(device_));
  launch_sleep_hold(base_, uploadEdit(indices, n), n, sleepIslands_, stream_);
  HIP_CHECK(hipGetLastError());
}

void Engine::setSleepEnabled(const uint32_t *indices, const uint8_t *flags, uint32_t n) {
  if (n == 0) return;
  if (!indices || !flags) throw std::invalid_argument("indices or flags is NULL");
  bool any = false;
  for (uint32_t k = 0; k < n; k++) {
    if (indices[k] >= nBodies_) throw std::invalid_argument("A live 3D GPU body index is required");
    if (flags[k] > 1) throw std::invalid_argument("Sleep flag must be boolean");
    any = any || flags[k] == 1;
  }
  std::vector<uint32_t> buf(2ull * n);
  for (uint32_t k = 0; k < n; k++) {
    sleepEligible_[indices[k]] = flags[k];
    buf[2 * k] = indices[k];
    buf[2 * k + 1] = flags[k];
  }
  if (!sleeping_) {
    if (any) enableSleeping(sleepSpeed_, sleepDelay_, -1, false);  // takes every body's eligibility
    return;
  }
  HIP_CHECK(hipSetDevice(device_));
  launch_sleep_eligible(base_, uploadEdit(buf.data(), buf.size()), n, stream_);
  wakeAll();
}

void Engine::sleepStats(uint32_t world, uint32_t *out) {
  if (world >= nWorlds_) throw std::out_of_range("world index");
  out[0] = out[1] = 0;
  if (!sleeping_) return;
  HIP_CHECK(hipSetDevice(device_));
  uint32_t g[4];
  HIP_CHECK(hipMemcpyAsync(g, base_.sleepGlobals + 4ull * world, sizeof g, hipMemcpyDeviceToHost, stream_));
  synchronize();
  out[0] = g[1];
  out[1] = g[0];
}

void Engine::cast(const float *queries, uint32_t n, const QueryFilter &filter, float *hits) {
  if (n == 0) return;
  if (!queries || !hits) throw std::invalid_argument("queries or hits is NULL");
  if (n > 65535) throw std::invalid_argument("Too many queries in one batch");
  if (filter.ignoreCount > 16) throw std::invalid_argument("Queries can ignore up to 16 live body indices");
  for (uint32_t k = 0; k < filter.ignoreCount; k++)
    if (filter.ignore[k] >= nBodies_) throw std::invalid_argument("Queries can ignore up to 16 live body indices");
  std::vector<QueryRec> recs(n);
  for (uint32_t k = 0; k < n; k++) {
    const float *q = queries + 8ull * k;
    bool ok = true;
    for (int c = 0; c < 8; c++) ok = ok && std::isfinite(q[c]);
    const double len = std::sqrt(double(q[4]) * q[4] + double(q[5]) * q[5] + double(q[6]) * q[6]);
    if (!ok || len == 0.0) throw std::invalid_argument("A query requires 3 finite components and a nonzero direction");
    if (q[3] < 0.f || q[7] < 0.f)
      throw std::invalid_argument("Query radius and maximum distance must be finite and nonnegative");
    recs[k].origin = make_float4(q[0], q[1], q[2], q[3]);
    recs[k].direction = make_float4(float(q[4] / len), float(q[5] / len), float(q[6] / len), q[7]);
  }
  HIP_CHECK(hipSetDevice(device_));
  if (n > queryCap_) {
    if (queryDev_) HIP_CHECK(hipFree(queryDev_));
    if (queryBest_) HIP_CHECK(hipFree(queryBest_));
    if (queryHits_) HIP_CHECK(hipFree(queryHits_));
    queryCap_ = std::max<uint32_t>(n, 64);
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&queryDev_), size_t(queryCap_) * sizeof(QueryRec)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&queryBest_), size_t(queryCap_) * 2 * sizeof(uint32_t)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&queryHits_), size_t(queryCap_) * 2 * sizeof(float4)));
    launch_query_reset(queryBest_, queryCap_, stream_);
  }
  HIP_CHECK(hipMemcpyAsync(queryDev_, recs.data(), size_t(n) * sizeof(QueryRec), hipMemcpyHostToDevice, stream_));
  launch_queries(base_, queryDev_, n, filter, queryBest_, queryHits_, stream_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(hits, queryHits_, size_t(n) * 2 * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void Engine::castNoAlloc(const float *queries, uint32_t n, const QueryFilter &filter, float *hits) {
  if (n == 0) return;
  if (!queries || !hits) throw std::invalid_argument("queries or hits is NULL");
  if (n > NOALLOC_MAX_QUERIES) throw std::invalid_argument("castNoAlloc takes at most 16384 queries per call");
  if (filter.ignoreCount > 16) throw std::invalid_argument("Queries can ignore up to 16 live body indices");
  for (uint32_t k = 0; k < filter.ignoreCount; k++)
    if (filter.ignore[k] >= nBodies_) throw std::invalid_argument("Queries can ignore up to 16 live body indices");
  for (uint32_t k = 0; k < n; k++) {  // every query is checked before anything runs
    const float *q = queries + 8ull * k;
    bool ok = true;
    for (int c = 0; c < 8; c++) ok = ok && std::isfinite(q[c]);
    const double len = std::sqrt(double(q[4]) * q[4] + double(q[5]) * q[5] + double(q[6]) * q[6]);
    if (!ok || len == 0.0) throw std::invalid_argument("A query requires 3 finite components and a nonzero direction");
    if (q[3] < 0.f || q[7] < 0.f)
      throw std::invalid_argument("Query radius and maximum distance must be finite and nonnegative");
  }
  HIP_CHECK(hipSetDevice(device_));
  uint8_t *base = static_cast<uint8_t *>(query_scratch(device_));
  QueryRec *recs = reinterpret_cast<QueryRec *>(base);
  uint32_t *best = reinterpret_cast<uint32_t *>(base + size_t(NOALLOC_MAX_QUERIES) * sizeof(QueryRec));
  float4 *out = reinterpret_cast<float4 *>(base + size_t(NOALLOC_MAX_QUERIES) * (sizeof(QueryRec) + 2 * sizeof(uint32_t)));
  // The best words start cleared in every call (the array keeps whatever the last call on this device left)
  launch_query_reset(best, n, stream_);
  // Records normalised as cast() does, through a fixed host chunk: each chunk's copy completes before it is refilled
  constexpr uint32_t CHUNK = 256;
  QueryRec chunk[CHUNK];
  for (uint32_t k0 = 0; k0 < n; k0 += CHUNK) {
    const uint32_t m = std::min(CHUNK, n - k0);
    for (uint32_t k = 0; k < m; k++) {
      const float *q = queries + 8ull * (k0 + k);
      const double len = std::sqrt(double(q[4]) * q[4] + double(q[5]) * q[5] + double(q[6]) * q[6]);
      chunk[k].origin = make_float4(q[0], q[1], q[2], q[3]);
      chunk[k].direction = make_float4(float(q[4] / len), float(q[5] / len), float(q[6] / len), q[7]);
    }
    HIP_CHECK(hipMemcpyAsync(recs + k0, chunk, size_t(m) * sizeof(QueryRec), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
  }
  launch_queries(base_, recs, n, filter, best, out, stream_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(hits, out, size_t(n) * 2 * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void Engine::watchContacts(uint32_t maxPairs, uint32_t maxEvents) {
  if (maxPairs < 1 || maxEvents < 1) throw std::invalid_argument("Contact capacities must be positive integers");
  if (cw_.table) return;
  HIP_CHECK(hipSetDevice(device_));
  uint32_t slots = 64;
  while (slots < 2ull * maxPairs) slots *= 2;
  cw_.mask = slots - 1;
  cw_.capacity = maxEvents;
  const size_t tableWords = CW_HEAD + 4ull * slots;
  HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&cw_.table), tableWords * 4));
  HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&cw_.events), size_t(maxEvents) * 3 * sizeof(float4)));
  HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&cw_.watch), size_t(nBodies_) * 4 + 4));
  HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&cw_.previousWatch), size_t(nBodies_) * 4 + 4));
  launch_fill_u32(cw_.table, 0u, tableWords, stream_);
  watchWords_.assign(nBodies_, 1u);
  uploadWatch();
  HIP_CHECK(hipMemcpyAsync(cw_.previousWatch, cw_.watch, size_t(nBodies_) * 4, hipMemcpyDeviceToDevice, stream_));
  synchronize();
  resetGraphs();  // the step gains its event passes
}

void Engine::uploadWatch() {
  HIP_CHECK(hipMemcpyAsync(cw_.watch, watchWords_.data(), size_t(nBodies_) * 4, hipMemcpyHostToDevice, stream_));
  synchronize();
}

void Engine::setWatched(const uint32_t *indices, const uint8_t *flags, uint32_t n) {
  if (!cw_.table) watchContacts(8192, 4096);
  for (uint32_t k = 0; k < n; k++)
    if (indices[k] >= nBodies_ || flags[k] > 1) throw std::invalid_argument("Watched indices must be live body indices");
  for (uint32_t k = 0; k < n; k++) watchWords_[indices[k]] = (watchWords_[indices[k]] & ~1u) | flags[k];
  HIP_CHECK(hipSetDevice(device_));
  uploadWatch();
}

void Engine::resetContactHistory(const uint32_t *indices, uint32_t n) {
  int64_t previous = -1;
  for (uint32_t k = 0; k < n; k++) {
    if (indices[k] >= nBodies_ || int64_t(indices[k]) <= previous)
      throw std::invalid_argument("Contact reset slots must be unique and ascending");
    previous = indices[k];
  }
  if (!cw_.table) watchContacts(8192, 4096);
  for (uint32_t k = 0; k < n; k++) watchWords_[indices[k]] += 2u;
  HIP_CHECK(hipSetDevice(device_));
  uploadWatch();
}

uint32_t Engine::readContactEvents(uint32_t *out, uint32_t cap, uint32_t *written, uint32_t *full) {
  if (!cw_.table) watchContacts(8192, 4096);
  HIP_CHECK(hipSetDevice(device_));
  uint32_t head[2];
  HIP_CHECK(hipMemcpyAsync(head, cw_.table, sizeof head, hipMemcpyDeviceToHost, stream_));
  synchronize();
  const uint32_t n = std::min(std::min(head[0], cw_.capacity), cap);
  if (n && out)
    HIP_CHECK(hipMemcpyAsync(out, cw_.events, size_t(n) * 3 * sizeof(float4), hipMemcpyDeviceToHost, stream_));
  launch_fill_u32(cw_.table, 0u, 2, stream_);
  synchronize();
  if (written) *written = head[0];
  if (full) *full = head[1];
  return n;
}

void Engine::readSleep(uint32_t *states, uint32_t *globals) {
  HIP_CHECK(hipSetDevice(device_));
  if (states)
    HIP_CHECK(hipMemcpyAsync(states, base_.sleep, size_t(nBodies_) * sizeof(SleepState), hipMemcpyDeviceToHost,
                             stream_));
  if (globals)
    HIP_CHECK(hipMemcpyAsync(globals, base_.sleepGlobals, 4ull * nWorlds_ * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             stream_));
  synchronize();
}

void Engine::setParams(uint32_t world, const ParamEdit &e) {
  if (world >= nWorlds_) throw std::invalid_argument("world index out of range");
  if (!(e.dt > 0.f) || !std::isfinite(e.dt)) throw std::invalid_argument("timeStep must be positive");
  const float values[9] = {e.gravity, e.betaLin, e.betaAng, e.gamma, e.alpha, e.up[0], e.up[1], e.up[2], 0.f};
  if (!finite(values, 8)) throw std::invalid_argument("parameters must be finite");
  const float upLen = std::sqrt(e.up[0] * e.up[0] + e.up[1] * e.up[1] + e.up[2] * e.up[2]);
  if (std::fabs(upLen - 1.f) > 1e-5f) throw std::invalid_argument("up must be a unit vector");
  HIP_CHECK(hipSetDevice(device_));
  launch_set_params(base_.worlds + world, e, stream_);
  HIP_CHECK(hipGetLastError());
}

// --- constraint edits between steps (appGpuSolver3D.js:500-563, 810-837, 922-957; solver.ts:1138-1204) ----------

void Engine::captureFrames(const std::vector<uint32_t> &slots) {
  if (slots.empty()) return;
  HIP_CHECK(hipSetDevice(device_));
  const uint32_t *dev = uploadEdit(slots.data(), slots.size());
  launch_capture_frames(base_.bodies, base_.joints, base_.info, dev, uint32_t(slots.size()), stream_);
  HIP_CHECK(hipGetLastError());
}

void Engine::editJointWords(const std::vector<uint32_t> &triples) {
  if (triples.empty()) return;
  HIP_CHECK(hipSetDevice(device_));
  const uint32_t *dev = uploadEdit(triples.data(), triples.size());
  launch_edit_joints(base_.joints, base_.info, dev, uint32_t(triples.size() / 3), stream_);
  HIP_CHECK(hipGetLastError());
}

int4 Engine::jointInfo(uint32_t slot) const {
  if (slot >= nJoints_) throw std::out_of_range("no such constraint slot");
  return infoHost_[slot];
}

void Engine::wakeConstraint(uint32_t slot) {
  if (!sleeping_) return;
  const int4 f = infoHost_[slot];
  uint32_t bodies[2];
  uint32_t n = 0;
  if (f.y >= 0) bodies[n++] = uint32_t(f.y);
  if (f.z >= 0) bodies[n++] = uint32_t(f.z);
  if (n) wakeBodies(bodies, n);
}

void Engine::releaseJoints(const uint32_t *slots, uint32_t n) {
  if (n && !slots) throw std::invalid_argument("slots is NULL");
  std::vector<uint32_t> order;
  for (uint32_t k = 0; k < n; k++)
    if (slots[k] < nJoints_ && infoHost_[slots[k]].x != T_NONE) order.push_back(slots[k]);
  std::sort(order.begin(), order.end());
  order.erase(std::unique(order.begin(), order.end()), order.end());
  if (order.empty()) return;
  for (uint32_t slot : order) wakeConstraint(slot);
  // The whole record zeroed and the slot's type none: no rows, no adjacency entry, and its bodies collide again
  std::vector<uint32_t> t;
  t.reserve(order.size() * 35 * 3);
  for (uint32_t slot : order) {
    for (uint32_t word = 0; word < 32; word++) t.insert(t.end(), {slot, word, 0u});
    t.insert(t.end(), {slot, 32u, uint32_t(T_NONE), slot, 33u, 0xffffffffu, slot, 34u, 0xffffffffu});
    infoHost_[slot] = make_int4(T_NONE, -1, -1, 0);
  }
  editJointWords(t);
}

void Engine::disableConstraint(uint32_t slot) {
  if (slot >= nJoints_) throw std::out_of_range("no such constraint slot");
  wakeConstraint(slot);
  // Penalties and stiffnesses (8 words) and both multipliers' xyz to zero: the constraint stops acting and its
  // bodies collide again; the thresholds in lamLin.w / lamAng.w stay (solver.ts:1200-1204)
  std::vector<uint32_t> t;
  for (uint32_t word : {0u, 1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u, 10u, 12u, 13u, 14u}) t.insert(t.end(), {slot, word, 0u});
  editJointWords(t);
}

static uint32_t fbitsHost(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

void Engine::setWorldAnchor(uint32_t slot, const float *point) {
  if (slot >= nJoints_ || infoHost_[slot].y != -1 || infoHost_[slot].x == T_NONE || !point ||
      !(std::isfinite(point[0]) && std::isfinite(point[1]) && std::isfinite(point[2])))
    throw std::invalid_argument("A world joint and finite xyz anchor are required");
  if (sleeping_) {
    const uint32_t b = uint32_t(infoHost_[slot].z);
    wakeBodies(&b, 1);
  }
  editJointWords({slot, 24u, fbitsHost(point[0]), slot, 25u, fbitsHost(point[1]), slot, 26u, fbitsHost(point[2])});
}

void Engine::setMotor(uint32_t slot, const float *speed, const float *maxTorque) {
  if (slot >= nJoints_ || infoHost_[slot].x != T_MOTOR) throw std::invalid_argument("Motor needs an active motor slot");
  if (speed && !std::isfinite(*speed)) throw std::invalid_argument("Motor speed must be finite");
  if (maxTorque && !(std::isfinite(*maxTorque) && *maxTorque >= 0.f))
    throw std::invalid_argument("maxTorque must be nonnegative");
  std::vector<uint32_t> t;
  if (speed) t.insert(t.end(), {slot, 27u, fbitsHost(*speed)});
  if (maxTorque) t.insert(t.end(), {slot, 31u, fbitsHost(*maxTorque)});
  if (sleeping_) wakeAll();
  editJointWords(t);
}

void Engine::setSpringMaterial(uint32_t slot, float breakStrain, float breakForce) {
  if (slot >= nJoints_ || infoHost_[slot].x != T_SPRING)
    throw std::invalid_argument("Spring material requires an active spring slot");
  for (float v : {breakStrain, breakForce})
    if (!(v == INFINITY || (std::isfinite(v) && v > 0.f)))
      throw std::invalid_argument("breakStrain and breakForce must be positive or Infinity");
  if (sleeping_) wakeAll();
  editJointWords({slot, 11u, fbitsHost(std::isinf(breakStrain) ? 0.f : breakStrain), slot, 15u,
                  fbitsHost(std::isinf(breakForce) ? 0.f : breakForce)});
}

void Engine::setAngularConstraint(uint32_t slot, int32_t type, const float *axisA, const float *axisB, float lo,
                                  float hi) {
  if (slot >= nJoints_ || infoHost_[slot].x == T_NONE) throw std::invalid_argument("Angular constraint needs a live slot");
  if (type != T_MOTOR && type != T_LIMIT) throw std::invalid_argument("Invalid angular constraint type");
  float axes[6];
  for (int s = 0; s < 2; s++) {
    const float *v = s ? axisB : axisA;
    if (!v) throw std::invalid_argument("Angular axis requires three finite, nonzero components");
    const double l = std::sqrt(double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2]);
    if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])) || !(l >= 1e-8))
      throw std::invalid_argument("Angular axis requires three finite, nonzero components");
    for (int c = 0; c < 3; c++) axes[3 * s + c] = float(v[c] / l);
  }
  if (type == T_MOTOR && !(std::isfinite(lo) && std::isfinite(hi) && hi >= 0.f))
    throw std::invalid_argument("Motor speed must be finite and maxTorque nonnegative");
  if (type == T_LIMIT && !(std::isfinite(lo) && std::isfinite(hi) && lo >= -3.14159265358979 && hi <= 3.14159265358979 && lo <= hi))
    throw std::invalid_argument("Hinge angles must be ordered within -PI..PI radians");
  // penLin = (1000, 1000, 0, 3e38), axes in rA / rB, speed or lower angle in rA.w, torque bound or upper angle in rB.w
  float rec[32] = {0};
  rec[0] = rec[1] = 1000.f;
  rec[3] = BIG;
  for (int c = 0; c < 3; c++) {
    rec[24 + c] = axes[c];
    rec[28 + c] = axes[3 + c];
  }
  rec[27] = lo;
  rec[31] = hi;
  std::vector<uint32_t> t;
  for (uint32_t word = 0; word < 32; word++) t.insert(t.end(), {slot, word, fbitsHost(rec[word])});
  t.insert(t.end(), {slot, 32u, uint32_t(type)});
  infoHost_[slot].x = type;
  editJointWords(t);
  captureFrames({slot});
  if (sleeping_) wakeAll();
}

void Engine::setIterations(uint32_t n) {
  if (n == 0) throw std::invalid_argument("iterations must be a positive integer");
  if (n != iterations_) {
    iterations_ = n;
    resetGraphs();
  }
}

void Engine::setColorCap(uint32_t cap) {
  cap = std::min<uint32_t>(std::max<uint32_t>(cap, 1), MAX_COLORS);
  if (cap != colorCap_) {
    colorCap_ = cap;
    base_.colorCap = cap;
    views_[0].colorCap = views_[1].colorCap = cap;
    resetGraphs();
  }
}
void Engine::setColorRounds(uint32_t rounds) {
  rounds += rounds & 1u;
  if (rounds != colorRounds_) {
    colorRounds_ = rounds;
    resetGraphs();
  }
}

bool Engine::adapt() { return adaptFrom(readCounters()); }

bool Engine::adaptFrom(const Counters &c) {
  bool changed = false;
  if (c.solveAborts && useDataflow_) {
    noteAborts(c);
    changed = true;
  }
  // Colour cap and rounds (solver.ts:1495-1511): grow at once when a spare runs out or the colouring clashed; shrink
  // to the colours in use + 3 after three quiet reads
  const uint32_t SPARE = 3;
  const uint32_t used = c.colors;
  uint32_t cap = colorCap_, rounds = colorRounds_;
  if (c.clashes > 0) rounds = std::min<uint32_t>(32, rounds * 2);
  if (c.clashes > 0 || cap < used + 2) {
    cap = std::min<uint32_t>(MAX_COLORS, std::max<uint32_t>(used + SPARE + 2, cap + 4));
    shrinkVotes_ = 0;
  } else if (used + SPARE + 2 < cap || rounds > 4) {
    if (++shrinkVotes_ >= 3) {
      cap = std::min<uint32_t>(MAX_COLORS, std::max<uint32_t>(used + SPARE, 1));
      rounds = std::max<uint32_t>(4, rounds - 4);
      shrinkVotes_ = 0;
    }
  } else {
    shrinkVotes_ = 0;
  }
  if (cap != colorCap_) {
    setColorCap(cap);
    changed = true;
  }
  if (rounds != colorRounds_) {
    setColorRounds(rounds);
    changed = true;
  }
  // Colouring rounds in one block while the mark pass leaves few bodies pending (three quiet reads), in one launch
  // per round again as soon as it leaves many (either way the same colours)
  const uint32_t LISTED_MAX = 2048;
  bool listed = listedRounds_;
  if (coloringRounds_ != 0) {
    listed = coloringRounds_ == 2;
  } else if (c.pending > LISTED_MAX) {
    listed = false;
    listedVotes_ = 0;
  } else if (!listed && ++listedVotes_ >= 3) {
    listed = true;
  }
  if (listed != listedRounds_) {
    listedRounds_ = listed;
    resetGraphs();
    changed = true;
  }
  // Storage: grow to twice the demand (the pairs found, also past the capacity) when full beyond 80% or overflowed,
  // and by at least a quarter (solver.ts:1512-1518)
  bool overflow = (c.overflow & 7u) != 0;
  if (overflow || c.pairs > 0.8 * pairCap_) {
    uint32_t want = uint32_t(std::max<double>(2.0 * c.pairs, 1.25 * pairCap_));
    allocateContacts(want);
    changed = true;
  }
  return changed;
}

}  // namespace avbd
