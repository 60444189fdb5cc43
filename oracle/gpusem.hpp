// ORACLE — test infrastructure only (see math.hpp).
//
// Sequential CPU restatement of the reference's GPU step semantics (the parity oracle for the HIP path), templated
// on the real type: World<float> is the FP32 mirror (bit-exact discrete decisions: pairs, feature keys, colours),
// World<double> the FP64 run used to state the FP32 error per quantity. Follows, in this order:
//   updateRefs            reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:20-47
//   pair predicate        wgsl-collision.ts:112-221 (testPair, faceSeparated, ignored; evaluated canonically with the
//                         higher-index body as the probe, so the set does not depend on the broadphase structure)
//   narrowphase           wgsl-collision.ts:617-808 (reuseContacts, SAT/sphere contacts, warm start by feature, then
//                         nearest anchor; contactPenaltyMin)
//   adjacency             reference/three-avbd/src/avbd2d/gpu/wgsl-topology.ts:61-151 (sorted per-body lists)
//   colouring             wgsl-topology.ts:155-322 (incremental Jones–Plassmann, lowbias32 priorities, cap + clashes)
//   warm starts           reference/three-avbd/src/avbd3d/gpu/wgsl-solve.ts:210-300
//   primal per colour     wgsl-solve.ts:53-206, 314-470 (outer-product accumulation, LDLᵀ with the 1e-5 pivot floor)
//   dual                  wgsl-solve.ts:474-535
//   velocities            wgsl-solve.ts:537-549
// Bodies of several independent worlds may share one World (batched stepping): a body carries its world id, each world
// its own parameters and step counter, pairs never cross worlds, and colouring priorities use the body's index within
// its world, so a world steps to the same result alone and in a batch.
#pragma once
#include <algorithm>
#include <cstdint>
#include <map>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "collide.hpp"
#include "hull.hpp"

namespace oracle {
namespace gpu {

constexpr uint32_t NO_COLOR = 255, PENDING = 256, MAX_COLORS = 64, STICK_BIT = 0x80000000u;
constexpr uint32_t T_NONE = 0, T_JOINT = 1, T_SPRING = 2, T_MOTOR = 3, T_LIMIT = 4;
constexpr uint32_t FLAG_MATCH_NEAREST = 1, FLAG_FACE_BIAS = 2, FLAG_REUSE_CONTACTS = 4, FLAG_START_AT_REST = 8,
                   FLAG_MASS_PENALTY = 16,
                   FLAG_CAPSULE_STEP_START_ARMS = 32;  // not the reference's: capsule lever arms at x⁻ too
constexpr float SHAPE_BOX = 0, SHAPE_SPHERE = 1, SHAPE_CAPSULE = -1;

/** Reference record layouts (reference/three-avbd/src/avbd3d/gpu/layout.ts:8-55, 127-169). */
struct BodyRec { float f[40]; };
struct JointRec { float f[32]; };
struct ManifoldRec { uint32_t ids[4]; float geo[4]; };
struct ContactRec { uint32_t w[16]; };

/** Per-world step parameters (layout.ts:171-207, the fields the 3D kernels read). */
struct WorldParams {
  float dt = 1.f / 60, gravity = -9.81f, betaLin = 10000, betaAng = 100, gamma = 0.999f, alpha = 0.99f;
  float up[3] = {0, 1, 0};
  uint32_t flags = FLAG_MATCH_NEAREST | FLAG_FACE_BIAS | FLAG_REUSE_CONTACTS | FLAG_START_AT_REST | FLAG_MASS_PENALTY;
  uint32_t step = 0;  // steps taken (the running step's number while it runs, from 1)
};

template <class R> struct Body {
  V3<R> pos; R friction = 0;
  Q4<R> rot;
  V3<R> ipos; R iposw = 0;
  Q4<R> irot;
  V3<R> size; R mass = 0;
  V3<R> moment; R radius = 0;
  V3<R> inertialPos; uint32_t stamp = 0;
  Q4<R> inertialRot;
  V3<R> vel; R prevUp = 0;
  V3<R> angVel; R shape = 0;
};

/** The 32-word constraint record (layout.ts:27-36). The w words of c0Lin, c0Ang, rA and rB carry a ball joint's rest
 * orientation (x, y, z, w); a spring keeps its rest length in rA.w and its break strain / force in frac / torqueArm;
 * a motor or limit keeps its warm-start values in c0LinW / c0AngW and speed, torque bound or angles in rest / rBW. */
template <class R> struct Joint {
  V3<R> penLin; R stiffLin = 0;
  V3<R> penAng; R stiffAng = 0;
  V3<R> lamLin; R frac = 0;
  V3<R> lamAng; R torqueArm = 0;
  V3<R> c0Lin; R c0LinW = 0;
  V3<R> c0Ang; R c0AngW = 0;
  V3<R> rA; R rest = 0;
  V3<R> rB; R rBW = 0;
  int32_t type = 0, a = -1, b = -1;
};

template <class R> struct Manifold {
  uint32_t a = 0, b = 0, first = 0, countw = 0;  // count | generation step << 4
  V3<R> normal; R friction = 0;
  uint32_t count() const { return countw & 15u; }
};

template <class R> struct Contact {
  V3<R> rA; uint32_t key = 0;
  V3<R> rB; R c0x = 0;
  V3<R> pen; R c0y = 0;
  V3<R> lam; R c0z = 0;
};

struct RefPose { float pos[4] = {0, 0, 0, 0}; float rot[4] = {0, 0, 0, 0}; };

template <class R> inline R toR(float x) { return R(x); }

/**
 * Contacts whose lever arms are taken at the step-start rotation (the Taylor point): the sphere key, as
 * wgsl-solve.ts:172-181 does (a lever arm turned with a rolling sphere puts a false separation into the normal row);
 * the capsule keys 4<<24 | j only with FLAG_CAPSULE_STEP_START_ARMS (an option of this build: the reference evaluates
 * capsule contacts at the current rotation, and a capsule rolling on its side then has the sphere's problem).
 */
inline bool stepStartArms(uint32_t key, uint32_t flags) {
  uint32_t k = key & ~STICK_BIT;
  return k == SPHERE_FEATURE || ((flags & FLAG_CAPSULE_STEP_START_ARMS) && (k >> 24) == (CAPSULE_FEATURE >> 24));
}

/** The GPU step's contact basis: rows (n, t1, t2) as orthonormal (layout.ts:236-242), with t1 scaled by the reciprocal
 * of its length (the HIP orthonormal's form). */
template <class R> inline M3<R> contactBasis(V3<R> n) {
  V3<R> t1 = std::fabs(n.x) > std::fabs(n.z) ? V3<R>(-n.y, n.x, 0) : V3<R>(0, -n.z, n.y);
  t1 = scale(t1, R(1) / length(t1));
  V3<R> t2 = cross(n, t1);
  M3<R> o;
  o.m[0] = n.x; o.m[1] = n.y; o.m[2] = n.z;
  o.m[3] = t1.x; o.m[4] = t1.y; o.m[5] = t1.z;
  o.m[6] = t2.x; o.m[7] = t2.y; o.m[8] = t2.z;
  return o;
}

/**
 * Contact points of one pair by shape (wgsl-collision.ts:684-726 with the hull branch :697-711 and the capsule splice
 * of capsuleCollision.js:17-29): capsules first (against a hull: capsulePoly), then two spheres, then pairs with a
 * hull (sphereHull or collidePoly, when a hull buffer is given), then sphere/box, then box/box SAT. sat.n points from
 * A towards B.
 */
template <class R>
inline Found<R> shapeContacts(const Body<R> &BA, const Body<R> &BB, const Box<R> &A, const Box<R> &B, bool faceBias,
                              Sat<R> &sat, const HullBuffer *H = nullptr) {
  Found<R> found;
  bool sphereA = BA.shape == R(SHAPE_SPHERE), sphereB = BB.shape == R(SHAPE_SPHERE);
  bool capsuleA = BA.shape == R(SHAPE_CAPSULE), capsuleB = BB.shape == R(SHAPE_CAPSULE);
  const bool hullA = H && BA.shape >= R(SHAPE_HULL_F), hullB = H && BB.shape >= R(SHAPE_HULL_F);
  auto poly = [&](const Body<R> &b) { return polyOf(*H, b.pos, b.rot, b.size, b.shape); };
  if (capsuleA || capsuleB) {
    // capsuleCollision.js:17-29: the capsule is the first argument; a capsule B flips the result back
    if (capsuleA && capsuleB) {
      found = capsuleCapsule(capsuleOf(BA.pos, BA.rot, BA.size), capsuleOf(BB.pos, BB.rot, BB.size), sat);
    } else if (capsuleA) {
      Capsule<R> C = capsuleOf(BA.pos, BA.rot, BA.size);
      if (sphereB) found = capsuleSphere(C, B.c, B.h.x, sat);
      else if (hullB) found = capsulePoly(*H, C, poly(BB), sat);
      else found = capsuleBox(C, B, BB.rot, sat);
    } else {
      Capsule<R> C = capsuleOf(BB.pos, BB.rot, BB.size);
      if (sphereA) found = capsuleSphere(C, A.c, A.h.x, sat);
      else if (hullA) found = capsulePoly(*H, C, poly(BA), sat);
      else found = capsuleBox(C, A, BA.rot, sat);
      sat.n = neg(sat.n);
      for (int c = 0; c < found.count; c++) std::swap(found.xA[c], found.xB[c]);
    }
  } else if (sphereA && sphereB) {
    found = sphereSphere(A, B, sat);
  } else if (hullA || hullB) {
    if (sphereA) {
      SphereBoxHit<R> h = sphereHull(*H, A.c, A.h.x, poly(BB));
      sat.n = h.n;
      if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
    } else if (sphereB) {
      SphereBoxHit<R> h = sphereHull(*H, B.c, B.h.x, poly(BA));
      sat.n = neg(h.n);
      if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
    } else {
      found = collidePoly(*H, poly(BA), poly(BB), faceBias, sat);
    }
  } else if (sphereA || sphereB) {
    if (sphereA) {
      SphereBoxHit<R> h = sphereBox(A, B, BB.rot);
      sat.n = h.n;
      if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
    } else {
      SphereBoxHit<R> h = sphereBox(B, A, BA.rot);
      sat.n = neg(h.n);
      if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
    }
  } else {
    found = collideBoxes(A, B, faceBias, 8, sat);
  }
  return found;
}

template <class R> class World {
 public:
  // Scene (in index order)
  std::vector<Body<R>> bodies;
  std::vector<uint32_t> world;        // per body
  std::vector<uint32_t> worldBase;    // per world: first body index
  std::vector<WorldParams> params;    // per world
  std::vector<Joint<R>> joints;
  std::vector<uint32_t> filters;      // 2 per body: groups, collides-with
  HullBuffer hulls;                   // hull records (shape codes 3 + header word)
  std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> noCollide;  // (hi, lo, joint slot or ~0), sorted
  // Persistent state
  std::vector<RefPose> refs;
  std::vector<uint32_t> colorA;       // colour state per body (colour | PENDING)
  std::vector<Manifold<R>> manifolds, prevManifolds;
  std::vector<Contact<R>> contacts, prevContacts;
  // Per-step products
  std::vector<std::pair<uint32_t, uint32_t>> pairs;  // (a > b), sorted
  std::vector<uint32_t> adjStart, adjList;
  std::vector<uint32_t> colorStart, colorBodies;
  uint32_t clashes = 0, numColors = 0;
  uint32_t pivotFallbacks = 0;  // bodies solved again in FP64 in the last step
  // Pair storage as the GPU world sizes it (0: unbounded); a step with more pairs overflows and is held
  uint32_t pairCapacity = 0, overflow = 0;
  // Global schedule
  uint32_t colorCap = 64, colorRounds = 16;
  int iterations = 10;
  bool bruteForcePairs = false;

  size_t n() const { return bodies.size(); }
  bool dynamicBody(int64_t i) const { return i >= 0 && bodies[size_t(i)].mass > 0; }
  const WorldParams &P(uint32_t i) const { return params[world[i]]; }
  uint32_t localIndex(uint32_t i) const { return i - worldBase[world[i]]; }

  // --- collision -----------------------------------------------------------------------------------------------

  void beginFrame() {
    prevManifolds.swap(manifolds);
    prevContacts.swap(contacts);
    manifolds.clear();
    contacts.clear();
  }

  void updateRefs() {
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n(); i++) {
      const WorldParams &p = P(i);
      if (!(p.flags & FLAG_REUSE_CONTACTS)) continue;
      Body<R> &b = bodies[i];
      RefPose &last = refs[i];
      Q4<R> lr(toR<R>(last.rot[0]), toR<R>(last.rot[1]), toR<R>(last.rot[2]), toR<R>(last.rot[3]));
      V3<R> lp(toR<R>(last.pos[0]), toR<R>(last.pos[1]), toR<R>(last.pos[2]));
      bool moved = (lr.x * lr.x + lr.y * lr.y + lr.z * lr.z + lr.w * lr.w) == 0 ||
                   length(sub(b.pos, lp)) > R(0.002f) || length(qsub(b.rot, lr)) > R(0.002f);
      if (moved) {
        last.pos[0] = float(b.pos.x); last.pos[1] = float(b.pos.y); last.pos[2] = float(b.pos.z); last.pos[3] = 0;
        last.rot[0] = float(b.rot.x); last.rot[1] = float(b.rot.y); last.rot[2] = float(b.rot.z); last.rot[3] = float(b.rot.w);
        b.stamp = p.step;
      }
    }
  }

  struct Shape {
    V3<R> ax[3];
    V3<R> h;
    bool sphere;
  };
  Shape shapeOf(uint32_t i) const {
    const Body<R> &b = bodies[i];
    Shape s;
    s.h = scale(b.size, R(0.5));
    s.sphere = b.shape == R(SHAPE_SPHERE);
    s.ax[0] = rotate(b.rot, V3<R>(1, 0, 0));
    s.ax[1] = rotate(b.rot, V3<R>(0, 1, 0));
    s.ax[2] = rotate(b.rot, V3<R>(0, 0, 1));
    return s;
  }
  static R extent(const Shape &s, V3<R> n) {
    if (s.sphere) return s.h.x;
    return s.h.x * std::fabs(dot(n, s.ax[0])) + s.h.y * std::fabs(dot(n, s.ax[1])) + s.h.z * std::fabs(dot(n, s.ax[2]));
  }
  static V3<R> aabbHalf(const Shape &s) {
    if (s.sphere) return V3<R>(s.h.x, s.h.x, s.h.x);
    return add(add(scale(vabs(s.ax[0]), s.h.x), scale(vabs(s.ax[1]), s.h.y)), scale(vabs(s.ax[2]), s.h.z));
  }
  static bool faceSeparated(const Shape &A, const Shape &B, V3<R> d) {
    for (int k = 0; k < 3; k++) {
      if (std::fabs(dot(d, A.ax[k])) - extent(A, A.ax[k]) - extent(B, A.ax[k]) > R(1e-3f)) return true;
      if (std::fabs(dot(d, B.ax[k])) - extent(A, B.ax[k]) - extent(B, B.ax[k]) > R(1e-3f)) return true;
    }
    return false;
  }
  bool ignored(uint32_t hi, uint32_t lo) const {
    auto it = std::lower_bound(noCollide.begin(), noCollide.end(), std::make_tuple(hi, lo, 0u));
    for (; it != noCollide.end(); ++it) {
      if (std::get<0>(*it) != hi || std::get<1>(*it) != lo) break;
      uint32_t j = std::get<2>(*it);
      if (j == 0xffffffffu) return true;
      if (joints[j].stiffLin != 0 || joints[j].stiffAng != 0) return true;
    }
    return false;
  }
  /** The pair predicate, with a = the higher index as the probe (testPair, wgsl-collision.ts:196-221). */
  bool acceptPair(uint32_t a, uint32_t b) const {
    if (world[a] != world[b]) return false;
    const Body<R> &A = bodies[a], &B = bodies[b];
    if (!(A.mass > 0) && B.mass <= 0) return false;
    uint32_t fa0 = filters[2 * a], fa1 = filters[2 * a + 1], fb0 = filters[2 * b], fb1 = filters[2 * b + 1];
    if ((fa0 & fb1) == 0 || (fb0 & fa1) == 0) return false;
    V3<R> d = sub(A.pos, B.pos);
    R r = A.radius + B.radius;
    if (dot(d, d) > r * r) return false;
    Shape sa = shapeOf(a), sb = shapeOf(b);
    V3<R> ha = aabbHalf(sa), hb = aabbHalf(sb);
    R pad = R(1e-4f);
    if (std::fabs(d.x) > ha.x + hb.x + pad || std::fabs(d.y) > ha.y + hb.y + pad || std::fabs(d.z) > ha.z + hb.z + pad)
      return false;
    if (faceSeparated(sa, sb, d)) return false;
    if (ignored(a, b)) return false;
    return true;
  }

  void findPairs() {
    pairs.clear();
    const uint32_t N = uint32_t(n());
    if (bruteForcePairs || N < 2048) {
      for (uint32_t a = 0; a < N; a++)
        for (uint32_t b = 0; b < a; b++)
          if (acceptPair(a, b)) pairs.push_back({a, b});
      return;
    }
    // A uniform grid over (world, cell): the largest bodies (above 4x the median radius) are tested against all
    // bodies of their world, the rest are binned in cells of twice the largest small radius.
    std::vector<R> radii(N);
    for (uint32_t i = 0; i < N; i++) radii[i] = bodies[i].radius;
    std::vector<R> sorted = radii;
    std::nth_element(sorted.begin(), sorted.begin() + N / 2, sorted.end());
    R median = sorted[N / 2];
    std::vector<char> large(N, 0);
    R maxSmall = 0;
    for (uint32_t i = 0; i < N; i++) {
      if (radii[i] > 4 * median) large[i] = 1;
      else maxSmall = rmax(maxSmall, radii[i]);
    }
    double cell = rmax(2.0 * double(maxSmall) * 1.0001, 1e-3);
    auto key = [&](uint32_t w, int64_t x, int64_t y, int64_t z) {
      return (uint64_t(w) << 48) ^ (uint64_t(x & 0xffff) << 32) ^ (uint64_t(y & 0xffff) << 16) ^ uint64_t(z & 0xffff);
    };
    auto cellOf = [&](uint32_t i, int64_t *c) {
      c[0] = int64_t(std::floor(double(bodies[i].pos.x) / cell));
      c[1] = int64_t(std::floor(double(bodies[i].pos.y) / cell));
      c[2] = int64_t(std::floor(double(bodies[i].pos.z) / cell));
    };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    std::vector<std::vector<uint32_t>> largeOf(worldBase.size());
    for (uint32_t i = 0; i < N; i++) {
      if (large[i]) {
        largeOf[world[i]].push_back(i);
        continue;
      }
      int64_t c[3];
      cellOf(i, c);
      grid[key(world[i], c[0], c[1], c[2])].push_back(i);
    }
    // Pairs (i, j), j < i, grouped by i: a small body's grid neighbours and the large bodies of its world below it; a
    // large body's whole world below it. Each group sorted and unique, so the groups in order are the sorted pair set.
    std::vector<std::vector<uint32_t>> below(N);
#pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t i = 0; i < N; i++) {
      std::vector<uint32_t> &js = below[i];
      if (large[i]) {
        for (uint32_t j = worldBase[world[i]]; j < i; j++) js.push_back(j);
      } else {
        int64_t c[3];
        cellOf(i, c);
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              auto it = grid.find(key(world[i], c[0] + dx, c[1] + dy, c[2] + dz));
              if (it == grid.end()) continue;
              for (uint32_t j : it->second)
                if (j < i) js.push_back(j);
            }
        for (uint32_t l : largeOf[world[i]])
          if (l < i) js.push_back(l);
        std::sort(js.begin(), js.end());
        js.erase(std::unique(js.begin(), js.end()), js.end());
      }
      size_t kept = 0;
      for (uint32_t j : js)
        if (acceptPair(i, j)) js[kept++] = j;
      js.resize(kept);
    }
    for (uint32_t i = 0; i < N; i++)
      for (uint32_t j : below[i]) pairs.push_back({i, j});
  }

  R contactPenaltyMin(uint32_t a, uint32_t b) const {
    const WorldParams &p = P(a);
    if (!(p.flags & FLAG_MASS_PENALTY)) return R(1);
    R mA = bodies[a].mass <= 0 ? R(3.4e38f) : bodies[a].mass;
    R mB = bodies[b].mass <= 0 ? R(3.4e38f) : bodies[b].mass;
    R dt = R(p.dt);
    return rclamp(rmin(mA, mB) / (dt * dt), R(1), R(1e10f));
  }

  /** The last step's manifolds by pair: (a << 32 | b, index) sorted by key; a repeated pair keeps its last index. */
  std::vector<std::pair<uint64_t, uint32_t>> prevIndex;

  /** Index of the last step's manifold of pair (a, b), or -1. */
  int64_t previousManifold(uint32_t a, uint32_t b) const {
    const uint64_t k = uint64_t(a) << 32 | b;
    auto it = std::upper_bound(prevIndex.begin(), prevIndex.end(), std::make_pair(k, UINT32_MAX));
    if (it == prevIndex.begin() || (it - 1)->first != k) return -1;
    return (it - 1)->second;
  }

  /** One pair's narrowphase output: its manifold (first filled in when appended) and points; count < 0: none. */
  struct PairOut {
    Manifold<R> m;
    Contact<R> c[8];
    int count = -1;
  };
  std::vector<PairOut> pairOuts;

  bool reuseContacts(uint32_t a, uint32_t b, PairOut &o) const {
    int64_t at = previousManifold(a, b);
    if (at < 0) return false;
    const Manifold<R> prev = prevManifolds[size_t(at)];
    uint32_t count = prev.count();
    uint32_t generated = prev.countw >> 4;
    uint32_t moved = rmax(bodies[a].stamp, bodies[b].stamp);
    if (count == 0 || generated < moved) return false;
    const WorldParams &p = P(a);
    o.m = prev;
    o.count = int(count);
    const Body<R> &A = bodies[a], &B = bodies[b];
    M3<R> basis = contactBasis(prev.normal);
    R alpha = R(p.alpha), gamma = R(p.gamma);
    R pmin = contactPenaltyMin(a, b);
    for (uint32_t i = 0; i < count; i++) {
      Contact<R> k = prevContacts[prev.first + i];
      V3<R> d = sub(add(rotate(A.rot, k.rA), A.pos), add(rotate(B.rot, k.rB), B.pos));
      k.c0x = dot(row(basis, 0), d) + R(0.01f);
      k.c0y = dot(row(basis, 1), d);
      k.c0z = dot(row(basis, 2), d);
      k.lam = scale(scale(k.lam, alpha), gamma);
      k.pen = V3<R>(rclampW(k.pen.x * gamma), rclampW(k.pen.y * gamma), rclampW(k.pen.z * gamma));
      k.pen.x = rmaxW(k.pen.x, pmin);
      o.c[i] = k;
    }
    return true;
  }
  // WGSL clamp(e, lo, hi) = min(max(e, lo), hi); max/min of two finite floats
  static R rclampW(R x) { return rmin(rmax(x, R(1)), R(1e10f)); }
  static R rmaxW(R a, R b) { return a > b ? a : b; }

  /** One pair's contacts (wgsl-collision.ts:617-808) into `o`, from the bodies and the last step's points only. */
  void pairContacts(uint32_t a, uint32_t b, PairOut &o) const {
    o.count = -1;
    const WorldParams &p = P(a);
    if ((p.flags & FLAG_REUSE_CONTACTS) && reuseContacts(a, b, o)) return;
    const Body<R> &BA = bodies[a], &BB = bodies[b];
    Box<R> A = makeBox(BA.pos, BA.rot, BA.size), B = makeBox(BB.pos, BB.rot, BB.size);
    Sat<R> sat;
    Found<R> found;
    found = shapeContacts(BA, BB, A, B, (p.flags & FLAG_FACE_BIAS) != 0, sat, &hulls);
    bool sphereA = BA.shape == R(SHAPE_SPHERE), sphereB = BB.shape == R(SHAPE_SPHERE);
    bool capsuleA = BA.shape == R(SHAPE_CAPSULE), capsuleB = BB.shape == R(SHAPE_CAPSULE);
    if (found.count == 0) return;
    Manifold<R> &mf = o.m;
    mf.a = a;
    mf.b = b;
    mf.first = 0;
    mf.countw = uint32_t(found.count) | (p.step << 4);
    mf.normal = neg(sat.n);
    mf.friction = std::sqrt(BA.friction * BB.friction);
    o.count = found.count;
    M3<R> basis = contactBasis(mf.normal);
    bool anySphere = sphereA || sphereB || capsuleA || capsuleB;
    uint32_t prevFirst = 0, prevCount = 0;
    int64_t at = previousManifold(a, b);
    if (at >= 0) {
      prevFirst = prevManifolds[size_t(at)].first;
      prevCount = prevManifolds[size_t(at)].count();
    }
    bool matchNearest = (p.flags & FLAG_MATCH_NEAREST) != 0;
    R minSide = rmin(rmin(rmin(A.h.x, A.h.y), A.h.z), rmin(rmin(B.h.x, B.h.y), B.h.z)) * R(2);
    R pmin = contactPenaltyMin(a, b);
    R alpha = R(p.alpha), gamma = R(p.gamma);
    for (int i = 0; i < found.count; i++) {
      V3<R> rA = rotate(conj(BA.rot), sub(found.xA[i], A.c));
      V3<R> rB = rotate(conj(BB.rot), sub(found.xB[i], B.c));
      V3<R> pen, lam;
      uint32_t stick = 0;
      int64_t j = -1;
      for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++)
        if ((prevContacts[c].key & ~STICK_BIT) == found.feature[i]) {
          j = c;
          break;
        }
      if (j < 0 && matchNearest) {
        R bestDist = R(0.05f) * minSide;
        for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++) {
          R dist = length(sub(prevContacts[c].rA, rA));
          if (dist <= bestDist) {
            bestDist = dist;
            j = c;
          }
        }
      }
      if (j >= 0) {
        const Contact<R> &pc = prevContacts[size_t(j)];
        pen = pc.pen;
        lam = pc.lam;
        stick = pc.key & STICK_BIT;
        if (stick != 0 && !anySphere) {
          rA = pc.rA;
          rB = pc.rB;
        }
      }
      V3<R> d = sub(add(rotate(BA.rot, rA), A.c), add(rotate(BB.rot, rB), B.c));
      Contact<R> k;
      k.rA = rA;
      k.rB = rB;
      k.key = found.feature[i] | stick;
      k.c0x = dot(row(basis, 0), d) + R(0.01f);
      k.c0y = dot(row(basis, 1), d);
      k.c0z = dot(row(basis, 2), d);
      k.lam = scale(scale(lam, alpha), gamma);
      k.pen = V3<R>(rclampW(pen.x * gamma), rclampW(pen.y * gamma), rclampW(pen.z * gamma));
      k.pen.x = rmaxW(k.pen.x, pmin);
      o.c[i] = k;
    }
  }

  /** Every pair's contacts (in parallel), then the manifolds and points appended in pair order. */
  void narrowphase() {
    prevIndex.resize(prevManifolds.size());
    for (uint32_t m = 0; m < prevManifolds.size(); m++)
      prevIndex[m] = {uint64_t(prevManifolds[m].a) << 32 | prevManifolds[m].b, m};
    std::sort(prevIndex.begin(), prevIndex.end());
    if (pairOuts.size() < pairs.size()) pairOuts.resize(pairs.size());
#pragma omp parallel for schedule(dynamic, 128)
    for (size_t q = 0; q < pairs.size(); q++) pairContacts(pairs[q].first, pairs[q].second, pairOuts[q]);
    for (size_t q = 0; q < pairs.size(); q++) {
      PairOut &o = pairOuts[q];
      if (o.count < 0) continue;
      o.m.first = uint32_t(contacts.size());
      manifolds.push_back(o.m);
      for (int i = 0; i < o.count; i++) contacts.push_back(o.c[i]);
    }
  }

  // --- topology --------------------------------------------------------------------------------------------------

  bool jointActive(uint32_t j) const {
    return joints[j].type != int32_t(T_NONE) && (joints[j].stiffLin != 0 || joints[j].stiffAng != 0);
  }
  std::pair<int64_t, int64_t> endpoints(uint32_t id) const {
    uint32_t J = uint32_t(joints.size());
    if (id < J) return {joints[id].a, joints[id].b};
    return {manifolds[id - J].a, manifolds[id - J].b};
  }
  void buildAdjacency() {
    const uint32_t N = uint32_t(n()), J = uint32_t(joints.size());
    std::vector<std::vector<uint32_t>> lists(N);
    auto addEnds = [&](uint32_t id) {
      auto e = endpoints(id);
      if (dynamicBody(e.first)) lists[size_t(e.first)].push_back(id);
      if (dynamicBody(e.second)) lists[size_t(e.second)].push_back(id);
    };
    for (uint32_t j = 0; j < J; j++)
      if (jointActive(j)) addEnds(j);
    for (uint32_t m = 0; m < manifolds.size(); m++) addEnds(J + m);
    adjStart.assign(N + 1, 0);
    adjList.clear();
    for (uint32_t b = 0; b < N; b++) {
      auto keyOf = [&](uint32_t id) -> std::pair<uint64_t, uint64_t> {
        if (id < J) return {0, id};
        auto e = endpoints(id);
        int64_t other = e.first == int64_t(b) ? e.second : e.first;
        return {uint64_t(other + 2), 0};
      };
      std::stable_sort(lists[b].begin(), lists[b].end(), [&](uint32_t x, uint32_t y) { return keyOf(x) < keyOf(y); });
      adjStart[b] = uint32_t(adjList.size());
      adjList.insert(adjList.end(), lists[b].begin(), lists[b].end());
    }
    adjStart[N] = uint32_t(adjList.size());
  }

  uint32_t priority(uint32_t i) const { return hash32(localIndex(i) + 0x9e3779b9u); }
  bool beats(uint32_t i, uint32_t j) const {
    uint32_t pi = priority(i), pj = priority(j);
    return pi > pj || (pi == pj && localIndex(i) > localIndex(j));
  }
  int64_t neighbour(uint32_t b, uint32_t e) const {
    auto ends = endpoints(adjList[e]);
    int64_t other = ends.first == int64_t(b) ? ends.second : ends.first;
    return dynamicBody(other) ? other : -1;
  }
  uint32_t smallestFree(uint32_t b, const std::vector<uint32_t> &src, bool skipPending) const {
    uint32_t lo = 0, hi = 0;
    for (uint32_t e = adjStart[b]; e < adjStart[b + 1]; e++) {
      int64_t j = neighbour(b, e);
      if (j < 0) continue;
      uint32_t s = src[size_t(j)];
      if (skipPending && (s & PENDING)) continue;
      uint32_t col = s & 0xffu;
      if (col == NO_COLOR) continue;
      if (col < 32) lo |= 1u << col;
      else hi |= 1u << (col - 32);
    }
    if (~lo != 0) return uint32_t(__builtin_ctz(~lo));
    if (~hi != 0) return 32u + uint32_t(__builtin_ctz(~hi));
    return MAX_COLORS - 1;
  }
  void colour() {
    const uint32_t N = uint32_t(n());
    std::vector<uint32_t> B(N);
    // 1. compaction (colorCompact)
#pragma omp parallel for schedule(static)
    for (uint32_t b = 0; b < N; b++) {
      if (!dynamicBody(b)) {
        B[b] = NO_COLOR;
        continue;
      }
      uint32_t col = colorA[b] & 0xffu;
      if (col >= colorCap) col = NO_COLOR;  // a colour at or above a lowered cap is coloured again
      if (col != NO_COLOR) {
        bool isMax = true;
        for (uint32_t e = adjStart[b]; e < adjStart[b + 1] && isMax; e++) {
          int64_t j = neighbour(b, e);
          if (j >= 0 && !beats(b, uint32_t(j))) isMax = false;
        }
        if (isMax) col = rmin(col, smallestFree(b, colorA, false));
      }
      B[b] = col;
    }
    // 2. pending (colorMark)
#pragma omp parallel for schedule(static)
    for (uint32_t b = 0; b < N; b++) {
      uint32_t col = B[b] & 0xffu;
      if (!dynamicBody(b)) {
        colorA[b] = NO_COLOR;
        continue;
      }
      bool pending = col == NO_COLOR;
      for (uint32_t e = adjStart[b]; e < adjStart[b + 1] && !pending; e++) {
        int64_t j = neighbour(b, e);
        if (j >= 0 && (B[size_t(j)] & 0xffu) == col && beats(uint32_t(j), b)) pending = true;
      }
      colorA[b] = col | (pending ? PENDING : 0);
    }
    // 3. Jones–Plassmann rounds, Jacobi between the two state buffers
    for (uint32_t r = 0; r < colorRounds; r++) {
      const std::vector<uint32_t> &src = (r % 2 == 0) ? colorA : B;
      std::vector<uint32_t> &dst = (r % 2 == 0) ? B : colorA;
#pragma omp parallel for schedule(static)
      for (uint32_t b = 0; b < N; b++) {  // Jacobi between the two buffers: bodies are independent
        uint32_t s = src[b];
        if (!(s & PENDING)) {
          dst[b] = s;
          continue;
        }
        bool isMax = true;
        for (uint32_t e = adjStart[b]; e < adjStart[b + 1] && isMax; e++) {
          int64_t j = neighbour(b, e);
          if (j >= 0 && (src[size_t(j)] & PENDING) && !beats(b, uint32_t(j))) isMax = false;
        }
        if (isMax) {
          uint32_t fr = smallestFree(b, src, true);
          dst[b] = fr >= colorCap ? ((colorCap - 1) | PENDING) : fr;
        } else {
          dst[b] = s;
        }
      }
    }
    // 4-6. counts, clashes and stable buckets
    clashes = 0;
    numColors = 0;
    colorStart.assign(MAX_COLORS + 1, 0);
    for (uint32_t b = 0; b < N; b++) {
      if (!dynamicBody(b)) continue;
      uint32_t s = colorA[b];
      uint32_t col = s & 0xffu;
      if (s & PENDING) clashes++;
      if (col == NO_COLOR) col = colorCap - 1;
      colorA[b] = col;
      numColors = rmax(numColors, col + 1);
      colorStart[col + 1]++;
    }
    for (uint32_t c = 0; c < MAX_COLORS; c++) colorStart[c + 1] += colorStart[c];
    colorBodies.assign(colorStart[MAX_COLORS], 0);
    std::vector<uint32_t> cursor(colorStart.begin(), colorStart.end() - 1);
    for (uint32_t b = 0; b < N; b++)
      if (dynamicBody(b)) colorBodies[cursor[colorA[b]]++] = b;
  }

  // --- solve -------------------------------------------------------------------------------------------------------

  V3<R> anchorA(const Joint<R> &k) const {
    if (k.a < 0) return k.rA;
    const Body<R> &A = bodies[size_t(k.a)];
    return add(rotate(A.rot, k.rA), A.pos);
  }
  V3<R> jointLinC(const Joint<R> &k) const {
    const Body<R> &B = bodies[size_t(k.b)];
    return sub(anchorA(k), add(rotate(B.rot, k.rB), B.pos));
  }
  /** withJointRestFrames (src/gpu/gpuSolverKernels.js:105-130): B is held at A's rotation times the captured rest
   * orientation (A's rotation alone for a record without one). */
  Q4<R> restTarget(const Joint<R> &k) const {
    const Q4<R> t = k.a < 0 ? Q4<R>(0, 0, 0, 1) : bodies[size_t(k.a)].rot;
    const Q4<R> r(k.c0LinW, k.c0AngW, k.rest, k.rBW);
    if (!(qlengthSq(r) > R(0.5))) return t;
    return Q4<R>((t.w * r.x + r.w * t.x) + (t.y * r.z - t.z * r.y), (t.w * r.y + r.w * t.y) + (t.z * r.x - t.x * r.z),
                 (t.w * r.z + r.w * t.z) + (t.x * r.y - t.y * r.x), t.w * r.w - ((t.x * r.x + t.y * r.y) + t.z * r.z));
  }
  V3<R> jointAngC(const Joint<R> &k) const {
    return scale(qsub(restTarget(k), bodies[size_t(k.b)].rot), k.torqueArm);
  }

  // Hinge motors and limits (src/gpu/angularConstraints.js:16-83)
  struct HingeFrame {
    V3<R> axis, ta, tb;
  };
  HingeFrame hingeFrame(const Joint<R> &k) const {
    const Q4<R> qa = k.a < 0 ? Q4<R>(0, 0, 0, 1) : bodies[size_t(k.a)].rot;
    return HingeFrame{rotate(qa, k.rA), rotate(qa, k.c0Lin), rotate(bodies[size_t(k.b)].rot, k.c0Ang)};
  }
  static R hingeAngle(const HingeFrame &f) { return atan2R(dot(f.axis, cross(f.ta, f.tb)), dot(f.ta, f.tb)); }
  static std::pair<R, R> hingeRows(const Joint<R> &k, R angle, R alpha, R dt) {
    if (k.type == int32_t(T_MOTOR)) return {wrapAngleR(angle - k.c0LinW) - k.rest * dt, R(0)};
    return {(angle - k.rest) - alpha * k.c0LinW, (angle - k.rBW) - alpha * k.c0AngW};
  }
  static std::pair<R, R> hingeForce(const Joint<R> &k, std::pair<R, R> C) {
    const R fx = k.penLin.x * C.first + k.lamLin.x, fy = k.penLin.y * C.second + k.lamLin.y;
    if (k.type == int32_t(T_MOTOR)) return {rmin(rmax(fx, -k.rBW), k.rBW), R(0)};
    return {rmin(R(0), fx), rmax(R(0), fy)};
  }
  void warmHinge(Joint<R> &k) const {
    if (k.stiffLin == 0) return;
    const WorldParams &p = P(uint32_t(k.b));
    const R angle = hingeAngle(hingeFrame(k));
    k.c0LinW = k.type == int32_t(T_MOTOR) ? angle : rmin(R(0), angle - k.rest);
    k.c0AngW = rmax(R(0), angle - k.rBW);
    k.lamLin = scale(k.lamLin, R(p.alpha) * R(p.gamma));
    const R g = R(p.gamma);
    k.penLin = V3<R>(rclampW(k.penLin.x * g), rclampW(k.penLin.y * g), rclampW(k.penLin.z * g));
  }
  void dualHinge(Joint<R> &k, R alpha) const {
    if (k.stiffLin == 0) return;
    const WorldParams &p = P(uint32_t(k.b));
    const std::pair<R, R> C = hingeRows(k, hingeAngle(hingeFrame(k)), alpha, R(p.dt));
    const std::pair<R, R> F = hingeForce(k, C);
    k.lamLin = V3<R>(F.first, F.second, R(0));
    const bool motor = k.type == int32_t(T_MOTOR);
    const R vx = motor ? std::fabs(C.first) : rmax(R(0), -C.first), vy = motor ? std::fabs(C.second) : rmax(R(0), C.second);
    k.penLin.x = rmin(k.penLin.x + vx * R(p.betaAng), R(1e10f));
    k.penLin.y = rmin(k.penLin.y + vy * R(p.betaAng), R(1e10f));
  }
  /** withSpringFracture (src/gpu/springMaterial.js:39-68): past its break strain or force a spring stops acting. */
  void springFracture(Joint<R> &k) const {
    const R strainLimit = k.frac, forceLimit = k.torqueArm;
    if (!(k.stiffLin > 0 && (strainLimit > 0 || forceLimit > 0))) return;
    const Body<R> &B = bodies[size_t(k.b)];
    const R lengthNow = length(sub(anchorA(k), add(rotate(B.rot, k.rB), B.pos)));
    const R extension = rmax(R(0), lengthNow - k.rest);
    const R strain = extension / rmax(k.rest, R(1.0e-6f));
    const R tension = k.stiffLin * extension;
    if ((strainLimit > 0 && k.rest > R(1.0e-6f) && strain > strainLimit) || (forceLimit > 0 && tension > forceLimit)) {
      k.penLin = V3<R>();
      k.stiffLin = 0;
      k.penAng = V3<R>();
      k.stiffAng = 0;
      k.lamLin = V3<R>();
      k.lamAng = V3<R>();
    }
  }

  /** The frames a new constraint takes from the current poses (constraintFrames.wgsl:11-19, angularFrames.wgsl): a
   * ball joint with angular stiffness its rest orientation, a motor or limit its reference tangents. */
  void captureFrame(Joint<R> &k) const {
    const Q4<R> qa = k.a < 0 ? Q4<R>(0, 0, 0, 1) : bodies[size_t(k.a)].rot;
    const Q4<R> qb = bodies[size_t(k.b)].rot;
    auto perpendicular = [](V3<R> v) {
      return std::fabs(v.x) > std::fabs(v.z) ? V3<R>(-v.y, v.x, 0) : V3<R>(0, -v.z, v.y);
    };
    if (k.type == int32_t(T_JOINT)) {
      const Q4<R> a = conj(qa), b = qb;
      const Q4<R> r((a.w * b.x + b.w * a.x) + (a.y * b.z - a.z * b.y), (a.w * b.y + b.w * a.y) + (a.z * b.x - a.x * b.z),
                    (a.w * b.z + b.w * a.z) + (a.x * b.y - a.y * b.x), a.w * b.w - ((a.x * b.x + a.y * b.y) + a.z * b.z));
      const R l = std::sqrt(qlengthSq(r));
      k.c0LinW = r.x / l;
      k.c0AngW = r.y / l;
      k.rest = r.z / l;
      k.rBW = r.w / l;
    } else if (k.type == int32_t(T_MOTOR) || k.type == int32_t(T_LIMIT)) {
      const V3<R> tangent = normalize(perpendicular(k.rA));
      const V3<R> world = rotate(qa, tangent);
      const V3<R> axisB = normalize(k.rB);
      V3<R> tangentB = rotate(conj(qb), world);
      tangentB = sub(tangentB, scale(axisB, dot(tangentB, axisB)));
      if (length(tangentB) < R(1.0e-6f)) tangentB = perpendicular(axisB);
      k.c0Lin = tangent;
      k.c0LinW = 0;
      k.c0Ang = normalize(tangentB);
      k.c0AngW = 0;
    }
  }
  /** Frames of every constraint that takes them at creation: ball joints with angular stiffness, motors, limits. */
  void captureFrames() {
    for (auto &k : joints)
      if ((k.type == int32_t(T_JOINT) && k.stiffAng > 0) || k.type == int32_t(T_MOTOR) || k.type == int32_t(T_LIMIT))
        captureFrame(k);
  }
  static bool hard(R stiffness) { return stiffness >= R(1e30f); }

  void warmStartJoints() {
#pragma omp parallel for schedule(static)
    for (size_t jj = 0; jj < joints.size(); jj++) {
      Joint<R> &k = joints[jj];
      if (k.type == int32_t(T_MOTOR) || k.type == int32_t(T_LIMIT)) {
        warmHinge(k);
        continue;
      }
      if (k.type != int32_t(T_JOINT)) continue;
      const WorldParams &p = P(uint32_t(k.b));
      k.c0Lin = jointLinC(k);
      k.c0Ang = jointAngC(k);
      R decay = R(p.alpha) * R(p.gamma);
      k.lamLin = scale(k.lamLin, decay);
      k.lamAng = scale(k.lamAng, decay);
      R g = R(p.gamma);
      k.penLin = V3<R>(rmin(rclampW(k.penLin.x * g), k.stiffLin), rmin(rclampW(k.penLin.y * g), k.stiffLin),
                       rmin(rclampW(k.penLin.z * g), k.stiffLin));
      k.penAng = V3<R>(rmin(rclampW(k.penAng.x * g), k.stiffAng), rmin(rclampW(k.penAng.y * g), k.stiffAng),
                       rmin(rclampW(k.penAng.z * g), k.stiffAng));
    }
  }

  void warmStartBodies() {
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n(); i++) {
      Body<R> &b = bodies[i];
      const WorldParams &p = P(i);
      R dt = R(p.dt), g = R(p.gravity);
      V3<R> up(R(p.up[0]), R(p.up[1]), R(p.up[2]));
      bool dyn = b.mass > 0;
      V3<R> inertialPos = add(b.pos, scale(b.vel, dt));
      if (dyn) inertialPos = add(inertialPos, scale(up, g * (dt * dt)));
      b.inertialPos = inertialPos;
      b.inertialRot = qadd(b.rot, scale(b.angVel, dt));
      R w = 0;
      if (std::fabs(g) > 0) {
        R x = (dot(b.vel, up) - b.prevUp) / dt * rsign(g) / std::fabs(g);
        w = rmin(rmax(x, R(0)), R(1));
      }
      bool slow = length(b.vel) + length(b.angVel) * b.radius < R(0.1f);
      bool atRest = (p.flags & FLAG_START_AT_REST) && slow;
      b.ipos = b.pos;
      b.iposw = b.friction;
      b.irot = b.rot;
      if (dyn && !atRest) {
        b.pos = add(add(b.pos, scale(b.vel, dt)), scale(up, g * (w * dt * dt)));
        b.rot = qadd(b.rot, scale(b.angVel, dt));
      }
      bool moving = b.vel.x != 0 || b.vel.y != 0 || b.vel.z != 0 || b.angVel.x != 0 || b.angVel.y != 0 || b.angVel.z != 0;
      if (!dyn && moving) {
        b.pos = add(b.pos, scale(b.vel, dt));
        b.rot = qadd(b.rot, scale(b.angVel, dt));
      }
    }
  }

  struct Acc {
    M3<R> lin, ang, cross;
    V3<R> rLin, rAng;
  };
  /** addRow (wgsl-solve.ts:53-59) with each accumulation as one fused multiply-add (the HIP primal's v_fma_f32):
   * lin += (l lᵀ) k, ang += (a aᵀ) k, cross += (a lᵀ) k, rhs += (l, a) f. */
  static void addRow(Acc &acc, V3<R> l, V3<R> a, R k, R f) {
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        acc.lin.m[r * 3 + c] = std::fma(l[r] * l[c], k, acc.lin.m[r * 3 + c]);
        acc.ang.m[r * 3 + c] = std::fma(a[r] * a[c], k, acc.ang.m[r * 3 + c]);
        acc.cross.m[r * 3 + c] = std::fma(a[r] * l[c], k, acc.cross.m[r * 3 + c]);
      }
    acc.rLin = V3<R>(std::fma(l.x, f, acc.rLin.x), std::fma(l.y, f, acc.rLin.y), std::fma(l.z, f, acc.rLin.z));
    acc.rAng = V3<R>(std::fma(a.x, f, acc.rAng.x), std::fma(a.y, f, acc.rAng.y), std::fma(a.z, f, acc.rAng.z));
  }

  // The contact rows' arithmetic as the HIP solve forms it: every sum of products one chain of fused multiply-adds
  // (solve.hip dotFused, crossFused, rotFused, dispFused).
  static R dotF(V3<R> a, V3<R> b) { return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z)); }
  static V3<R> crossF(V3<R> a, V3<R> b) {
    return V3<R>(std::fma(a.y, b.z, -(a.z * b.y)), std::fma(a.z, b.x, -(a.x * b.z)), std::fma(a.x, b.y, -(a.y * b.x)));
  }
  /** q v q* = v + 2 (w u + q⃗ × u), u = q⃗ × v (the quaternion rotation of layout.ts). */
  static V3<R> rotF(Q4<R> q, V3<R> v) {
    const V3<R> qv = q.xyz(), u = crossF(qv, v), t = crossF(qv, u);
    return V3<R>(std::fma(R(2), std::fma(q.w, u.x, t.x), v.x), std::fma(R(2), std::fma(q.w, u.y, t.y), v.y),
                 std::fma(R(2), std::fma(q.w, u.z, t.z), v.z));
  }
  /** qsub(a, b) = 2 (a b⁻¹)⃗ from bi = b⁻¹ (qinverse), each component of the product one chain. */
  static V3<R> dispF(Q4<R> a, Q4<R> bi) {
    using std::fma;
    return V3<R>(R(2) * fma(a.w, bi.x, fma(a.x, bi.w, fma(a.y, bi.z, -(a.z * bi.y)))),
                 R(2) * fma(a.w, bi.y, fma(a.y, bi.w, fma(a.z, bi.x, -(a.x * bi.z)))),
                 R(2) * fma(a.w, bi.z, fma(a.z, bi.w, fma(a.x, bi.y, -(a.y * bi.x)))));
  }

  struct PairBody {
    uint32_t index;
    Q4<R> rot;
    V3<R> dLin, dAng;
  };
  PairBody pairBody(uint32_t i) const {
    const Body<R> &b = bodies[i];
    return {i, b.rot, sub(b.pos, b.ipos), dispF(b.rot, qinverse(b.irot))};
  }
  struct Eval {
    V3<R> rAW, rBW, C, F;
    R frictionScale, bounds;
  };
  Eval evalContact(const Contact<R> &k, const M3<R> &basis, R friction, const PairBody &A, const PairBody &B,
                   R alpha) const {
    Q4<R> rotA = A.rot, rotB = B.rot;
    if (stepStartArms(k.key, P(A.index).flags)) {
      rotA = bodies[A.index].irot;
      rotB = bodies[B.index].irot;
    }
    Eval e;
    e.rAW = rotF(rotA, k.rA);
    e.rBW = rotF(rotB, k.rB);
    V3<R> n = row(basis, 0), t1 = row(basis, 1), t2 = row(basis, 2);
    R keep = R(1) - alpha;
    // The anchors' relative displacement once, D = (Δx_A − Δx_B) + (Δθ_A × r_A − Δθ_B × r_B), each row
    // C0 (1 − α) + l·D (the reference's per-row sum l·Δx_A − l·Δx_B + (r_A × l)·Δθ_A − (r_B × l)·Δθ_B, rounded as the
    // HIP step does; exchanging A and B negates D exactly, so the HIP step's rows from B's side have these bits)
    const V3<R> D = add(sub(A.dLin, B.dLin), sub(crossF(A.dAng, e.rAW), crossF(B.dAng, e.rBW)));
    e.C = V3<R>(std::fma(k.c0x, keep, dotF(n, D)), std::fma(k.c0y, keep, dotF(t1, D)), std::fma(k.c0z, keep, dotF(t2, D)));
    V3<R> F(std::fma(k.pen.x, e.C.x, k.lam.x), std::fma(k.pen.y, e.C.y, k.lam.y), std::fma(k.pen.z, e.C.z, k.lam.z));
    F.x = rmin(F.x, R(0));
    e.bounds = std::fabs(F.x) * friction;
    e.frictionScale = std::sqrt(std::fma(F.y, F.y, F.z * F.z));
    if (e.frictionScale > e.bounds && e.frictionScale > 0) {
      R s = e.bounds / e.frictionScale;
      F.y = F.y * s;
      F.z = F.z * s;
    }
    e.F = F;
    return e;
  }

  V3<R> jointLinRows(const Joint<R> &k, R alpha) const {
    V3<R> c = jointLinC(k);
    return sub(c, hard(k.stiffLin) ? scale(k.c0Lin, alpha) : V3<R>(0, 0, 0));
  }
  V3<R> jointAngRows(const Joint<R> &k, R alpha) const {
    V3<R> c = jointAngC(k);
    return sub(c, hard(k.stiffAng) ? scale(k.c0Ang, alpha) : V3<R>(0, 0, 0));
  }

  void addJoint(Acc &acc, const Joint<R> &k, R alpha, uint32_t i) const {
    bool isA = int64_t(i) == k.a;
    R sg = isA ? R(1) : R(-1);
    if (k.type == int32_t(T_MOTOR) || k.type == int32_t(T_LIMIT)) {  // addAngular (angularConstraints.js:36-44)
      if (k.stiffLin == 0 || (k.type == int32_t(T_MOTOR) && k.rBW == 0)) return;
      const HingeFrame hf = hingeFrame(k);
      const V3<R> axis = scale(hf.axis, isA ? R(-1) : R(1));
      const std::pair<R, R> C = hingeRows(k, hingeAngle(hf), alpha, R(P(uint32_t(k.b)).dt));
      const std::pair<R, R> F = hingeForce(k, C);
      if (k.type == int32_t(T_MOTOR) || F.first != 0) addRow(acc, V3<R>(0, 0, 0), axis, k.penLin.x, F.first);
      if (k.type == int32_t(T_LIMIT) && F.second != 0) addRow(acc, V3<R>(0, 0, 0), axis, k.penLin.y, F.second);
      return;
    }
    if (k.type == int32_t(T_SPRING)) {
      const Body<R> &A = bodies[size_t(k.a)], &B = bodies[size_t(k.b)];
      V3<R> rAW = rotate(A.rot, k.rA), rBW = rotate(B.rot, k.rB);
      V3<R> d = sub(add(rAW, A.pos), add(rBW, B.pos));
      R len = length(d);
      if (len <= R(1.0e-6f)) return;
      V3<R> n = divs(d, len);
      R stiffness = k.stiffLin;
      V3<R> r = isA ? rAW : rBW;
      addRow(acc, scale(n, sg), scale(cross(r, n), sg), stiffness, stiffness * (len - k.rest));
      return;
    }
    if (dot(k.penLin, k.penLin) > 0) {
      V3<R> F = add(mulv3(k.penLin, jointLinRows(k, alpha)), k.lamLin);
      V3<R> r = isA ? rotate(bodies[size_t(k.a)].rot, k.rA) : neg(rotate(bodies[size_t(k.b)].rot, k.rB));
      V3<R> v = neg(r);
      addRow(acc, V3<R>(sg, 0, 0), V3<R>(0, -v.z, v.y), k.penLin.x, F.x);
      addRow(acc, V3<R>(0, sg, 0), V3<R>(v.z, 0, -v.x), k.penLin.y, F.y);
      addRow(acc, V3<R>(0, 0, sg), V3<R>(-v.y, v.x, 0), k.penLin.z, F.z);
      R rf = dot(r, F);
      V3<R> g(length(sub(scale(r, F.x), V3<R>(rf, 0, 0))), length(sub(scale(r, F.y), V3<R>(0, rf, 0))),
              length(sub(scale(r, F.z), V3<R>(0, 0, rf))));
      acc.ang.m[0] += g.x;
      acc.ang.m[4] += g.y;
      acc.ang.m[8] += g.z;
    }
    if (dot(k.penAng, k.penAng) > 0) {
      V3<R> F = add(mulv3(k.penAng, jointAngRows(k, alpha)), k.lamAng);
      R s = sg * k.torqueArm;
      addRow(acc, V3<R>(0, 0, 0), V3<R>(s, 0, 0), k.penAng.x, F.x);
      addRow(acc, V3<R>(0, 0, 0), V3<R>(0, s, 0), k.penAng.y, F.y);
      addRow(acc, V3<R>(0, 0, 0), V3<R>(0, 0, s), k.penAng.z, F.z);
    }
  }

  /** Lanes per body: W = LANES for a body with at most LANES adjacency entries, else HEAVY_LANES. Entries are taken
   * in chunks of W; a chunk's work items (a joint, or one contact point) are numbered in entry order and item t goes to
   * lane t mod W (lane 0 also holds the inertia terms); the partial systems are then summed by the fixed tree
   * k += k + W/2, ..., k += k + 1 (the HIP primal's order). */
  static constexpr uint32_t LANES = 8, HEAVY_LANES = 32;

  /** Work items of adjacency entry `id`: a joint is one item, a manifold one item per contact point. */
  uint32_t itemCount(uint32_t id) const {
    const uint32_t J = uint32_t(joints.size());
    return id < J ? 1u : manifolds[id - J].count();
  }

  /** Item `p` of entry `id` of body i: the joint's rows, or the three rows of contact point p of the manifold. */
  void addItem(Acc &acc, uint32_t i, uint32_t id, uint32_t p, R alpha) const {
    const uint32_t J = uint32_t(joints.size());
    if (id < J) {
      addJoint(acc, joints[id], alpha, i);
      return;
    }
    const Manifold<R> &mf = manifolds[id - J];
    bool isA = i == mf.a;
    PairBody A = pairBody(mf.a), B = pairBody(mf.b);
    M3<R> basis = contactBasis(mf.normal);
    const Contact<R> &k = contacts[mf.first + p];
    Eval ev = evalContact(k, basis, mf.friction, A, B, alpha);
    addContactRows(acc, basis, isA ? R(1) : R(-1), isA ? ev.rAW : ev.rBW, k.pen, ev.F);
  }

  /**
   * The three rows of one contact point at once (the HIP addContactRows): M = Σ k_q b_q b_qᵀ, g = sg Σ f_q b_q; the
   * linear block takes M, the cross block [r]× M, the angular block [r]× M [r]×ᵀ (lower triangle; the upper one is
   * not read), the right-hand side (g, r × g), each entry with the same fused multiply-adds.
   */
  static void addContactRows(Acc &acc, const M3<R> &basis, R sg, V3<R> r, V3<R> k, V3<R> f) {
    const V3<R> n = row(basis, 0), t1 = row(basis, 1), t2 = row(basis, 2);
    const V3<R> kn = scale(n, k.x), k1 = scale(t1, k.y), k2 = scale(t2, k.z);
    using std::fma;
    const R m00 = fma(k2.x, t2.x, fma(k1.x, t1.x, kn.x * n.x));
    const R m10 = fma(k2.y, t2.x, fma(k1.y, t1.x, kn.y * n.x));
    const R m11 = fma(k2.y, t2.y, fma(k1.y, t1.y, kn.y * n.y));
    const R m20 = fma(k2.z, t2.x, fma(k1.z, t1.x, kn.z * n.x));
    const R m21 = fma(k2.z, t2.y, fma(k1.z, t1.y, kn.z * n.y));
    const R m22 = fma(k2.z, t2.z, fma(k1.z, t1.z, kn.z * n.z));
    const R c00 = fma(r.y, m20, -(r.z * m10)), c01 = fma(r.y, m21, -(r.z * m11)), c02 = fma(r.y, m22, -(r.z * m21));
    const R c10 = fma(r.z, m00, -(r.x * m20)), c11 = fma(r.z, m10, -(r.x * m21)), c12 = fma(r.z, m20, -(r.x * m22));
    const R c20 = fma(r.x, m10, -(r.y * m00)), c21 = fma(r.x, m11, -(r.y * m10)), c22 = fma(r.x, m21, -(r.y * m20));
    R *a = acc.ang.m, *l = acc.lin.m, *c = acc.cross.m;
    a[0] = fma(r.y, c02, fma(-r.z, c01, a[0]));
    a[3] = fma(r.y, c12, fma(-r.z, c11, a[3]));
    a[4] = fma(r.z, c10, fma(-r.x, c12, a[4]));
    a[6] = fma(r.y, c22, fma(-r.z, c21, a[6]));
    a[7] = fma(r.z, c20, fma(-r.x, c22, a[7]));
    a[8] = fma(r.x, c21, fma(-r.y, c20, a[8]));
    l[0] = l[0] + m00;
    l[3] = l[3] + m10;
    l[4] = l[4] + m11;
    l[6] = l[6] + m20;
    l[7] = l[7] + m21;
    l[8] = l[8] + m22;
    const R cs[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
    for (int q = 0; q < 9; q++) c[q] = c[q] + cs[q];
    const V3<R> g = scale(V3<R>(fma(f.z, t2.x, fma(f.y, t1.x, f.x * n.x)), fma(f.z, t2.y, fma(f.y, t1.y, f.x * n.y)),
                                fma(f.z, t2.z, fma(f.y, t1.z, f.x * n.z))),
                          sg);
    acc.rLin = add(acc.rLin, g);
    acc.rAng = V3<R>(fma(r.y, g.z, fma(-r.z, g.y, acc.rAng.x)), fma(r.z, g.x, fma(-r.x, g.z, acc.rAng.y)),
                     fma(r.x, g.y, fma(-r.y, g.x, acc.rAng.z)));
  }

  static void mergeInto(Acc &a, const Acc &o) {
    for (int q = 0; q < 9; q++) {
      a.lin.m[q] = a.lin.m[q] + o.lin.m[q];
      a.ang.m[q] = a.ang.m[q] + o.ang.m[q];
      a.cross.m[q] = a.cross.m[q] + o.cross.m[q];
    }
    a.rLin = add(a.rLin, o.rLin);
    a.rAng = add(a.rAng, o.rAng);
  }

  Acc accumulate(uint32_t i, R alpha) const {
    const Body<R> &b = bodies[i];
    const WorldParams &p = P(i);
    R dt = R(p.dt);
    R dt2 = dt * dt;
    R m = b.mass / dt2;
    V3<R> I = divs(b.moment, dt2);
    const uint32_t W = adjStart[i + 1] - adjStart[i] > LANES ? HEAVY_LANES : LANES;
    Acc lanes[HEAVY_LANES];
    lanes[0].lin = diag(m, m, m);
    lanes[0].ang = diag(I.x, I.y, I.z);
    lanes[0].rLin = scale(sub(b.pos, b.inertialPos), m);
    lanes[0].rAng = mulv3(I, qsub(b.rot, b.inertialRot));
    // Entries in chunks of W; the chunk's work items (joints, contact points, in entry order) go to lanes t mod W
    for (uint32_t c0 = adjStart[i]; c0 < adjStart[i + 1]; c0 += W) {
      uint32_t t = 0;
      for (uint32_t e = c0; e < std::min(c0 + W, adjStart[i + 1]); e++)
        for (uint32_t p = 0; p < itemCount(adjList[e]); p++, t++) addItem(lanes[t % W], i, adjList[e], p, alpha);
    }
    for (uint32_t s = W / 2; s > 0; s >>= 1)
      for (uint32_t lane = 0; lane < s; lane++) mergeInto(lanes[lane], lanes[lane + s]);
    return lanes[0];
  }

  /**
   * The HIP primal's LDLᵀ solve of the 6x6 system [lin crossᵀ; cross ang] x = b (solve.hip ldlt6): pivots after the
   * first floored at pivotFloor × their diagonal (wgsl-solve.ts:416-470), each pivot inverted once, and every sum of
   * products (D_k = A_kk − Σ_j L_kj² D_j, L_ik = (A_ik − Σ_j L_ij L_kj D_j) / D_k, both substitutions) one chain of
   * fused multiply-adds in increasing j.
   */
  static void ldltFused(const M3<R> &aLin, const M3<R> &aAng, const M3<R> &aCross, V3<R> bLin, V3<R> bAng, V3<R> &xLin,
                        V3<R> &xAng, R pivotFloor, bool *floored) {
    // Lower triangle row by row: rows 1-3 from lin, rows 4-6 from cross (columns 1-3) and ang (columns 4-6)
    const R A[21] = {aLin.m[0], aLin.m[3], aLin.m[4], aLin.m[6], aLin.m[7], aLin.m[8],
                     aCross.m[0], aCross.m[1], aCross.m[2], aAng.m[0],
                     aCross.m[3], aCross.m[4], aCross.m[5], aAng.m[3], aAng.m[4],
                     aCross.m[6], aCross.m[7], aCross.m[8], aAng.m[6], aAng.m[7], aAng.m[8]};
    const R b[6] = {bLin.x, bLin.y, bLin.z, bAng.x, bAng.y, bAng.z};
    R L[6][6] = {}, Dg[6], rc[6], y[6], x[6];
    for (int k = 0; k < 6; k++) {
      R d = A[k * (k + 1) / 2 + k];
      for (int j = 0; j < k; j++) d = std::fma(-(L[k][j] * L[k][j]), Dg[j], d);
      if (k > 0) {
        const R f = A[k * (k + 1) / 2 + k] * pivotFloor;
        if (floored && !(d > f)) *floored = true;
        d = d > f ? d : f;
      }
      Dg[k] = d;
      rc[k] = R(1) / d;
      for (int i = k + 1; i < 6; i++) {
        R a = A[i * (i + 1) / 2 + k];
        for (int j = 0; j < k; j++) a = std::fma(-(L[i][j] * L[k][j]), Dg[j], a);
        L[i][k] = a * rc[k];
      }
    }
    for (int k = 0; k < 6; k++) {
      y[k] = b[k];
      for (int j = 0; j < k; j++) y[k] = std::fma(-L[k][j], y[j], y[k]);
    }
    for (int k = 5; k >= 0; k--) {
      x[k] = y[k] * rc[k];
      for (int j = k + 1; j < 6; j++) x[k] = std::fma(-L[j][k], x[j], x[k]);
    }
    xLin = V3<R>(x[0], x[1], x[2]);
    xAng = V3<R>(x[3], x[4], x[5]);
  }

  /**
   * finishBody (wgsl-solve.ts:416-470) with the FP64 pivot fallback: when an FP32 pivot takes its floor the system is
   * solved again in FP64 from the same FP32 entries (same floor), and the result rounded to FP32.
   */
  void finishBody(uint32_t i, const Acc &acc) {
    Body<R> &b = bodies[i];
    V3<R> xl, xa;
    bool floored = false;
    ldltFused(acc.lin, acc.ang, acc.cross, acc.rLin, acc.rAng, xl, xa, R(1e-5f), &floored);
    if (floored && sizeof(R) == sizeof(float)) {
      auto up = [](const M3<R> &m) {
        M3<double> d;
        for (int q = 0; q < 9; q++) d.m[q] = double(m.m[q]);
        return d;
      };
      auto upv = [](V3<R> v) { return V3<double>(double(v.x), double(v.y), double(v.z)); };
      V3<double> dl, da;
      World<double>::ldltFused(up(acc.lin), up(acc.ang), up(acc.cross), upv(acc.rLin), upv(acc.rAng), dl, da,
                               double(1e-5f), nullptr);
      xl = V3<R>(R(dl.x), R(dl.y), R(dl.z));
      xa = V3<R>(R(da.x), R(da.y), R(da.z));
#pragma omp atomic
      pivotFallbacks++;
    }
    b.pos = sub(b.pos, xl);
    b.rot = qadd(b.rot, neg(xa));
  }

  void dualJoint(Joint<R> &k, R alpha) {
    if (k.type == int32_t(T_SPRING)) {
      springFracture(k);
      return;
    }
    if (k.type == int32_t(T_MOTOR) || k.type == int32_t(T_LIMIT)) {
      dualHinge(k, alpha);
      return;
    }
    if (k.type != int32_t(T_JOINT)) return;
    const WorldParams &p = P(uint32_t(k.b));
    if (dot(k.penLin, k.penLin) > 0) {
      V3<R> C = jointLinRows(k, alpha);
      if (hard(k.stiffLin)) k.lamLin = add(mulv3(k.penLin, C), k.lamLin);
      R cap = rmin(k.stiffLin, R(1e10f));
      V3<R> q = add(k.penLin, scale(vabs(C), R(p.betaLin)));
      k.penLin = V3<R>(rmin(q.x, cap), rmin(q.y, cap), rmin(q.z, cap));
    }
    if (dot(k.penAng, k.penAng) > 0) {
      V3<R> C = jointAngRows(k, alpha);
      if (hard(k.stiffAng)) k.lamAng = add(mulv3(k.penAng, C), k.lamAng);
      R cap = rmin(k.stiffAng, R(1e10f));
      V3<R> q = add(k.penAng, scale(vabs(C), R(p.betaAng)));
      k.penAng = V3<R>(rmin(q.x, cap), rmin(q.y, cap), rmin(q.z, cap));
    }
    R frac = k.frac;
    R limit = std::fabs(frac);
    bool linear = frac < 0 && dot(k.lamLin, k.lamLin) > limit * limit;
    if (limit < R(3e38f) && (dot(k.lamAng, k.lamAng) > limit * limit || linear)) {
      k.penLin = V3<R>();
      k.stiffLin = 0;
      k.penAng = V3<R>();
      k.stiffAng = 0;
      k.lamLin = V3<R>();
      k.lamAng = V3<R>();
    }
  }

  void dualManifold(uint32_t m, R alpha) {
    const Manifold<R> &mf = manifolds[m];
    const WorldParams &p = P(mf.a);
    PairBody A = pairBody(mf.a), B = pairBody(mf.b);
    M3<R> basis = contactBasis(mf.normal);
    for (uint32_t c = mf.first; c < mf.first + mf.count(); c++) {
      Contact<R> &k = contacts[c];
      Eval e = evalContact(k, basis, mf.friction, A, B, alpha);
      k.lam = e.F;
      V3<R> pen = k.pen;
      if (e.F.x < 0) pen.x = rmin(pen.x + R(p.betaLin) * std::fabs(e.C.x), R(1e10f));
      if (e.frictionScale <= e.bounds) {
        pen.y = rmin(pen.y + R(p.betaLin) * std::fabs(e.C.y), R(1e10f));
        pen.z = rmin(pen.z + R(p.betaLin) * std::fabs(e.C.z), R(1e10f));
        bool stick = std::sqrt(e.C.y * e.C.y + e.C.z * e.C.z) < R(0.00001f);
        k.key = (k.key & ~STICK_BIT) | (stick ? STICK_BIT : 0u);
      }
      k.pen = pen;
    }
  }

  void updateVelocities() {
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n(); i++) {
      Body<R> &b = bodies[i];
      const WorldParams &p = P(i);
      V3<R> up(R(p.up[0]), R(p.up[1]), R(p.up[2]));
      R dt = R(p.dt);
      R prevUp = dot(b.vel, up);
      if (b.mass > 0) {
        b.vel = divs(sub(b.pos, b.ipos), dt);
        b.prevUp = prevUp;
        b.angVel = divs(qsub(b.rot, b.irot), dt);
      } else {
        b.prevUp = prevUp;
      }
    }
  }

  // --- the step ----------------------------------------------------------------------------------------------------

  /** Collision, adjacency and colouring (everything before the warm starts). */
  void collisionAndTopology() {
    for (auto &p : params) p.step++;
    beginFrame();
    updateRefs();
    findPairs();
    narrowphase();
    // sleepers woken by an impact: their pairs with other sleepers were left out above, so collide again
    if (sleepOn && sleepImpact()) {
      manifolds.clear();
      contacts.clear();
      findPairs();
      narrowphase();
    }
    overflow = (pairCapacity != 0 && pairs.size() > pairCapacity) ? 1u : 0u;
    buildAdjacency();
    colour();
  }

  /**
   * Warm starts, iterations × (primal per colour, dual), velocities. A step whose colouring clashed (bodies still
   * pending after the rounds) or whose pairs overflowed the pair capacity is held: bodies and joints stay as they are
   * (protectSolverStep's solverStepValid, src/gpu/gpuSolverKernels.js:72-103).
   */
  void solve() {
    if (clashes != 0 || overflow != 0) return;
    warmStartJoints();
    warmStartBodies();
    for (int it = 0; it < iterations; it++) {
      for (uint32_t c = 0; c < colorCap; c++)
#pragma omp parallel for schedule(dynamic, 64)
        for (uint32_t s = colorStart[c]; s < colorStart[c + 1]; s++) {  // a colour's bodies are independent
          uint32_t i = colorBodies[s];
          R alpha = R(P(i).alpha);
          Acc acc = accumulate(i, alpha);
          finishBody(i, acc);
        }
#pragma omp parallel for schedule(static)
      for (size_t j = 0; j < joints.size(); j++)
        if (joints[j].type != int32_t(T_NONE)) dualJoint(joints[j], R(P(uint32_t(joints[j].b)).alpha));
#pragma omp parallel for schedule(dynamic, 256)
      for (uint32_t m = 0; m < manifolds.size(); m++) dualManifold(m, R(P(manifolds[m].a).alpha));
    }
    updateVelocities();
  }

  // --- restitution around the step (src/gpu/restitution.js:16-88) -----------------------------------------------

  std::vector<R> restitution;  // per body, in [0, 1]
  struct BounceVel {
    V3<R> inLin, inAng, lin, ang;
    R linW = 0, angW = 0;
  };
  std::vector<BounceVel> bounceVel;
  struct Bounce {
    R impulse = 0, delta = 0, rebound = 0, mass = 0;
    V3<R> ra, rb;
  };

  bool bouncing() const {
    for (R r : restitution)
      if (r > 0) return true;
    return false;
  }
  static R inverseMassOf(const Body<R> &b) { return b.mass > 0 ? R(1) / rmax(b.mass, R(1e-20f)) : R(0); }
  /** restitution.js:29-34: the world-frame inverse inertia applied to t. */
  static V3<R> inverseInertiaOf(const Body<R> &b, V3<R> t) {
    if (b.mass <= 0) return V3<R>(0, 0, 0);
    V3<R> local = rotate(conj(b.rot), t);
    V3<R> m = b.moment;
    V3<R> scaled(m.x > 0 ? local.x / rmax(m.x, R(1e-20f)) : R(0), m.y > 0 ? local.y / rmax(m.y, R(1e-20f)) : R(0),
                 m.z > 0 ? local.z / rmax(m.z, R(1e-20f)) : R(0));
    return rotate(b.rot, scaled);
  }
  V3<R> pointVelocity(uint32_t i, V3<R> r, bool incoming) const {
    const BounceVel &v = bounceVel[i];
    return add(incoming ? v.inLin : v.lin, cross(incoming ? v.inAng : v.ang, r));
  }
  /** captureIncoming: every body's velocities before the step. */
  void restitutionBefore() {
    bounceVel.resize(n());
    for (uint32_t i = 0; i < n(); i++) {
      bounceVel[i].inLin = bodies[i].vel;
      bounceVel[i].inAng = bodies[i].angVel;
    }
  }
  /** captureSolved, initialize, 12 × (pairs, apply), writeBack. */
  void restitutionAfter() {
    const uint32_t N = uint32_t(n()), M = uint32_t(manifolds.size()), J = uint32_t(joints.size());
    for (uint32_t i = 0; i < N; i++) {
      bounceVel[i].lin = bodies[i].vel;
      bounceVel[i].linW = bodies[i].prevUp;
      bounceVel[i].ang = bodies[i].angVel;
      bounceVel[i].angW = bodies[i].shape;
    }
    std::vector<Bounce> bounce(M);
#pragma omp parallel for schedule(static)
    for (uint32_t m = 0; m < M; m++) {
      const Manifold<R> &mf = manifolds[m];
      const uint32_t count = mf.count(), a = mf.a, b = mf.b;
      const R e = rmax(restitution[a], restitution[b]);
      if (count == 0 || e <= 0 || (mf.countw & 0x80000000u) != 0 || clashes != 0 || overflow != 0) continue;
      const Body<R> &A = bodies[a], &B = bodies[b];
      V3<R> ra(0, 0, 0), rb(0, 0, 0);
      for (uint32_t c = 0; c < count; c++) {
        ra = add(ra, rotate(A.rot, contacts[mf.first + c].rA));
        rb = add(rb, rotate(B.rot, contacts[mf.first + c].rB));
      }
      const R cn = R(count);
      ra = V3<R>(ra.x / cn, ra.y / cn, ra.z / cn);
      rb = V3<R>(rb.x / cn, rb.y / cn, rb.z / cn);
      const R vn = dot(mf.normal, sub(pointVelocity(a, ra, true), pointVelocity(b, rb, true)));
      if (vn >= R(-1)) continue;
      const V3<R> ta = cross(ra, mf.normal), tb = cross(rb, mf.normal);
      const R k = inverseMassOf(A) + inverseMassOf(B) + dot(ta, inverseInertiaOf(A, ta)) + dot(tb, inverseInertiaOf(B, tb));
      if (k <= 0) continue;
      Bounce &s = bounce[m];
      s.rebound = -e * vn;
      s.mass = k;
      s.ra = ra;
      s.rb = rb;
    }
    for (int round = 0; round < 12; round++) {
#pragma omp parallel for schedule(static)
      for (uint32_t m = 0; m < M; m++) {
        Bounce &s = bounce[m];
        if (s.mass <= 0) continue;
        const Manifold<R> &mf = manifolds[m];
        const R vn = dot(mf.normal, sub(pointVelocity(mf.a, s.ra, false), pointVelocity(mf.b, s.rb, false)));
        const uint32_t degree = rmax(1u, rmax(adjStart[mf.a + 1] - adjStart[mf.a], adjStart[mf.b + 1] - adjStart[mf.b]));
        const R impulse = rmax(R(0), s.impulse + (s.rebound - vn) / (s.mass * R(degree)));
        s.delta = impulse - s.impulse;
        s.impulse = impulse;
      }
#pragma omp parallel for schedule(static)
      for (uint32_t i = 0; i < N; i++) {
        const Body<R> &B = bodies[i];
        if (B.mass <= 0) continue;
        V3<R> dv(0, 0, 0), dw(0, 0, 0);
        for (uint32_t e = adjStart[i]; e < adjStart[i + 1]; e++) {
          const uint32_t id = adjList[e];
          if (id < J) continue;
          const Manifold<R> &mf = manifolds[id - J];
          const Bounce &s = bounce[id - J];
          if (s.mass <= 0) continue;
          const bool isA = i == mf.a;
          const V3<R> impulse = scale(scale(mf.normal, s.delta), isA ? R(1) : R(-1));
          dv = add(dv, scale(impulse, inverseMassOf(B)));
          dw = add(dw, inverseInertiaOf(B, cross(isA ? s.ra : s.rb, impulse)));
        }
        BounceVel &v = bounceVel[i];
        v.lin = add(v.lin, dv);
        v.linW = v.linW + R(0);
        v.ang = add(v.ang, dw);
        v.angW = v.angW + R(0);
        if (round == 11) {
          bodies[i].vel = v.lin;
          bodies[i].prevUp = v.linW;
          bodies[i].angVel = v.ang;
          bodies[i].shape = v.angW;
        }
      }
    }
  }

  // --- sleeping (src/gpu/worldSleeping.js, worldSleeping.wgsl, islandSleeping.wgsl, sleepContact.wgsl) --------------
  // A sleeping body is a fixed support: its mass (size.w) is 0 and its velocities are 0; its mass while awake is kept in
  // its sleep state and restored when it wakes. Whole-world mode: a body sleeps once it has been quiet (linear speed +
  // angular speed × bounding radius at most `speed`) and supported (gravity 0, or resting on a static, sleeping or
  // quiet partner below it) for `delay` seconds; any impact on a sleeper wakes the world. Island mode: bodies joined by
  // contacts and joints (links among sleepers are kept while they sleep) form islands in a union-find forest whose
  // roots are the islands' smallest indices; an island sleeps only when no member is busy and wakes as a whole.

  struct Sleep {
    R mass = 0, quiet = 0;
    uint32_t asleep = 0, support = 0, eligible = 1, parent = 0, wake = 0, flags = 0;
  };
  enum : uint32_t { SF_READY = 1, SF_HELD = 2, SF_START_QUIET = 4, SF_GATHERED = 8 };  // Sleep::flags
  enum : uint32_t { SW_IMPACT = 1, SW_GROUNDED = 2, SW_BUSY = 4, SW_REQUEST = 8 };     // Sleep::wake (at a root)
  std::vector<Sleep> sleep;
  std::vector<uint32_t> sleepGlobals;  // per world: wake all, sleeping, woken by an impact, local wake signal
  bool sleepOn = false, sleepIslands = false;
  uint32_t sleepForce = 0;  // the next step's force bits: 1 wake all, 2 start asleep (island mode)
  R sleepSpeed = R(0.03f), sleepDelay = R(0.5f);
  static constexpr uint32_t SLEEP_NONE = 0xffffffffu;

  /** worldSleeping.js:17-103, register(): every body's state from its current mass and eligibility. */
  void sleepEnable(float speed, float delay, bool islands, bool startAsleep, const uint8_t *eligible) {
    sleepOn = true;
    sleepIslands = islands;
    sleepSpeed = R(speed);
    sleepDelay = R(delay);
    sleep.assign(n(), Sleep{});
    for (uint32_t i = 0; i < n(); i++) {
      sleep[i].mass = bodies[i].mass;
      sleep[i].eligible = eligible ? eligible[i] : 1u;
      sleep[i].parent = i;
    }
    sleepGlobals.assign(4 * params.size(), 0u);
    sleepForce = startAsleep && islands ? 2u : 0u;
  }
  /** disableSleeping(): wake every body (restoring masses), then drop the state. */
  void sleepDisable() {
    if (!sleepOn) return;
    sleepWakeAll();
    sleepOn = false;
    sleep.clear();
  }
  R motion(uint32_t i) const { return length(bodies[i].vel) + length(bodies[i].angVel) * bodies[i].radius; }
  uint32_t *globalsOf(uint32_t i) { return &sleepGlobals[4 * world[i]]; }
  void wakeOne(uint32_t i) {
    Sleep &s = sleep[i];
    if (s.mass > 0) {
      bodies[i].mass = s.mass;
      s.asleep = 0;
      s.quiet = 0;
      s.flags = 0;
    }
  }
  /** islandSleeping.wgsl:104-118 root() (its path halving changes no root, so it is left out here). */
  uint32_t sleepRoot(uint32_t i) const {
    while (sleep[i].parent != i) i = sleep[i].parent;
    return i;
  }
  /** islandSleeping.wgsl:119-128: link the higher root below the lower one. */
  void sleepConnect(uint32_t a, uint32_t b) {
    if (a >= n() || b >= n() || sleep[a].mass <= 0 || sleep[b].mass <= 0) return;
    const uint32_t ra = sleepRoot(a), rb = sleepRoot(b);
    if (ra != rb) sleep[std::max(ra, rb)].parent = std::min(ra, rb);
  }
  /** before(): both shaders' before(), then clear(). */
  void sleepBefore(uint32_t force) {
    for (uint32_t i = 0; i < n(); i++) {
      Sleep &s = sleep[i];
      const bool all = (force & 1u) || globalsOf(i)[0] != 0;
      if (sleepIslands) {
        if (all) {
          wakeOne(i);
          s.parent = i;
        } else if (s.flags & (SF_HELD | SF_GATHERED)) {
          const uint32_t held = s.flags & SF_HELD;
          wakeOne(i);
          s.flags |= held;
        }
        if (!s.asleep) s.parent = i;
        if ((force & 2u) && motion(i) <= sleepSpeed && s.eligible) s.flags |= SF_START_QUIET;
        s.wake = 0;
      } else if (all || (s.flags & SF_HELD)) {
        wakeOne(i);
      }
      s.support = 0;
    }
    std::fill(sleepGlobals.begin(), sleepGlobals.end(), 0u);
  }
  void sleepWakeAll() { sleepBefore(1u); }
  /** wakeBody() + requestWake/gatherWake (islandSleeping.wgsl:23-30): hold bodies awake; in island mode their
   * islands wake at the next step's start. */
  void sleepHold(const uint32_t *indices, uint32_t count) {
    for (uint32_t k = 0; k < count; k++) sleep[indices[k]].flags = SF_HELD;
    if (!sleepIslands) return;
    for (uint32_t i = 0; i < n(); i++)
      if (sleep[i].asleep && (sleep[i].flags & SF_HELD)) sleep[sleepRoot(i)].wake |= SW_REQUEST;
    for (uint32_t i = 0; i < n(); i++)
      if (sleep[sleepRoot(i)].wake & SW_REQUEST) sleep[i].flags |= SF_GATHERED;
  }
  void signalWake(uint32_t i) {
    uint32_t *g = globalsOf(i);
    if (sleepIslands) {
      g[3] = 1;
      sleep[sleepRoot(i)].wake |= SW_IMPACT;
    } else {
      g[0] = 1;
    }
  }
  void sleepImpactBodies(uint32_t a, uint32_t b) {
    if (sleep[a].asleep && !sleep[b].asleep && motion(b) > R(2) * sleepSpeed) signalWake(a);
  }
  /** sleepContact.wgsl:14-30 contactVelocity(): a point's velocity now, or over the step just solved (from the
   * step-start and inertial poses, less this step's gravity for a dynamic body). */
  V3<R> sleepPointVelocity(uint32_t i, V3<R> r, bool previous) const {
    if (sleep[i].asleep) return V3<R>(0, 0, 0);
    const Body<R> &B = bodies[i];
    V3<R> v = B.vel, w = B.angVel;
    Q4<R> q = B.rot;
    if (previous) {
      const WorldParams &p = P(i);
      const R dt = R(p.dt);
      v = V3<R>((B.inertialPos.x - B.ipos.x) / dt, (B.inertialPos.y - B.ipos.y) / dt, (B.inertialPos.z - B.ipos.z) / dt);
      if (B.mass > 0) v = sub(v, scale(V3<R>(R(p.up[0]), R(p.up[1]), R(p.up[2])), R(p.gravity) * dt));
      const Q4<R> a = B.inertialRot, b = B.irot;
      const V3<R> av(a.x, a.y, a.z), bv(b.x, b.y, b.z);
      const V3<R> t = scale(sub(add(scale(bv, -a.w), scale(av, b.w)), cross(av, bv)), R(2));
      const R d = (((b.x * b.x + b.y * b.y) + b.z * b.z) + b.w * b.w) * dt;
      w = V3<R>(t.x / d, t.y / d, t.z / d);
      q = b;
    }
    return add(v, cross(w, rotate(q, r)));
  }
  void sleepContactImpact(uint32_t a, uint32_t b, V3<R> nrm, V3<R> rA, V3<R> rB, bool beforeSolve) {
    if (!sleep[a].asleep || sleep[b].asleep) return;
    R incoming = 0;
    if (!beforeSolve) incoming = dot(sub(sleepPointVelocity(b, rB, true), sleepPointVelocity(a, rA, true)), nrm);
    const R outgoing = dot(sub(sleepPointVelocity(b, rB, false), sleepPointVelocity(a, rA, false)), nrm);
    if (rmax(incoming, outgoing) > R(2) * sleepSpeed) signalWake(a);
  }
  /** sleepContact.wgsl:31-40 manifoldImpact(): each point, both ways, for a manifold between a sleeper and a waker. */
  void sleepManifoldImpact(const Manifold<R> &m, bool beforeSolve) {
    const uint32_t a = m.a, b = m.b;
    if (sleep[a].asleep == sleep[b].asleep) return;
    for (uint32_t k = 0; k < m.count(); k++) {
      const Contact<R> &c = contacts[m.first + k];
      sleepContactImpact(a, b, m.normal, c.rA, c.rB, beforeSolve);
      sleepContactImpact(b, a, neg(m.normal), c.rB, c.rA, beforeSolve);
    }
  }
  bool sleepLiveManifold(const Manifold<R> &m) const { return m.count() != 0 && (m.countw & 0x80000000u) == 0; }
  bool sleepLiveJoint(const Joint<R> &j) const { return j.type != 0 && j.a >= 0; }
  /** encodeImpact(): impacts on sleepers from this step's contacts and joints; the islands (or world) they hit wake.
   * Returns whether a body woke (the pairs are then found again). */
  bool sleepImpact() {
    for (const Manifold<R> &m : manifolds)
      if (sleepLiveManifold(m)) sleepManifoldImpact(m, true);
    for (const Joint<R> &j : joints)
      if (sleepLiveJoint(j)) {
        sleepImpactBodies(uint32_t(j.a), uint32_t(j.b));
        sleepImpactBodies(uint32_t(j.b), uint32_t(j.a));
      }
    uint32_t woken = 0;
    for (uint32_t i = 0; i < n(); i++) {
      uint32_t *g = globalsOf(i);
      if (!sleep[i].asleep || (g[3] == 0 && g[0] == 0)) continue;
      if (g[0] != 0 || (sleepIslands && (sleep[sleepRoot(i)].wake & SW_IMPACT))) {
        wakeOne(i);
        g[2]++;
        woken++;
      }
    }
    return woken != 0;
  }
  /** support(): a's contact below with a static or sleeping partner, or one that was quiet and supported and is still
   * quiet, supports a. */
  void sleepSupport(uint32_t a, uint32_t b) {
    const Sleep &s = sleep[b];
    if (s.mass <= 0 || s.asleep || ((s.flags & SF_READY) && motion(b) <= sleepSpeed)) sleep[a].support = 1;
  }
  /** after(): islands (connect, flattened to their roots), contacts, joints, activity (island mode), rest. */
  void sleepAfter(uint32_t force) {
    const uint32_t N = uint32_t(n());
    if (sleepIslands) {
      for (const Manifold<R> &m : manifolds)
        if (sleepLiveManifold(m)) sleepConnect(m.a, m.b);
      for (const Joint<R> &j : joints)
        if (j.type != 0 && j.a >= 0 && j.b >= 0) sleepConnect(uint32_t(j.a), uint32_t(j.b));
      for (uint32_t i = 0; i < N; i++) sleep[i].parent = sleepRoot(i);
    }
    for (const Manifold<R> &m : manifolds) {
      if (!sleepLiveManifold(m)) continue;
      const uint32_t a = m.a, b = m.b;
      sleepManifoldImpact(m, false);
      const WorldParams &p = P(a);
      const R up = dot(m.normal, V3<R>(R(p.up[0]), R(p.up[1]), R(p.up[2])));
      if (up > R(0.3f)) sleepSupport(a, b);
      if (up < R(-0.3f)) sleepSupport(b, a);
      if (sleepIslands && (force & 2u)) {
        if (sleep[b].mass <= 0 && motion(b) <= sleepSpeed && up > R(0.3f) && sleep[a].mass > 0)
          sleep[sleepRoot(a)].wake |= SW_GROUNDED;
        if (sleep[a].mass <= 0 && motion(a) <= sleepSpeed && up < R(-0.3f) && sleep[b].mass > 0)
          sleep[sleepRoot(b)].wake |= SW_GROUNDED;
      }
    }
    for (const Joint<R> &j : joints)
      if (sleepLiveJoint(j)) {
        sleepImpactBodies(uint32_t(j.a), uint32_t(j.b));
        sleepImpactBodies(uint32_t(j.b), uint32_t(j.a));
      }
    if (sleepIslands)
      for (uint32_t i = 0; i < N; i++) {
        const Sleep &s = sleep[i];
        if (s.mass <= 0) continue;
        const WorldParams &p = P(i);
        bool busy = false;
        if (force & 2u) {
          busy = (s.flags & SF_START_QUIET) == 0;
        } else if (!s.asleep) {
          busy = motion(i) > sleepSpeed;
          if (s.eligible && !(s.flags & SF_HELD))
            busy = busy || s.quiet + R(p.dt) < sleepDelay || (p.gravity != 0.f && !s.support);
        }
        if (busy) sleep[sleepRoot(i)].wake |= SW_BUSY;
      }
    for (uint32_t i = 0; i < N; i++) {
      Sleep &s = sleep[i];
      if (s.mass <= 0) continue;
      if (!s.eligible) {
        wakeOne(i);
        continue;
      }
      uint32_t *g = globalsOf(i);
      const uint32_t islandWake = sleepIslands ? sleep[sleepRoot(i)].wake : 0u;
      if ((islandWake & SW_IMPACT) || g[0] != 0 || clashes != 0 || overflow != 0) {
        wakeOne(i);
        continue;
      }
      const WorldParams &p = P(i);
      const bool startQuiet = (s.flags & SF_START_QUIET) != 0, held = (s.flags & SF_HELD) != 0;
      const bool ready = motion(i) <= sleepSpeed && (p.gravity == 0.f || s.support != 0);
      s.flags = ready ? SF_READY : 0u;
      if (!s.asleep) {
        s.quiet = ready ? s.quiet + R(p.dt) : R(0);
        const bool now = sleepIslands ? !held && ((s.quiet >= sleepDelay && !(islandWake & SW_BUSY)) ||
                                                  ((force & 2u) && startQuiet && (islandWake & 6u) == SW_GROUNDED))
                                      : s.quiet >= sleepDelay;
        if (now) {
          s.asleep = 1;
          Body<R> &B = bodies[i];
          B.mass = 0;
          B.vel = V3<R>(0, 0, 0);
          B.prevUp = 0;
          B.angVel = V3<R>(0, 0, 0);
        }
      }
      if (s.asleep) g[1]++;
    }
  }

  void step() {
    pivotFallbacks = 0;
    const uint32_t force = sleepForce;
    if (sleepOn) {
      sleepBefore(force);
      sleepForce = 0;
    }
    const bool bounce = bouncing();
    if (bounce) restitutionBefore();
    collisionAndTopology();
    solve();
    if (bounce) restitutionAfter();
    if (sleepOn) sleepAfter(force);
  }

  // --- edits between steps -----------------------------------------------------------------------------------------

  /** ω change of an impulse about the centre: world → body frame, ÷ max(moment, 1e-10), back (bodyCommands3D.wgsl:9-12). */
  static V3<R> angularImpulse(const Body<R> &b, V3<R> impulse) {
    V3<R> local = rotate(conj(b.rot), impulse);
    const R floor = R(1e-10f);
    V3<R> scaled{local.x / rmax(b.moment.x, floor), local.y / rmax(b.moment.y, floor), local.z / rmax(b.moment.z, floor)};
    return rotate(b.rot, scaled);
  }

  /**
   * Body commands, 12 words each: body, kind, has point, 0 | value xyzw (f32) | point xyzw (f32), applied in
   * submission order (bodyCommands3D.wgsl:14-36 applies each body's commands in order, and commands of different
   * bodies commute). Kinds: 0 set linear velocity, 1 set angular velocity, 2 impulse (at the point if given),
   * 3 force × value.w (= dt), 4 teleport (position; rotation = point if given), 5 angular impulse, 6 torque × value.w.
   * Impulses leave static bodies (mass 0) unchanged. A teleport resets the step-start and inertial poses, and the last
   * step's manifolds that touch a teleported body lose their warm start (count and generation word = 0,
   * bodyCommands3D.wgsl:38-44).
   */
  void applyCommands(const uint32_t *words, uint32_t count) {
    std::vector<uint8_t> teleported(n(), 0);
    for (uint32_t k = 0; k < count; k++) {
      const uint32_t *c = words + 12 * k;
      float f[8];
      std::memcpy(f, c + 4, sizeof f);
      Body<R> &b = bodies[c[0]];
      const uint32_t kind = c[1];
      const bool hasPoint = c[2] != 0;
      V3<R> value{R(f[0]), R(f[1]), R(f[2])};
      if (kind == 0) {
        b.vel = value;
      } else if (kind == 1) {
        b.angVel = value;
      } else if (kind == 2 || kind == 3) {
        if (b.mass > 0) {
          V3<R> impulse = scale(value, kind == 3 ? R(f[3]) : R(1));
          b.vel = add(b.vel, divs(impulse, b.mass));
          if (hasPoint) {
            V3<R> point{R(f[4]), R(f[5]), R(f[6])};
            b.angVel = add(b.angVel, angularImpulse(b, cross(sub(point, b.pos), impulse)));
          }
        }
      } else if (kind == 5 || kind == 6) {
        if (b.mass > 0) b.angVel = add(b.angVel, angularImpulse(b, scale(value, kind == 6 ? R(f[3]) : R(1))));
      } else if (kind == 4) {
        b.pos = value;
        if (hasPoint) b.rot = Q4<R>(R(f[4]), R(f[5]), R(f[6]), R(f[7]));
        b.ipos = b.pos;
        b.irot = b.rot;
        b.inertialPos = b.pos;
        b.inertialRot = b.rot;
        teleported[c[0]] = 1;
      }
    }
    for (Manifold<R> &m : manifolds)
      if (m.count() != 0 && (teleported[m.a] || teleported[m.b])) m.countw = 0;
  }
};

}  // namespace gpu
}  // namespace oracle
