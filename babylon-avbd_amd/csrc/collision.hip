// Collision phase of the AVBD 3D step on gfx950: frame bookkeeping, reference poses for contact reuse, the uniform-grid
// broadphase with deterministic pair emission, the narrowphase with feature-keyed warm starts, and manifold compaction.
// Semantics: reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts (beginFrame 65-76, updateRefs 32-47, grid 86-110,
// findPairs/testPair 196-257, hash 278-313, reuseContacts 650-678, narrowphase 682-808).
//
// Design for MI355X: pair emission is count → scan → write (no atomic append), each body's pairs sorted by partner, so
// the pair list, and every manifold and contact slot derived from it, is the same every run. The narrowphase writes a
// pair's manifold and its contacts into fixed slots (contacts at 8 × pair index), and a scan compacts the manifolds.
#include <algorithm>

#include "engine_types.h"
#include "kernels.h"
#include "hullnarrow.h"
#include "narrowphase.h"
#include "pairtest.h"

namespace avbd {

// --- frame -----------------------------------------------------------------------------------------------------------

__global__ void k_begin_frame(DeviceWorld w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < w.nWorlds) w.worlds[i].step += 1;
  if (i == 0) {
    uint32_t *c = w.counters;
    c[C_PREV_MANIFOLDS] = c[C_MANIFOLDS];
    c[C_PREV_CONTACTS] = c[C_CONTACTS];
    c[C_MANIFOLDS] = 0;
    c[C_CONTACTS] = 0;
    c[C_PAIRS] = 0;
    c[C_OVERFLOW] = 0;
    c[C_CLASHES] = 0;
    c[C_NUM_COLORS] = 0;
    c[C_CANDIDATES] = 0;
    c[C_PIVOT64] = 0;
    c[C_DF_ABORT] = 0;
    c[C_PENDING] = 0;
    c[C_HULL_PAIRS] = 0;
  }
}

AVBD_D void updateRef(const DeviceWorld &w, uint32_t i) {
  const WorldParams &p = w.worlds[w.bodyWorld[i]];
  if (!(p.flags & FLAG_REUSE_CONTACTS)) return;
  f3 pos = xyz(w.bodies[i].pos);
  float4 rot = w.bodies[i].rot;
  float4 lp = w.refs[2 * i];
  float4 lr = w.refs[2 * i + 1];
  bool moved = (lr.x * lr.x + lr.y * lr.y + lr.z * lr.z + lr.w * lr.w) == 0.f ||
               length(sub(pos, xyz(lp))) > REUSE_LIN_TOL || length(qsub(rot, lr)) > REUSE_ANG_TOL;
  if (moved) {
    w.refs[2 * i] = mk4(pos, 0.f);
    w.refs[2 * i + 1] = rot;
    w.bodies[i].inertialPos.w = bitsf(p.step);
  }
}

// --- last step's manifolds by pair (wgsl-collision.ts:278-313) --------------------------------------------------------

AVBD_D uint32_t pairHash(uint32_t a, uint32_t b) { return hash32((a * 0x9e3779b1u) ^ hash32(b)); }

/**
 * Each body's contact-reuse reference (above), the cleared pair table, and the cleared adjacency counters (the
 * manifold compaction counts degrees into them), in one pass.
 */
__global__ void k_refs_and_clear(DeviceWorld w) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= w.nBodies; i += stride) {
    if (i < w.nBodies) {
      updateRef(w, i);
      w.adjFill[i] = 0u;
    }
    w.adjStart[i] = 0u;
  }
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h <= w.hashMask; h += stride) w.table[h] = 0;
}

__global__ void k_hash_insert(DeviceWorld w) {
  uint32_t count = w.counters[C_PREV_MANIFOLDS];
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < count; m += gridDim.x * blockDim.x) {
    uint4 ids = w.prevManifolds[m].ids;
    uint32_t h = pairHash(ids.x, ids.y) & w.hashMask;
    uint32_t value = m + 1;
    for (uint32_t probe = 0; probe <= w.hashMask; probe++) {
      value = atomicExch(&w.table[h], value);
      if (value == 0) break;
      h = (h + 1) & w.hashMask;
    }
  }
}

/** The first slot of pair (a, b)'s probe sequence and the value it holds: read ahead of the collision test, so the
 * load's latency is hidden behind it. */
struct HashProbe {
  uint32_t h, v;
};
AVBD_D HashProbe hashProbe(const DeviceWorld &w, uint32_t a, uint32_t b) {
  const uint32_t h = pairHash(a, b) & w.hashMask;
  return HashProbe{h, __atomic_load_n(&w.table[h], __ATOMIC_RELAXED)};
}
/** The previous manifold of pair (a, b), or -1, probing on from the first slot `pr` already read. */
AVBD_D int hashFind(const DeviceWorld &w, uint32_t a, uint32_t b, HashProbe pr) {
  uint32_t h = pr.h, v = pr.v;
  for (uint32_t probe = 0; probe <= w.hashMask; probe++) {
    if (probe != 0) v = __atomic_load_n(&w.table[h], __ATOMIC_RELAXED);
    if (v == 0) return -1;
    uint4 ids = w.prevManifolds[v - 1].ids;
    if (ids.x == a && ids.y == b) return int(v - 1);
    h = (h + 1) & w.hashMask;
  }
  return -1;
}
AVBD_D int hashFind(const DeviceWorld &w, uint32_t a, uint32_t b) { return hashFind(w, a, b, hashProbe(w, a, b)); }

// --- broadphase ------------------------------------------------------------------------------------------------------

AVBD_D bool isLarge(const DeviceWorld &w, uint32_t i) { return w.bodies[i].moment.w > w.worlds[w.bodyWorld[i]].maxSmall; }

AVBD_D int4 cellCoords(const DeviceWorld &w, uint32_t i) {
  float4 p = w.bodies[i].pos;
  return make_int4(int(floorf(p.x / w.cellSize)), int(floorf(p.y / w.cellSize)), int(floorf(p.z / w.cellSize)),
                   int(w.bodyWorld[i]));
}

AVBD_D uint32_t cellHash(const DeviceWorld &w, int4 c) {
  return ((uint32_t(c.x) * 73856093u) ^ (uint32_t(c.y) * 19349663u) ^ (uint32_t(c.z) * 83492791u) ^
          (uint32_t(c.w) * 2654435761u)) &
         w.tableMask;
}

__global__ void k_grid_count(DeviceWorld w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  if (isLarge(w, i)) {
    w.cellOf[i] = make_int4(0, 0, 0, -1);
    return;
  }
  int4 c = cellCoords(w, i);
  w.cellOf[i] = c;
  atomicAdd(&w.gridCount[cellHash(w, c)], 1u);
}

__global__ void k_grid_scatter(DeviceWorld w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  int4 c = w.cellOf[i];
  if (c.w < 0) return;
  uint32_t h = cellHash(w, c);
  uint32_t slot = w.gridCount[h] + atomicAdd(&w.gridCursor[h], 1u);
  w.gridSorted[slot] = i;
  float4 p = w.bodies[i].pos;
  w.sortedPR[slot] = make_float4(p.x, p.y, p.z, w.bodies[i].moment.w);
  w.sortedCell[slot] = c;
}

/**
 * One body's pairs: small bodies scan the 27 cells around theirs for lower-index small bodies and test their world's
 * large bodies; a large body tests the higher-index large bodies of its world (findPairs, wgsl-collision.ts:223-257).
 * The count pass stores how many; the write pass stores them at the scanned offset, then sorts them by partner (each
 * partner appears once, so the order is total and the list is the same whatever order the grid buckets hold).
 */
template <bool STASH> __global__ void k_pairs(DeviceWorld w) {
  if (gatedOff(w)) return;
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= w.nBodies) return;
  // Threads take bodies in cell order (small bodies) then the large ones: neighbouring threads scan the same cells
  const uint32_t nSmall = w.nBodies - w.largeCount;
  uint32_t i = q < nSmall ? w.gridSorted[q] : w.largeList[q - nSmall];
  uint32_t base = STASH ? 0 : w.pairCount[i];
  uint32_t total = STASH ? 0 : w.pairCount[i + 1] - base;
  if (!STASH && total <= PAIR_STASH) {
    writeStashedPairs(w, i, base, total);
    return;
  }
  uint32_t n = 0;
  auto emit = [&](uint32_t a, uint32_t b) {
    if (!acceptPair(w, a, b)) return;
    if (STASH) {
      if (n < PAIR_STASH) w.pairStash[size_t(i) * PAIR_STASH + n] = a == i ? b : a;
    } else if (base + n < w.pairCapacity) {
      w.pairs[base + n] = make_uint2(a, b);
    }
    n++;
  };
  int4 c = w.cellOf[i];
  bool small = c.w >= 0;
  uint32_t wid = w.bodyWorld[i];
  if (small) {
    // Half shell: the body's own cell (lower indices only) and the 13 cells lexicographically after it, so each
    // pair of small bodies is visited once; the cheap bounding-sphere test runs on the cell-sorted copies first
    const float4 pi = make_float4(w.bodies[i].pos.x, w.bodies[i].pos.y, w.bodies[i].pos.z, w.bodies[i].moment.w);
    for (int o = 0; o < 14; o++) {
      int ox, oy, oz;
      if (o == 0) {
        ox = oy = oz = 0;
      } else if (o < 10) {
        oz = 1;
        oy = (o - 1) / 3 - 1;
        ox = (o - 1) % 3 - 1;
      } else if (o < 13) {
        oz = 0;
        oy = 1;
        ox = o - 11;
      } else {
        oz = 0;
        oy = 0;
        ox = 1;
      }
      int4 cell = make_int4(c.x + ox, c.y + oy, c.z + oz, c.w);
      uint32_t h = cellHash(w, cell);
      uint32_t kEnd = w.gridCount[h + 1];
      for (uint32_t k = w.gridCount[h]; k < kEnd; k++) {
        int4 cj = w.sortedCell[k];
        if (cj.x != cell.x || cj.y != cell.y || cj.z != cell.z || cj.w != cell.w) continue;
        uint32_t j = w.gridSorted[k];
        if (o == 0 && j >= i) continue;
        float4 pj = w.sortedPR[k];
        f3 d = j < i ? mk3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) : mk3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
        float r = j < i ? pi.w + pj.w : pj.w + pi.w;
        if (dot(d, d) > r * r) continue;
        emit(umax(i, j), umin(i, j));
      }
    }
  }
  for (uint32_t k = w.largeStart[wid]; k < w.largeStart[wid + 1]; k++) {
    uint32_t l = w.largeList[k];
    if (l == i || (!small && l < i)) continue;
    emit(umax(i, l), umin(i, l));
  }
  if (STASH) {
    w.pairCount[i] = n;
    return;
  }
  sortBodyPairs(w, i, base, n);
}

__global__ void k_pairs_total(DeviceWorld w) {
  if (gatedOff(w)) return;
  uint32_t total = w.pairCount[w.nBodies];
  w.counters[C_PAIR_DEMAND] = total;
  if (total > w.pairCapacity) {
    w.counters[C_OVERFLOW] |= 1u;
    total = w.pairCapacity;
  }
  w.counters[C_PAIRS] = total;
}

// --- narrowphase -------------------------------------------------------------------------------------------------------

AVBD_D float contactPenaltyMin(const DeviceWorld &w, const WorldParams &p, uint32_t a, uint32_t b) {
  if (!(p.flags & FLAG_MASS_PENALTY)) return PENALTY_MIN;
  float mA = w.bodies[a].size.w <= 0.f ? 3.4e38f : w.bodies[a].size.w;
  float mB = w.bodies[b].size.w <= 0.f ? 3.4e38f : w.bodies[b].size.w;
  return rclamp(rmin(mA, mB) / (p.dt * p.dt), PENALTY_MIN, PENALTY_MAX);
}

AVBD_D f3 clampPen(f3 pen, float gamma) {
  return mk3(wclamp(pen.x * gamma, PENALTY_MIN, PENALTY_MAX), wclamp(pen.y * gamma, PENALTY_MIN, PENALTY_MAX),
             wclamp(pen.z * gamma, PENALTY_MIN, PENALTY_MAX));
}

/** Keep last step's points of a pair whose bodies have not moved since they were made (wgsl-collision.ts:650-678). */
AVBD_D bool reuseContacts(const DeviceWorld &w, const WorldParams &p, uint32_t pslot, uint32_t a, uint32_t b) {
  int pm = hashFind(w, a, b);
  if (pm < 0) return false;
  Manifold prev = w.prevManifolds[pm];
  uint32_t count = prev.ids.w & 15u;
  uint32_t generated = prev.ids.w >> 4;
  uint32_t moved = umax(fbits(w.bodies[a].inertialPos.w), fbits(w.bodies[b].inertialPos.w));
  if (count == 0 || generated < moved) return false;
  uint32_t base = pslot * MAX_CONTACTS;
  w.mfSlots[pslot] = Manifold{make_uint4(a, b, base, prev.ids.w), prev.geo};
  f3 pA = xyz(w.bodies[a].pos), pB = xyz(w.bodies[b].pos);
  float4 qA = w.bodies[a].rot, qB = w.bodies[b].rot;
  Basis basis = orthonormal(xyz(prev.geo));
  float pmin = contactPenaltyMin(w, p, a, b);
  for (uint32_t i = 0; i < count; i++) {
    Contact k = w.prevContacts[prev.ids.z + i];
    f3 d = sub(add(qrotate(qA, xyz(k.rA)), pA), add(qrotate(qB, xyz(k.rB)), pB));
    k.rB.w = dot(basis.n, d) + COLLISION_MARGIN;
    k.pen.w = dot(basis.t1, d);
    k.lam.w = dot(basis.t2, d);
    f3 lam = scale(scale(xyz(k.lam), p.alpha), p.gamma);
    f3 pen = clampPen(xyz(k.pen), p.gamma);
    pen.x = rmax(pen.x, pmin);
    k.lam = mk4(lam, k.lam.w);
    k.pen = mk4(pen, k.pen.w);
    w.contacts[base + i] = k;
  }
  return true;
}

/**
 * A pair slot's manifold and contact points from the narrowphase's points (wgsl-collision.ts:734-807): the manifold
 * word (count | generation step), the normal (B towards A, -satN) and friction; each point's body-local anchors, warm
 * start from last step's point with the same feature (or the nearest anchor with FLAG_MATCH_NEAREST; static friction
 * keeps the old anchors except for rounded shapes), its C0 rows and the scaled multipliers and penalties.
 */
AVBD_D void writePairContacts(const DeviceWorld &w, const WorldParams &p, uint32_t pslot, uint32_t a, uint32_t b,
                              const Body &BA, const Body &BB, const Box &A, const Box &B, const Found &found, f3 satN,
                              bool anySphere, HashProbe probe) {
  uint32_t base = pslot * MAX_CONTACTS;
  if (found.count == 0) {
    w.mfSlots[pslot].ids = make_uint4(a, b, base, 0u);
    return;
  }
  f3 n = neg(satN);
  w.mfSlots[pslot] = Manifold{make_uint4(a, b, base, uint32_t(found.count) | (p.step << 4)),
                              mk4(n, sqrtf(BA.pos.w * BB.pos.w))};
  float4 qA = BA.rot, qB = BB.rot;
  Basis basis = orthonormal(n);
  uint32_t prevFirst = 0, prevCount = 0;
  int pm = hashFind(w, a, b, probe);
  if (pm >= 0) {
    prevFirst = w.prevManifolds[pm].ids.z;
    prevCount = w.prevManifolds[pm].ids.w & 15u;
  }
  // The previous points' feature keys, read together once: each new point is matched against them in registers
  uint32_t prevKey[MAX_CONTACTS];
#pragma unroll
  for (uint32_t c = 0; c < uint32_t(MAX_CONTACTS); c++)
    prevKey[c] = c < prevCount ? fbits(w.prevContacts[prevFirst + c].rA.w) & ~STICK_BIT : 0u;
  bool matchNearest = (p.flags & FLAG_MATCH_NEAREST) != 0;
  float minSide = rmin(rmin(rmin(A.h.x, A.h.y), A.h.z), rmin(rmin(B.h.x, B.h.y), B.h.z)) * 2.f;
  float pmin = contactPenaltyMin(w, p, a, b);
  // Unrolled over the slots, so the found points are read at constant indices and stay in registers
#pragma unroll
  for (int i = 0; i < MAX_CONTACTS; i++) {
    if (i >= found.count) break;
    const f3 xA = found.xA[i], xB = found.xB[i];
    const uint32_t feature = found.feature[i];
    f3 rA = qrotate(qconj(qA), sub(xA, A.c));
    f3 rB = qrotate(qconj(qB), sub(xB, B.c));
    f3 pen = mk3(0.f, 0.f, 0.f), lam = mk3(0.f, 0.f, 0.f);
    uint32_t stick = 0;
    int j = -1;  // the first previous point with the same key
#pragma unroll
    for (int c = MAX_CONTACTS - 1; c >= 0; c--)
      if (uint32_t(c) < prevCount && prevKey[c] == feature) j = int(prevFirst) + c;
    if (j < 0 && matchNearest) {
      float bestDist = NEAREST_FRACTION * minSide;
      for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++) {
        float dist = length(sub(xyz(w.prevContacts[c].rA), rA));
        if (dist <= bestDist) {
          bestDist = dist;
          j = int(c);
        }
      }
    }
    if (j >= 0) {
      Contact pc = w.prevContacts[j];
      pen = xyz(pc.pen);
      lam = xyz(pc.lam);
      stick = fbits(pc.rA.w) & STICK_BIT;
      if (stick != 0 && !anySphere) {
        rA = xyz(pc.rA);
        rB = xyz(pc.rB);
      }
    }
    f3 d = sub(add(qrotate(qA, rA), A.c), add(qrotate(qB, rB), B.c));
    float c0x = dot(basis.n, d) + COLLISION_MARGIN;
    float c0y = dot(basis.t1, d);
    float c0z = dot(basis.t2, d);
    lam = scale(scale(lam, p.alpha), p.gamma);
    pen = clampPen(pen, p.gamma);
    pen.x = rmax(pen.x, pmin);
    Contact k;
    k.rA = mk4(rA, bitsf(feature | stick));
    k.rB = mk4(rB, c0x);
    k.pen = mk4(pen, c0y);
    k.lam = mk4(lam, c0z);
    w.contacts[base + i] = k;
  }
}

/**
 * Contacts of every pair slot (wgsl-collision.ts:617-808). kRounded = false compiles the box/box path only (a batch
 * without spheres or capsules), which keeps the kernel's code small enough to stay in the instruction cache. With
 * kHulls, a pair with a hull on either side is listed for k_narrowphase_hull instead.
 */
template <bool kRounded, bool kHulls> __global__ void __launch_bounds__(NARROW_BLOCK) k_narrowphase(DeviceWorld w) {
  if (gatedOff(w)) return;
  uint32_t P = w.counters[C_PAIRS];
  for (uint32_t pslot = blockIdx.x * blockDim.x + threadIdx.x; pslot < P; pslot += gridDim.x * blockDim.x) {
    uint2 pr = w.pairs[pslot];
    uint32_t a = pr.x, b = pr.y;
    const HashProbe probe = hashProbe(w, a, b);
    const WorldParams &p = w.worlds[w.bodyWorld[a]];
    if ((p.flags & FLAG_REUSE_CONTACTS) && reuseContacts(w, p, pslot, a, b)) continue;
    const Body &BA = w.bodies[a];
    const Body &BB = w.bodies[b];
    if (kHulls && (isHullShape(BA.angVel.w) || isHullShape(BB.angVel.w))) {
      w.hullPairs[atomicAdd(&w.counters[C_HULL_PAIRS], 1u)] = pslot;  // at most one entry per pair slot
      continue;
    }
    Box A = makeBox(xyz(BA.pos), BA.rot, xyz(BA.size));
    Box B = makeBox(xyz(BB.pos), BB.rot, xyz(BB.size));
    Sat sat;
    sat.n = mk3(0.f, 0.f, 0.f);
    Found found;
    found.count = 0;
    const bool sphereA = kRounded && BA.angVel.w == SHAPE_SPHERE, sphereB = kRounded && BB.angVel.w == SHAPE_SPHERE;
    const bool capsuleA = kRounded && BA.angVel.w == SHAPE_CAPSULE, capsuleB = kRounded && BB.angVel.w == SHAPE_CAPSULE;
    if (capsuleA || capsuleB) {
      // The capsule is the first argument; a capsule B flips the result back (capsuleCollision.js:17-29)
      if (capsuleA && capsuleB) {
        found = capsuleCapsule(capsuleOf(BA), capsuleOf(BB), sat);
      } else if (capsuleA) {
        Capsule C = capsuleOf(BA);
        found = sphereB ? capsuleSphere(C, B.c, B.h.x, sat) : capsuleBox(C, B, BB.rot, sat);
      } else {
        Capsule C = capsuleOf(BB);
        found = sphereA ? capsuleSphere(C, A.c, A.h.x, sat) : capsuleBox(C, A, BA.rot, sat);
        sat.n = neg(sat.n);
        swapFoundSides(found);
      }
    } else if (sphereA && sphereB) {
      found = sphereSphere(A, B, sat);
    } else if (sphereA || sphereB) {
      if (sphereA) {
        SphereBoxHit h = sphereBox(A, B, BB.rot);
        sat.n = h.n;
        if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
      } else {
        SphereBoxHit h = sphereBox(B, A, BA.rot);
        sat.n = neg(h.n);
        if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
      }
    } else {
      found = collideBoxes(A, B, (p.flags & FLAG_FACE_BIAS) != 0, sat);
    }
    writePairContacts(w, p, pslot, a, b, BA, BB, A, B, found, sat.n, sphereA || sphereB || capsuleA || capsuleB,
                      probe);
  }
}

/**
 * Contacts of the pair slots with a hull (the hull branch of wgsl-collision.ts:697-711 and capsuleCollision.js:17-29),
 * one wavefront per pair: capsule against hull (capsulePoly, the capsule as the first argument), sphere against hull
 * (sphereHull), hull or box against hull (collidePoly); then the same manifold and warm start as every pair.
 */
__global__ void __launch_bounds__(HULL_BLOCK) k_narrowphase_hull(DeviceWorld w) {
  if (gatedOff(w)) return;
  __shared__ HullScratch sc;
  const uint32_t count = w.counters[C_HULL_PAIRS];
  for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
    const uint32_t pslot = w.hullPairs[k];
    const uint2 pr = w.pairs[pslot];
    const uint32_t a = pr.x, b = pr.y;
    const WorldParams &p = w.worlds[w.bodyWorld[a]];
    const Body &BA = w.bodies[a];
    const Body &BB = w.bodies[b];
    const Box A = makeBox(xyz(BA.pos), BA.rot, xyz(BA.size));
    const Box B = makeBox(xyz(BB.pos), BB.rot, xyz(BB.size));
    Sat sat;
    sat.n = mk3(0.f, 0.f, 0.f);
    Found found;
    found.count = 0;
    const bool sphereA = BA.angVel.w == SHAPE_SPHERE, sphereB = BB.angVel.w == SHAPE_SPHERE;
    const bool capsuleA = BA.angVel.w == SHAPE_CAPSULE, capsuleB = BB.angVel.w == SHAPE_CAPSULE;
    if (capsuleA) {
      found = capsulePoly(w, capsuleOf(BA), polyOf(w, b), sat);
    } else if (capsuleB) {
      found = capsulePoly(w, capsuleOf(BB), polyOf(w, a), sat);
      sat.n = neg(sat.n);
      swapFoundSides(found);
    } else if (sphereA) {
      const SphereBoxHit h = sphereHull(w, A.c, A.h.x, polyOf(w, b));
      sat.n = h.n;
      if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
    } else if (sphereB) {
      const SphereBoxHit h = sphereHull(w, B.c, B.h.x, polyOf(w, a));
      sat.n = neg(h.n);
      if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
    } else {
      found = collidePoly(w, polyOf(w, a), polyOf(w, b), (p.flags & FLAG_FACE_BIAS) != 0, sat, sc);
    }
    if (laneId() == 0)
      writePairContacts(w, p, pslot, a, b, BA, BB, A, B, found, sat.n, sphereA || sphereB || capsuleA || capsuleB,
                        hashProbe(w, a, b));
    __syncthreads();  // the scratch is free for the next pair
  }
}

constexpr uint32_t COMPACT_BLOCK = 256, COMPACT_GRID = 512;

/** Exclusive prefix of v over the 256 threads of a block, and the block total (every thread must call it). */
AVBD_D uint32_t blockExclusive(uint32_t v, uint32_t &total, uint32_t *waveTot) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  __syncthreads();  // waveTot is free (an earlier call has read it)
  if (lane == 63u) waveTot[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
#pragma unroll
  for (uint32_t k = 0; k < COMPACT_BLOCK / 64; k++) {
    if (k < wave) before += waveTot[k];
    total += waveTot[k];
  }
  return before + incl - v;
}

/**
 * Dense manifolds and the dense point list (manifold << 3 | point). Each block takes one contiguous range of pair
 * slots: it counts the range's manifolds and points, takes their ranges of the two lists with one atomic each, then
 * writes tile by tile (a tile's manifold and point offsets from one packed block scan). The order of the blocks' ranges
 * varies from run to run; nothing downstream depends on it (adjacency lists are sorted, the hash finds manifolds by
 * pair, every point is updated on its own). Each manifold and active joint adds to its dynamic ends' adjacency
 * degrees (cleared by k_refs_and_clear).
 */
__global__ void __launch_bounds__(COMPACT_BLOCK) k_manifold_compact(DeviceWorld w) {
  if (gatedOff(w)) return;
  __shared__ uint32_t waveTot[COMPACT_BLOCK / 64];
  __shared__ uint32_t mFirst, pFirst;
  const uint32_t P = w.counters[C_PAIRS];
  const uint32_t per = (P + gridDim.x - 1) / gridDim.x;
  const uint32_t chunk = (per + COMPACT_BLOCK - 1) / COMPACT_BLOCK * COMPACT_BLOCK;
  const uint32_t lo = blockIdx.x * chunk;
  if (lo < P) {  // the whole block
    const uint32_t hi = umin(P, lo + chunk);
    uint32_t packed = 0;  // points << 16 | manifolds of this thread's slots
    for (uint32_t p = lo + threadIdx.x; p < hi; p += COMPACT_BLOCK) {
      const uint32_t n = w.mfSlots[p].ids.w & 15u;
      if (n) packed += (n << 16) | 1u;
    }
    uint32_t total = 0;
    blockExclusive(packed, total, waveTot);
    if (threadIdx.x == 0) {
      mFirst = (total & 0xffffu) ? atomicAdd(&w.counters[C_MANIFOLDS], total & 0xffffu) : 0u;
      pFirst = (total >> 16) ? atomicAdd(&w.counters[C_CONTACTS], total >> 16) : 0u;
    }
    __syncthreads();
    uint32_t mRun = mFirst, pRun = pFirst;
    for (uint32_t base = lo; base < hi; base += COMPACT_BLOCK) {  // the same turns for every thread of the block
      const uint32_t p = base + threadIdx.x;
      Manifold m;
      uint32_t points = 0;
      if (p < hi) {
        m = w.mfSlots[p];
        points = m.ids.w & 15u;
      }
      uint32_t tileTotal = 0;
      const uint32_t off = blockExclusive(points ? (points << 16) | 1u : 0u, tileTotal, waveTot);
      if (points) {
        const uint32_t o = mRun + (off & 0xffffu), first = pRun + (off >> 16);
        w.manifolds[o] = m;
        for (uint32_t k = 0; k < points; k++) w.pointMap[first + k] = (o << 3) | k;
        // The adjacency degree of each dynamic end
        if (w.bodies[m.ids.x].size.w > 0.f) atomicAdd(&w.adjStart[m.ids.x], 1u);
        if (w.bodies[m.ids.y].size.w > 0.f) atomicAdd(&w.adjStart[m.ids.y], 1u);
      }
      mRun += tileTotal & 0xffffu;
      pRun += tileTotal >> 16;
    }
  }
  // Active joints' degrees (a joint is active while either stiffness is non-zero)
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < w.nJoints; j += gridDim.x * blockDim.x) {
    const int4 f = w.info[j];
    if (f.x == T_NONE || (w.joints[j].penLin.w == 0.f && w.joints[j].penAng.w == 0.f)) continue;
    if (f.y >= 0 && w.bodies[f.y].size.w > 0.f) atomicAdd(&w.adjStart[f.y], 1u);
    if (f.z >= 0 && w.bodies[f.z].size.w > 0.f) atomicAdd(&w.adjStart[f.z], 1u);
  }
}

// --- launch wrappers ---------------------------------------------------------------------------------------------------

static inline uint32_t blocks(uint64_t n, uint32_t bs, uint32_t cap = 1u << 20) {
  uint64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return uint32_t(g);
}

void launch_begin_frame(const DeviceWorld &w, hipStream_t s) {
  k_begin_frame<<<blocks(w.nWorlds, 256), 256, 0, s>>>(w);
}
void launch_refs_and_hash(const DeviceWorld &w, uint32_t prevCapacity, hipStream_t s) {
  const uint64_t items = std::max<uint64_t>(w.nBodies, uint64_t(w.hashMask) + 1);
  k_refs_and_clear<<<blocks(items, 256, 4096), 256, 0, s>>>(w);
  k_hash_insert<<<blocks(prevCapacity, 256, 2048), 256, 0, s>>>(w);
}
void launch_grid(const DeviceWorld &w, hipStream_t s) {
  launch_fill_u32(w.gridCount, 0u, uint64_t(w.tableMask) + 2, s);
  launch_fill_u32(w.gridCursor, 0u, uint64_t(w.tableMask) + 1, s);
  k_grid_count<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
  launch_scan_exclusive(w.gridCount, w.tableMask + 2, w.scanScratch, s, w.counters + C_DF_ABORT);
  k_grid_scatter<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
}
void launch_pairs(const DeviceWorld &w, hipStream_t s) {
  k_pairs<true><<<blocks(w.nBodies, 128), 128, 0, s>>>(w);
  launch_scan_exclusive(w.pairCount, w.nBodies + 1, w.scanScratch, s, w.counters + C_DF_ABORT, w.gate);
  k_pairs<false><<<blocks(w.nBodies, 128), 128, 0, s>>>(w);
  k_pairs_total<<<1, 1, 0, s>>>(w);
}
void launch_pairs_total(const DeviceWorld &w, hipStream_t s) { k_pairs_total<<<1, 1, 0, s>>>(w); }
void launch_narrowphase(const DeviceWorld &w, hipStream_t s) {
  const uint32_t grid = blocks(w.pairCapacity, NARROW_BLOCK, 8192);
  if (w.hullShapes) {
    k_narrowphase<true, true><<<grid, NARROW_BLOCK, 0, s>>>(w);
    k_narrowphase_hull<<<2048, HULL_BLOCK, 0, s>>>(w);
  } else if (w.roundedShapes) {
    k_narrowphase<true, false><<<grid, NARROW_BLOCK, 0, s>>>(w);
  } else {
    k_narrowphase<false, false><<<grid, NARROW_BLOCK, 0, s>>>(w);
  }
  // Ranges of at most 4096 slots, so a range's manifold and point counts each fit 16 bits of the packed scan
  const uint32_t compactGrid = std::max<uint32_t>(COMPACT_GRID, (w.pairCapacity + 4095u) / 4096u);
  k_manifold_compact<<<compactGrid, COMPACT_BLOCK, 0, s>>>(w);
}

}  // namespace avbd
