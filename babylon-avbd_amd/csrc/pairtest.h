// The broadphase pair predicate shared by the grid and the BVH: testPair of
// reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:112-221 (filters, bounding spheres, world AABBs with a 1e-4
// pad, face axes separating by more than 1e-3, the no-collide list with broken joints re-enabling), evaluated with
// the higher index as the probe so a pair's acceptance does not depend on which structure found it.
#pragma once
#include "engine_types.h"

namespace avbd {

/** Body i as the pair tests see it (wgsl-collision.ts:138-171). */
struct Shape {
  f3 ax[3];
  f3 h;
  bool sphere;
};

AVBD_D Shape shapeOf(const Body &b) {
  Shape s;
  s.h = scale(xyz(b.size), 0.5f);
  s.sphere = b.angVel.w == SHAPE_SPHERE;
  s.ax[0] = qrotate(b.rot, mk3(1.f, 0.f, 0.f));
  s.ax[1] = qrotate(b.rot, mk3(0.f, 1.f, 0.f));
  s.ax[2] = qrotate(b.rot, mk3(0.f, 0.f, 1.f));
  return s;
}
AVBD_D float extent(const Shape &s, f3 n) {
  if (s.sphere) return s.h.x;
  return s.h.x * fabsf(dot(n, s.ax[0])) + s.h.y * fabsf(dot(n, s.ax[1])) + s.h.z * fabsf(dot(n, s.ax[2]));
}
AVBD_D f3 aabbHalf(const Shape &s) {
  if (s.sphere) return mk3(s.h.x, s.h.x, s.h.x);
  return add(add(scale(vabs(s.ax[0]), s.h.x), scale(vabs(s.ax[1]), s.h.y)), scale(vabs(s.ax[2]), s.h.z));
}
AVBD_D bool faceSeparated(const Shape &A, const Shape &B, f3 d) {
  for (int k = 0; k < 3; k++) {
    if (fabsf(dot(d, A.ax[k])) - extent(A, A.ax[k]) - extent(B, A.ax[k]) > 1e-3f) return true;
    if (fabsf(dot(d, B.ax[k])) - extent(A, B.ax[k]) - extent(B, B.ax[k]) > 1e-3f) return true;
  }
  return false;
}

/**
 * The pair is on the no-collide list (an ignored pair, or a joint that has not broken; wgsl-collision.ts:112). The
 * list is grouped by the higher body, so this scans hi's few entries instead of searching the whole list.
 */
AVBD_D bool ignored(const DeviceWorld &w, uint32_t hi, uint32_t lo) {
  const uint32_t end = w.ncStart[hi + 1];
  for (uint32_t k = w.ncStart[hi]; k < end; k++) {
    if (w.noCollide[3 * k + 1] != lo) continue;
    uint32_t j = w.noCollide[3 * k + 2];
    if (j == 0xffffffffu) return true;
    if (w.joints[j].penLin.w != 0.f || w.joints[j].penAng.w != 0.f) return true;
  }
  return false;
}

/**
 * The pair predicate of testPair (wgsl-collision.ts:196-221), evaluated with the higher index a as the probe, so a
 * pair's acceptance does not depend on which thread or broadphase structure found it.
 */
AVBD_D bool acceptPair(const DeviceWorld &w, uint32_t a, uint32_t b) {
  if (w.bodyWorld[a] != w.bodyWorld[b]) return false;
  const Body &A = w.bodies[a];
  const Body &B = w.bodies[b];
  if (!(A.size.w > 0.f) && B.size.w <= 0.f) return false;
  uint2 fa = w.filters[a], fb = w.filters[b];
  if ((fa.x & fb.y) == 0u || (fb.x & fa.y) == 0u) return false;
  f3 d = sub(xyz(A.pos), xyz(B.pos));
  float r = A.moment.w + B.moment.w;
  if (dot(d, d) > r * r) return false;
  Shape sa = shapeOf(A), sb = shapeOf(B);
  f3 ha = aabbHalf(sa), hb = aabbHalf(sb);
  if (fabsf(d.x) > ha.x + hb.x + 1e-4f || fabsf(d.y) > ha.y + hb.y + 1e-4f || fabsf(d.z) > ha.z + hb.z + 1e-4f)
    return false;
  if (faceSeparated(sa, sb, d)) return false;
  if (ignored(w, a, b)) return false;
  return true;
}


/** Partners a body keeps from the counting pass; bodies with more are queried again in the write pass. */
constexpr uint32_t PAIR_STASH = 16;

/** Write body i's stashed partners at `base` as (a > b) pairs, sorted by partner. */
AVBD_D void writeStashedPairs(const DeviceWorld &w, uint32_t i, uint32_t base, uint32_t n) {
  uint32_t p[PAIR_STASH];
  for (uint32_t x = 0; x < n; x++) {
    uint32_t v = w.pairStash[size_t(i) * PAIR_STASH + x];
    uint32_t y = x;
    while (y > 0 && p[y - 1] > v) {
      p[y] = p[y - 1];
      y--;
    }
    p[y] = v;
  }
  for (uint32_t x = 0; x < n; x++)
    if (base + x < w.pairCapacity) w.pairs[base + x] = make_uint2(umax(i, p[x]), umin(i, p[x]));
}

/** Sort body i's n pairs written at `base` by partner, in place (each partner appears once: a total order). */
AVBD_D void sortBodyPairs(const DeviceWorld &w, uint32_t i, uint32_t base, uint32_t n) {
  uint32_t lim = base + n < w.pairCapacity ? base + n : w.pairCapacity;
  for (uint32_t x = base + 1; x < lim; x++) {
    uint2 v = w.pairs[x];
    uint32_t vo = v.x == i ? v.y : v.x;
    uint32_t y = x;
    while (y > base) {
      uint2 u = w.pairs[y - 1];
      uint32_t uo = u.x == i ? u.y : u.x;
      if (uo <= vo) break;
      w.pairs[y] = u;
      y--;
    }
    w.pairs[y] = v;
  }
}

}  // namespace avbd
