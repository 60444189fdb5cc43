// Device narrowphase geometry: box/box separating axes (3 + 3 faces, 9 edges), incident-face clipping with feature
// keys, the edge/edge closest-points contact, and the sphere contacts. Follows
// reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:317-626 (a port of ref/collide.ts) with FLAG_FACE_BIAS;
// testAxis takes 1/sqrtf (correctly rounded) where the WGSL has inverseSqrt, so the axis choice is reproducible.
#pragma once
#include "avbd_device.h"

namespace avbd {

constexpr uint32_t AXIS_FACE_A = 0, AXIS_FACE_B = 1, AXIS_EDGE = 2;

/** An oriented box: centre, half extents and the three world axes (named members, so that choosing an axis by a
 * runtime index is a register select and never an indexed stack access). */
struct Box {
  f3 c, h;
  f3 ax0, ax1, ax2;
};

AVBD_D Box makeBox(f3 pos, float4 rot, f3 size) {
  Box b;
  b.c = pos;
  b.h = scale(size, 0.5f);
  b.ax0 = qrotate(rot, mk3(1.f, 0.f, 0.f));
  b.ax1 = qrotate(rot, mk3(0.f, 1.f, 0.f));
  b.ax2 = qrotate(rot, mk3(0.f, 0.f, 1.f));
  return b;
}

struct Sat {
  uint32_t kind, ia, ib;
  float sep;
  f3 n;  // from A towards B
  bool valid;
};

struct Found {
  uint32_t feature[MAX_CONTACTS];
  f3 xA[MAX_CONTACTS];
  f3 xB[MAX_CONTACTS];
  int count;
};

/**
 * Append a point unless it lies within the merge distance of one already found. The loops run over every slot with a
 * guard, so the arrays are indexed by constants and stay in registers.
 */
AVBD_D void addFound(Found &f, f3 xA, f3 xB, uint32_t feature) {
  f3 mid = scale(add(xA, xB), 0.5f);
  bool near = false;
#pragma unroll
  for (int i = 0; i < MAX_CONTACTS; i++) {
    if (i < f.count) {
      f3 d = sub(mid, scale(add(f.xA[i], f.xB[i]), 0.5f));
      if (dot(d, d) < CONTACT_MERGE_DIST_SQ) near = true;
    }
  }
  if (near || f.count >= MAX_CONTACTS) return;
#pragma unroll
  for (int i = 0; i < MAX_CONTACTS; i++) {
    if (i == f.count) {
      f.feature[i] = feature;
      f.xA[i] = xA;
      f.xB[i] = xB;
    }
  }
  f.count++;
}

/** Swap the two sides of every point (a capsule argument reversed). */
AVBD_D void swapFoundSides(Found &f) {
#pragma unroll
  for (int k = 0; k < MAX_CONTACTS; k++) {
    if (k < f.count) {
      f3 t = f.xA[k];
      f.xA[k] = f.xB[k];
      f.xB[k] = t;
    }
  }
}

AVBD_D float projRadius(const Box &b, f3 n) {
  return b.h.x * fabsf(dot(n, b.ax0)) + b.h.y * fabsf(dot(n, b.ax1)) + b.h.z * fabsf(dot(n, b.ax2));
}

AVBD_D bool testAxis(const Box &A, const Box &B, f3 delta, f3 axis, uint32_t kind, uint32_t ia, uint32_t ib,
                     Sat &best) {
  float lenSq = dot(axis, axis);
  if (lenSq < SAT_AXIS_EPSILON) return true;
  float invLen = 1.f / sqrtf(lenSq);
  f3 n = scale(axis, invLen);
  if (dot(n, delta) < 0.f) n = neg(n);
  float sep = fabsf(dot(delta, n)) - (projRadius(A, n) + projRadius(B, n));
  if (sep > 0.f) return false;
  if (!best.valid || sep > best.sep) {
    best.valid = true;
    best.kind = kind;
    best.ia = ia;
    best.ib = ib;
    best.sep = sep;
    best.n = n;
  }
  return true;
}

AVBD_D f3 support(const Box &b, f3 dir) {
  float sx = dot(dir, b.ax0) >= 0.f ? 1.f : -1.f;
  float sy = dot(dir, b.ax1) >= 0.f ? 1.f : -1.f;
  float sz = dot(dir, b.ax2) >= 0.f ? 1.f : -1.f;
  f3 p = addScaled(b.c, b.ax0, b.h.x * sx);
  p = addScaled(p, b.ax1, b.h.y * sy);
  return addScaled(p, b.ax2, b.h.z * sz);
}

/**
 * Clipping polygons live in LDS, one slice per lane (the narrowphase runs 64-lane blocks): component c of vertex k of
 * polygon buffer b at clipPoly[b][c][k][lane], so the per-vertex reads and writes of varying index stay out of
 * scratch memory.
 */
constexpr uint32_t NARROW_BLOCK = 64;
__shared__ float clipPoly[2][3][MAX_POLY][NARROW_BLOCK];
AVBD_D f3 polyGet(int buf, int k) {
  const uint32_t l = threadIdx.x;
  return mk3(clipPoly[buf][0][k][l], clipPoly[buf][1][k][l], clipPoly[buf][2][k][l]);
}
AVBD_D void polySet(int buf, int k, f3 v) {
  const uint32_t l = threadIdx.x;
  clipPoly[buf][0][k][l] = v.x;
  clipPoly[buf][1][k][l] = v.y;
  clipPoly[buf][2][k][l] = v.z;
}

/** Sutherland–Hodgman against one plane: polygon buffer src (n vertices) into buffer dst; returns the vertex count. */
AVBD_D int clipPlane(int src, int n, int dst, f3 pn, float offset) {
  if (n == 0) return 0;
  int count = 0;
  f3 a = polyGet(src, n - 1);
  float da = dot(pn, a) - offset;
  for (int i = 0; i < n; i++) {
    f3 b = polyGet(src, i);
    float db = dot(pn, b) - offset;
    bool aIn = da <= PLANE_EPSILON;
    bool bIn = db <= PLANE_EPSILON;
    if (aIn != bIn) {
      float t = 0.f;
      float denom = da - db;
      if (fabsf(denom) > SAT_AXIS_EPSILON) t = rclamp(da / denom, 0.f, 1.f);
      if (count < MAX_POLY) polySet(dst, count++, addScaled(a, sub(b, a), t));
    }
    if (bIn && count < MAX_POLY) polySet(dst, count++, b);
    a = b;
    da = db;
  }
  return count;
}

AVBD_D int faceU(int k) { return k == 0 ? 1 : 0; }
AVBD_D int faceV(int k) { return k != 2 ? 2 : 1; }
/** c ? x : y member by member (a select of values; a select of references would keep both boxes in memory). */
AVBD_D f3 pick3(bool c, f3 x, f3 y) { return mk3(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z); }
AVBD_D Box pickBox(bool c, const Box &x, const Box &y) {
  Box b;
  b.c = pick3(c, x.c, y.c);
  b.h = pick3(c, x.h, y.h);
  b.ax0 = pick3(c, x.ax0, y.ax0);
  b.ax1 = pick3(c, x.ax1, y.ax1);
  b.ax2 = pick3(c, x.ax2, y.ax2);
  return b;
}

AVBD_D f3 boxAxis(const Box &b, int k) {
  const f3 a = k == 0 ? b.ax0 : b.ax1;
  return k == 2 ? b.ax2 : a;
}

AVBD_D Found faceManifold(const Box &A, const Box &B, bool refIsA, int refAxis, f3 nAB) {
  Found found;
  found.count = 0;
  const Box R = pickBox(refIsA, A, B);
  const Box I = pickBox(refIsA, B, A);
  f3 outward = refIsA ? nAB : neg(nAB);
  f3 rax = boxAxis(R, refAxis);
  float rs = dot(outward, rax) >= 0.f ? 1.f : -1.f;
  f3 rn = scale(rax, rs);
  f3 rc = addScaled(R.c, rn, comp(R.h, refAxis));
  f3 ru = boxAxis(R, faceU(refAxis)), rv = boxAxis(R, faceV(refAxis));
  float eu = comp(R.h, faceU(refAxis)), ev = comp(R.h, faceV(refAxis));

  int incAxis = 0;
  float bestDot = -3.4028234663852886e38f;
  for (int k = 0; k < 3; k++) {
    float d = fabsf(dot(boxAxis(I, k), rn));
    if (d > bestDot) {
      bestDot = d;
      incAxis = k;
    }
  }
  f3 iax = boxAxis(I, incAxis);
  float is = dot(iax, rn) > 0.f ? -1.f : 1.f;
  f3 fn = scale(iax, is);
  f3 ic = addScaled(I.c, fn, comp(I.h, incAxis));
  f3 u = boxAxis(I, faceU(incAxis)), v = boxAxis(I, faceV(incAxis));
  float eU = comp(I.h, faceU(incAxis)), eV = comp(I.h, faceV(incAxis));
  polySet(0, 0, addScaled(addScaled(ic, u, 1.f * eU), v, 1.f * eV));
  polySet(0, 1, addScaled(addScaled(ic, u, -1.f * eU), v, 1.f * eV));
  polySet(0, 2, addScaled(addScaled(ic, u, -1.f * eU), v, -1.f * eV));
  polySet(0, 3, addScaled(addScaled(ic, u, 1.f * eU), v, -1.f * eV));
  int n = 4;
  n = clipPlane(0, n, 1, ru, dot(ru, rc) + eu);
  n = clipPlane(1, n, 0, neg(ru), dot(neg(ru), rc) + eu);
  n = clipPlane(0, n, 1, rv, dot(rv, rc) + ev);
  n = clipPlane(1, n, 0, neg(rv), dot(neg(rv), rc) + ev);
  if (n == 0) return found;

  uint32_t prefix = ((refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24) | (uint32_t(refAxis) << 16) | (uint32_t(incAxis) << 8);
  for (int i = 0; i < n && found.count < MAX_CONTACTS; i++) {
    f3 p = polyGet(0, i);
    float dist = dot(sub(p, rc), rn);
    if (dist > PLANE_EPSILON) continue;
    f3 onRef = addScaled(p, rn, -dist);
    addFound(found, refIsA ? onRef : p, refIsA ? p : onRef, prefix | uint32_t(i));
  }
  if (found.count == 0) addFound(found, support(A, nAB), support(B, neg(nAB)), prefix);
  return found;
}

AVBD_D void supportEdge(const Box &b, int k, f3 dir, f3 &e0, f3 &e1) {
  int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
  float s1 = dot(dir, boxAxis(b, k1)) >= 0.f ? 1.f : -1.f;
  float s2 = dot(dir, boxAxis(b, k2)) >= 0.f ? 1.f : -1.f;
  f3 c = addScaled(b.c, boxAxis(b, k1), comp(b.h, k1) * s1);
  c = addScaled(c, boxAxis(b, k2), comp(b.h, k2) * s2);
  f3 e = scale(boxAxis(b, k), comp(b.h, k));
  e0 = sub(c, e);
  e1 = add(c, e);
}

AVBD_D Found edgeContact(const Box &A, const Box &B, int ia, int ib, f3 nAB) {
  Found found;
  found.count = 0;
  f3 p0, p1, q0, q1;
  supportEdge(A, ia, nAB, p0, p1);
  supportEdge(B, ib, neg(nAB), q0, q1);
  f3 d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
  float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  float s = 0.f, t = 0.f;
  if (a <= SAT_AXIS_EPSILON && e <= SAT_AXIS_EPSILON) {
  } else if (a <= SAT_AXIS_EPSILON) {
    t = rclamp(f / e, 0.f, 1.f);
  } else {
    float c = dot(d1, r);
    if (e <= SAT_AXIS_EPSILON) {
      s = rclamp(-c / a, 0.f, 1.f);
    } else {
      float b = dot(d1, d2);
      float denom = a * e - b * b;
      if (fabsf(denom) > SAT_AXIS_EPSILON) s = rclamp((b * f - c * e) / denom, 0.f, 1.f);
      t = (b * s + f) / e;
      if (t < 0.f) {
        t = 0.f;
        s = rclamp(-c / a, 0.f, 1.f);
      } else if (t > 1.f) {
        t = 1.f;
        s = rclamp((b - c) / a, 0.f, 1.f);
      }
    }
  }
  uint32_t key = (AXIS_EDGE << 24) | (uint32_t(ia) << 8) | uint32_t(ib);
  addFound(found, addScaled(p0, d1, s), addScaled(q0, d2, t), key);
  if (found.count == 0) addFound(found, support(A, nAB), support(B, neg(nAB)), key);
  return found;
}

AVBD_D Found collideBoxes(const Box &A, const Box &B, bool faceBias, Sat &sat) {
  Found none;
  none.count = 0;
  f3 delta = sub(B.c, A.c);
  Sat face, edge;
  face.valid = false;
  edge.valid = false;
  face.kind = edge.kind = face.ia = face.ib = edge.ia = edge.ib = 0;
  face.sep = edge.sep = 0.f;
  face.n = edge.n = mk3(0.f, 0.f, 0.f);
#pragma unroll
  for (int i = 0; i < 3; i++)
    if (!testAxis(A, B, delta, boxAxis(A, i), AXIS_FACE_A, uint32_t(i), 0u, face)) return none;
#pragma unroll
  for (int i = 0; i < 3; i++)
    if (!testAxis(A, B, delta, boxAxis(B, i), AXIS_FACE_B, 0u, uint32_t(i), face)) return none;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (!testAxis(A, B, delta, cross(boxAxis(A, i), boxAxis(B, j)), AXIS_EDGE, uint32_t(i), uint32_t(j), edge))
        return none;
  if (!face.valid) return none;
  Sat best = face;
  if (edge.valid) {
    bool edgeWins = faceBias ? (edge.sep > 0.95f * face.sep + 0.01f) : (0.95f * edge.sep > face.sep + 0.01f);
    if (edgeWins) best = edge;
  }
  sat = best;
  if (best.kind == AXIS_EDGE) return edgeContact(A, B, int(best.ia), int(best.ib), best.n);
  bool refIsA = best.kind == AXIS_FACE_A;
  return faceManifold(A, B, refIsA, int(refIsA ? best.ia : best.ib), best.n);
}

// --- spheres (wgsl-collision.ts:536-585) ---------------------------------------------------------------------------

AVBD_D Found sphereSphere(const Box &A, const Box &B, Sat &sat) {
  Found found;
  found.count = 0;
  f3 d = sub(B.c, A.c);
  float dist = length(d);
  if (dist > A.h.x + B.h.x) return found;
  f3 n = mk3(0.f, 0.f, 1.f);
  if (dist > 0.f) n = divs(d, dist);
  sat.n = n;
  addFound(found, addScaled(A.c, n, A.h.x), sub(B.c, scale(n, B.h.x)), SPHERE_FEATURE);
  return found;
}

struct SphereBoxHit {
  bool hit;
  f3 onSphere, onBox, n;  // n: from the sphere towards the box
};

AVBD_D SphereBoxHit sphereBox(const Box &S, const Box &X, float4 qX) {
  SphereBoxHit out;
  out.hit = false;
  out.n = out.onBox = out.onSphere = mk3(0.f, 0.f, 0.f);
  f3 p = qrotate(qconj(qX), sub(S.c, X.c));
  f3 q = mk3(rclamp(p.x, -X.h.x, X.h.x), rclamp(p.y, -X.h.y, X.h.y), rclamp(p.z, -X.h.z, X.h.z));
  float r = S.h.x;
  f3 nLocal, onBox = q;
  if (p.x != q.x || p.y != q.y || p.z != q.z) {
    f3 d = sub(p, q);
    float dist = length(d);
    if (dist > r) return out;
    nLocal = divs(d, dist);
  } else {
    f3 depth = sub(X.h, vabs(p));
    int k = 0;
    if (depth.y < comp(depth, k)) k = 1;
    if (depth.z < comp(depth, k)) k = 2;
    float s = comp(p, k) >= 0.f ? 1.f : -1.f;
    nLocal = mk3(0.f, 0.f, 0.f);
    setComp(nLocal, k, s);
    setComp(onBox, k, s * comp(X.h, k));
  }
  f3 n = qrotate(qX, nLocal);
  out.hit = true;
  out.n = neg(n);
  out.onBox = add(X.c, qrotate(qX, onBox));
  out.onSphere = sub(S.c, scale(n, r));
  return out;
}

// --- capsules (src/gpu/capsuleCollision.wgsl:1-84; dispatch as capsuleCollision.js:17-29) ---------------------------
// Shape code -1: radius = size.x / 2, segment along local Y with half-length size.y / 2 - radius. Constants are the
// double literals rounded to float, as the CPU restatement rounds them.

struct Capsule {
  f3 a, b;
  float r;
};
struct SegmentPair {
  f3 a, b;
};

AVBD_D Capsule capsuleOf(const Body &B) {
  float radius = B.size.x * 0.5f;
  f3 axis = qrotate(B.rot, mk3(0.f, rmax(0.f, B.size.y * 0.5f - radius), 0.f));
  f3 c = xyz(B.pos);
  return Capsule{sub(c, axis), add(c, axis), radius};
}

AVBD_D f3 capSegmentPoint(f3 a, f3 b, f3 p) {
  f3 d = sub(b, a);
  float t = wclamp(dot(sub(p, a), d) / rmax(dot(d, d), float(1e-20)), 0.f, 1.f);
  return add(a, scale(d, t));
}

AVBD_D SegmentPair capSegments(f3 a, f3 b, f3 c, f3 d) {
  f3 u = sub(b, a), v = sub(d, c), w = sub(a, c);
  float uu = dot(u, u), vv = dot(v, v), uv = dot(u, v), uw = dot(u, w), vw = dot(v, w);
  float s = 0.f, t = 0.f;
  if (uu <= float(1e-16)) {
    t = wclamp(vw / rmax(vv, float(1e-20)), 0.f, 1.f);
  } else if (vv <= float(1e-16)) {
    s = wclamp(-uw / uu, 0.f, 1.f);
  } else {
    float denominator = uu * vv - uv * uv;
    if (denominator > float(1e-8) * uu * vv) s = wclamp((uv * vw - uw * vv) / denominator, 0.f, 1.f);
    t = (uv * s + vw) / vv;
    if (t < 0.f) {
      t = 0.f;
      s = wclamp(-uw / uu, 0.f, 1.f);
    } else if (t > 1.f) {
      t = 1.f;
      s = wclamp((uv - uw) / uu, 0.f, 1.f);
    }
  }
  return SegmentPair{add(a, scale(u, s)), add(c, scale(v, t))};
}

AVBD_D f3 capFallback(f3 axis) {
  if (length(axis) < float(1e-8)) return mk3(1.f, 0.f, 0.f);
  f3 n = normalize(axis);
  return normalize(fabsf(n.x) > fabsf(n.z) ? mk3(-n.y, n.x, 0.f) : mk3(0.f, -n.z, n.y));
}

AVBD_D Found capsuleSphere(const Capsule &A, f3 c, float r, Sat &sat) {
  Found found;
  found.count = 0;
  f3 p = capSegmentPoint(A.a, A.b, c);
  f3 delta = sub(c, p);
  float distance = length(delta);
  if (distance > A.r + r) return found;
  f3 n = capFallback(sub(A.b, A.a));
  if (distance > float(1e-8)) n = divs(delta, distance);
  sat.n = n;
  addFound(found, addScaled(p, n, A.r), sub(c, scale(n, r)), CAPSULE_FEATURE);
  return found;
}

AVBD_D Found capsuleCapsule(const Capsule &A, const Capsule &B, Sat &sat) {
  Found found;
  found.count = 0;
  SegmentPair points = capSegments(A.a, A.b, B.a, B.b);
  f3 delta = sub(points.b, points.a);
  float distance = length(delta);
  if (distance > A.r + B.r) return found;
  f3 n = capFallback(sub(A.b, A.a));
  f3 crossing = cross(sub(A.b, A.a), sub(B.b, B.a));
  if (distance > float(1e-8)) n = divs(delta, distance);
  else if (length(crossing) > float(1e-8)) n = normalize(crossing);
  sat.n = n;
  f3 da = sub(A.b, A.a), db = sub(B.b, B.a);
  if (length(da) > float(1e-8) && length(db) > float(1e-8) &&
      length(cross(normalize(da), normalize(db))) < float(1e-4)) {
    f3 axis = normalize(da);
    float lo = rmax(dot(A.a, axis), rmin(dot(B.a, axis), dot(B.b, axis)));
    float hi = rmin(dot(A.b, axis), rmax(dot(B.a, axis), dot(B.b, axis)));
    if (hi - lo > float(1e-5)) {
#pragma unroll
      for (uint32_t j = 0; j < 2; j++) {
        f3 p = add(A.a, scale(axis, (j == 1 ? hi : lo) - dot(A.a, axis)));
        f3 q = capSegmentPoint(B.a, B.b, p);
        addFound(found, addScaled(p, n, A.r), sub(q, scale(n, B.r)), CAPSULE_FEATURE | j);
      }
      return found;
    }
  }
  addFound(found, addScaled(points.a, n, A.r), sub(points.b, scale(n, B.r)), CAPSULE_FEATURE);
  return found;
}

/**
 * Closest points of segment ab and the box [-h, h]: the segment is cut at every face-plane crossing and the piecewise
 * quadratic distance minimised on each piece (capSegmentBox, capsuleCollision.wgsl:43-60). The breaks are kept in
 * fixed slots (2.0 = unused) and put in order by a 19-comparator sorting network, so the array stays in registers;
 * the ordered values are the same as an insertion sort's.
 */
AVBD_D SegmentPair capSegmentBox(f3 a, f3 b, f3 h) {
  f3 d = sub(b, a);
  float br[8];
  br[0] = 0.f;
  br[1] = 1.f;
  uint32_t count = 2;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float dk = comp(d, k), ak = comp(a, k), hk = comp(h, k);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      float v = 2.f;
      if (fabsf(dk) > float(1e-12)) {
        float t = ((side == 1 ? hk : -hk) - ak) / dk;
        if (t > 0.f && t < 1.f) {
          v = t;
          count++;
        }
      }
      br[2 + 2 * k + side] = v;
    }
  }
  constexpr int NET[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 1}, {2, 3},
                              {4, 5}, {6, 7}, {2, 4}, {3, 5}, {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
  for (int c = 0; c < 19; c++) {
    float x = br[NET[c][0]], y = br[NET[c][1]];
    br[NET[c][0]] = x < y ? x : y;
    br[NET[c][1]] = x < y ? y : x;
  }
  float best = float(3.4e38), bestT = 0.5f;
  SegmentPair result{mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
#pragma unroll
  for (uint32_t j = 0; j < 7; j++) {
    if (j + 1 < count) {
      float lo = br[j], hi = br[j + 1];
      float mid = 0.5f * (lo + hi);
      f3 p = add(a, scale(d, mid));
      float numerator = 0.f, denominator = 0.f;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        float pk = comp(p, k), hk = comp(h, k), dk = comp(d, k);
        if (pk < -hk || pk > hk) {
          float edge = wclamp(pk, -hk, hk);
          numerator += dk * (comp(a, k) - edge);
          denominator += dk * dk;
        }
      }
      float t = mid;
      if (denominator > float(1e-20)) t = wclamp(-numerator / denominator, lo, hi);
      f3 point = add(a, scale(d, t));
      f3 box = mk3(wclamp(point.x, -h.x, h.x), wclamp(point.y, -h.y, h.y), wclamp(point.z, -h.z, h.z));
      f3 gap = sub(point, box);
      float dist = dot(gap, gap);
      bool exactIntersection = dist == 0.f, previousIntersection = best == 0.f;
      if ((exactIntersection && !previousIntersection) || dist < best - float(1e-12) ||
          (exactIntersection == previousIntersection && fabsf(dist - best) <= float(1e-12) &&
           fabsf(t - 0.5f) < fabsf(bestT - 0.5f))) {
        best = dist;
        bestT = t;
        result = SegmentPair{point, box};
      }
    }
  }
  return result;
}

AVBD_D bool capBoxAxis(const Capsule &A, f3 h, f3 axis, float &best, f3 &normal) {
  float l = length(axis);
  if (l < float(1e-8)) return true;
  f3 n = divs(axis, l);
  float lo = rmin(dot(A.a, n), dot(A.b, n)) - A.r;
  float hi = rmax(dot(A.a, n), dot(A.b, n)) + A.r;
  float extent = dot(vabs(n), h);
  float positive = hi + extent, negative = extent - lo;
  if (positive < 0.f || negative < 0.f) return false;
  float depth = rmin(positive, negative);
  if (depth < best) {
    best = depth;
    normal = positive < negative ? n : neg(n);
  }
  return true;
}

AVBD_D Found capsuleBox(const Capsule &A, const Box &B, float4 q, Sat &sat) {
  Found found;
  found.count = 0;
  float4 inv = qconj(q);
  Capsule local{qrotate(inv, sub(A.a, B.c)), qrotate(inv, sub(A.b, B.c)), A.r};
  SegmentPair closest = capSegmentBox(local.a, local.b, B.h);
  f3 delta = sub(closest.b, closest.a);
  float distance = length(delta);
  if (distance > A.r) return found;
  f3 nLocal = mk3(0.f, 0.f, 0.f), center = closest.a, onBox = closest.b;
  if (distance > float(1e-8)) {
    nLocal = divs(delta, distance);
  } else {
    float depth = float(3.4e38);
    for (int k = 0; k < 3; k++) {
      f3 axis = mk3(0.f, 0.f, 0.f);
      setComp(axis, k, 1.f);
      if (!capBoxAxis(local, B.h, axis, depth, nLocal)) return found;
      if (!capBoxAxis(local, B.h, cross(sub(local.b, local.a), axis), depth, nLocal)) return found;
    }
    for (uint32_t k = 0; k < 8; k++) {
      f3 vertex = mk3((k & 1u) ? B.h.x : -B.h.x, (k & 2u) ? B.h.y : -B.h.y, (k & 4u) ? B.h.z : -B.h.z);
      if (!capBoxAxis(local, B.h, sub(vertex, capSegmentPoint(local.a, local.b, vertex)), depth, nLocal))
        return found;
    }
    center = dot(local.b, nLocal) > dot(local.a, nLocal) ? local.b : local.a;
    f3 point = addScaled(center, nLocal, local.r);
    onBox = sub(point, scale(nLocal, depth));
  }
  f3 n = qrotate(q, nLocal);
  sat.n = n;
  Found ends;
  ends.count = 0;
  for (uint32_t j = 0; j < 2; j++) {
    Box S;
    S.c = j == 1 ? A.b : A.a;
    S.h = mk3(A.r, A.r, A.r);
    SphereBoxHit hit = sphereBox(S, B, q);
    if (hit.hit && dot(hit.n, n) > float(0.9999)) addFound(ends, hit.onSphere, hit.onBox, CAPSULE_FEATURE | j);
  }
  if (ends.count == 2) return ends;
  addFound(found, addScaled(add(B.c, qrotate(q, center)), n, A.r), add(B.c, qrotate(q, onBox)), CAPSULE_FEATURE);
  return found;
}

}  // namespace avbd
