// ORACLE — test infrastructure only (see math.hpp).
//
// Box/box narrowphase: separating-axis test over 3 + 3 face axes and 9 edge axes, incident-face clipping against the
// reference face's side planes (Sutherland–Hodgman) and the edge/edge closest-points contact, with the feature keys
// the warm start matches on. Restates reference/three-avbd/src/avbd3d/ref/collide.ts:1-345 (the upstream demo's
// collide.cpp) and its GPU twin reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:317-679 (which adds the
// face-bias rule, FLAG_FACE_BIAS, and holds 8 clipped vertices instead of 16). Also the GPU-only sphere contacts
// (wgsl-collision.ts:536-585) and the sphere/capsule feature constants.
#pragma once
#include "math.hpp"

namespace oracle {

constexpr int MAX_CONTACTS = 8;
constexpr double SAT_AXIS_EPSILON = 1.0e-6;
constexpr double PLANE_EPSILON = 1.0e-5;
constexpr double CONTACT_MERGE_DIST_SQ = 1.0e-6;
constexpr uint32_t AXIS_FACE_A = 0, AXIS_FACE_B = 1, AXIS_EDGE = 2;
constexpr uint32_t SPHERE_FEATURE = 3u << 24;
constexpr uint32_t CAPSULE_FEATURE = 4u << 24;

template <class R> struct Box {
  V3<R> c, h;
  V3<R> ax[3];
};

template <class R> inline Box<R> makeBox(V3<R> pos, Q4<R> rot, V3<R> size) {
  Box<R> b;
  b.c = pos;
  b.h = scale(size, R(0.5));
  b.ax[0] = rotate(rot, V3<R>(1, 0, 0));
  b.ax[1] = rotate(rot, V3<R>(0, 1, 0));
  b.ax[2] = rotate(rot, V3<R>(0, 0, 1));
  return b;
}

template <class R> struct Sat {
  uint32_t kind = 0, ia = 0, ib = 0;
  R sep = 0;
  V3<R> n;  // from A towards B
  bool valid = false;
};

/** Up to 8 contact points as world points on A and B with their feature keys (wgsl-collision.ts:341-346). */
template <class R> struct Found {
  uint32_t feature[MAX_CONTACTS];
  V3<R> xA[MAX_CONTACTS], xB[MAX_CONTACTS];
  int count = 0;
};

/** ContactBuilder.add (collide.ts:152-166) = addFound (wgsl-collision.ts:373-384). */
template <class R> inline void addFound(Found<R> &f, V3<R> xA, V3<R> xB, uint32_t feature) {
  V3<R> mid = scale(add(xA, xB), R(0.5));
  for (int i = 0; i < f.count; i++) {
    V3<R> d = sub(mid, scale(add(f.xA[i], f.xB[i]), R(0.5)));
    if (dot(d, d) < R(CONTACT_MERGE_DIST_SQ)) return;
  }
  if (f.count >= MAX_CONTACTS) return;
  f.feature[f.count] = feature;
  f.xA[f.count] = xA;
  f.xB[f.count] = xB;
  f.count++;
}

template <class R> inline R projRadius(const Box<R> &b, V3<R> n) {
  return b.h.x * std::fabs(dot(n, b.ax[0])) + b.h.y * std::fabs(dot(n, b.ax[1])) + b.h.z * std::fabs(dot(n, b.ax[2]));
}

/**
 * collide.ts:169-190 / wgsl-collision.ts:362-371. The WGSL uses inverseSqrt; both sides of this build use the
 * correctly rounded 1/sqrt(lenSq) instead, so the GPU and the FP32 mirror pick the same axis.
 */
template <class R>
inline bool testAxis(const Box<R> &A, const Box<R> &B, V3<R> delta, V3<R> axis, uint32_t kind, uint32_t ia, uint32_t ib,
                     Sat<R> &best) {
  R lenSq = lengthSq(axis);
  if (lenSq < R(SAT_AXIS_EPSILON)) return true;
  R invLen = R(1) / std::sqrt(lenSq);
  V3<R> n = scale(axis, invLen);
  if (dot(n, delta) < 0) n = neg(n);
  R sep = std::fabs(dot(delta, n)) - (projRadius(A, n) + projRadius(B, n));
  if (sep > 0) return false;
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

/** collide.ts:81-88 / wgsl-collision.ts:373-376 (support). */
template <class R> inline V3<R> support(const Box<R> &b, V3<R> dir) {
  R sx = dot(dir, b.ax[0]) >= 0 ? R(1) : R(-1);
  R sy = dot(dir, b.ax[1]) >= 0 ? R(1) : R(-1);
  R sz = dot(dir, b.ax[2]) >= 0 ? R(1) : R(-1);
  V3<R> p = addScaled(b.c, b.ax[0], b.h.x * sx);
  p = addScaled(p, b.ax[1], b.h.y * sy);
  return addScaled(p, b.ax[2], b.h.z * sz);
}

/** Sutherland–Hodgman against dot(pn, x) <= offset (collide.ts:124-145 / wgsl-collision.ts:391-410). */
template <class R>
inline int clipPlane(const V3<R> *src, int n, V3<R> *dst, int maxPoly, V3<R> pn, R offset) {
  if (n == 0) return 0;
  int count = 0;
  V3<R> a = src[n - 1];
  R da = dot(pn, a) - offset;
  for (int i = 0; i < n; i++) {
    V3<R> b = src[i];
    R db = dot(pn, b) - offset;
    bool aIn = da <= R(PLANE_EPSILON);
    bool bIn = db <= R(PLANE_EPSILON);
    if (aIn != bIn) {
      R t = 0;
      R denom = da - db;
      if (std::fabs(denom) > R(SAT_AXIS_EPSILON)) t = rclamp(da / denom, R(0), R(1));
      if (count < maxPoly) dst[count++] = addScaled(a, sub(b, a), t);
    }
    if (bIn && count < maxPoly) dst[count++] = b;
    a = b;
    da = db;
  }
  return count;
}

inline int faceU(int k) { return k == 0 ? 1 : 0; }
inline int faceV(int k) { return k != 2 ? 2 : 1; }

/** Reference-face clipping manifold (collide.ts:233-273 / wgsl-collision.ts:416-468). */
template <class R>
inline Found<R> faceManifold(const Box<R> &A, const Box<R> &B, bool refIsA, int refAxis, V3<R> nAB, int maxPoly) {
  Found<R> found;
  const Box<R> &Rb = refIsA ? A : B;
  const Box<R> &Ib = refIsA ? B : A;
  V3<R> outward = refIsA ? nAB : neg(nAB);
  R rs = dot(outward, Rb.ax[refAxis]) >= 0 ? R(1) : R(-1);
  V3<R> rn = scale(Rb.ax[refAxis], rs);
  V3<R> rc = addScaled(Rb.c, rn, Rb.h[refAxis]);
  V3<R> ru = Rb.ax[faceU(refAxis)], rv = Rb.ax[faceV(refAxis)];
  R eu = Rb.h[faceU(refAxis)], ev = Rb.h[faceV(refAxis)];

  int incAxis = 0;
  R bestDot = -R(3.4028234663852886e38);
  for (int k = 0; k < 3; k++) {
    R d = std::fabs(dot(Ib.ax[k], rn));
    if (d > bestDot) {
      bestDot = d;
      incAxis = k;
    }
  }
  R is = dot(Ib.ax[incAxis], rn) > 0 ? R(-1) : R(1);
  V3<R> fn = scale(Ib.ax[incAxis], is);
  V3<R> ic = addScaled(Ib.c, fn, Ib.h[incAxis]);
  V3<R> u = Ib.ax[faceU(incAxis)], v = Ib.ax[faceV(incAxis)];
  R eU = Ib.h[faceU(incAxis)], eV = Ib.h[faceV(incAxis)];
  auto vert = [&](R su, R sv) { return addScaled(addScaled(ic, u, su * eU), v, sv * eV); };
  V3<R> poly[16], tmp[16];
  poly[0] = vert(1, 1);
  poly[1] = vert(-1, 1);
  poly[2] = vert(-1, -1);
  poly[3] = vert(1, -1);
  int n = 4;
  n = clipPlane(poly, n, tmp, maxPoly, ru, dot(ru, rc) + eu);
  if (!n) return found;
  n = clipPlane(tmp, n, poly, maxPoly, neg(ru), dot(neg(ru), rc) + eu);
  if (!n) return found;
  n = clipPlane(poly, n, tmp, maxPoly, rv, dot(rv, rc) + ev);
  if (!n) return found;
  n = clipPlane(tmp, n, poly, maxPoly, neg(rv), dot(neg(rv), rc) + ev);
  if (!n) return found;

  uint32_t prefix = ((refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24) | (uint32_t(refAxis) << 16) | (uint32_t(incAxis) << 8);
  for (int i = 0; i < n && found.count < MAX_CONTACTS; i++) {
    V3<R> p = poly[i];
    R dist = dot(sub(p, rc), rn);
    if (dist > R(PLANE_EPSILON)) continue;
    V3<R> onRef = addScaled(p, rn, -dist);
    addFound(found, refIsA ? onRef : p, refIsA ? p : onRef, prefix | uint32_t(i));
  }
  if (found.count == 0) addFound(found, support(A, nAB), support(B, neg(nAB)), prefix);
  return found;
}

/** collide.ts:192-202 / wgsl-collision.ts:470-478. */
template <class R> inline void supportEdge(const Box<R> &b, int k, V3<R> dir, V3<R> &e0, V3<R> &e1) {
  int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
  R s1 = dot(dir, b.ax[k1]) >= 0 ? R(1) : R(-1);
  R s2 = dot(dir, b.ax[k2]) >= 0 ? R(1) : R(-1);
  V3<R> c = addScaled(b.c, b.ax[k1], b.h[k1] * s1);
  c = addScaled(c, b.ax[k2], b.h[k2] * s2);
  V3<R> e = scale(b.ax[k], b.h[k]);
  e0 = sub(c, e);
  e1 = add(c, e);
}

/** Closest points between segments (collide.ts:204-231 / wgsl-collision.ts:480-516). */
template <class R> inline void closestSegments(V3<R> p0, V3<R> p1, V3<R> q0, V3<R> q1, V3<R> &xA, V3<R> &xB) {
  V3<R> d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
  R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  R s = 0, t = 0;
  const R eps = R(SAT_AXIS_EPSILON);
  if (a <= eps && e <= eps) {
  } else if (a <= eps) {
    t = rclamp(f / e, R(0), R(1));
  } else {
    R c = dot(d1, r);
    if (e <= eps) {
      s = rclamp(-c / a, R(0), R(1));
    } else {
      R b = dot(d1, d2);
      R denom = a * e - b * b;
      if (std::fabs(denom) > eps) s = rclamp((b * f - c * e) / denom, R(0), R(1));
      t = (b * s + f) / e;
      if (t < 0) {
        t = 0;
        s = rclamp(-c / a, R(0), R(1));
      } else if (t > 1) {
        t = 1;
        s = rclamp((b - c) / a, R(0), R(1));
      }
    }
  }
  xA = addScaled(p0, d1, s);
  xB = addScaled(q0, d2, t);
}

template <class R> inline Found<R> edgeContact(const Box<R> &A, const Box<R> &B, int ia, int ib, V3<R> nAB) {
  Found<R> found;
  V3<R> a0, a1, b0, b1, xA, xB;
  supportEdge(A, ia, nAB, a0, a1);
  supportEdge(B, ib, neg(nAB), b0, b1);
  closestSegments(a0, a1, b0, b1, xA, xB);
  uint32_t key = (AXIS_EDGE << 24) | (uint32_t(ia) << 8) | uint32_t(ib);
  addFound(found, xA, xB, key);
  if (found.count == 0) addFound(found, support(A, nAB), support(B, neg(nAB)), key);
  return found;
}

/**
 * Collide boxes A and B (collide.ts:305-345 / wgsl-collision.ts:593-626). Returns the contacts; `sat.n` is the
 * normal from A towards B (the manifold normal, B to A, is its negation). faceBias: the GPU's FLAG_FACE_BIAS rule
 * (edge wins when edge.sep > 0.95·face.sep + 0.01); otherwise the demo's 0.95·edge.sep > face.sep + 0.01.
 */
template <class R>
inline Found<R> collideBoxes(const Box<R> &A, const Box<R> &B, bool faceBias, int maxPoly, Sat<R> &sat) {
  Found<R> none;
  V3<R> delta = sub(B.c, A.c);
  Sat<R> face, edge;
  for (int i = 0; i < 3; i++)
    if (!testAxis(A, B, delta, A.ax[i], AXIS_FACE_A, uint32_t(i), 0u, face)) return none;
  for (int i = 0; i < 3; i++)
    if (!testAxis(A, B, delta, B.ax[i], AXIS_FACE_B, 0u, uint32_t(i), face)) return none;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (!testAxis(A, B, delta, cross(A.ax[i], B.ax[j]), AXIS_EDGE, uint32_t(i), uint32_t(j), edge)) return none;
  if (!face.valid) return none;
  Sat<R> best = face;
  if (edge.valid) {
    bool edgeWins = faceBias ? (edge.sep > R(0.95) * face.sep + R(0.01)) : (R(0.95) * edge.sep > face.sep + R(0.01));
    if (edgeWins) best = edge;
  }
  sat = best;
  if (best.kind == AXIS_EDGE) return edgeContact(A, B, int(best.ia), int(best.ib), best.n);
  bool refIsA = best.kind == AXIS_FACE_A;
  return faceManifold(A, B, refIsA, int(refIsA ? best.ia : best.ib), best.n, maxPoly);
}

// --- Spheres (GPU-only extension; wgsl-collision.ts:536-585) ---------------------------------------------------

template <class R> inline Found<R> sphereSphere(const Box<R> &A, const Box<R> &B, Sat<R> &sat) {
  Found<R> found;
  V3<R> d = sub(B.c, A.c);
  R dist = length(d);
  if (dist > A.h.x + B.h.x) return found;
  V3<R> n(0, 0, 1);
  if (dist > 0) n = divs(d, dist);
  sat.n = n;
  addFound(found, addScaled(A.c, n, A.h.x), sub(B.c, scale(n, B.h.x)), SPHERE_FEATURE);
  return found;
}

template <class R> struct SphereBoxHit {
  bool hit = false;
  V3<R> onSphere, onBox, n;  // n: from the sphere towards the box
};

/** Sphere S against box X (wgsl-collision.ts:555-585). */
template <class R> inline SphereBoxHit<R> sphereBox(const Box<R> &S, const Box<R> &X, Q4<R> qX) {
  SphereBoxHit<R> out;
  V3<R> p = rotate(conj(qX), sub(S.c, X.c));
  V3<R> q(rclamp(p.x, -X.h.x, X.h.x), rclamp(p.y, -X.h.y, X.h.y), rclamp(p.z, -X.h.z, X.h.z));
  R r = S.h.x;
  V3<R> nLocal, onBox = q;
  if (p.x != q.x || p.y != q.y || p.z != q.z) {
    V3<R> d = sub(p, q);
    R dist = length(d);
    if (dist > r) return out;
    nLocal = divs(d, dist);
  } else {
    V3<R> depth = sub(X.h, vabs(p));
    int k = 0;
    if (depth.y < depth[k]) k = 1;
    if (depth.z < depth[k]) k = 2;
    R s = p[k] >= 0 ? R(1) : R(-1);
    nLocal = V3<R>(0, 0, 0);
    nLocal[k] = s;
    onBox[k] = s * X.h[k];
  }
  V3<R> n = rotate(qX, nLocal);
  out.hit = true;
  out.n = neg(n);
  out.onBox = add(X.c, rotate(qX, onBox));
  out.onSphere = sub(S.c, scale(n, r));
  return out;
}

// --- Capsules (analytic; src/gpu/capsuleCollision.wgsl:1-84, spliced into the narrowphase by
// src/gpu/capsuleCollision.js:2-32). Shape code -1: radius = size.x / 2, segment along local Y with half-length
// size.y / 2 - radius. Feature keys: CAPSULE_FEATURE, or CAPSULE_FEATURE | j for two-point parallel contacts. --------

/** WGSL clamp(x, lo, hi) = min(max(x, lo), hi). */
template <class R> inline R wclamp(R x, R lo, R hi) { return rmin(rmax(x, lo), hi); }

template <class R> struct Capsule {
  V3<R> a, b;
  R r;
};
template <class R> struct SegmentPair {
  V3<R> a, b;
};

/** capsuleOf (capsuleCollision.wgsl:6-9). */
template <class R> inline Capsule<R> capsuleOf(V3<R> pos, Q4<R> rot, V3<R> size) {
  R radius = size.x * R(0.5);
  V3<R> axis = rotate(rot, V3<R>(0, rmax(R(0), size.y * R(0.5) - radius), 0));
  return {sub(pos, axis), add(pos, axis), radius};
}

/** Closest point of segment ab to p (capSegmentPoint, capsuleCollision.wgsl:10). */
template <class R> inline V3<R> capSegmentPoint(V3<R> a, V3<R> b, V3<R> p) {
  V3<R> d = sub(b, a);
  R t = wclamp(dot(sub(p, a), d) / rmax(dot(d, d), R(1e-20)), R(0), R(1));
  return add(a, scale(d, t));
}

/** Closest points of segments ab and cd (capSegments, capsuleCollision.wgsl:11-18). */
template <class R> inline SegmentPair<R> capSegments(V3<R> a, V3<R> b, V3<R> c, V3<R> d) {
  V3<R> u = sub(b, a), v = sub(d, c), w = sub(a, c);
  R uu = dot(u, u), vv = dot(v, v), uv = dot(u, v), uw = dot(u, w), vw = dot(v, w);
  R s = 0, t = 0;
  if (uu <= R(1e-16)) {
    t = wclamp(vw / rmax(vv, R(1e-20)), R(0), R(1));
  } else if (vv <= R(1e-16)) {
    s = wclamp(-uw / uu, R(0), R(1));
  } else {
    R denominator = uu * vv - uv * uv;
    if (denominator > R(1e-8) * uu * vv) s = wclamp((uv * vw - uw * vv) / denominator, R(0), R(1));
    t = (uv * s + vw) / vv;
    if (t < R(0)) {
      t = 0;
      s = wclamp(-uw / uu, R(0), R(1));
    } else if (t > R(1)) {
      t = 1;
      s = wclamp((uv - uw) / uu, R(0), R(1));
    }
  }
  return {add(a, scale(u, s)), add(c, scale(v, t))};
}

/** A normal perpendicular to the axis for coincident segments (capFallback, capsuleCollision.wgsl:19-22). */
template <class R> inline V3<R> capFallback(V3<R> axis) {
  if (length(axis) < R(1e-8)) return V3<R>(1, 0, 0);
  V3<R> n = normalize(axis);
  return normalize(std::fabs(n.x) > std::fabs(n.z) ? V3<R>(-n.y, n.x, 0) : V3<R>(0, -n.z, n.y));
}

/** Capsule A against sphere (c, r); n from A towards the sphere (capsuleSphere, capsuleCollision.wgsl:23-27). */
template <class R> inline Found<R> capsuleSphere(const Capsule<R> &A, V3<R> c, R r, Sat<R> &sat) {
  Found<R> found;
  V3<R> p = capSegmentPoint(A.a, A.b, c);
  V3<R> delta = sub(c, p);
  R distance = length(delta);
  if (distance > A.r + r) return found;
  V3<R> n = capFallback(sub(A.b, A.a));
  if (distance > R(1e-8)) n = divs(delta, distance);
  sat.n = n;
  addFound(found, addScaled(p, n, A.r), sub(c, scale(n, r)), CAPSULE_FEATURE);
  return found;
}

/** Capsule against capsule; two points when the axes are parallel and overlap (capsuleCollision.wgsl:28-40). */
template <class R> inline Found<R> capsuleCapsule(const Capsule<R> &A, const Capsule<R> &B, Sat<R> &sat) {
  Found<R> found;
  SegmentPair<R> points = capSegments(A.a, A.b, B.a, B.b);
  V3<R> delta = sub(points.b, points.a);
  R distance = length(delta);
  if (distance > A.r + B.r) return found;
  V3<R> n = capFallback(sub(A.b, A.a));
  V3<R> crossing = cross(sub(A.b, A.a), sub(B.b, B.a));
  if (distance > R(1e-8)) n = divs(delta, distance);
  else if (length(crossing) > R(1e-8)) n = normalize(crossing);
  sat.n = n;
  V3<R> da = sub(A.b, A.a), db = sub(B.b, B.a);
  if (length(da) > R(1e-8) && length(db) > R(1e-8) && length(cross(normalize(da), normalize(db))) < R(1e-4)) {
    V3<R> axis = normalize(da);
    R lo = rmax(dot(A.a, axis), rmin(dot(B.a, axis), dot(B.b, axis)));
    R hi = rmin(dot(A.b, axis), rmax(dot(B.a, axis), dot(B.b, axis)));
    if (hi - lo > R(1e-5)) {
      for (uint32_t j = 0; j < 2; j++) {
        V3<R> p = add(A.a, scale(axis, (j == 1 ? hi : lo) - dot(A.a, axis)));
        V3<R> q = capSegmentPoint(B.a, B.b, p);
        addFound(found, addScaled(p, n, A.r), sub(q, scale(n, B.r)), CAPSULE_FEATURE | j);
      }
      return found;
    }
  }
  addFound(found, addScaled(points.a, n, A.r), sub(points.b, scale(n, B.r)), CAPSULE_FEATURE);
  return found;
}

/**
 * Closest points of segment ab and the box [-h, h] (capSegmentBox, capsuleCollision.wgsl:43-60): the segment is cut
 * at every face-plane crossing and the piecewise quadratic distance is minimised on each piece; an exact intersection
 * wins over an exterior sample, ties go to the parameter nearest the middle.
 */
template <class R> inline SegmentPair<R> capSegmentBox(V3<R> a, V3<R> b, V3<R> h) {
  V3<R> d = sub(b, a);
  R breaks[8];
  breaks[0] = 0;
  breaks[1] = 1;
  uint32_t count = 2;
  for (int k = 0; k < 3; k++) {
    if (std::fabs(d[k]) > R(1e-12)) {
      for (int side = 0; side < 2; side++) {
        R t = ((side == 1 ? h[k] : -h[k]) - a[k]) / d[k];
        if (t > R(0) && t < R(1)) breaks[count++] = t;
      }
    }
  }
  for (uint32_t i = 1; i < count; i++) {
    R value = breaks[i];
    uint32_t j = i;
    while (j > 0 && breaks[j - 1] > value) {
      breaks[j] = breaks[j - 1];
      j--;
    }
    breaks[j] = value;
  }
  R best = R(3.4e38), bestT = R(0.5);
  SegmentPair<R> result{V3<R>(0, 0, 0), V3<R>(0, 0, 0)};
  for (uint32_t j = 0; j + 1 < count; j++) {
    R lo = breaks[j], hi = breaks[j + 1];
    R mid = R(0.5) * (lo + hi);
    V3<R> p = add(a, scale(d, mid));
    R numerator = 0, denominator = 0;
    for (int k = 0; k < 3; k++) {
      if (p[k] < -h[k] || p[k] > h[k]) {
        R edge = wclamp(p[k], -h[k], h[k]);
        numerator += d[k] * (a[k] - edge);
        denominator += d[k] * d[k];
      }
    }
    R t = mid;
    if (denominator > R(1e-20)) t = wclamp(-numerator / denominator, lo, hi);
    V3<R> point = add(a, scale(d, t));
    V3<R> box(wclamp(point.x, -h.x, h.x), wclamp(point.y, -h.y, h.y), wclamp(point.z, -h.z, h.z));
    V3<R> gap = sub(point, box);
    R dist = dot(gap, gap);
    bool exactIntersection = dist == R(0), previousIntersection = best == R(0);
    if ((exactIntersection && !previousIntersection) || dist < best - R(1e-12) ||
        (exactIntersection == previousIntersection && std::fabs(dist - best) <= R(1e-12) &&
         std::fabs(t - R(0.5)) < std::fabs(bestT - R(0.5)))) {
      best = dist;
      bestT = t;
      result = {point, box};
    }
  }
  return result;
}

/** One SAT axis of a capsule against the box [-h, h] (capBoxAxis, capsuleCollision.wgsl:61-66). */
template <class R> inline bool capBoxAxis(const Capsule<R> &A, V3<R> h, V3<R> axis, R &best, V3<R> &normal) {
  R l = length(axis);
  if (l < R(1e-8)) return true;
  V3<R> n = divs(axis, l);
  R lo = rmin(dot(A.a, n), dot(A.b, n)) - A.r;
  R hi = rmax(dot(A.a, n), dot(A.b, n)) + A.r;
  R extent = dot(vabs(n), h);
  R positive = hi + extent, negative = extent - lo;
  if (positive < R(0) || negative < R(0)) return false;
  R depth = rmin(positive, negative);
  if (depth < best) {
    best = depth;
    normal = positive < negative ? n : neg(n);
  }
  return true;
}

/**
 * Capsule A against box B with rotation q; n from A towards B (capsuleBox, capsuleCollision.wgsl:67-84). A segment
 * inside the box takes the shallowest of the box faces, segment × face axes and vertex axes; end caps both in
 * parallel support give a two-point manifold.
 */
template <class R> inline Found<R> capsuleBox(const Capsule<R> &A, const Box<R> &B, Q4<R> q, Sat<R> &sat) {
  Found<R> found;
  Q4<R> inv = conj(q);
  Capsule<R> local{rotate(inv, sub(A.a, B.c)), rotate(inv, sub(A.b, B.c)), A.r};
  SegmentPair<R> closest = capSegmentBox(local.a, local.b, B.h);
  V3<R> delta = sub(closest.b, closest.a);
  R distance = length(delta);
  if (distance > A.r) return found;
  V3<R> nLocal(0, 0, 0), center = closest.a, onBox = closest.b;
  if (distance > R(1e-8)) {
    nLocal = divs(delta, distance);
  } else {
    R depth = R(3.4e38);
    for (int k = 0; k < 3; k++) {
      V3<R> axis(0, 0, 0);
      axis[k] = 1;
      if (!capBoxAxis(local, B.h, axis, depth, nLocal)) return found;
      if (!capBoxAxis(local, B.h, cross(sub(local.b, local.a), axis), depth, nLocal)) return found;
    }
    for (uint32_t k = 0; k < 8; k++) {
      V3<R> vertex((k & 1u) ? B.h.x : -B.h.x, (k & 2u) ? B.h.y : -B.h.y, (k & 4u) ? B.h.z : -B.h.z);
      if (!capBoxAxis(local, B.h, sub(vertex, capSegmentPoint(local.a, local.b, vertex)), depth, nLocal))
        return found;
    }
    center = dot(local.b, nLocal) > dot(local.a, nLocal) ? local.b : local.a;
    V3<R> point = addScaled(center, nLocal, local.r);
    onBox = sub(point, scale(nLocal, depth));
  }
  V3<R> n = rotate(q, nLocal);
  sat.n = n;
  Found<R> ends;
  for (uint32_t j = 0; j < 2; j++) {
    Box<R> S;
    S.c = j == 1 ? A.b : A.a;
    S.h = V3<R>(A.r, A.r, A.r);
    SphereBoxHit<R> hit = sphereBox(S, B, q);
    if (hit.hit && dot(hit.n, n) > R(0.9999)) addFound(ends, hit.onSphere, hit.onBox, CAPSULE_FEATURE | j);
  }
  if (ends.count == 2) return ends;
  addFound(found, addScaled(add(B.c, rotate(q, center)), n, A.r), add(B.c, rotate(q, onBox)), CAPSULE_FEATURE);
  return found;
}

}  // namespace oracle
