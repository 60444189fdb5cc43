// Device narrowphase geometry: box/box separating axes (3 + 3 faces, 9 edges), incident-face clipping with feature
// keys, the edge/edge closest-points contact, and the sphere contacts. Follows
// reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:317-626 (a port of ref/collide.ts) with FLAG_FACE_BIAS;
// testAxis takes 1/sqrtf (correctly rounded) where the WGSL has inverseSqrt, so the axis choice is reproducible.
#pragma once
#include "avbd_device.h"

namespace avbd {

constexpr uint32_t AXIS_FACE_A = 0, AXIS_FACE_B = 1, AXIS_EDGE = 2;

struct Box {
  f3 c, h;
  f3 ax[3];
};

AVBD_D Box makeBox(f3 pos, float4 rot, f3 size) {
  Box b;
  b.c = pos;
  b.h = scale(size, 0.5f);
  b.ax[0] = qrotate(rot, mk3(1.f, 0.f, 0.f));
  b.ax[1] = qrotate(rot, mk3(0.f, 1.f, 0.f));
  b.ax[2] = qrotate(rot, mk3(0.f, 0.f, 1.f));
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

AVBD_D void addFound(Found &f, f3 xA, f3 xB, uint32_t feature) {
  f3 mid = scale(add(xA, xB), 0.5f);
  for (int i = 0; i < f.count; i++) {
    f3 d = sub(mid, scale(add(f.xA[i], f.xB[i]), 0.5f));
    if (dot(d, d) < CONTACT_MERGE_DIST_SQ) return;
  }
  if (f.count >= MAX_CONTACTS) return;
  f.feature[f.count] = feature;
  f.xA[f.count] = xA;
  f.xB[f.count] = xB;
  f.count++;
}

AVBD_D float projRadius(const Box &b, f3 n) {
  return b.h.x * fabsf(dot(n, b.ax[0])) + b.h.y * fabsf(dot(n, b.ax[1])) + b.h.z * fabsf(dot(n, b.ax[2]));
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
  float sx = dot(dir, b.ax[0]) >= 0.f ? 1.f : -1.f;
  float sy = dot(dir, b.ax[1]) >= 0.f ? 1.f : -1.f;
  float sz = dot(dir, b.ax[2]) >= 0.f ? 1.f : -1.f;
  f3 p = addScaled(b.c, b.ax[0], b.h.x * sx);
  p = addScaled(p, b.ax[1], b.h.y * sy);
  return addScaled(p, b.ax[2], b.h.z * sz);
}

AVBD_D int clipPlane(const f3 *src, int n, f3 *dst, f3 pn, float offset) {
  if (n == 0) return 0;
  int count = 0;
  f3 a = src[n - 1];
  float da = dot(pn, a) - offset;
  for (int i = 0; i < n; i++) {
    f3 b = src[i];
    float db = dot(pn, b) - offset;
    bool aIn = da <= PLANE_EPSILON;
    bool bIn = db <= PLANE_EPSILON;
    if (aIn != bIn) {
      float t = 0.f;
      float denom = da - db;
      if (fabsf(denom) > SAT_AXIS_EPSILON) t = rclamp(da / denom, 0.f, 1.f);
      if (count < MAX_POLY) dst[count++] = addScaled(a, sub(b, a), t);
    }
    if (bIn && count < MAX_POLY) dst[count++] = b;
    a = b;
    da = db;
  }
  return count;
}

AVBD_D int faceU(int k) { return k == 0 ? 1 : 0; }
AVBD_D int faceV(int k) { return k != 2 ? 2 : 1; }
AVBD_D f3 boxAxis(const Box &b, int k) { return k == 0 ? b.ax[0] : (k == 1 ? b.ax[1] : b.ax[2]); }

AVBD_D Found faceManifold(const Box &A, const Box &B, bool refIsA, int refAxis, f3 nAB) {
  Found found;
  found.count = 0;
  const Box &R = refIsA ? A : B;
  const Box &I = refIsA ? B : A;
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
  f3 poly[MAX_POLY], tmp[MAX_POLY];
  poly[0] = addScaled(addScaled(ic, u, 1.f * eU), v, 1.f * eV);
  poly[1] = addScaled(addScaled(ic, u, -1.f * eU), v, 1.f * eV);
  poly[2] = addScaled(addScaled(ic, u, -1.f * eU), v, -1.f * eV);
  poly[3] = addScaled(addScaled(ic, u, 1.f * eU), v, -1.f * eV);
  int n = 4;
  n = clipPlane(poly, n, tmp, ru, dot(ru, rc) + eu);
  n = clipPlane(tmp, n, poly, neg(ru), dot(neg(ru), rc) + eu);
  n = clipPlane(poly, n, tmp, rv, dot(rv, rc) + ev);
  n = clipPlane(tmp, n, poly, neg(rv), dot(neg(rv), rc) + ev);
  if (n == 0) return found;

  uint32_t prefix = ((refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24) | (uint32_t(refAxis) << 16) | (uint32_t(incAxis) << 8);
  for (int i = 0; i < n && found.count < MAX_CONTACTS; i++) {
    f3 p = poly[i];
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
  for (int i = 0; i < 3; i++)
    if (!testAxis(A, B, delta, A.ax[i], AXIS_FACE_A, uint32_t(i), 0u, face)) return none;
  for (int i = 0; i < 3; i++)
    if (!testAxis(A, B, delta, B.ax[i], AXIS_FACE_B, 0u, uint32_t(i), face)) return none;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (!testAxis(A, B, delta, cross(A.ax[i], B.ax[j]), AXIS_EDGE, uint32_t(i), uint32_t(j), edge)) return none;
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

}  // namespace avbd
