// ORACLE — test infrastructure only (see math.hpp).
//
// Convex hull contacts, sequentially, as the reference writes them: reference/three-avbd/src/avbd3d/gpu/wgsl-hull.ts
// (Poly :41-150, queryFaces :154-163, minkowskiFace :166-175, queryEdges :188-218, clipHull :221-240, hullFaceContact
// :243-325, hullEdgeContact :327-350, collidePoly :353-377, segmentPoint :380-383, sphereHull :386-441) and capsulePoly
// (src/gpu/capsuleCollision.js:35-67), over the hull buffer layout of wgsl-hull.ts:13-19 (4 u32 per word). One loop per
// pair, in the reference's order, so "first best" choices are the loops' own.
#pragma once
#include <cstring>
#include <vector>

#include "collide.hpp"

namespace oracle {

constexpr float SHAPE_HULL_F = 3.0f;
constexpr uint32_t NO_HULL = 0xffffffffu;
constexpr uint32_t MAX_CLIP = 64;

/** The hull buffer: 4 u32 per 16-byte word. */
struct HullBuffer {
  std::vector<uint32_t> w;
  uint32_t u(uint32_t word, int c) const { return w[4ull * word + c]; }
  float f(uint32_t word, int c) const {
    float x;
    std::memcpy(&x, &w[4ull * word + c], 4);
    return x;
  }
};

inline const uint32_t BOX_FACE_VERTS[24] = {1, 3, 7, 5, 0, 2, 6, 4, 2, 3, 7, 6, 0, 1, 5, 4, 4, 5, 7, 6, 0, 1, 3, 2};
inline const uint32_t BOX_EDGES[12][4] = {{0, 1, 3, 5}, {2, 3, 2, 5}, {4, 5, 3, 4}, {6, 7, 2, 4},
                                          {0, 2, 1, 5}, {1, 3, 0, 5}, {4, 6, 1, 4}, {5, 7, 0, 4},
                                          {0, 4, 1, 3}, {1, 5, 0, 3}, {2, 6, 1, 2}, {3, 7, 0, 2}};

template <class R> struct Poly {
  V3<R> c;
  Q4<R> q;
  V3<R> h;
  uint32_t hull = NO_HULL, vs = 0, nv = 8, fs = 0, nf = 6, es = 0, ne = 12, is = 0;
};

/** polyOf (wgsl-hull.ts:49-66): a box, or the hull whose header the shape code names. */
template <class R>
inline Poly<R> polyOf(const HullBuffer &H, V3<R> pos, Q4<R> rot, V3<R> size, R shape) {
  Poly<R> P;
  P.c = pos;
  P.q = rot;
  P.h = scale(size, R(0.5));
  if (shape >= R(SHAPE_HULL_F)) {
    const uint32_t h = uint32_t(float(shape) - SHAPE_HULL_F + 0.5f);
    P.hull = h;
    P.vs = H.u(h, 0);
    P.nv = H.u(h, 1);
    P.fs = H.u(h, 2);
    P.nf = H.u(h, 3);
    P.es = H.u(h + 1, 0);
    P.ne = H.u(h + 1, 1);
    P.is = H.u(h + 1, 2);
  }
  return P;
}

template <class R> inline V3<R> hvec(const HullBuffer &H, uint32_t word) {
  return V3<R>(R(H.f(word, 0)), R(H.f(word, 1)), R(H.f(word, 2)));
}

template <class R> inline V3<R> vertLocal(const HullBuffer &H, const Poly<R> &P, uint32_t v) {
  if (P.hull == NO_HULL) return V3<R>((v & 1u) ? P.h.x : -P.h.x, (v & 2u) ? P.h.y : -P.h.y, (v & 4u) ? P.h.z : -P.h.z);
  return hvec<R>(H, P.vs + v);
}
template <class R> inline V3<R> vert(const HullBuffer &H, const Poly<R> &P, uint32_t v) {
  return add(P.c, rotate(P.q, vertLocal(H, P, v)));
}

/** Face f's world plane: outward normal and offset (wgsl-hull.ts:83-102). */
template <class R> inline void facePlane(const HullBuffer &H, const Poly<R> &P, uint32_t f, V3<R> &n, R &w) {
  V3<R> nl;
  R d;
  if (P.hull == NO_HULL) {
    const int k = int(f / 2u);
    nl = V3<R>(0, 0, 0);
    nl[k] = (f & 1u) ? R(-1) : R(1);
    d = P.h[k];
  } else {
    nl = hvec<R>(H, P.fs + 2u * f);
    d = R(H.f(P.fs + 2u * f, 3));
  }
  n = rotate(P.q, nl);
  w = d + dot(n, P.c);
}
template <class R> inline V3<R> faceNormal(const HullBuffer &H, const Poly<R> &P, uint32_t f) {
  V3<R> n;
  R w;
  facePlane(H, P, f, n, w);
  return n;
}

template <class R> inline uint32_t faceVertCount(const HullBuffer &H, const Poly<R> &P, uint32_t f) {
  return P.hull == NO_HULL ? 4u : H.u(P.fs + 2u * f + 1u, 1);
}
template <class R> inline uint32_t faceVert(const HullBuffer &H, const Poly<R> &P, uint32_t f, uint32_t j) {
  if (P.hull == NO_HULL) return BOX_FACE_VERTS[f * 4u + j];
  const uint32_t k = H.u(P.fs + 2u * f + 1u, 0) + j;
  return H.u(P.is + k / 4u, int(k % 4u));
}
template <class R> inline void edgeOf(const HullBuffer &H, const Poly<R> &P, uint32_t e, uint32_t out[4]) {
  for (int c = 0; c < 4; c++) out[c] = P.hull == NO_HULL ? BOX_EDGES[e][c] : H.u(P.es + e, c);
}

/** polySupport (wgsl-hull.ts:139-150). */
template <class R> inline V3<R> polySupport(const HullBuffer &H, const Poly<R> &P, V3<R> dir) {
  const V3<R> l = rotate(conj(P.q), dir);
  if (P.hull == NO_HULL)
    return add(P.c, rotate(P.q, V3<R>(l.x >= 0 ? P.h.x : -P.h.x, l.y >= 0 ? P.h.y : -P.h.y, l.z >= 0 ? P.h.z : -P.h.z)));
  R best = R(-3.4e38f);
  V3<R> at(0, 0, 0);
  for (uint32_t v = 0; v < P.nv; v++) {
    const V3<R> p = hvec<R>(H, P.vs + v);
    const R s = dot(p, l);
    if (s > best) {
      best = s;
      at = p;
    }
  }
  return add(P.c, rotate(P.q, at));
}

template <class R> struct FaceQuery {
  R sep;
  uint32_t face;
};

template <class R> inline FaceQuery<R> queryFaces(const HullBuffer &H, const Poly<R> &A, const Poly<R> &B) {
  FaceQuery<R> out{R(-3.4e38f), 0u};
  for (uint32_t f = 0; f < A.nf; f++) {
    V3<R> n;
    R w;
    facePlane(H, A, f, n, w);
    const R s = dot(n, polySupport(H, B, neg(n))) - w;
    if (s > out.sep) out = {s, f};
  }
  return out;
}

template <class R> inline bool minkowskiFace(V3<R> a, V3<R> b, V3<R> c, V3<R> d) {
  const V3<R> bxa = cross(b, a), dxc = cross(d, c);
  const R cba = dot(c, bxa), dba = dot(d, bxa), adc = dot(a, dxc), bdc = dot(b, dxc);
  return cba * dba < 0 && adc * bdc < 0 && cba * bdc > 0;
}

template <class R> inline V3<R> faceNormalLocal(const HullBuffer &H, const Poly<R> &P, uint32_t f) {
  if (P.hull == NO_HULL) {
    V3<R> n(0, 0, 0);
    n[int(f / 2u)] = (f & 1u) ? R(-1) : R(1);
    return n;
  }
  return hvec<R>(H, P.fs + 2u * f);
}

template <class R> struct EdgeQuery {
  R sep = R(-3.4e38f);
  uint32_t ea = 0, eb = 0;
  V3<R> n;
  bool valid = false;
};

template <class R> inline EdgeQuery<R> queryEdges(const HullBuffer &H, const Poly<R> &A, const Poly<R> &B) {
  EdgeQuery<R> out;
  const Q4<R> toB = qmul(conj(B.q), A.q);
  for (uint32_t i = 0; i < A.ne; i++) {
    uint32_t ea[4];
    edgeOf(H, A, i, ea);
    const V3<R> a = rotate(toB, faceNormalLocal(H, A, ea[2]));
    const V3<R> b = rotate(toB, faceNormalLocal(H, A, ea[3]));
    for (uint32_t j = 0; j < B.ne; j++) {
      uint32_t eb[4];
      edgeOf(H, B, j, eb);
      if (!minkowskiFace(a, b, neg(faceNormalLocal(H, B, eb[2])), neg(faceNormalLocal(H, B, eb[3])))) continue;
      const V3<R> pa = vert(H, A, ea[0]), qa = vert(H, A, ea[1]);
      const V3<R> pb = vert(H, B, eb[0]), qb = vert(H, B, eb[1]);
      V3<R> n = cross(sub(qa, pa), sub(qb, pb));
      const R l = length(n);
      if (l < R(1e-5f) * length(sub(qa, pa)) * length(sub(qb, pb))) continue;
      n = divs(n, l);
      if (dot(n, sub(pa, A.c)) < 0) n = neg(n);
      const R s = dot(n, sub(pb, pa));
      if (s > out.sep) {
        out.sep = s;
        out.ea = i;
        out.eb = j;
        out.n = n;
        out.valid = true;
      }
    }
  }
  return out;
}

/** clipHull (wgsl-hull.ts:221-240): src (n points) clipped to dot(pn, x) <= offset. */
template <class R>
inline uint32_t clipHull(const V3<R> *src, uint32_t n, V3<R> *dst, V3<R> pn, R offset) {
  if (n == 0) return 0;
  uint32_t count = 0;
  V3<R> a = src[n - 1];
  R da = dot(pn, a) - offset;
  for (uint32_t i = 0; i < n; i++) {
    const V3<R> b = src[i];
    const R db = dot(pn, b) - offset;
    if ((da <= R(PLANE_EPSILON)) != (db <= R(PLANE_EPSILON))) {
      R t = 0;
      if (std::fabs(da - db) > R(SAT_AXIS_EPSILON)) t = wclamp(da / (da - db), R(0), R(1));
      if (count < MAX_CLIP) dst[count++] = add(a, scale(sub(b, a), t));
    }
    if (db <= R(PLANE_EPSILON) && count < MAX_CLIP) dst[count++] = b;
    a = b;
    da = db;
  }
  return count;
}

template <class R>
inline Found<R> hullFaceContact(const HullBuffer &H, const Poly<R> &A, const Poly<R> &B, bool refIsA, uint32_t refFace) {
  Found<R> found;
  const Poly<R> &Rf = refIsA ? A : B;
  const Poly<R> &I = refIsA ? B : A;
  V3<R> rn;
  R pw;
  facePlane(H, Rf, refFace, rn, pw);
  uint32_t deepest = 0;
  R low = R(3.4e38f);
  for (uint32_t v = 0; v < I.nv; v++) {
    const R d = dot(vert(H, I, v), rn);
    if (d < low) {
      low = d;
      deepest = v;
    }
  }
  uint32_t inc = 0;
  R least = R(3.4e38f);
  for (uint32_t f = 0; f < I.nf; f++) {
    bool touches = false;
    for (uint32_t j = 0; j < faceVertCount(H, I, f); j++)
      if (faceVert(H, I, f, j) == deepest) {
        touches = true;
        break;
      }
    if (!touches) continue;
    const R d = dot(faceNormal(H, I, f), rn);
    if (d < least) {
      least = d;
      inc = f;
    }
  }
  V3<R> poly[MAX_CLIP], tmp[MAX_CLIP];
  uint32_t n = std::min(faceVertCount(H, I, inc), MAX_CLIP);
  for (uint32_t j = 0; j < n; j++) poly[j] = vert(H, I, faceVert(H, I, inc, j));
  const uint32_t k = faceVertCount(H, Rf, refFace);
  V3<R> centre(0, 0, 0);
  for (uint32_t j = 0; j < k; j++) centre = add(centre, vert(H, Rf, faceVert(H, Rf, refFace, j)));
  centre = divs(centre, R(float(k)));
  for (uint32_t j = 0; j < k && n > 0; j++) {
    const V3<R> v0 = vert(H, Rf, faceVert(H, Rf, refFace, j));
    const V3<R> v1 = vert(H, Rf, faceVert(H, Rf, refFace, (j + 1u) % k));
    V3<R> s = cross(sub(v1, v0), rn);
    const R l = length(s);
    if (l < R(SAT_AXIS_EPSILON)) continue;
    s = divs(s, l);
    if (dot(s, sub(centre, v0)) > 0) s = neg(s);
    n = clipHull(poly, n, tmp, s, dot(s, v0));
    for (uint32_t m = 0; m < n; m++) poly[m] = tmp[m];
  }
  R depth[MAX_CLIP];
  uint32_t index[MAX_CLIP];
  uint32_t keep = 0;
  for (uint32_t m = 0; m < n; m++) {
    const R dist = dot(rn, poly[m]) - pw;
    if (dist <= R(PLANE_EPSILON)) {
      tmp[keep] = poly[m];
      depth[keep] = dist;
      index[keep] = m;
      keep++;
    }
  }
  const uint32_t prefix = ((refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24) | (std::min(refFace, 255u) << 16) |
                          (std::min(inc, 255u) << 8);
  bool chosen[MAX_CLIP] = {false};
  for (uint32_t c = 0; c < std::min<uint32_t>(keep, MAX_CONTACTS); c++) {
    uint32_t pick = 0;
    R bestScore = R(-3.4e38f);
    for (uint32_t m = 0; m < keep; m++) {
      if (chosen[m]) continue;
      R score = -depth[m];
      if (c > 0) {
        score = R(3.4e38f);
        for (uint32_t o = 0; o < keep; o++)
          if (chosen[o]) score = rmin(score, dot(sub(tmp[m], tmp[o]), sub(tmp[m], tmp[o])));
      }
      if (score > bestScore) {
        bestScore = score;
        pick = m;
      }
    }
    chosen[pick] = true;
    const V3<R> p = tmp[pick];
    const V3<R> onRef = sub(p, scale(rn, depth[pick]));
    addFound(found, refIsA ? onRef : p, refIsA ? p : onRef, prefix | std::min(index[pick], 255u));
  }
  if (found.count == 0) {
    const V3<R> nAB = refIsA ? rn : neg(rn);
    addFound(found, polySupport(H, A, nAB), polySupport(H, B, neg(nAB)), prefix);
  }
  return found;
}

template <class R>
inline Found<R> hullEdgeContact(const HullBuffer &H, const Poly<R> &A, const Poly<R> &B, const EdgeQuery<R> &q) {
  Found<R> found;
  uint32_t ea[4], eb[4];
  edgeOf(H, A, q.ea, ea);
  edgeOf(H, B, q.eb, eb);
  const V3<R> p0 = vert(H, A, ea[0]), q0 = vert(H, B, eb[0]);
  const V3<R> d1 = sub(vert(H, A, ea[1]), p0), d2 = sub(vert(H, B, eb[1]), q0);
  const V3<R> r = sub(p0, q0);
  const R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), c = dot(d1, r), b = dot(d1, d2);
  const R denom = a * e - b * b;
  R s = 0;
  if (std::fabs(denom) > R(SAT_AXIS_EPSILON) * a * e) s = wclamp((b * f - c * e) / denom, R(0), R(1));
  R t = (b * s + f) / rmax(e, R(SAT_AXIS_EPSILON));
  if (t < 0) {
    t = 0;
    s = wclamp(-c / rmax(a, R(SAT_AXIS_EPSILON)), R(0), R(1));
  } else if (t > 1) {
    t = 1;
    s = wclamp((b - c) / rmax(a, R(SAT_AXIS_EPSILON)), R(0), R(1));
  }
  addFound(found, add(p0, scale(d1, s)), add(q0, scale(d2, t)),
           (AXIS_EDGE << 24) | (std::min(q.ea, 255u) << 8) | std::min(q.eb, 255u));
  return found;
}

/** The three queries of collidePoly, for checks of the axes. */
template <class R> struct PolyQueries {
  FaceQuery<R> fa, fb;
  EdgeQuery<R> eq;
};

template <class R>
inline Found<R> collidePoly(const HullBuffer &H, const Poly<R> &A, const Poly<R> &B, bool faceBias, Sat<R> &sat,
                            PolyQueries<R> *queries = nullptr) {
  Found<R> none;
  const FaceQuery<R> fa = queryFaces(H, A, B);
  if (queries) queries->fa = fa;
  if (fa.sep > 0) return none;
  const FaceQuery<R> fb = queryFaces(H, B, A);
  if (queries) queries->fb = fb;
  if (fb.sep > 0) return none;
  const EdgeQuery<R> eq = queryEdges(H, A, B);
  if (queries) queries->eq = eq;
  if (eq.valid && eq.sep > 0) return none;
  const R faceSep = rmax(fa.sep, fb.sep);
  bool edgeWins = eq.valid && eq.sep > R(0.95f) * faceSep + R(0.01f);
  if (!faceBias) edgeWins = eq.valid && R(0.95f) * eq.sep > faceSep + R(0.01f);
  sat.valid = true;
  if (edgeWins) {
    sat.kind = AXIS_EDGE;
    sat.ia = eq.ea;
    sat.ib = eq.eb;
    sat.sep = eq.sep;
    sat.n = eq.n;
    return hullEdgeContact(H, A, B, eq);
  }
  if (fb.sep > fa.sep) {
    sat.kind = AXIS_FACE_B;
    sat.ia = 0;
    sat.ib = fb.face;
    sat.sep = fb.sep;
    sat.n = neg(faceNormal(H, B, fb.face));
    return hullFaceContact(H, A, B, false, fb.face);
  }
  sat.kind = AXIS_FACE_A;
  sat.ia = fa.face;
  sat.ib = 0;
  sat.sep = fa.sep;
  sat.n = faceNormal(H, A, fa.face);
  return hullFaceContact(H, A, B, true, fa.face);
}

template <class R> inline V3<R> segmentPoint(V3<R> a, V3<R> b, V3<R> p) {
  const V3<R> d = sub(b, a);
  return add(a, scale(d, wclamp(dot(sub(p, a), d) / rmax(dot(d, d), R(1e-20f)), R(0), R(1))));
}

template <class R> inline SphereBoxHit<R> sphereHull(const HullBuffer &H, V3<R> c, R r, const Poly<R> &X) {
  SphereBoxHit<R> out;
  R deepest = R(-3.4e38f);
  uint32_t deepFace = 0;
  for (uint32_t f = 0; f < X.nf; f++) {
    V3<R> n;
    R w;
    facePlane(H, X, f, n, w);
    const R s = dot(n, c) - w;
    if (s > deepest) {
      deepest = s;
      deepFace = f;
    }
  }
  if (deepest > r) return out;
  if (deepest <= 0) {
    const V3<R> n = faceNormal(H, X, deepFace);
    out.hit = true;
    out.n = neg(n);
    out.onBox = sub(c, scale(n, deepest));
    out.onSphere = sub(c, scale(n, r));
    return out;
  }
  V3<R> best(0, 0, 0);
  R bestDist = R(3.4e38f);
  for (uint32_t f = 0; f < X.nf; f++) {
    V3<R> pn;
    R pw;
    facePlane(H, X, f, pn, pw);
    const R s = dot(pn, c) - pw;
    if (s <= 0) continue;
    V3<R> p = sub(c, scale(pn, s));
    const uint32_t k = faceVertCount(H, X, f);
    V3<R> centre(0, 0, 0);
    for (uint32_t j = 0; j < k; j++) centre = add(centre, vert(H, X, faceVert(H, X, f, j)));
    centre = divs(centre, R(float(k)));
    bool inside = true;
    V3<R> edgeBest(0, 0, 0);
    R edgeDist = R(3.4e38f);
    for (uint32_t j = 0; j < k; j++) {
      const V3<R> v0 = vert(H, X, faceVert(H, X, f, j));
      const V3<R> v1 = vert(H, X, faceVert(H, X, f, (j + 1u) % k));
      V3<R> side = cross(sub(v1, v0), pn);
      if (dot(side, sub(centre, v0)) > 0) side = neg(side);
      if (dot(side, sub(p, v0)) > 0) inside = false;
      const V3<R> q = segmentPoint(v0, v1, c);
      const R dq = dot(sub(q, c), sub(q, c));
      if (dq < edgeDist) {
        edgeDist = dq;
        edgeBest = q;
      }
    }
    if (!inside) p = edgeBest;
    const R dist = length(sub(p, c));
    if (dist < bestDist) {
      bestDist = dist;
      best = p;
    }
  }
  if (bestDist > r) return out;
  V3<R> n(0, 0, 1);
  if (bestDist > 0) n = divs(sub(c, best), bestDist);
  out.hit = true;
  out.n = neg(n);
  out.onBox = best;
  out.onSphere = sub(c, scale(n, r));
  return out;
}

template <class R>
inline bool capPolyAxis(const HullBuffer &H, const Capsule<R> &A, const Poly<R> &P, V3<R> axis, R &best, V3<R> &normal) {
  const R l = length(axis);
  if (l < R(1e-8f)) return true;
  const V3<R> n = divs(axis, l);
  const R lo = rmin(dot(A.a, n), dot(A.b, n)) - A.r;
  const R hi = rmax(dot(A.a, n), dot(A.b, n)) + A.r;
  const R low = dot(polySupport(H, P, neg(n)), n);
  const R high = dot(polySupport(H, P, n), n);
  const R positive = hi - low, negative = high - lo;
  if (positive < 0 || negative < 0) return false;
  const R depth = rmin(positive, negative);
  if (depth < best) {
    best = depth;
    normal = positive < negative ? n : neg(n);
  }
  return true;
}

template <class R> inline Found<R> capsulePoly(const HullBuffer &H, const Capsule<R> &A, const Poly<R> &P, Sat<R> &sat) {
  Found<R> found;
  R best = R(3.4e38f);
  SegmentPair<R> points{V3<R>(0, 0, 0), V3<R>(0, 0, 0)};
  const V3<R> direction = sub(A.b, A.a);
  R insideLo = 0, insideHi = 1;
  for (uint32_t f = 0; f < P.nf; f++) {
    V3<R> n;
    R pw;
    facePlane(H, P, f, n, pw);
    const R s = dot(n, A.a) - pw, slope = dot(n, direction);
    if (std::fabs(slope) < R(1e-12f)) {
      if (s > 0) insideLo = 2;
    } else if (slope > 0) {
      insideHi = rmin(insideHi, -s / slope);
    } else {
      insideLo = rmax(insideLo, -s / slope);
    }
    const uint32_t count = faceVertCount(H, P, f);
    V3<R> center(0, 0, 0);
    for (uint32_t j = 0; j < count; j++) center = add(center, vert(H, P, faceVert(H, P, f, j)));
    center = divs(center, R(float(count)));
    R lo = 0, hi = 1;
    const V3<R> projectedA = sub(A.a, scale(n, s)), projectedD = sub(direction, scale(n, slope));
    for (uint32_t j = 0; j < count; j++) {
      const V3<R> a = vert(H, P, faceVert(H, P, f, j));
      const V3<R> b = vert(H, P, faceVert(H, P, f, (j + 1u) % count));
      V3<R> side = cross(sub(b, a), n);
      if (dot(side, sub(center, a)) > 0) side = neg(side);
      const R edgeS = dot(side, sub(projectedA, a)), edgeD = dot(side, projectedD);
      if (std::fabs(edgeD) < R(1e-12f)) {
        if (edgeS > R(1e-7f)) lo = 2;
      } else if (edgeD > 0) {
        hi = rmin(hi, -edgeS / edgeD);
      } else {
        lo = rmax(lo, -edgeS / edgeD);
      }
      const SegmentPair<R> pair = capSegments(A.a, A.b, a, b);
      const R dist = dot(sub(pair.a, pair.b), sub(pair.a, pair.b));
      if (dist < best) {
        best = dist;
        points = pair;
      }
    }
    if (lo <= hi) {
      R t = R(0.5f) * (lo + hi);
      if (std::fabs(slope) > R(1e-12f)) t = wclamp(-s / slope, lo, hi);
      const V3<R> p = add(A.a, scale(direction, t));
      const V3<R> q = sub(p, scale(n, s + t * slope));
      const R dist = dot(sub(p, q), sub(p, q));
      if (dist < best) {
        best = dist;
        points = {p, q};
      }
    }
  }
  const bool inside = insideLo <= insideHi;
  const R distance = std::sqrt(best);
  if (!inside && distance > A.r) return found;
  V3<R> n(0, 0, 0);
  if (!inside && distance > R(1e-8f)) {
    n = divs(sub(points.b, points.a), distance);
  } else {
    R depth = R(3.4e38f);
    for (uint32_t f = 0; f < P.nf; f++)
      if (!capPolyAxis(H, A, P, faceNormal(H, P, f), depth, n)) return found;
    for (uint32_t e = 0; e < P.ne; e++) {
      uint32_t edge[4];
      edgeOf(H, P, e, edge);
      if (!capPolyAxis(H, A, P, cross(direction, sub(vert(H, P, edge[1]), vert(H, P, edge[0]))), depth, n)) return found;
    }
    for (uint32_t v = 0; v < P.nv; v++) {
      const V3<R> vertex = vert(H, P, v);
      if (!capPolyAxis(H, A, P, sub(vertex, capSegmentPoint(A.a, A.b, vertex)), depth, n)) return found;
    }
    points.a = dot(A.b, n) > dot(A.a, n) ? A.b : A.a;
    points.b = add(points.a, scale(n, A.r - depth));
  }
  sat.n = n;
  addFound(found, add(points.a, scale(n, A.r)), points.b, CAPSULE_FEATURE);
  return found;
}

}  // namespace oracle
