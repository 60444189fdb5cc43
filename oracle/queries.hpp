// ORACLE — test infrastructure only (see math.hpp).
//
// CPU restatement of the reference's shape queries (src/gpu/shapeQueries.wgsl: turn, closest, raySphere, rayCapsule,
// roundedPoly for boxes, trace, ignored; src/gpu/shapeQueries.js:99-258 cast) in FP32, in the order the HIP kernels
// evaluate them: every body is traced, the nearest non-negative distance wins and, on a tie, the smallest body index.
// Pinned in tests by an independent signed-distance ray march (sphere, capsule and box distance functions) and by
// closed forms (a ray at a sphere's centre, at an axis-aligned box face, a sphere cast at a face).
#pragma once
#include <cstring>
#include <cmath>
#include <cstdint>

namespace oracle_queries {

struct V {
  float x, y, z;
};
inline V mk(float x, float y, float z) { return V{x, y, z}; }
inline V add(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V sub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V mul(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
inline V quo(V a, float s) { return V{a.x / s, a.y / s, a.z / s}; }
inline V neg(V a) { return V{-a.x, -a.y, -a.z}; }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len(V a) { return std::sqrt(dot(a, a)); }
inline float fmaxr(float a, float b) { return a > b ? a : b; }
inline float fminr(float a, float b) { return a < b ? a : b; }
inline float axis(V v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

struct Hit {
  float t;
  V n;
};
inline Hit none() { return Hit{-1.f, mk(0, 0, 0)}; }

/** The rotation of v by the unit quaternion (ux, uy, uz, w): v + w·t + u×t with t = 2 u×v. */
inline V turn(const float q[4], V v) {
  const V u = mk(q[0], q[1], q[2]);
  const V t = mul(cross(u, v), 2.f);
  return add(add(v, mul(t, q[3])), cross(u, t));
}
inline V closest(V a, V b, V p) {
  const V e = sub(b, a);
  const float s = dot(sub(p, a), e) / fmaxr(dot(e, e), 1e-20f);
  return add(a, mul(e, fminr(fmaxr(s, 0.f), 1.f)));
}
inline Hit sphere(V o, V d, V c, float r) {
  const V v = sub(o, c);
  const float b = dot(v, d);
  const float disc = (b * b - dot(v, v)) + r * r;
  if (disc < 0.f) return none();
  const float t = -b - std::sqrt(disc);
  if (t < 0.f) return none();
  const V p = sub(add(o, mul(d, t)), c);
  const float l = len(p);
  return Hit{t, l > 1e-10f ? quo(p, fmaxr(l, 1e-20f)) : neg(d)};
}
inline Hit capsule(V o, V d, V a, V b, float r) {
  const V ba = sub(b, a), oa = sub(o, a);
  const float bb = dot(ba, ba);
  if (bb < 1e-16f) return sphere(o, d, a, r);
  const float bd = dot(ba, d), bo = dot(ba, oa), od = dot(oa, d);
  const float A = bb - bd * bd, B = bb * od - bo * bd, C = (bb * dot(oa, oa) - bo * bo) - (r * r) * bb;
  Hit hit = none();
  const float disc = B * B - A * C;
  if (A > 1e-12f * bb && disc >= 0.f) {
    const float t = (-B - std::sqrt(disc)) / A;
    const float y = bo + t * bd;
    if (t >= 0.f && y >= 0.f && y <= bb) {
      const V w = sub(add(o, mul(d, t)), add(a, quo(mul(ba, y), bb)));
      hit = Hit{t, quo(w, len(w))};
    }
  }
  for (int end = 0; end < 2; end++) {
    const Hit h = sphere(o, d, end ? b : a, r);
    if (h.t >= 0.f && (hit.t < 0.f || h.t < hit.t)) hit = h;
  }
  return hit;
}

// Box corners by bits (x, y, z), each face's corners in order, the twelve edges (shapeQueries.wgsl FACE_VERTS, EDGES)
inline V boxCorner(unsigned v, V h) { return mk(v & 1 ? h.x : -h.x, v & 2 ? h.y : -h.y, v & 4 ? h.z : -h.z); }
constexpr unsigned kFace[24] = {1, 3, 7, 5, 0, 2, 6, 4, 2, 3, 7, 6, 0, 1, 5, 4, 4, 5, 7, 6, 0, 1, 3, 2};
constexpr unsigned kEdge[24] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 2, 1, 3, 4, 6, 5, 7, 0, 4, 1, 5, 2, 6, 3, 7};
inline V faceNormal(unsigned f) {
  const float s = f & 1 ? -1.f : 1.f;
  const unsigned k = f / 2;
  return mk(k == 0 ? s : 0.f, k == 1 ? s : 0.f, k == 2 ? s : 0.f);
}
/** A box of half extents h, or a hull record at word `header` of the hull buffer (4 u32 per word). */
struct Shape {
  V h;
  const uint32_t *H = nullptr;
  unsigned header = 0;
  bool hull = false;
  unsigned u(unsigned word, int c) const { return H[4ull * word + c]; }
  float f(unsigned word, int c) const {
    float x;
    std::memcpy(&x, &H[4ull * word + c], 4);
    return x;
  }
};
inline V vertexOf(const Shape &S, unsigned v) {
  if (!S.hull) return boxCorner(v, S.h);
  const unsigned w = S.u(S.header, 0) + v;
  return mk(S.f(w, 0), S.f(w, 1), S.f(w, 2));
}
inline V planeNormal(const Shape &S, unsigned f) {
  if (!S.hull) return faceNormal(f);
  const unsigned w = S.u(S.header, 2) + 2 * f;
  return mk(S.f(w, 0), S.f(w, 1), S.f(w, 2));
}
inline float planeOffset(const Shape &S, unsigned f) {
  return S.hull ? S.f(S.u(S.header, 2) + 2 * f, 3) : axis(S.h, int(f / 2));
}
inline unsigned faceCount(const Shape &S, unsigned f) { return S.hull ? S.u(S.u(S.header, 2) + 2 * f + 1, 1) : 4u; }
inline unsigned faceVertex(const Shape &S, unsigned f, unsigned j) {
  if (!S.hull) return kFace[4 * f + j];
  const unsigned index = S.u(S.u(S.header, 2) + 2 * f + 1, 0) + j;
  return S.u(S.u(S.header + 1, 2) + index / 4, int(index % 4));
}
inline void edgeOf(const Shape &S, unsigned e, unsigned &a, unsigned &b) {
  if (!S.hull) {
    a = kEdge[2 * e];
    b = kEdge[2 * e + 1];
    return;
  }
  const unsigned w = S.u(S.header + 1, 0) + e;
  a = S.u(w, 0);
  b = S.u(w, 1);
}
inline bool insideFace(V p, unsigned f, const Shape &S) {
  const V n = planeNormal(S, f);
  const unsigned count = faceCount(S, f);
  V c = mk(0, 0, 0);
  for (unsigned j = 0; j < count; j++) c = add(c, vertexOf(S, faceVertex(S, f, j)));
  c = quo(c, float(count));
  for (unsigned j = 0; j < count; j++) {
    const V a = vertexOf(S, faceVertex(S, f, j)), b = vertexOf(S, faceVertex(S, f, (j + 1) % count));
    const V side = cross(sub(b, a), n);
    if (dot(side, sub(p, a)) * (dot(side, sub(c, a)) >= 0.f ? 1.f : -1.f) < -1e-6f) return false;
  }
  return true;
}
/** roundedPoly (shapeQueries.wgsl:54-76) for a box or a hull. */
inline Hit poly(V o, V d, const Shape &S, float r) {
  const unsigned nf = S.hull ? S.u(S.header, 3) : 6u, ne = S.hull ? S.u(S.header + 1, 1) : 12u;
  if (r == 0.f) {
    float enter = -3.4e38f, exit = 3.4e38f;
    V normal = neg(d);
    for (unsigned f = 0; f < nf; f++) {
      const V n = planeNormal(S, f);
      const float w = planeOffset(S, f);
      const float denom = dot(n, d), gap = w - dot(n, o);
      if (std::fabs(denom) < 1e-12f) {
        if (gap < 0.f) return none();
      } else {
        const float t = gap / denom;
        if (denom < 0.f) {
          if (t > enter) {
            enter = t;
            normal = n;
          }
        } else {
          exit = fminr(exit, t);
        }
      }
      if (enter > exit) return none();
    }
    if (exit < 0.f) return none();
    if (enter <= 0.f) return Hit{0.f, neg(d)};
    return Hit{enter, normal};
  }
  bool inside = true;
  float distance = 3.4e38f;
  Hit hit = none();
  for (unsigned f = 0; f < nf; f++) {
    const V n = planeNormal(S, f);
    const float w = planeOffset(S, f);
    const float s = dot(n, o) - w;
    if (s > 0.f) inside = false;
    if (insideFace(sub(o, mul(n, s)), f, S)) distance = fminr(distance, s * s);
    const float denom = dot(n, d);
    if (denom < -1e-12f) {
      const float t = ((w + r) - dot(n, o)) / denom;
      if (t >= 0.f && (hit.t < 0.f || t < hit.t) && insideFace(sub(add(o, mul(d, t)), mul(n, r)), f, S))
        hit = Hit{t, n};
    }
  }
  for (unsigned e = 0; e < ne; e++) {
    unsigned ia, ib;
    edgeOf(S, e, ia, ib);
    const V a = vertexOf(S, ia), b = vertexOf(S, ib);
    const V q = sub(o, closest(a, b, o));
    distance = fminr(distance, dot(q, q));
    const Hit c = capsule(o, d, a, b, r);
    if (c.t >= 0.f && (hit.t < 0.f || c.t < hit.t)) hit = c;
  }
  if (inside || distance <= r * r) return Hit{0.f, neg(d)};
  return hit;
}

/** One query (origin, max distance, unit direction, radius) against one body record (40 floats); hull bodies read
 * their record from the hull buffer H. */
inline Hit trace(const float *rec, const float *q, const uint32_t *H = nullptr) {
  const float rot[4] = {rec[4], rec[5], rec[6], rec[7]};
  const float inv[4] = {-rec[4], -rec[5], -rec[6], rec[7]};
  const V o = turn(inv, sub(mk(q[0], q[1], q[2]), mk(rec[0], rec[1], rec[2])));
  const V d = turn(inv, mk(q[4], q[5], q[6]));
  const float r = q[7];
  const V size = mk(rec[16], rec[17], rec[18]);
  const float shape = rec[39];
  Hit hit;
  if (shape == 1.f) {
    const float radius = size.x * 0.5f + r;
    hit = dot(o, o) <= radius * radius ? Hit{0.f, neg(d)} : sphere(o, d, mk(0, 0, 0), radius);
  } else if (shape == -1.f) {
    const float half = fmaxr(0.f, size.y * 0.5f - size.x * 0.5f);
    const V a = mk(0, -half, 0), b = mk(0, half, 0);
    const float radius = size.x * 0.5f + r;
    const V delta = sub(o, closest(a, b, o));
    hit = dot(delta, delta) <= radius * radius ? Hit{0.f, neg(d)} : capsule(o, d, a, b, radius);
  } else {
    Shape S;
    S.h = mul(size, 0.5f);
    if (shape >= 3.f && H) {
      S.hull = true;
      S.H = H;
      S.header = unsigned(fmaxr(0.f, shape - 3.f) + 0.5f);
    }
    hit = poly(o, d, S, r);
  }
  if (hit.t > q[3]) return none();
  hit.n = turn(rot, hit.n);
  return hit;
}

}  // namespace oracle_queries
