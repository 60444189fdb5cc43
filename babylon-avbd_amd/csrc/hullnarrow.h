// Convex hull contacts on the device, one 64-lane wavefront per pair: hull/hull, hull/box, sphere/hull and
// capsule/hull. Follows reference/three-avbd/src/avbd3d/gpu/wgsl-hull.ts (Gregorius's SAT over face normals and the
// edge pairs whose Gauss-map arcs cross; reference-face clipping of the incident face; edge/edge closest points; the
// sphere's closest surface point) and capsulePoly (src/gpu/capsuleCollision.js:35-67). The reference runs a pair in one
// thread; here the lanes share it: faces, edge pairs, vertices and clip-polygon points are spread over the lanes and
// every "first best" choice of the sequential loops becomes a wave reduction that keeps the larger value and, on a tie,
// the smaller loop index, which picks the same element. Sums whose order matters (face centres) run in loop order on
// every lane. Every value a lane holds after a reduction is the same on all 64 lanes, so control flow stays uniform
// and the block (one wavefront) can use __syncthreads around its LDS exchanges.
#pragma once
#include "narrowphase.h"

namespace avbd {

constexpr uint32_t NO_HULL = 0xffffffffu, NO_INDEX = 0xffffffffu;
constexpr uint32_t MAX_CLIP = 64;
constexpr uint32_t HULL_BLOCK = 64;  // one wavefront per block: a pair per block at a time
constexpr float BIGF = 3.4e38f;

AVBD_D bool isHullShape(float s) { return s >= SHAPE_HULL; }

// Box topology (wgsl-hull.ts:33-40): vertex v has signs (bit 0: x, 1: y, 2: z); faces +x -x +y -y +z -z as loops;
// edges as vertex a, vertex b and the faces on each side.
__constant__ uint32_t BOX_FACE_VERTS[24] = {1, 3, 7, 5, 0, 2, 6, 4, 2, 3, 7, 6, 0, 1, 5, 4, 4, 5, 7, 6, 0, 1, 3, 2};
__constant__ uint32_t BOX_EDGES[48] = {0, 1, 3, 5, 2, 3, 2, 5, 4, 5, 3, 4, 6, 7, 2, 4, 0, 2, 1, 5, 1, 3, 0, 5,
                                       4, 6, 1, 4, 5, 7, 0, 4, 0, 4, 1, 3, 1, 5, 0, 3, 2, 6, 1, 2, 3, 7, 0, 2};

/** A convex shape as the hull code sees it: a box (hull = NO_HULL, half extents h) or a hull record. */
struct Poly {
  f3 c;
  float4 q;
  f3 h;
  uint32_t hull, vs, nv, fs, nf, es, ne, is;
};

AVBD_D Poly polyOf(const DeviceWorld &w, uint32_t i) {
  const Body &b = w.bodies[i];
  Poly P;
  P.c = xyz(b.pos);
  P.q = b.rot;
  P.h = scale(xyz(b.size), 0.5f);
  P.hull = NO_HULL;
  P.vs = P.fs = P.es = P.is = 0;
  P.nv = 8;
  P.nf = 6;
  P.ne = 12;
  if (isHullShape(b.angVel.w)) {
    const uint32_t h = uint32_t(b.angVel.w - SHAPE_HULL + 0.5f);
    const uint4 h0 = w.hulls[h], h1 = w.hulls[h + 1];
    P.hull = h;
    P.vs = h0.x;
    P.nv = h0.y;
    P.fs = h0.z;
    P.nf = h0.w;
    P.es = h1.x;
    P.ne = h1.y;
    P.is = h1.z;
  }
  return P;
}

AVBD_D f3 hullVec(const DeviceWorld &w, uint32_t word) {
  const uint4 u = w.hulls[word];
  return mk3(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z));
}

AVBD_D f3 vertLocal(const DeviceWorld &w, const Poly &P, uint32_t v) {
  if (P.hull == NO_HULL)
    return mk3((v & 1u) ? P.h.x : -P.h.x, (v & 2u) ? P.h.y : -P.h.y, (v & 4u) ? P.h.z : -P.h.z);
  return hullVec(w, P.vs + v);
}
AVBD_D f3 polyVert(const DeviceWorld &w, const Poly &P, uint32_t v) { return add(P.c, qrotate(P.q, vertLocal(w, P, v))); }

/** Face f's outward normal in the shape's own frame. */
AVBD_D f3 faceNormalLocal(const DeviceWorld &w, const Poly &P, uint32_t f) {
  if (P.hull == NO_HULL) {
    f3 n = mk3(0.f, 0.f, 0.f);
    setComp(n, int(f / 2u), (f & 1u) ? -1.f : 1.f);
    return n;
  }
  return hullVec(w, P.fs + 2u * f);
}

/** Face f's world plane: outward normal (xyz) and offset (w). */
AVBD_D float4 facePlane(const DeviceWorld &w, const Poly &P, uint32_t f) {
  f3 nl;
  float d;
  if (P.hull == NO_HULL) {
    const int k = int(f / 2u);
    nl = mk3(0.f, 0.f, 0.f);
    setComp(nl, k, (f & 1u) ? -1.f : 1.f);
    d = comp(P.h, k);
  } else {
    const uint4 u = w.hulls[P.fs + 2u * f];
    nl = mk3(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z));
    d = __uint_as_float(u.w);
  }
  const f3 n = qrotate(P.q, nl);
  return mk4(n, d + dot(n, P.c));
}

AVBD_D uint32_t faceVertCount(const DeviceWorld &w, const Poly &P, uint32_t f) {
  return P.hull == NO_HULL ? 4u : w.hulls[P.fs + 2u * f + 1u].y;
}
AVBD_D uint32_t faceVert(const DeviceWorld &w, const Poly &P, uint32_t f, uint32_t j) {
  if (P.hull == NO_HULL) return BOX_FACE_VERTS[f * 4u + j];
  const uint32_t k = w.hulls[P.fs + 2u * f + 1u].x + j;
  const uint4 u = w.hulls[P.is + k / 4u];
  const uint32_t r = k % 4u;
  return r == 0 ? u.x : (r == 1 ? u.y : (r == 2 ? u.z : u.w));
}
AVBD_D uint4 edgeOf(const DeviceWorld &w, const Poly &P, uint32_t e) {
  if (P.hull == NO_HULL) return make_uint4(BOX_EDGES[4 * e], BOX_EDGES[4 * e + 1], BOX_EDGES[4 * e + 2], BOX_EDGES[4 * e + 3]);
  return w.hulls[P.es + e];
}

/** The shape's farthest point along dir (the first vertex of the largest projection). */
AVBD_D f3 polySupport(const DeviceWorld &w, const Poly &P, f3 dir) {
  const f3 l = qrotate(qconj(P.q), dir);
  if (P.hull == NO_HULL)
    return add(P.c, qrotate(P.q, mk3(l.x >= 0.f ? P.h.x : -P.h.x, l.y >= 0.f ? P.h.y : -P.h.y,
                                     l.z >= 0.f ? P.h.z : -P.h.z)));
  float best = -BIGF;
  f3 at = mk3(0.f, 0.f, 0.f);
  for (uint32_t v = 0; v < P.nv; v++) {
    const f3 p = hullVec(w, P.vs + v);
    const float s = dot(p, l);
    if (s > best) {
      best = s;
      at = p;
    }
  }
  return add(P.c, qrotate(P.q, at));
}

// --- wave helpers ---------------------------------------------------------------------------------------------------

AVBD_D uint32_t laneId() { return threadIdx.x & 63u; }

/** (v, i) with the largest v over the wave, the smallest i among equal v; every lane gets it. */
AVBD_D void waveArgMax(float &v, uint32_t &i) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const float ov = __shfl_xor(v, d, 64);
    const uint32_t oi = __shfl_xor(i, d, 64);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

/** Exclusive prefix sum of c over the wave; total on every lane. */
AVBD_D uint32_t waveExclusive(uint32_t c, uint32_t &total) {
  const uint32_t lane = laneId();
  uint32_t incl = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= uint32_t(d)) incl += o;
  }
  total = __shfl(incl, 63, 64);
  return incl - c;
}

AVBD_D f3 shfl3(f3 v, uint32_t src) {
  return mk3(__shfl(v.x, int(src), 64), __shfl(v.y, int(src), 64), __shfl(v.z, int(src), 64));
}

/** LDS of one wavefront: a clip polygon's scatter target and the kept points. */
struct HullScratch {
  float x[MAX_CLIP], y[MAX_CLIP], z[MAX_CLIP];
};

// --- separating axes (wgsl-hull.ts:152-226) ---------------------------------------------------------------------------

struct FaceQuery {
  float sep;
  uint32_t face;
};

/** The face of A whose plane B sticks out of least (largest separation); lanes take faces. */
AVBD_D FaceQuery queryFaces(const DeviceWorld &w, const Poly &A, const Poly &B) {
  float best = -BIGF;
  uint32_t face = NO_INDEX;
  for (uint32_t f = laneId(); f < A.nf; f += 64) {
    const float4 p = facePlane(w, A, f);
    const f3 n = xyz(p);
    const float s = dot(n, polySupport(w, B, neg(n))) - p.w;
    if (s > best) {
      best = s;
      face = f;
    }
  }
  waveArgMax(best, face);
  return FaceQuery{best, face == NO_INDEX ? 0u : face};
}

/** The edges' Gauss-map arcs (a-b on A's, c-d on B's negated) cross: their cross product is a candidate axis. */
AVBD_D bool minkowskiFace(f3 a, f3 b, f3 c, f3 d) {
  const f3 bxa = cross(b, a), dxc = cross(d, c);
  const float cba = dot(c, bxa), dba = dot(d, bxa), adc = dot(a, dxc), bdc = dot(b, dxc);
  return cba * dba < 0.f && adc * bdc < 0.f && cba * bdc > 0.f;
}

struct EdgeQuery {
  float sep;
  uint32_t ea, eb;
  f3 n;
  bool valid;
};

/** One edge pair's axis and separation (valid false: not a Minkowski face, or parallel edges). */
AVBD_D bool edgeAxis(const DeviceWorld &w, const Poly &A, const Poly &B, float4 toB, uint32_t i, uint32_t j, f3 &n,
                     float &s) {
  const uint4 ea = edgeOf(w, A, i);
  const f3 a = qrotate(toB, faceNormalLocal(w, A, ea.z));
  const f3 b = qrotate(toB, faceNormalLocal(w, A, ea.w));
  const uint4 eb = edgeOf(w, B, j);
  if (!minkowskiFace(a, b, neg(faceNormalLocal(w, B, eb.z)), neg(faceNormalLocal(w, B, eb.w)))) return false;
  const f3 pa = polyVert(w, A, ea.x), qa = polyVert(w, A, ea.y);
  const f3 pb = polyVert(w, B, eb.x), qb = polyVert(w, B, eb.y);
  n = cross(sub(qa, pa), sub(qb, pb));
  const float l = length(n);
  // Parallel edges: their faces' normals already cover the axis
  if (l < 1e-5f * length(sub(qa, pa)) * length(sub(qb, pb))) return false;
  n = divs(n, l);
  if (dot(n, sub(pa, A.c)) < 0.f) n = neg(n);
  s = dot(n, sub(pb, pa));
  return true;
}

/** The best edge-pair axis; lanes take the pairs in (A edge, B edge) order. */
AVBD_D EdgeQuery queryEdges(const DeviceWorld &w, const Poly &A, const Poly &B) {
  // The Gauss-map test in B's frame: A's normals turned into it, B's read as stored
  const float4 toB = qmul(qconj(B.q), A.q);
  const uint32_t total = A.ne * B.ne;
  float best = -BIGF;
  uint32_t at = NO_INDEX;
  for (uint32_t k = laneId(); k < total; k += 64) {
    f3 n;
    float s;
    if (!edgeAxis(w, A, B, toB, k / B.ne, k % B.ne, n, s)) continue;
    if (s > best) {
      best = s;
      at = k;
    }
  }
  waveArgMax(best, at);
  EdgeQuery out;
  out.sep = best;
  out.valid = at != NO_INDEX;
  out.ea = out.valid ? at / B.ne : 0u;
  out.eb = out.valid ? at % B.ne : 0u;
  out.n = mk3(0.f, 0.f, 0.f);
  if (out.valid) {
    float s;
    edgeAxis(w, A, B, toB, out.ea, out.eb, out.n, s);
  }
  return out;
}

// --- contacts ---------------------------------------------------------------------------------------------------------

/**
 * Face contact (wgsl-hull.ts:250-325): the reference face (A's when refIsA) and the other shape's incident face (the
 * most anti-parallel among the faces that touch its deepest vertex), clipped by the reference face's side planes;
 * points on or below the reference plane are kept, at most 8: the deepest first, then each time the one farthest from
 * those chosen. Lanes hold the clip polygon, one point each.
 */
AVBD_D Found hullFaceContact(const DeviceWorld &w, const Poly &A, const Poly &B, bool refIsA, uint32_t refFace,
                             HullScratch &sc) {
  Found found;
  found.count = 0;
  const Poly &R = refIsA ? A : B;
  const Poly &I = refIsA ? B : A;
  const uint32_t lane = laneId();
  const float4 plane = facePlane(w, R, refFace);
  const f3 rn = xyz(plane);
  // Deepest vertex of I along rn (first of the smallest)
  float low = -BIGF;
  uint32_t deepest = NO_INDEX;
  for (uint32_t v = lane; v < I.nv; v += 64) {
    const float d = -dot(polyVert(w, I, v), rn);
    if (d > low) {
      low = d;
      deepest = v;
    }
  }
  waveArgMax(low, deepest);
  if (deepest == NO_INDEX) deepest = 0;
  // Incident face: the most anti-parallel of the faces touching that vertex
  float least = -BIGF;
  uint32_t inc = NO_INDEX;
  for (uint32_t f = lane; f < I.nf; f += 64) {
    bool touches = false;
    const uint32_t k = faceVertCount(w, I, f);
    for (uint32_t j = 0; j < k; j++)
      if (faceVert(w, I, f, j) == deepest) {
        touches = true;
        break;
      }
    if (!touches) continue;
    const float d = -dot(xyz(facePlane(w, I, f)), rn);
    if (d > least) {
      least = d;
      inc = f;
    }
  }
  waveArgMax(least, inc);
  if (inc == NO_INDEX) inc = 0;
  // The incident polygon, a point per lane
  uint32_t n = umin(faceVertCount(w, I, inc), MAX_CLIP);
  f3 mine = mk3(0.f, 0.f, 0.f);
  if (lane < n) mine = polyVert(w, I, faceVert(w, I, inc, lane));
  // Side planes of the reference face, outward from its centre (summed in loop order)
  const uint32_t k = faceVertCount(w, R, refFace);
  f3 centre = mk3(0.f, 0.f, 0.f);
  for (uint32_t j = 0; j < k; j++) centre = add(centre, polyVert(w, R, faceVert(w, R, refFace, j)));
  centre = divs(centre, float(k));
  for (uint32_t j = 0; j < k && n > 0; j++) {
    const f3 v0 = polyVert(w, R, faceVert(w, R, refFace, j));
    const f3 v1 = polyVert(w, R, faceVert(w, R, refFace, (j + 1u) % k));
    f3 s = cross(sub(v1, v0), rn);
    const float l = length(s);
    if (l < SAT_AXIS_EPSILON) continue;
    s = divs(s, l);
    if (dot(s, sub(centre, v0)) > 0.f) s = neg(s);
    const float offset = dot(s, v0);
    // Sutherland–Hodgman, a point per lane: lane i takes the edge from point i-1 to point i and emits the crossing
    // (if the edge crosses the plane) and then point i (if inside), at its place in the output order
    const f3 a = shfl3(mine, (lane + n - 1u) % n);
    const float da = dot(s, a) - offset, db = dot(s, mine) - offset;
    const bool inA = da <= PLANE_EPSILON, inB = db <= PLANE_EPSILON;
    const bool active = lane < n;
    const uint32_t emit = active ? uint32_t(inA != inB) + uint32_t(inB) : 0u;
    uint32_t total;
    uint32_t pos = waveExclusive(emit, total);
    __syncthreads();
    if (active && inA != inB) {
      float t = 0.f;
      if (fabsf(da - db) > SAT_AXIS_EPSILON) t = wclamp(da / (da - db), 0.f, 1.f);
      const f3 x = add(a, scale(sub(mine, a), t));
      if (pos < MAX_CLIP) {
        sc.x[pos] = x.x;
        sc.y[pos] = x.y;
        sc.z[pos] = x.z;
      }
      pos++;
    }
    if (active && inB && pos < MAX_CLIP) {
      sc.x[pos] = mine.x;
      sc.y[pos] = mine.y;
      sc.z[pos] = mine.z;
    }
    __syncthreads();
    n = umin(total, MAX_CLIP);
    if (lane < n) mine = mk3(sc.x[lane], sc.y[lane], sc.z[lane]);
  }
  // Points on or below the reference plane, compacted in order (their polygon index is the feature's low byte)
  const float dist = dot(rn, mine) - plane.w;
  const bool kept = lane < n && dist <= PLANE_EPSILON;
  uint32_t keep;
  const uint32_t slot = waveExclusive(kept ? 1u : 0u, keep);
  __syncthreads();
  if (kept) {
    sc.x[slot] = __uint_as_float(lane);  // the point's index in the clipped polygon
    sc.y[slot] = dist;
  }
  __syncthreads();
  float depthK = 0.f;
  uint32_t indexK = 0;
  if (lane < keep) {
    depthK = sc.y[lane];
    indexK = __float_as_uint(sc.x[lane]);
  }
  const f3 pointK = shfl3(mine, lane < keep ? indexK : 0u);
  __syncthreads();
  const uint32_t prefix = ((refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24) | (umin(refFace, 255u) << 16) |
                          (umin(inc, 255u) << 8);
  // Greedy choice: the deepest, then the farthest (squared distance to the nearest chosen) each time
  bool chosen = false;
  const uint32_t want = umin(keep, uint32_t(MAX_CONTACTS));
  for (uint32_t c = 0; c < want; c++) {
    float score = -BIGF;
    uint32_t cand = NO_INDEX;
    if (lane < keep && !chosen) {
      float s = -depthK;
      if (c > 0) {
        s = BIGF;
        for (uint32_t o = 0; o < c; o++) {
          const f3 q = mk3(sc.x[o], sc.y[o], sc.z[o]);
          const f3 d = sub(pointK, q);
          s = rmin(s, dot(d, d));
        }
      }
      if (s > score) {
        score = s;
        cand = lane;
      }
    }
    waveArgMax(score, cand);
    const uint32_t pick = cand == NO_INDEX ? 0u : cand;
    if (lane == pick) chosen = true;
    const f3 p = shfl3(pointK, pick);
    __syncthreads();
    if (lane == 0) {  // the chosen points, for the next turn's distances
      sc.x[c] = p.x;
      sc.y[c] = p.y;
      sc.z[c] = p.z;
    }
    __syncthreads();
    const float depthP = __shfl(depthK, int(pick), 64);
    const uint32_t indexP = __shfl(indexK, int(pick), 64);
    const f3 onRef = sub(p, scale(rn, depthP));
    addFound(found, refIsA ? onRef : p, refIsA ? p : onRef, prefix | umin(indexP, 255u));
  }
  __syncthreads();
  if (found.count == 0) {
    const f3 nAB = refIsA ? rn : neg(rn);
    addFound(found, polySupport(w, A, nAB), polySupport(w, B, neg(nAB)), prefix);
  }
  return found;
}

/** Edge contact (wgsl-hull.ts:327-350): the closest points of the two edges. */
AVBD_D Found hullEdgeContact(const DeviceWorld &w, const Poly &A, const Poly &B, const EdgeQuery &q) {
  Found found;
  found.count = 0;
  const uint4 ea = edgeOf(w, A, q.ea), eb = edgeOf(w, B, q.eb);
  const f3 p0 = polyVert(w, A, ea.x), q0 = polyVert(w, B, eb.x);
  const f3 d1 = sub(polyVert(w, A, ea.y), p0), d2 = sub(polyVert(w, B, eb.y), q0);
  const f3 r = sub(p0, q0);
  const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), c = dot(d1, r), b = dot(d1, d2);
  const float denom = a * e - b * b;
  float s = 0.f;
  if (fabsf(denom) > SAT_AXIS_EPSILON * a * e) s = wclamp((b * f - c * e) / denom, 0.f, 1.f);
  float t = (b * s + f) / rmax(e, SAT_AXIS_EPSILON);
  if (t < 0.f) {
    t = 0.f;
    s = wclamp(-c / rmax(a, SAT_AXIS_EPSILON), 0.f, 1.f);
  } else if (t > 1.f) {
    t = 1.f;
    s = wclamp((b - c) / rmax(a, SAT_AXIS_EPSILON), 0.f, 1.f);
  }
  addFound(found, add(p0, scale(d1, s)), add(q0, scale(d2, t)),
           (AXIS_EDGE << 24) | (umin(q.ea, 255u) << 8) | umin(q.eb, 255u));
  return found;
}

/** Two convex shapes, at least one a hull (wgsl-hull.ts:353-377); sat.n from A towards B. */
AVBD_D Found collidePoly(const DeviceWorld &w, const Poly &A, const Poly &B, bool faceBias, Sat &sat, HullScratch &sc) {
  Found none;
  none.count = 0;
  const FaceQuery fa = queryFaces(w, A, B);
  if (fa.sep > 0.f) return none;
  const FaceQuery fb = queryFaces(w, B, A);
  if (fb.sep > 0.f) return none;
  const EdgeQuery eq = queryEdges(w, A, B);
  if (eq.valid && eq.sep > 0.f) return none;
  // Faces are preferred (stabler manifolds) unless an edge axis is clearly better
  const float faceSep = rmax(fa.sep, fb.sep);
  const bool edgeWins = faceBias ? (eq.valid && eq.sep > 0.95f * faceSep + 0.01f)
                                 : (eq.valid && 0.95f * eq.sep > faceSep + 0.01f);
  sat.valid = true;
  if (edgeWins) {
    sat.kind = AXIS_EDGE;
    sat.ia = eq.ea;
    sat.ib = eq.eb;
    sat.sep = eq.sep;
    sat.n = eq.n;
    return hullEdgeContact(w, A, B, eq);
  }
  if (fb.sep > fa.sep) {
    sat.kind = AXIS_FACE_B;
    sat.ia = 0;
    sat.ib = fb.face;
    sat.sep = fb.sep;
    sat.n = neg(xyz(facePlane(w, B, fb.face)));
    return hullFaceContact(w, A, B, false, fb.face, sc);
  }
  sat.kind = AXIS_FACE_A;
  sat.ia = fa.face;
  sat.ib = 0;
  sat.sep = fa.sep;
  sat.n = xyz(facePlane(w, A, fa.face));
  return hullFaceContact(w, A, B, true, fa.face, sc);
}

/** Closest point of segment ab to p. */
AVBD_D f3 segmentPoint(f3 a, f3 b, f3 p) {
  const f3 d = sub(b, a);
  return add(a, scale(d, wclamp(dot(sub(p, a), d) / rmax(dot(d, d), 1e-20f), 0.f, 1.f)));
}

/**
 * Sphere (centre c, radius r) against shape X (wgsl-hull.ts:385-441): the closest point of X's surface, or out through
 * the face the centre is least deep behind when it is inside. n points from the sphere towards X. Lanes take faces.
 */
AVBD_D SphereBoxHit sphereHull(const DeviceWorld &w, f3 c, float r, const Poly &X) {
  SphereBoxHit out;
  out.hit = false;
  out.n = out.onBox = out.onSphere = mk3(0.f, 0.f, 0.f);
  const uint32_t lane = laneId();
  float deepest = -BIGF;
  uint32_t deepFace = NO_INDEX;
  for (uint32_t f = lane; f < X.nf; f += 64) {
    const float4 p = facePlane(w, X, f);
    const float s = dot(xyz(p), c) - p.w;
    if (s > deepest) {
      deepest = s;
      deepFace = f;
    }
  }
  waveArgMax(deepest, deepFace);
  if (deepFace == NO_INDEX) deepFace = 0;
  if (deepest > r) return out;
  if (deepest <= 0.f) {
    const f3 n = xyz(facePlane(w, X, deepFace));
    out.hit = true;
    out.n = neg(n);
    out.onBox = sub(c, scale(n, deepest));
    out.onSphere = sub(c, scale(n, r));
    return out;
  }
  // Outside: the closest point lies on a face the centre is in front of (each lane its faces' nearest points)
  float bestNeg = -BIGF;
  uint32_t bestFace = NO_INDEX;
  f3 bestPoint = mk3(0.f, 0.f, 0.f);
  for (uint32_t f = lane; f < X.nf; f += 64) {
    const float4 plane = facePlane(w, X, f);
    const f3 pn = xyz(plane);
    const float s = dot(pn, c) - plane.w;
    if (s <= 0.f) continue;
    f3 p = sub(c, scale(pn, s));
    const uint32_t k = faceVertCount(w, X, f);
    f3 centre = mk3(0.f, 0.f, 0.f);
    for (uint32_t j = 0; j < k; j++) centre = add(centre, polyVert(w, X, faceVert(w, X, f, j)));
    centre = divs(centre, float(k));
    bool inside = true;
    f3 edgeBest = mk3(0.f, 0.f, 0.f);
    float edgeDist = BIGF;
    for (uint32_t j = 0; j < k; j++) {
      const f3 v0 = polyVert(w, X, faceVert(w, X, f, j));
      const f3 v1 = polyVert(w, X, faceVert(w, X, f, (j + 1u) % k));
      f3 side = cross(sub(v1, v0), pn);
      if (dot(side, sub(centre, v0)) > 0.f) side = neg(side);
      if (dot(side, sub(p, v0)) > 0.f) inside = false;
      const f3 q = segmentPoint(v0, v1, c);
      const float dq = dot(sub(q, c), sub(q, c));
      if (dq < edgeDist) {
        edgeDist = dq;
        edgeBest = q;
      }
    }
    if (!inside) p = edgeBest;
    const float dist = length(sub(p, c));
    if (-dist > bestNeg) {
      bestNeg = -dist;
      bestFace = f;
      bestPoint = p;
    }
  }
  waveArgMax(bestNeg, bestFace);
  if (bestFace == NO_INDEX) return out;  // (no face in front: cannot happen when deepest > 0)
  // The winning lane holds its point: the lane that took that face is bestFace % 64
  const f3 best = shfl3(bestPoint, bestFace % 64u);
  const float bestDist = -bestNeg;
  if (bestDist > r) return out;
  f3 n = mk3(0.f, 0.f, 1.f);
  if (bestDist > 0.f) n = divs(sub(c, best), bestDist);  // from the hull towards the sphere
  out.hit = true;
  out.n = neg(n);
  out.onBox = best;
  out.onSphere = sub(c, scale(n, r));
  return out;
}

/** One axis of the capsule's fallback SAT (capPolyAxis, capsuleCollision.js:38-44): false when it separates. */
AVBD_D bool capPolyAxis(const DeviceWorld &w, const Capsule &A, const Poly &P, f3 axis, float &depth, f3 &normal,
                        bool &used) {
  used = false;
  const float l = length(axis);
  if (l < 1e-8f) return true;
  const f3 n = divs(axis, l);
  const float lo = rmin(dot(A.a, n), dot(A.b, n)) - A.r;
  const float hi = rmax(dot(A.a, n), dot(A.b, n)) + A.r;
  const float low = dot(polySupport(w, P, neg(n)), n);
  const float high = dot(polySupport(w, P, n), n);
  const float positive = hi - low, negative = high - lo;
  if (positive < 0.f || negative < 0.f) return false;
  depth = rmin(positive, negative);
  normal = positive < negative ? n : neg(n);
  used = true;
  return true;
}

/**
 * Capsule A against shape P (capsulePoly, capsuleCollision.js:45-66): the segment's closest approach to P's surface
 * (every face's clipped projection and every face edge), or, when the segment is inside or touching, the shallowest of
 * the face, edge × segment and vertex axes. Lanes take faces (and the fallback's axes).
 */
AVBD_D Found capsulePoly(const DeviceWorld &w, const Capsule &A, const Poly &P, Sat &sat) {
  Found found;
  found.count = 0;
  const uint32_t lane = laneId();
  const f3 direction = sub(A.b, A.a);
  float insideLo = 0.f, insideHi = 1.f;
  float bestNeg = -BIGF;
  uint32_t bestAt = NO_INDEX;
  SegmentPair points{mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
  for (uint32_t f = lane; f < P.nf; f += 64) {
    const float4 plane = facePlane(w, P, f);
    const f3 n = xyz(plane);
    const float s = dot(n, A.a) - plane.w, slope = dot(n, direction);
    if (fabsf(slope) < 1e-12f) {
      if (s > 0.f) insideLo = 2.f;  // (any value above 1 keeps the segment outside: insideHi never exceeds 1)
    } else if (slope > 0.f) {
      insideHi = rmin(insideHi, -s / slope);
    } else {
      insideLo = rmax(insideLo, -s / slope);
    }
    const uint32_t count = faceVertCount(w, P, f);
    f3 center = mk3(0.f, 0.f, 0.f);
    for (uint32_t j = 0; j < count; j++) center = add(center, polyVert(w, P, faceVert(w, P, f, j)));
    center = divs(center, float(count));
    float lo = 0.f, hi = 1.f;
    const f3 projectedA = sub(A.a, scale(n, s)), projectedD = sub(direction, scale(n, slope));
    // candidates of this face in loop order: its edges j, then the projected segment; k = 2 f · 33 + position
    for (uint32_t j = 0; j < count; j++) {
      const f3 a = polyVert(w, P, faceVert(w, P, f, j));
      const f3 b = polyVert(w, P, faceVert(w, P, f, (j + 1u) % count));
      f3 side = cross(sub(b, a), n);
      if (dot(side, sub(center, a)) > 0.f) side = neg(side);
      const float edgeS = dot(side, sub(projectedA, a)), edgeD = dot(side, projectedD);
      if (fabsf(edgeD) < 1e-12f) {
        if (edgeS > 1e-7f) lo = 2.f;
      } else if (edgeD > 0.f) {
        hi = rmin(hi, -edgeS / edgeD);
      } else {
        lo = rmax(lo, -edgeS / edgeD);
      }
      const SegmentPair pair = capSegments(A.a, A.b, a, b);
      const float dist = dot(sub(pair.a, pair.b), sub(pair.a, pair.b));
      if (-dist > bestNeg) {
        bestNeg = -dist;
        bestAt = f * 64u + j;
        points = pair;
      }
    }
    if (lo <= hi) {
      float t = 0.5f * (lo + hi);
      if (fabsf(slope) > 1e-12f) t = wclamp(-s / slope, lo, hi);
      const f3 p = add(A.a, scale(direction, t));
      const f3 q = sub(p, scale(n, s + t * slope));
      const float dist = dot(sub(p, q), sub(p, q));
      if (-dist > bestNeg) {
        bestNeg = -dist;
        bestAt = f * 64u + 63u;
        points = SegmentPair{p, q};
      }
    }
  }
  // insideLo is the largest lower bound, insideHi the smallest upper one (min and max are exact in any order)
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    insideLo = rmax(insideLo, __shfl_xor(insideLo, d, 64));
    insideHi = rmin(insideHi, __shfl_xor(insideHi, d, 64));
  }
  waveArgMax(bestNeg, bestAt);
  const uint32_t owner = bestAt == NO_INDEX ? 0u : (bestAt / 64u) % 64u;  // the lane that took that face
  points.a = shfl3(points.a, owner);
  points.b = shfl3(points.b, owner);
  const float best = bestAt == NO_INDEX ? BIGF : -bestNeg;
  const bool inside = insideLo <= insideHi;
  const float distance = sqrtf(best);
  if (!inside && distance > A.r) return found;
  f3 n = mk3(0.f, 0.f, 0.f);
  if (!inside && distance > 1e-8f) {
    n = divs(sub(points.b, points.a), distance);
  } else {
    // The fallback SAT over faces, then edges × segment, then vertices: the first shallowest; none if any separates
    const uint32_t axes = P.nf + P.ne + P.nv;
    float shallowNeg = -BIGF;
    uint32_t shallowAt = NO_INDEX;
    f3 shallowN = mk3(0.f, 0.f, 0.f);
    bool separated = false;
    for (uint32_t k = lane; k < axes; k += 64) {
      f3 axis;
      if (k < P.nf) {
        axis = xyz(facePlane(w, P, k));
      } else if (k < P.nf + P.ne) {
        const uint4 e = edgeOf(w, P, k - P.nf);
        axis = cross(direction, sub(polyVert(w, P, e.y), polyVert(w, P, e.x)));
      } else {
        const f3 vertex = polyVert(w, P, k - P.nf - P.ne);
        axis = sub(vertex, capSegmentPoint(A.a, A.b, vertex));
      }
      float depth;
      f3 an;
      bool used;
      if (!capPolyAxis(w, A, P, axis, depth, an, used)) {
        separated = true;
        continue;
      }
      if (used && -depth > shallowNeg) {
        shallowNeg = -depth;
        shallowAt = k;
        shallowN = an;
      }
    }
    if (__any(separated)) return found;
    waveArgMax(shallowNeg, shallowAt);
    float depth = BIGF;
    if (shallowAt != NO_INDEX) {
      n = shfl3(shallowN, shallowAt % 64u);
      depth = -shallowNeg;
    }
    points.a = dot(A.b, n) > dot(A.a, n) ? A.b : A.a;
    points.b = add(points.a, scale(n, A.r - depth));
  }
  sat.n = n;
  addFound(found, add(points.a, scale(n, A.r)), points.b, CAPSULE_FEATURE);
  return found;
}

}  // namespace avbd
