// Shape queries (src/gpu/shapeQueries.js, shapeQueries.wgsl): ray casts and sphere casts against the current poses of
// every body. One launch traces every (body, query) pair and keeps each query's nearest distance (an atomic min on its
// bits: distances are non-negative, so their bit patterns order like the values); a second launch keeps the smallest
// body index at that distance; a third traces the winner again for its normal. Boxes are tested as rounded boxes
// (offset faces and edge capsules, so a sphere cast never hits a corner of an inflated bounding box), spheres and
// capsules analytically. A query that starts inside or touching a shape hits it at distance 0 with normal −direction.
#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels.h"

namespace avbd {

namespace {

struct Hit {
  float t;  // < 0: no hit
  f3 n;
};

AVBD_D Hit miss() { return Hit{-1.f, mk3(0.f, 0.f, 0.f)}; }

/** q · v · q⁻¹ as v + w·t + q×t with t = 2 q×v (the shader's turn()). */
AVBD_D f3 turn(float4 q, f3 v) {
  const f3 u = xyz(q);
  const f3 t = scale(cross(u, v), 2.f);
  return add(add(v, scale(t, q.w)), cross(u, t));
}

AVBD_D float clamp01(float x) { return rmin(rmax(x, 0.f), 1.f); }
/** A hit distance as reported and compared: negative zero becomes zero (its bits must order like the value). */
AVBD_D float nonneg(float t) { return t > 0.f ? t : 0.f; }

/** The point of segment ab closest to p. */
AVBD_D f3 closest(f3 a, f3 b, f3 p) {
  const f3 e = sub(b, a);
  return add(a, scale(e, clamp01(dot(sub(p, a), e) / rmax(dot(e, e), 1e-20f))));
}

AVBD_D Hit raySphere(f3 o, f3 d, f3 c, float r) {
  const f3 v = sub(o, c);
  const float b = dot(v, d);
  const float disc = (b * b - dot(v, v)) + r * r;
  if (disc < 0.f) return miss();
  const float t = -b - sqrtf(disc);
  if (t < 0.f) return miss();
  const f3 p = sub(add(o, scale(d, t)), c);
  const float l = length(p);
  return Hit{t, l > 1e-10f ? divs(p, rmax(l, 1e-20f)) : neg(d)};
}

AVBD_D Hit rayCapsule(f3 o, f3 d, f3 a, f3 b, float r) {
  const f3 ba = sub(b, a), oa = sub(o, a);
  const float bb = dot(ba, ba);
  if (bb < 1e-16f) return raySphere(o, d, a, r);
  const float bd = dot(ba, d), bo = dot(ba, oa), od = dot(oa, d);
  const float A = bb - bd * bd;
  const float B = bb * od - bo * bd;
  const float C = (bb * dot(oa, oa) - bo * bo) - (r * r) * bb;
  Hit hit = miss();
  const float disc = B * B - A * C;
  if (A > 1e-12f * bb && disc >= 0.f) {
    const float t = (-B - sqrtf(disc)) / A;
    const float y = bo + t * bd;
    if (t >= 0.f && y >= 0.f && y <= bb) {
      const f3 p = add(o, scale(d, t));
      hit = Hit{t, normalize(sub(p, add(a, divs(scale(ba, y), bb))))};
    }
  }
  for (int end = 0; end < 2; end++) {
    const Hit h = raySphere(o, d, end == 1 ? b : a, r);
    if (h.t >= 0.f && (hit.t < 0.f || h.t < hit.t)) hit = h;
  }
  return hit;
}

/** Box corner v (bit k set: +h along axis k). */
AVBD_D f3 corner(uint32_t v, f3 h) {
  return mk3((v & 1u) ? h.x : -h.x, (v & 2u) ? h.y : -h.y, (v & 4u) ? h.z : -h.z);
}
/** Box face f: outward normal along axis f / 2 (negative for odd f), offset h. */
AVBD_D float4 facePlane(uint32_t f, f3 h) {
  const uint32_t k = f / 2u;
  const float s = (f & 1u) ? -1.f : 1.f;
  return make_float4(k == 0 ? s : 0.f, k == 1 ? s : 0.f, k == 2 ? s : 0.f, comp(h, int(k)));
}
// Corners of each face in order around it, and the twelve edges (pairs of corners)
__constant__ uint8_t FACE_CORNERS[24] = {1, 3, 7, 5, 0, 2, 6, 4, 2, 3, 7, 6, 0, 1, 5, 4, 4, 5, 7, 6, 0, 1, 3, 2};
__constant__ uint8_t EDGE_CORNERS[24] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 2, 1, 3, 4, 6, 5, 7, 0, 4, 1, 5, 2, 6, 3, 7};

/**
 * The convex shape a query traces in the body's frame: a box of half extents h, or a hull record (its header word;
 * shapeQueries.wgsl:34-46 vertex, plane, faceCount, faceVertex and the edge list).
 */
struct QShape {
  f3 h;
  const uint4 *H;
  uint32_t header;
  bool hull;
};
AVBD_D f3 qVertex(const QShape &S, uint32_t v) {
  if (!S.hull) return corner(v, S.h);
  const uint4 u = S.H[S.H[S.header].x + v];
  return mk3(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z));
}
AVBD_D float4 qPlane(const QShape &S, uint32_t f) {
  if (!S.hull) return facePlane(f, S.h);
  const uint4 u = S.H[S.H[S.header].z + 2u * f];
  return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
}
AVBD_D uint32_t qFaceCount(const QShape &S, uint32_t f) { return S.hull ? S.H[S.H[S.header].z + 2u * f + 1u].y : 4u; }
AVBD_D uint32_t qFaceVertex(const QShape &S, uint32_t f, uint32_t j) {
  if (!S.hull) return FACE_CORNERS[f * 4 + j];
  const uint32_t index = S.H[S.H[S.header].z + 2u * f + 1u].x + j;
  const uint4 u = S.H[S.H[S.header + 1u].z + index / 4u];
  const uint32_t k = index % 4u;
  return k == 0 ? u.x : (k == 1 ? u.y : (k == 2 ? u.z : u.w));
}
AVBD_D uint2 qEdge(const QShape &S, uint32_t e) {
  if (!S.hull) return make_uint2(EDGE_CORNERS[2 * e], EDGE_CORNERS[2 * e + 1]);
  const uint4 u = S.H[S.H[S.header + 1u].x + e];
  return make_uint2(u.x, u.y);
}

/** p (on face f's plane) lies inside the face, with a 1e-6 tolerance. */
AVBD_D bool onFace(f3 p, uint32_t f, const QShape &S) {
  const f3 normal = xyz(qPlane(S, f));
  const uint32_t count = qFaceCount(S, f);
  f3 center = mk3(0.f, 0.f, 0.f);
  for (uint32_t j = 0; j < count; j++) center = add(center, qVertex(S, qFaceVertex(S, f, j)));
  center = divs(center, float(count));
  for (uint32_t j = 0; j < count; j++) {
    const f3 a = qVertex(S, qFaceVertex(S, f, j)), b = qVertex(S, qFaceVertex(S, f, (j + 1) % count));
    const f3 side = cross(sub(b, a), normal);
    if (dot(side, sub(p, a)) * (dot(side, sub(center, a)) >= 0.f ? 1.f : -1.f) < -1e-6f) return false;
  }
  return true;
}

/**
 * A ray (r = 0: the faces' slab test) or a sphere of radius r swept along d against a box or hull (roundedPoly,
 * shapeQueries.wgsl:54-76): the offset faces, the edges' capsules (whose end spheres cover the corners), and an
 * overlap at the start (centre inside or within r of the surface) as a hit at 0.
 */
AVBD_D Hit roundedPoly(f3 o, f3 d, const QShape &S, float r) {
  const uint32_t nf = S.hull ? S.H[S.header].w : 6u, ne = S.hull ? S.H[S.header + 1u].y : 12u;
  if (r == 0.f) {
    float enter = -3.4e38f, exit = 3.4e38f;
    f3 normal = neg(d);
    for (uint32_t f = 0; f < nf; f++) {
      const float4 pl = qPlane(S, f);
      const float denom = dot(xyz(pl), d), gap = pl.w - dot(xyz(pl), o);
      if (fabsf(denom) < 1e-12f) {
        if (gap < 0.f) return miss();
      } else {
        const float t = gap / denom;
        if (denom < 0.f) {
          if (t > enter) {
            enter = t;
            normal = xyz(pl);
          }
        } else {
          exit = rmin(exit, t);
        }
      }
      if (enter > exit) return miss();
    }
    if (exit < 0.f) return miss();
    if (enter <= 0.f) return Hit{0.f, neg(d)};
    return Hit{enter, normal};
  }
  bool inside = true;
  float distance = 3.4e38f;
  Hit hit = miss();
  for (uint32_t f = 0; f < nf; f++) {
    const float4 pl = qPlane(S, f);
    const f3 n = xyz(pl);
    const float signed_ = dot(n, o) - pl.w;
    if (signed_ > 0.f) inside = false;
    if (onFace(sub(o, scale(n, signed_)), f, S)) distance = rmin(distance, signed_ * signed_);
    const float denom = dot(n, d);
    if (denom < -1e-12f) {
      const float t = ((pl.w + r) - dot(n, o)) / denom;
      if (t >= 0.f && (hit.t < 0.f || t < hit.t) && onFace(sub(add(o, scale(d, t)), scale(n, r)), f, S))
        hit = Hit{t, n};
    }
  }
  for (uint32_t e = 0; e < ne; e++) {
    const uint2 edge = qEdge(S, e);
    const f3 a = qVertex(S, edge.x), b = qVertex(S, edge.y);
    const f3 q = sub(o, closest(a, b, o));
    distance = rmin(distance, dot(q, q));
    const Hit c = rayCapsule(o, d, a, b, r);
    if (c.t >= 0.f && (hit.t < 0.f || c.t < hit.t)) hit = c;
  }
  if (inside || distance <= r * r) return Hit{0.f, neg(d)};
  return hit;
}

/** Query q against body B in the body's frame; the normal comes back in world space. */
AVBD_D Hit trace(const Body &B, const QueryRec &q, const uint4 *hulls) {
  const float4 rot = B.rot;
  const float4 inv = make_float4(-rot.x, -rot.y, -rot.z, rot.w);
  const f3 o = turn(inv, sub(xyz(q.origin), xyz(B.pos)));
  const f3 d = turn(inv, xyz(q.direction));
  const float r = q.direction.w;
  const f3 size = xyz(B.size);
  const float shape = B.angVel.w;
  Hit hit;
  if (shape == SHAPE_SPHERE) {
    const float radius = size.x * 0.5f + r;
    hit = dot(o, o) <= radius * radius ? Hit{0.f, neg(d)} : raySphere(o, d, mk3(0.f, 0.f, 0.f), radius);
  } else if (shape == SHAPE_CAPSULE) {
    const float half = rmax(0.f, size.y * 0.5f - size.x * 0.5f);
    const f3 a = mk3(0.f, -half, 0.f), b = mk3(0.f, half, 0.f);
    const float radius = size.x * 0.5f + r;
    const f3 delta = sub(o, closest(a, b, o));
    hit = dot(delta, delta) <= radius * radius ? Hit{0.f, neg(d)} : rayCapsule(o, d, a, b, radius);
  } else {
    const bool isHull = shape >= SHAPE_HULL;
    const QShape S{scale(size, 0.5f), hulls, isHull ? uint32_t(rmax(0.f, shape - SHAPE_HULL) + 0.5f) : 0u, isHull};
    hit = roundedPoly(o, d, S, r);
  }
  if (hit.t > q.origin.w) return miss();
  hit.n = turn(rot, hit.n);
  return hit;
}

/** Left out of a query: its group misses the query's mask, the query's membership misses its mask (when given), or
 * it is on the ignore list. */
AVBD_D bool ignored(const DeviceWorld &w, uint32_t i, const QueryFilter &f) {
  const uint2 flt = w.filters[i];
  if ((flt.x & f.mask) == 0u) return true;
  if (f.useMembership && (f.membership == 0u || (flt.y & f.membership) == 0u)) return true;
  for (uint32_t k = 0; k < f.ignoreCount; k++)
    if (f.ignore[k] == i) return true;
  return false;
}

}  // namespace

/** Each query's nearest distance over all bodies (grid: bodies × queries). */
__global__ void k_query_nearest(DeviceWorld w, const QueryRec *queries, QueryFilter f, uint32_t *best) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
  if (i >= w.nBodies || ignored(w, i, f)) return;
  const Hit h = trace(w.bodies[i], queries[q], w.hulls);
  if (h.t >= 0.f) atomicMin(&best[2 * q], __float_as_uint(nonneg(h.t)));
}

/** The smallest body index at that distance. */
__global__ void k_query_owner(DeviceWorld w, const QueryRec *queries, QueryFilter f, uint32_t *best) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
  if (i >= w.nBodies || ignored(w, i, f)) return;
  const Hit h = trace(w.bodies[i], queries[q], w.hulls);
  if (h.t >= 0.f && __float_as_uint(nonneg(h.t)) == __hip_atomic_load(&best[2 * q], __ATOMIC_RELAXED,
                                                                          __HIP_MEMORY_SCOPE_AGENT))
    atomicMin(&best[2 * q + 1], i);
}

/** Per query: the winner's normal, distance and index (8 words: normal xyz, distance or −1, index bits, 0, 0, 0);
 * resets the query's best words for the next batch. */
__global__ void k_query_finish(DeviceWorld w, const QueryRec *queries, uint32_t n, uint32_t *best, float4 *hits) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t index = best[2 * q + 1];
  best[2 * q] = 0x7f800000u;
  best[2 * q + 1] = 0xffffffffu;
  if (index == 0xffffffffu) {
    hits[2 * q] = make_float4(0.f, 0.f, 0.f, -1.f);
    hits[2 * q + 1] = make_float4(__uint_as_float(0xffffffffu), 0.f, 0.f, 0.f);
    return;
  }
  const Hit h = trace(w.bodies[index], queries[q], w.hulls);
  hits[2 * q] = mk4(h.n, nonneg(h.t));
  hits[2 * q + 1] = make_float4(__uint_as_float(index), 0.f, 0.f, 0.f);
}

__global__ void k_query_reset(uint32_t *best, uint32_t n) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  best[2 * q] = 0x7f800000u;
  best[2 * q + 1] = 0xffffffffu;
}

// The scratch of castNoAlloc: NOALLOC_MAX_QUERIES query records, then their best words, then their hits. Each
// device holds its own copy at its own address until the process ends; host code reaches it only through the address
// query_scratch() looked up for that device (the array's name is not a device address in host code).
static __device__ __align__(256) uint8_t scratch[size_t(NOALLOC_MAX_QUERIES) * NOALLOC_QUERY_BYTES];

void *query_scratch(int device) {
  constexpr int MAX_DEVICES = 64;
  static std::mutex mu;
  static void *address[MAX_DEVICES] = {};
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("device ordinal out of range for query scratch");
  std::lock_guard<std::mutex> lk(mu);
  if (!address[device]) {
    int current = 0;
    if (hipGetDevice(&current) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    if (current != device && hipSetDevice(device) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
    void *ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (err != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(err));
    address[device] = ptr;
  }
  return address[device];
}

void launch_query_reset(uint32_t *best, uint32_t capacity, hipStream_t s) {
  k_query_reset<<<(capacity + 255) / 256, 256, 0, s>>>(best, capacity);
}

void launch_queries(const DeviceWorld &w, const QueryRec *queries, uint32_t n, const QueryFilter &f, uint32_t *best,
                    float4 *hits, hipStream_t s) {
  if (n == 0) return;
  const dim3 grid((w.nBodies + 255) / 256, n);
  k_query_nearest<<<grid, 256, 0, s>>>(w, queries, f, best);
  k_query_owner<<<grid, 256, 0, s>>>(w, queries, f, best);
  k_query_finish<<<(n + 255) / 256, 256, 0, s>>>(w, queries, n, best, hits);
}

}  // namespace avbd
