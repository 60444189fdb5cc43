// Restitution around the step (src/gpu/restitution.js:16-88): velocity-level normal impulses after the position solve
// for manifolds whose bodies have a restitution coefficient. The incoming velocities are captured before the step; after
// it each manifold with a normal approach faster than 1 m/s gets a rebound speed −e·vn (e = the larger coefficient of
// its two bodies) and its effective mass; then 12 rounds of a per-manifold impulse update (relaxed by the larger
// adjacency degree of its bodies) and a per-body velocity update (summing its manifolds' impulse changes in adjacency
// order); the last round writes the velocities back. Each manifold owns its impulse and each body its velocity, so
// no round needs atomics and the result does not depend on scheduling.
#include <algorithm>

#include "kernels.h"

namespace avbd {

namespace {

constexpr float BOUNCE_THRESHOLD = 1.f;  // normal approach speed (m/s) below which a contact does not bounce

AVBD_D float inverseMass(const Body &b) { return b.size.w > 0.f ? 1.f / rmax(b.size.w, 1e-20f) : 0.f; }

/** World-frame inverse inertia applied to t (restitution.js:29-34). */
AVBD_D f3 inverseInertia(const Body &b, f3 t) {
  if (b.size.w <= 0.f) return mk3(0.f, 0.f, 0.f);
  const f3 local = qrotate(qconj(b.rot), t);
  const f3 m = xyz(b.moment);
  const f3 scaled = mk3(m.x > 0.f ? local.x / rmax(m.x, 1e-20f) : 0.f, m.y > 0.f ? local.y / rmax(m.y, 1e-20f) : 0.f,
                        m.z > 0.f ? local.z / rmax(m.z, 1e-20f) : 0.f);
  return qrotate(b.rot, scaled);
}

/** Velocity of body i at lever r: incoming (captured before the step) or current (solved plus applied bounces). */
AVBD_D f3 pointVelocity(const DeviceWorld &w, uint32_t i, f3 r, bool incoming) {
  const float4 *v = w.bounceVel + 4ull * i;
  return add(xyz(v[incoming ? 0 : 2]), cross(xyz(v[incoming ? 1 : 3]), r));
}

AVBD_D uint32_t manifoldCount(const DeviceWorld &w) { return umin(w.counters[C_MANIFOLDS], w.pairCapacity); }

}  // namespace

/** Before the step: every body's incoming velocities (captureIncoming). */
__global__ void k_rest_incoming(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  w.bounceVel[4ull * i] = w.bodies[i].vel;
  w.bounceVel[4ull * i + 1] = w.bodies[i].angVel;
}

/**
 * After the step: every body's solved velocities (captureSolved) and every manifold's bounce record (initialize):
 * (impulse, delta, rebound speed, effective mass), mean lever arms on A and on B. A manifold that does not bounce keeps
 * an effective mass of 0. Nothing bounces in a step that overflowed or clashed.
 */
__global__ void k_rest_begin(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < w.nBodies) {
    w.bounceVel[4ull * i + 2] = w.bodies[i].vel;
    w.bounceVel[4ull * i + 3] = w.bodies[i].angVel;
  }
  if (i >= manifoldCount(w)) return;
  float4 *s = w.bounce + 3ull * i;
  s[0] = make_float4(0.f, 0.f, 0.f, 0.f);
  s[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  s[2] = make_float4(0.f, 0.f, 0.f, 0.f);
  const Manifold m = w.manifolds[i];
  const uint32_t count = m.ids.w & 15u;
  const uint32_t a = m.ids.x, b = m.ids.y;
  const float e = rmax(w.restitution[a], w.restitution[b]);
  if (count == 0u || e <= 0.f || (m.ids.w & 0x80000000u) != 0u || w.counters[C_OVERFLOW] != 0u ||
      w.counters[C_CLASHES] != 0u)
    return;
  const Body &A = w.bodies[a], &B = w.bodies[b];
  f3 ra = mk3(0.f, 0.f, 0.f), rb = mk3(0.f, 0.f, 0.f);
  for (uint32_t c = 0; c < count; c++) {
    const Contact k = w.contacts[m.ids.z + c];
    ra = add(ra, qrotate(A.rot, xyz(k.rA)));
    rb = add(rb, qrotate(B.rot, xyz(k.rB)));
  }
  const float n = float(count);
  ra = mk3(ra.x / n, ra.y / n, ra.z / n);
  rb = mk3(rb.x / n, rb.y / n, rb.z / n);
  const f3 normal = xyz(m.geo);
  const float vn = dot(normal, sub(pointVelocity(w, a, ra, true), pointVelocity(w, b, rb, true)));
  if (vn >= -BOUNCE_THRESHOLD) return;
  const f3 ta = cross(ra, normal), tb = cross(rb, normal);
  const float k = inverseMass(A) + inverseMass(B) + dot(ta, inverseInertia(A, ta)) + dot(tb, inverseInertia(B, tb));
  if (k <= 0.f) return;
  s[0] = make_float4(0.f, 0.f, -e * vn, k);
  s[1] = mk4(ra, 0.f);
  s[2] = mk4(rb, 0.f);
}

/** One round's impulse per bouncing manifold (pairs): toward the rebound speed, relaxed by the larger degree. */
__global__ void k_rest_pairs(DeviceWorld w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= manifoldCount(w)) return;
  float4 *sp = w.bounce + 3ull * i;
  float4 s = sp[0];
  if (s.w <= 0.f) return;
  const Manifold m = w.manifolds[i];
  const uint32_t a = m.ids.x, b = m.ids.y;
  const f3 ra = xyz(sp[1]), rb = xyz(sp[2]);
  const float vn = dot(xyz(m.geo), sub(pointVelocity(w, a, ra, false), pointVelocity(w, b, rb, false)));
  const uint32_t degree = umax(1u, umax(w.adjStart[a + 1] - w.adjStart[a], w.adjStart[b + 1] - w.adjStart[b]));
  const float impulse = rmax(0.f, s.x + (s.z - vn) / (s.w * float(degree)));
  s.y = impulse - s.x;
  s.x = impulse;
  sp[0] = s;
}

/** One round's velocity change per dynamic body (apply), summed over its manifolds in adjacency order; the last round
 * also writes the velocities back to the body records (writeBack). */
__global__ void k_rest_apply(DeviceWorld w, uint32_t last) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  const Body &B = w.bodies[i];
  if (B.size.w <= 0.f) return;
  f3 dv = mk3(0.f, 0.f, 0.f), dw = mk3(0.f, 0.f, 0.f);
  const uint32_t end = umin(w.adjStart[i + 1], w.adjCapacity);
  for (uint32_t j = w.adjStart[i]; j < end; j++) {
    const uint32_t key = w.adjList[j];
    if (key < w.nJoints) continue;
    const Manifold m = w.manifolds[key - w.nJoints];
    const float4 *sp = w.bounce + 3ull * (key - w.nJoints);
    const float4 s = sp[0];
    if (s.w <= 0.f) continue;
    const bool isA = i == m.ids.x;
    const f3 impulse = scale(scale(xyz(m.geo), s.y), isA ? 1.f : -1.f);
    dv = add(dv, scale(impulse, inverseMass(B)));
    dw = add(dw, inverseInertia(B, cross(xyz(isA ? sp[1] : sp[2]), impulse)));
  }
  float4 *v = w.bounceVel + 4ull * i;
  const float4 lin = mk4(add(xyz(v[2]), dv), v[2].w + 0.f), ang = mk4(add(xyz(v[3]), dw), v[3].w + 0.f);
  v[2] = lin;
  v[3] = ang;
  if (last) {
    w.bodies[i].vel = lin;
    w.bodies[i].angVel = ang;
  }
}

static inline uint32_t blocks(uint64_t n, uint32_t bs) {
  uint64_t g = (n + bs - 1) / bs;
  return uint32_t(g < 1 ? 1 : g);
}

void launch_restitution_before(const DeviceWorld &w, hipStream_t s) {
  k_rest_incoming<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
}

void launch_restitution_after(const DeviceWorld &w, hipStream_t s) {
  constexpr uint32_t ROUNDS = 12;
  k_rest_begin<<<blocks(std::max<uint64_t>(w.nBodies, w.pairCapacity), 256), 256, 0, s>>>(w);
  for (uint32_t r = 0; r < ROUNDS; r++) {
    k_rest_pairs<<<blocks(w.pairCapacity, 256), 256, 0, s>>>(w);
    k_rest_apply<<<blocks(w.nBodies, 256), 256, 0, s>>>(w, r + 1 == ROUNDS ? 1u : 0u);
  }
}

}  // namespace avbd
