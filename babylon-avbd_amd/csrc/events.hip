// Contact events (src/vendor/three-avbd-lib/contacts.ts ContactWatch; src/native/solver3D.js:252-311): after each
// step, the pairs touching this step with a watched body go into this step's hash set (two sets, this step's and the
// last one's, by step parity); a pair missing from the last set began touching this step (event: where — the mean of
// its points on body a —, how hard — |Σ normal multiplier| × dt —, and the normal), and a pair of the last set missing
// from this one stopped. A body's watch word carries a lifetime epoch above bit 0: a pair whose body changed epoch
// since the last step begins again and does not end. Events append to a bounded log read and cleared on demand; the
// log's head counts every event (past the capacity too) and flags a full set.
#include "kernels.h"

namespace avbd {

namespace {

AVBD_D uint32_t eventHash(uint32_t a, uint32_t b) { return hash32((a * 0x9e3779b1u) ^ hash32(b)); }

AVBD_D uint32_t *slotOf(const ContactWatchDev &c, uint32_t set, uint32_t h) {
  return c.table + CW_HEAD + 2ull * (uint64_t(set) * (c.mask + 1) + h);
}

AVBD_D void insertPair(const ContactWatchDev &c, uint32_t set, uint32_t a, uint32_t b) {
  uint32_t h = eventHash(a, b) & c.mask;
  for (uint32_t probe = 0; probe <= c.mask; probe++) {
    uint32_t *s = slotOf(c, set, h);
    if (atomicCAS(s, 0u, a + 1u) == 0u) {
      __hip_atomic_store(s + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    h = (h + 1u) & c.mask;
  }
  atomicOr(&c.table[1], 1u);
}

AVBD_D bool hasPair(const ContactWatchDev &c, uint32_t set, uint32_t a, uint32_t b) {
  uint32_t h = eventHash(a, b) & c.mask;
  for (uint32_t probe = 0; probe <= c.mask; probe++) {
    const uint32_t *s = slotOf(c, set, h);
    const uint32_t k = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == 0u) return false;
    if (k == a + 1u && __hip_atomic_load(s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == b) return true;
    h = (h + 1u) & c.mask;
  }
  return false;
}

AVBD_D void emit(const ContactWatchDev &c, uint32_t kind, uint32_t a, uint32_t b, uint32_t step, f3 point,
                 float impulse, f3 normal) {
  const uint32_t e = atomicAdd(&c.table[0], 1u);
  if (e >= c.capacity) return;
  c.events[3ull * e] = make_float4(__uint_as_float(kind), __uint_as_float(a), __uint_as_float(b), __uint_as_float(step));
  c.events[3ull * e + 1] = mk4(point, impulse);
  c.events[3ull * e + 2] = mk4(normal, 0.f);
}

/** v turned by q: v + w·t + q×t with t = 2 q×v. */
AVBD_D f3 turnBy(float4 q, f3 v) {
  const f3 u = xyz(q);
  const f3 t = scale(cross(u, v), 2.f);
  return add(add(v, scale(t, q.w)), cross(u, t));
}

}  // namespace

/** This step's set is cleared (before the touching pass). */
__global__ void k_events_clear(ContactWatchDev c, uint32_t set) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 2 * (c.mask + 1)) c.table[CW_HEAD + 2ull * uint64_t(set) * (c.mask + 1) + k] = 0u;
}

/** Per manifold of this step: into this step's set; a begin if it was not in the last one (or a body's epoch moved). */
__global__ void k_events_touching(DeviceWorld w, ContactWatchDev c, uint32_t now) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= umin(w.counters[C_MANIFOLDS], w.pairCapacity)) return;
  const Manifold mf = w.manifolds[m];
  const uint32_t count = mf.ids.w & 15u, a = mf.ids.x, b = mf.ids.y;
  if (count == 0u || ((c.watch[a] | c.watch[b]) & 1u) == 0u) return;
  insertPair(c, now, a, b);
  if (hasPair(c, 1u - now, a, b) && c.watch[a] == c.previousWatch[a] && c.watch[b] == c.previousWatch[b]) return;
  const Body &A = w.bodies[a];
  const f3 pA = xyz(A.pos);
  f3 point = mk3(0.f, 0.f, 0.f);
  float force = 0.f;
  for (uint32_t i = 0; i < count; i++) {
    const Contact k = w.contacts[mf.ids.z + i];
    point = add(point, add(turnBy(A.rot, xyz(k.rA)), pA));
    force += k.lam.x;
  }
  const WorldParams &p = w.worlds[w.bodyWorld[a]];
  emit(c, CW_BEGIN, a, b, p.step, divs(point, float(count)), fabsf(force) * p.dt, xyz(mf.geo));
}

/** Per slot of the last step's set: an end if the pair is not in this step's (and neither body's epoch moved). */
__global__ void k_events_ended(DeviceWorld w, ContactWatchDev c, uint32_t now) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h > c.mask) return;
  const uint32_t *s = slotOf(c, 1u - now, h);
  const uint32_t k = s[0];
  if (k == 0u) return;
  const uint32_t a = k - 1u, b = s[1];
  if (c.watch[a] != c.previousWatch[a] || c.watch[b] != c.previousWatch[b]) return;
  if (!hasPair(c, now, a, b))
    emit(c, CW_END, a, b, w.worlds[w.bodyWorld[a]].step, mk3(0.f, 0.f, 0.f), 0.f, mk3(0.f, 0.f, 0.f));
}

/** After both passes: this step's watch words become the last step's. */
__global__ void k_events_copy_watch(ContactWatchDev c, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c.previousWatch[i] = c.watch[i];
}

void launch_contact_events(const DeviceWorld &w, const ContactWatchDev &c, uint32_t now, hipStream_t s) {
  const uint32_t slots = c.mask + 1;
  k_events_clear<<<(2 * slots + 255) / 256, 256, 0, s>>>(c, now);
  k_events_touching<<<(w.pairCapacity + 255) / 256, 256, 0, s>>>(w, c, now);
  k_events_ended<<<(slots + 255) / 256, 256, 0, s>>>(w, c, now);
  k_events_copy_watch<<<(w.nBodies + 255) / 256, 256, 0, s>>>(c, w.nBodies);
}

}  // namespace avbd
