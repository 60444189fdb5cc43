// Per-step topology on gfx950: per-body constraint adjacency (degree → scan → fill → per-body sort into a fixed order, so
// the primal sums a body's rows in the same order every run) and incremental Jones–Plassmann colouring with lowbias32
// priorities, colour cap and clash counting, then stable per-colour buckets.
// Semantics: reference/three-avbd/src/avbd2d/gpu/wgsl-topology.ts:61-322 (with the 3D accessors of
// reference/three-avbd/src/avbd3d/gpu/layout.ts:245-269). Priorities use a body's index within its world, so a world
// colours the same way alone and inside a batch.
#include "engine_types.h"
#include "kernels.h"
#include "stepsetup.h"

namespace avbd {

AVBD_D bool dynamicBody(const DeviceWorld &w, int i) { return i >= 0 && w.bodies[i].size.w > 0.f; }

AVBD_D bool jointActive(const DeviceWorld &w, uint32_t j) {
  return w.info[j].x != T_NONE && (w.joints[j].penLin.w != 0.f || w.joints[j].penAng.w != 0.f);
}

AVBD_D int2 endpoints(const DeviceWorld &w, uint32_t id) {
  if (id < w.nJoints) {
    int4 f = w.info[id];
    return make_int2(f.y, f.z);
  }
  uint4 ids = w.manifolds[id - w.nJoints].ids;
  return make_int2(int(ids.x), int(ids.y));
}

// --- adjacency ---------------------------------------------------------------------------------------------------------

__global__ void k_adj_fill(DeviceWorld w) {
  uint32_t total = w.nJoints + w.counters[C_MANIFOLDS];
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < total; id += gridDim.x * blockDim.x) {
    if (id < w.nJoints && !jointActive(w, id)) continue;
    int2 e = endpoints(w, id);
    int ends[2] = {e.x, e.y};
    for (int s = 0; s < 2; s++) {
      int x = ends[s];
      if (!dynamicBody(w, x)) continue;
      uint32_t slot = w.adjStart[x] + atomicAdd(&w.adjFill[x], 1u);
      if (slot < w.adjCapacity) w.adjList[slot] = id;
    }
  }
}

/** Joints first by id, then contact pairs by the other body (wgsl-topology.ts:114-123; one manifold per pair in 3D). */
AVBD_D uint64_t adjKey(const DeviceWorld &w, uint32_t b, uint32_t id) {
  if (id < w.nJoints) return uint64_t(id);
  int2 e = endpoints(w, id);
  int other = e.x == int(b) ? e.y : e.x;
  return uint64_t(uint32_t(other + 2)) << 32;
}

constexpr uint32_t ADJ_SORT_BLOCK = 64, ADJ_SORT_LDS = 32;

/** Entry records of body b's slot e (adjacency id `id`): what the colouring and the primal read. */
AVBD_D void writeEntry(const DeviceWorld &w, uint32_t b, uint32_t e, uint32_t id) {
  uint4 head, tail;
  if (id < w.nJoints) {
    int4 f = w.info[id];
    int other = f.y == int(b) ? f.z : f.y;
    head = make_uint4(other < 0 ? ENTRY_NONE : uint32_t(other), id,
                      ENTRY_JOINT | (dynamicBody(w, other) ? ENTRY_DYNAMIC : 0u), 0u);
    tail = make_uint4(0u, 0u, 0u, 0u);
  } else {
    const Manifold mf = w.manifolds[id - w.nJoints];
    bool isA = mf.ids.x == b;
    uint32_t other = isA ? mf.ids.y : mf.ids.x;
    head = make_uint4(other, mf.ids.z,
                      (mf.ids.w & 15u) | (isA ? ENTRY_IS_A : 0u) | (dynamicBody(w, int(other)) ? ENTRY_DYNAMIC : 0u),
                      __float_as_uint(mf.geo.w));
    tail = make_uint4(__float_as_uint(mf.geo.x), __float_as_uint(mf.geo.y), __float_as_uint(mf.geo.z),
                      __float_as_uint(tangentScale(xyz(mf.geo))));
  }
  w.adjEntry[2ull * e] = head;
  w.adjEntry[2ull * e + 1] = tail;
}

/**
 * Sort body b's list into key order (keys are distinct, so the order is unique), then write the entry records. A list
 * of up to ADJ_SORT_LDS entries is sorted in this thread's LDS column with its keys read once; a longer one in place.
 */
__global__ void __launch_bounds__(ADJ_SORT_BLOCK) k_adj_sort(DeviceWorld w) {
  __shared__ uint64_t keys[ADJ_SORT_LDS][ADJ_SORT_BLOCK];
  __shared__ uint32_t ids[ADJ_SORT_LDS][ADJ_SORT_BLOCK];
  const uint32_t t = threadIdx.x;
  uint32_t b = blockIdx.x * blockDim.x + t;
  if (b >= w.nBodies) return;  // no block-wide synchronisation below
  uint32_t lo = w.adjStart[b], hi = w.adjStart[b + 1];
  if (hi > w.adjCapacity) hi = w.adjCapacity;
  if (hi <= lo) return;
  const uint32_t n = hi - lo;
  if (n <= ADJ_SORT_LDS) {
    constexpr uint32_t BATCH = 8;  // loads of a batch are independent, so they are in flight together
    for (uint32_t k0 = 0; k0 < n; k0 += BATCH) {
      uint32_t id[BATCH];
      uint64_t key[BATCH];
#pragma unroll
      for (uint32_t q = 0; q < BATCH; q++) id[q] = k0 + q < n ? w.adjList[lo + k0 + q] : 0u;
#pragma unroll
      for (uint32_t q = 0; q < BATCH; q++) key[q] = k0 + q < n ? adjKey(w, b, id[q]) : 0ull;
#pragma unroll
      for (uint32_t q = 0; q < BATCH; q++) {
        if (k0 + q >= n) break;
        uint32_t j = k0 + q;
        while (j > 0 && key[q] < keys[j - 1][t]) {
          keys[j][t] = keys[j - 1][t];
          ids[j][t] = ids[j - 1][t];
          j--;
        }
        keys[j][t] = key[q];
        ids[j][t] = id[q];
      }
    }
    for (uint32_t k0 = 0; k0 < n; k0 += BATCH) {
#pragma unroll
      for (uint32_t q = 0; q < BATCH; q++) {
        if (k0 + q >= n) break;
        const uint32_t id = ids[k0 + q][t];
        w.adjList[lo + k0 + q] = id;
        writeEntry(w, b, lo + k0 + q, id);
      }
    }
    return;
  }
  for (uint32_t i = lo + 1; i < hi; i++) {
    uint32_t id = w.adjList[i];
    uint64_t key = adjKey(w, b, id);
    uint32_t j = i;
    while (j > lo) {
      uint32_t prev = w.adjList[j - 1];
      if (!(key < adjKey(w, b, prev))) break;
      w.adjList[j] = prev;
      j--;
    }
    w.adjList[j] = id;
  }
  for (uint32_t e = lo; e < hi; e++) writeEntry(w, b, e, w.adjList[e]);
}

// --- colouring ---------------------------------------------------------------------------------------------------------

/** Colouring priority of body i of the world whose first body is `first` (its index within its world, hashed). */
AVBD_D uint32_t priorityOf(uint32_t i, uint32_t first) { return hash32(i - first + 0x9e3779b9u); }
/** i beats j (same world): higher priority, ties by index (within a world, local and global index order agree). */
AVBD_D bool beats(uint32_t i, uint32_t j, uint32_t first) {
  uint32_t pi = priorityOf(i, first), pj = priorityOf(j, first);
  return pi > pj || (pi == pj && i > j);
}
AVBD_D uint32_t firstOf(const DeviceWorld &w, uint32_t b) { return w.worlds[w.bodyWorld[b]].firstBody; }

/** The dynamic body at the other end of adjacency entry e, or -1. */
AVBD_D int neighbour(const DeviceWorld &w, uint32_t e) {
  uint4 head = w.adjEntry[2ull * e];
  return (head.z & ENTRY_DYNAMIC) ? int(head.x) : -1;
}

/** Neighbours read per batch by the colouring passes. */
constexpr uint32_t NB_BATCH = 8;

/**
 * Adjacency entries [e0, min(e0 + NB_BATCH, end)) of body b: the dynamic neighbour of each (-1 for a static partner or
 * a slot past the list) and its state in `src`. Every slot loads (a slot past the list re-reads the list's last entry,
 * a slot without a neighbour reads b's own state), so a batch's loads are in flight together rather than one
 * dependent pair per entry.
 */
struct NeighbourBatch {
  int j[NB_BATCH];
  uint32_t s[NB_BATCH];
};
AVBD_D NeighbourBatch neighbourBatch(const DeviceWorld &w, uint32_t b, uint32_t e0, uint32_t end, const uint32_t *src) {
  NeighbourBatch n;
#pragma unroll
  for (uint32_t q = 0; q < NB_BATCH; q++) {
    const int j = neighbour(w, umin(e0 + q, end - 1u));
    n.j[q] = e0 + q < end ? j : -1;
  }
#pragma unroll
  for (uint32_t q = 0; q < NB_BATCH; q++) n.s[q] = src[n.j[q] >= 0 ? uint32_t(n.j[q]) : b];
  return n;
}

/** Colour bit c of a 64-colour set (lo: colours 0-31, hi: 32-63). */
AVBD_D void addColour(uint32_t &lo, uint32_t &hi, uint32_t c) {
  if (c < 32) lo |= 1u << c;
  else hi |= 1u << (c - 32);
}
/** The lowest colour missing from the set; MAX_COLORS - 1 when all 64 are taken. */
AVBD_D uint32_t lowestMissing(uint32_t lo, uint32_t hi) {
  if (~lo != 0u) return uint32_t(__builtin_ctz(~lo));
  if (~hi != 0u) return 32u + uint32_t(__builtin_ctz(~hi));
  return MAX_COLORS - 1;
}

// Step 1 (state A to state B), lowering: a coloured body that outranks every dynamic neighbour moves down to the
// lowest colour none of them holds, if that is below its own
__global__ void k_color_compact(DeviceWorld w) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < COLOR_FLAG_WORDS) w.colorFlags[b] = 0u;  // the rounds' flags start clear (current state: colorA)
  if (b >= w.nBodies) return;
  if (!dynamicBody(w, int(b))) {
    w.colorB[b] = NO_COLOR;
    return;
  }
  uint32_t col = w.colorA[b] & 0xffu;
  // A colour at or above a (lowered) cap is no colour: the body is coloured again below the cap
  if (col >= w.colorCap) col = NO_COLOR;
  if (col != NO_COLOR) {
    bool isMax = true;
    uint32_t lo = 0, hi = 0;
    const uint32_t first = firstOf(w, b), end = w.adjStart[b + 1];
    for (uint32_t e0 = w.adjStart[b]; e0 < end && isMax; e0 += NB_BATCH) {
      const NeighbourBatch n = neighbourBatch(w, b, e0, end, w.colorA);
#pragma unroll
      for (uint32_t q = 0; q < NB_BATCH; q++) {
        if (n.j[q] < 0) continue;
        if (!beats(b, uint32_t(n.j[q]), first)) isMax = false;
        const uint32_t c = n.s[q] & 0xffu;
        if (c != NO_COLOR) addColour(lo, hi, c);
      }
    }
    if (isMax) col = umin(col, lowestMissing(lo, hi));
  }
  w.colorB[b] = col;
}

// Step 2 (state B to state A), conflicts: a body is left to the rounds (pending, and listed) when it has no colour or
// shares its colour with a neighbour that outranks it
__global__ void k_color_mark(DeviceWorld w) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= w.nBodies) return;
  uint32_t col = w.colorB[b] & 0xffu;
  if (!dynamicBody(w, int(b))) {
    w.colorA[b] = NO_COLOR;
    return;
  }
  bool pending = col == NO_COLOR;
  const uint32_t first = firstOf(w, b), end = w.adjStart[b + 1];
  for (uint32_t e0 = w.adjStart[b]; e0 < end && !pending; e0 += NB_BATCH) {
    const NeighbourBatch n = neighbourBatch(w, b, e0, end, w.colorB);
#pragma unroll
    for (uint32_t q = 0; q < NB_BATCH; q++)
      if (n.j[q] >= 0 && (n.s[q] & 0xffu) == col && beats(uint32_t(n.j[q]), b, first)) pending = true;
  }
  w.colorA[b] = col | (pending ? PENDING : 0u);
  if (pending) {
    w.colorFlags[1] = 1u;  // round 0 has work
    w.colorPending[atomicAdd(&w.counters[C_PENDING], 1u)] = b;
  }
}

// Step 3, the rounds (Jones–Plassmann, reading one state buffer and writing the other): a pending body that outranks
// each pending neighbour settles on the lowest colour its settled neighbours leave free. With every colour below the
// cap taken it takes the cap's last colour and stays pending, which is counted as a clash. Once a round's input has no
// pending body, that round and all later ones would only copy their input: they return at once, and the count pass
// reads the state from the buffer the last working round wrote (colorFlags[0]).
/** Body b's state after one round, from the round's input state src. */
AVBD_D uint32_t roundState(const DeviceWorld &w, const uint32_t *src, uint32_t b) {
  uint32_t s = src[b];
  if (!(s & PENDING)) return s;
  bool isMax = true;
  uint32_t lo = 0, hi = 0;  // colours of the settled neighbours
  const uint32_t first = firstOf(w, b), end = w.adjStart[b + 1];
  for (uint32_t e0 = w.adjStart[b]; e0 < end && isMax; e0 += NB_BATCH) {
    const NeighbourBatch n = neighbourBatch(w, b, e0, end, src);
#pragma unroll
    for (uint32_t q = 0; q < NB_BATCH; q++) {
      if (n.j[q] < 0) continue;
      if (n.s[q] & PENDING) {
        if (!beats(b, uint32_t(n.j[q]), first)) isMax = false;
      } else {
        const uint32_t c = n.s[q] & 0xffu;
        if (c != NO_COLOR) addColour(lo, hi, c);
      }
    }
  }
  if (!isMax) return s;
  const uint32_t fr = lowestMissing(lo, hi);
  return fr >= w.colorCap ? ((w.colorCap - 1) | PENDING) : fr;
}

__global__ void k_color_round(DeviceWorld w, const uint32_t *src, uint32_t *dst, uint32_t r) {
  const bool flagged = r + 2 < COLOR_FLAG_WORDS;
  if (flagged && w.colorFlags[1 + r] == 0u) return;
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) w.colorFlags[0] = (r % 2 == 0) ? 1u : 0u;  // dst is colorB for even rounds
  if (b >= w.nBodies) return;
  const uint32_t out = roundState(w, src, b);
  dst[b] = out;
  if (flagged && (out & PENDING)) w.colorFlags[2 + r] = 1u;  // the next round has work
}

/**
 * All rounds in one block over the mark pass's pending list: a body off the list is settled and holds the same state
 * in both buffers, so a round only has to update the listed bodies. The same result as the per-round launches, in one
 * launch; the engine uses it while few bodies are left pending (any count is handled, a large one slowly).
 */
__global__ void __launch_bounds__(1024) k_color_rounds_listed(DeviceWorld w, uint32_t rounds) {
  __shared__ uint32_t more;
  const uint32_t P = w.counters[C_PENDING];
  uint32_t current = 0;  // the buffer holding the state: 0 colorA, 1 colorB
  for (uint32_t r = 0; r < rounds && P != 0; r++) {
    const uint32_t *src = (r % 2 == 0) ? w.colorA : w.colorB;
    uint32_t *dst = (r % 2 == 0) ? w.colorB : w.colorA;
    if (threadIdx.x == 0) more = 0;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < P; k += blockDim.x) {
      const uint32_t b = w.colorPending[k];
      const uint32_t out = roundState(w, src, b);
      dst[b] = out;
      if (out & PENDING) more = 1u;
    }
    __syncthreads();  // this round's states are written (block scope) before the next reads them
    current = (r % 2 == 0) ? 1u : 0u;
    const bool again = more != 0u;
    __syncthreads();  // every thread has read `more` before the next round clears it
    if (!again) break;
  }
  if (threadIdx.x == 0) w.colorFlags[0] = current;
}

// Step 4, counting: each block of COLOR_WG bodies counts its bodies per bucket and ranks each body among the block's
// bodies of its bucket, so a bucket lists its bodies in index order
/** Bucket key of a coloured body: its colour's heavy or light half (engine_types.h, COLOR_KEYS). */
AVBD_D uint32_t bucketKey(const DeviceWorld &w, uint32_t b, uint32_t col) {
  return 2u * col + (w.adjStart[b + 1] - w.adjStart[b] > LANES ? 0u : 1u);
}

__global__ void __launch_bounds__(COLOR_WG) k_color_count(DeviceWorld w) {
  __shared__ uint32_t hist[COLOR_KEYS];
  __shared__ uint32_t chunkKeys[COLOR_WG];
  uint32_t lid = threadIdx.x;
  if (lid < COLOR_KEYS) hist[lid] = 0;
  uint32_t b = blockIdx.x * COLOR_WG + lid;
  uint32_t key = ~0u;
  bool clash = false;
  if (b < w.nBodies && dynamicBody(w, int(b))) {
    uint32_t s = (w.colorFlags[0] ? w.colorB : w.colorA)[b];
    uint32_t col = s & 0xffu;
    clash = (s & PENDING) != 0u;
    if (col == NO_COLOR) col = w.colorCap - 1;
    w.colorA[b] = col;
    key = bucketKey(w, b, col);
  }
  // One atomic per wavefront that holds a clash (the colour count comes from the buckets, k_color_starts)
  const uint32_t clashes = uint32_t(__popcll(__ballot(clash)));
  if ((lid & 63u) == 0u && clashes) atomicAdd(&w.counters[C_CLASHES], clashes);
  chunkKeys[lid] = key;
  __syncthreads();
  if (key != ~0u) {
    uint32_t rank = 0;
    for (uint32_t j = 0; j < lid; j++) rank += chunkKeys[j] == key ? 1u : 0u;
    w.colorRank[b] = rank;
    atomicAdd(&hist[key], 1u);
  }
  __syncthreads();
  if (lid < COLOR_KEYS) w.colorHist[lid * w.colorGroups + blockIdx.x] = hist[lid];
  if (blockIdx.x == 0 && lid == 0) w.colorHist[COLOR_KEYS * w.colorGroups] = 0;
}

// Step 5, after the scan of the counts: where each bucket begins (the scanned count of its first block), how many
// colours are in use (one past the highest colour with a body in either of its buckets) and, when planIterations > 0,
// the dataflow solve's work list
__global__ void __launch_bounds__(COLOR_KEYS) k_color_starts(DeviceWorld w, uint32_t planIterations) {
  __shared__ uint32_t used;
  __shared__ uint32_t starts[COLOR_KEYS + 1];
  uint32_t k = threadIdx.x;
  if (k == 0) used = 0;
  __syncthreads();
  const uint32_t start = w.colorHist[k * w.colorGroups], next = w.colorHist[(k + 1) * w.colorGroups];
  w.colorStart[k] = start;
  starts[k] = start;
  if (k == 0) {
    w.colorStart[COLOR_KEYS] = w.colorHist[COLOR_KEYS * w.colorGroups];
    starts[COLOR_KEYS] = w.colorStart[COLOR_KEYS];
  }
  if (next > start) atomicMax(&used, k / 2 + 1);
  __syncthreads();
  if (k == 0) {
    w.counters[C_NUM_COLORS] = used;
    if (planIterations) writeDataflowPlan(w, starts, planIterations);
  }
}

// Step 6, placing: a body's slot is its block's scanned offset in its bucket plus its rank; with `warm`, also every
// body's warm start (placing reads no pose, and no joint has to be warm-started first)
template <bool warm> __global__ void __launch_bounds__(COLOR_WG) k_color_scatter(DeviceWorld w) {
  uint32_t b = blockIdx.x * COLOR_WG + threadIdx.x;
  if (b >= w.nBodies) return;
  if (dynamicBody(w, int(b))) {
    uint32_t col = w.colorA[b] & 0xffu;
    uint32_t slot = w.colorHist[bucketKey(w, b, col) * w.colorGroups + blockIdx.x] + w.colorRank[b];
    w.colorBodies[slot] = b;
    w.colorSlot[slot] = make_uint4(b, w.adjStart[b], umin(w.adjStart[b + 1], w.adjCapacity), w.bodyWorld[b]);
  }
  if (warm && !held(w)) warmBody(w, b);
}

// Step 7, for the dataflow solve: each dynamic partner's colour into bits 8-15 of its entry's flag word (ENTRY_COLOR_SHIFT),
// so the solve reads it with the entry; one thread per entry
__global__ void k_entry_colors(DeviceWorld w) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= umin(w.adjStart[w.nBodies], w.adjCapacity)) return;
  const uint4 head = w.adjEntry[2ull * e];
  if (head.z & ENTRY_DYNAMIC)
    w.adjEntry[2ull * e].z = (head.z & ((1u << ENTRY_COLOR_SHIFT) - 1u)) |
                             ((w.colorA[head.x] & 0xffu) << ENTRY_COLOR_SHIFT);
}

// --- launch wrappers ---------------------------------------------------------------------------------------------------

static inline uint32_t blocks(uint64_t n, uint32_t bs, uint32_t cap = 1u << 20) {
  uint64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return uint32_t(g);
}

void launch_adjacency(const DeviceWorld &w, uint32_t manifoldCapacity, hipStream_t s) {
  // adjStart holds the degrees (counted by the manifold compaction) and adjFill is clear
  uint64_t items = uint64_t(w.nJoints) + manifoldCapacity;
  launch_scan_exclusive(w.adjStart, w.nBodies + 1, w.scanScratch, s, w.counters + C_DF_ABORT);
  k_adj_fill<<<blocks(items, 256, 4096), 256, 0, s>>>(w);
  k_adj_sort<<<blocks(w.nBodies, ADJ_SORT_BLOCK), ADJ_SORT_BLOCK, 0, s>>>(w);
}

void launch_coloring(const DeviceWorld &w, uint32_t rounds, uint32_t planIterations, bool listedRounds,
                     hipStream_t s) {
  uint32_t g = blocks(w.nBodies, 256);
  k_color_compact<<<g, 256, 0, s>>>(w);
  k_color_mark<<<g, 256, 0, s>>>(w);
  if (listedRounds) {
    k_color_rounds_listed<<<1, 1024, 0, s>>>(w, rounds);
  } else {
    for (uint32_t r = 0; r < rounds; r++) {
      if (r % 2 == 0) k_color_round<<<g, 256, 0, s>>>(w, w.colorA, w.colorB, r);
      else k_color_round<<<g, 256, 0, s>>>(w, w.colorB, w.colorA, r);
    }
  }
  k_color_count<<<w.colorGroups, COLOR_WG, 0, s>>>(w);
  launch_scan_exclusive(w.colorHist, COLOR_KEYS * w.colorGroups + 1, w.scanScratch, s,
                        w.counters + C_DF_ABORT);
  k_color_starts<<<1, COLOR_KEYS, 0, s>>>(w, planIterations);
  if (w.nJoints) k_color_scatter<false><<<w.colorGroups, COLOR_WG, 0, s>>>(w);
  else k_color_scatter<true><<<w.colorGroups, COLOR_WG, 0, s>>>(w);
  if (planIterations) k_entry_colors<<<blocks(w.adjCapacity, 256), 256, 0, s>>>(w);
}

}  // namespace avbd
