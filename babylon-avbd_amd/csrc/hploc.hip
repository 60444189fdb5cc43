// H-PLOC bounding-volume hierarchy broadphase on gfx950: per-step leaf refit, Morton ordering (world id above the
// 30-bit code, so each world's leaves stay contiguous), a stable LSD radix sort, 16-cluster treelets merged by mutual
// nearest neighbours under the surface-area cost, level by level until one root per world, child-pair node records
// (both children's boxes in one 64-byte load) for a depth-first traversal with an LDS stack, and queries that feed the
// same pair predicate as the grid (so the pair set does not depend on the broadphase). Semantics of the reference's portable variant: src/gpu/hplocpp.wgsl:35-268, hplocSort.wgsl:11-48,
// gpuHploc.js:24-56 (findTreePairs), 207-375; rebuild every 64 steps, refit every step (docs/HPLOC.md:14-33).
//
// The level plan (which clusters each treelet merges, where its nodes go) depends only on the body count of each
// world, so the host computes it once and every level is one launch with 4 treelets per 64-lane wavefront.
#include "engine_types.h"
#include "hploc.h"
#include "kernels.h"
#include "pairtest.h"

#include <utility>

namespace avbd {

constexpr uint32_t END = 0xffffffffu;
constexpr uint32_t WIDTH = 16;

AVBD_D f3 fmin3(f3 a, f3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
AVBD_D f3 fmax3(f3 a, f3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }

/** World-space AABB half extents of a body (a sphere: its radius), as the pair test computes them. */
AVBD_D f3 bodyHalf(const Body &b) {
  f3 h = scale(xyz(b.size), 0.5f);
  if (b.angVel.w == SHAPE_SPHERE) return mk3(h.x, h.x, h.x);
  f3 a0 = qrotate(b.rot, mk3(1.f, 0.f, 0.f)), a1 = qrotate(b.rot, mk3(0.f, 1.f, 0.f)),
     a2 = qrotate(b.rot, mk3(0.f, 0.f, 1.f));
  return add(add(scale(vabs(a0), h.x), scale(vabs(a1), h.y)), scale(vabs(a2), h.z));
}

/** Leaf boxes from the current poses, padded so that float rounding never drops a pair the exact test accepts. */
AVBD_D void leafBox(const DeviceWorld &w, const Bvh &t, uint32_t i) {
  const Body &b = w.bodies[i];
  f3 p = xyz(b.pos);
  f3 h = bodyHalf(b);
  float pad = 1e-3f + 1e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z) + h.x + h.y + h.z);
  h = add(h, mk3(pad, pad, pad));
  t.nodes[i] = BvhNode{mk4(sub(p, h), __uint_as_float(i)), mk4(add(p, h), __uint_as_float(END))};
}

__global__ void k_bvh_leaves(DeviceWorld w, Bvh t) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < w.nBodies) leafBox(w, t, i);
}

AVBD_D uint32_t expandBits(uint32_t v) {
  uint32_t x = v & 1023u;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  return (x | (x << 2)) & 0x09249249u;
}

/** Per-world scene bounds of the leaf centres (atomic min/max on order-preserving integer encodings). */
AVBD_D uint32_t orderedBits(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
AVBD_D float fromOrdered(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__global__ void k_bvh_bounds_init(Bvh t, uint32_t nWorlds) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nWorlds) return;
  for (int c = 0; c < 3; c++) {
    t.bounds[6 * k + c] = 0xffffffffu;  // min
    t.bounds[6 * k + 3 + c] = 0u;       // max
  }
}

/** Threads per block of the bounds pass. */
constexpr uint32_t BOUNDS_BLOCK = 1024;

/**
 * Per-world bounds of the body centres (order-preserving bits, atomic min/max). A wavefront whose bodies share one
 * world reduces in registers, and a block whose wavefronts all share one then reduces in LDS: one block of a large
 * world makes 6 atomics instead of 6 per wavefront (the atomics on a world's six words serialise). Other wavefronts
 * make theirs per wavefront or per body.
 */
__global__ void __launch_bounds__(BOUNDS_BLOCK) k_bvh_bounds(DeviceWorld w, Bvh t) {
  __shared__ uint32_t waveWorld[BOUNDS_BLOCK / 64], waveLo[3][BOUNDS_BLOCK / 64], waveHi[3][BOUNDS_BLOCK / 64];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < w.nBodies;
  uint32_t k = live ? w.bodyWorld[i] : 0xffffffffu;
  float4 p = live ? w.bodies[i].pos : make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t lo[3] = {orderedBits(p.x), orderedBits(p.y), orderedBits(p.z)};
  uint32_t hi[3] = {lo[0], lo[1], lo[2]};
  if (!live) {
    lo[0] = lo[1] = lo[2] = 0xffffffffu;
    hi[0] = hi[1] = hi[2] = 0u;
  }
  const uint32_t wave = threadIdx.x / 64, waves = blockDim.x / 64;
  uint32_t k0 = __shfl(k, 0, 64);
  const bool uniform = __all(k == k0 || !live);
  if (uniform) {
    for (int d = 32; d > 0; d >>= 1)
      for (int a = 0; a < 3; a++) {
        lo[a] = umin(lo[a], __shfl_xor(lo[a], d, 64));
        hi[a] = umax(hi[a], __shfl_xor(hi[a], d, 64));
      }
  } else if (live) {
    for (int a = 0; a < 3; a++) {
      atomicMin(&t.bounds[6 * k + a], lo[a]);
      atomicMax(&t.bounds[6 * k + 3 + a], hi[a]);
    }
  }
  // Wavefront results to LDS (a mixed wavefront, or one without a live body, has world ~0: it is done)
  if ((threadIdx.x & 63) == 0) {
    waveWorld[wave] = uniform ? k0 : 0xffffffffu;
    for (int a = 0; a < 3; a++) {
      waveLo[a][wave] = lo[a];
      waveHi[a][wave] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x >= waves) return;
  const uint32_t wk = waveWorld[threadIdx.x];
  // The block's first lanes, one per wavefront: do all uniform wavefronts share one world?
  bool same = true;
  for (uint32_t v = 0; v < waves; v++) same = same && waveWorld[v] == waveWorld[0];
  if (same) {
    if (threadIdx.x != 0 || wk == 0xffffffffu) return;
    for (int a = 0; a < 3; a++) {
      uint32_t l = waveLo[a][0], h = waveHi[a][0];
      for (uint32_t v = 1; v < waves; v++) {
        l = umin(l, waveLo[a][v]);
        h = umax(h, waveHi[a][v]);
      }
      atomicMin(&t.bounds[6 * wk + a], l);
      atomicMax(&t.bounds[6 * wk + 3 + a], h);
    }
  } else if (wk != 0xffffffffu) {
    for (int a = 0; a < 3; a++) {
      atomicMin(&t.bounds[6 * wk + a], waveLo[a][threadIdx.x]);
      atomicMax(&t.bounds[6 * wk + 3 + a], waveHi[a][threadIdx.x]);
    }
  }
}

/** Key = world << 30 | 30-bit Morton code of the body's centre within its world's bounds; value = body index. */
__global__ void k_bvh_morton(DeviceWorld w, Bvh t) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  uint32_t k = w.bodyWorld[i];
  float4 p = w.bodies[i].pos;
  float c[3] = {p.x, p.y, p.z};
  uint32_t q[3];
  for (int a = 0; a < 3; a++) {
    float lo = fromOrdered(t.bounds[6 * k + a]), hi = fromOrdered(t.bounds[6 * k + 3 + a]);
    float ext = fmaxf(hi * 0.5f - lo * 0.5f, 1e-20f);
    float u = fminf(fmaxf((c[a] * 0.5f - lo * 0.5f) / ext, 0.f), 0.999999f);
    q[a] = uint32_t(u * 1024.f);
  }
  uint32_t code = (expandBits(q[0]) << 2) | (expandBits(q[1]) << 1) | expandBits(q[2]);
  t.keys[i] = (uint64_t(k) << 30) | code;
  t.vals[i] = i;
}

// --- stable LSD radix sort, 8-bit digits, 256 keys per tile ------------------------------------------------------

constexpr uint32_t SORT_TILE = 256;

__global__ void __launch_bounds__(SORT_TILE) k_sort_count(const uint64_t *keys, uint32_t n, uint32_t shift,
                                                          uint32_t tiles, uint32_t *hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t i = blockIdx.x * SORT_TILE + threadIdx.x;
  if (i < n) atomicAdd(&h[uint32_t(keys[i] >> shift) & 255u], 1u);
  __syncthreads();
  hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0) hist[256 * tiles] = 0;
}

__global__ void __launch_bounds__(SORT_TILE) k_sort_scatter(const uint64_t *keys, const uint32_t *vals, uint32_t n,
                                                            uint32_t shift, uint32_t tiles, const uint32_t *hist,
                                                            uint64_t *keysOut, uint32_t *valsOut) {
  __shared__ uint32_t digits[SORT_TILE];
  uint32_t i = blockIdx.x * SORT_TILE + threadIdx.x;
  uint32_t d = 256;  // past every digit: padding lanes never match
  uint64_t key = 0;
  if (i < n) {
    key = keys[i];
    d = uint32_t(key >> shift) & 255u;
  }
  digits[threadIdx.x] = d;
  __syncthreads();
  if (i >= n) return;
  uint32_t rank = 0;
  for (uint32_t j = 0; j < threadIdx.x; j++) rank += digits[j] == d ? 1u : 0u;
  uint32_t dst = hist[d * tiles + blockIdx.x] + rank;
  keysOut[dst] = key;
  valsOut[dst] = vals[i];
}

// --- treelets ---------------------------------------------------------------------------------------------------------

AVBD_D float area(f3 lo, f3 hi) {
  f3 d = mk3(fminf(fmaxf(hi.x - lo.x, 0.f), 1e18f), fminf(fmaxf(hi.y - lo.y, 0.f), 1e18f),
             fminf(fmaxf(hi.z - lo.z, 0.f), 1e18f));
  return d.x * d.y + d.y * d.z + d.z * d.x;
}

/**
 * One level: each group of up to 16 clusters (consecutive in Morton order, within one world) merges mutual nearest
 * neighbours (smallest union surface area, ties to the lower lane) until one cluster is left. A 64-lane block holds
 * four groups; the loop runs until every group of the block is done.
 */
__global__ void __launch_bounds__(64) k_bvh_level(Bvh t, const BvhGroup *groups, uint32_t nGroups,
                                                  const uint32_t *in, uint32_t *out) {
  __shared__ uint32_t alive[64], id[64], best[64], flag[64];
  __shared__ float4 lo[64], hi[64];
  __shared__ uint32_t created[4], survivors[4];
  const uint32_t lane = threadIdx.x % WIDTH, g = threadIdx.x / WIDTH;
  const uint32_t gi = blockIdx.x * 4 + g;
  const bool groupLive = gi < nGroups;
  BvhGroup grp = groupLive ? groups[gi] : BvhGroup{0, 0, 0, 0};
  const bool live = groupLive && lane < grp.count;
  alive[threadIdx.x] = live ? 1u : 0u;
  if (live) {
    uint32_t c = in[grp.inStart + lane];
    id[threadIdx.x] = c;
    BvhNode nd = t.nodes[c];
    lo[threadIdx.x] = nd.lo;
    hi[threadIdx.x] = nd.hi;
  }
  if (lane == 0) {
    created[g] = 0;
    survivors[g] = groupLive ? grp.count : 0;
  }
  __syncthreads();
  const uint32_t base = g * WIDTH;
  for (int round = 0; round < 64; round++) {
    bool any = false;
    for (int q = 0; q < 4; q++) any |= survivors[q] > 1;
    if (!any) break;
    uint32_t winner = END;
    float cost = 3.402823e38f;
    if (alive[threadIdx.x] && survivors[g] > 1) {
      f3 alo = xyz(lo[threadIdx.x]), ahi = xyz(hi[threadIdx.x]);
      for (uint32_t j = 0; j < WIDTH; j++) {
        if (j == lane || !alive[base + j]) continue;
        float c = area(fmin3(alo, xyz(lo[base + j])), fmax3(ahi, xyz(hi[base + j])));
        if (c < cost || (c == cost && j < winner)) {
          cost = c;
          winner = j;
        }
      }
    }
    best[threadIdx.x] = winner;
    __syncthreads();
    flag[threadIdx.x] = (alive[threadIdx.x] && winner != END && lane < winner && best[base + winner] == lane) ? 1u : 0u;
    __syncthreads();
    uint32_t rank = 0, merged = 0;
    for (uint32_t j = 0; j < WIDTH; j++) {
      if (j < lane) rank += flag[base + j];
      merged += flag[base + j];
    }
    uint32_t aID = 0, bID = 0;
    float4 nlo = make_float4(0, 0, 0, 0), nhi = nlo;
    if (flag[threadIdx.x]) {
      aID = id[threadIdx.x];
      bID = id[base + winner];
      f3 l = fmin3(xyz(lo[threadIdx.x]), xyz(lo[base + winner]));
      f3 h = fmax3(xyz(hi[threadIdx.x]), xyz(hi[base + winner]));
      nlo = mk4(l, __uint_as_float(aID));
      nhi = mk4(h, __uint_as_float(bID));
    }
    __syncthreads();
    if (flag[threadIdx.x]) {
      uint32_t nid = grp.nodeBase + created[g] + rank;
      t.nodes[nid] = BvhNode{nlo, nhi};
      t.parent[aID] = nid;
      t.parent[bID] = nid;
      id[threadIdx.x] = nid;
      lo[threadIdx.x] = nlo;
      hi[threadIdx.x] = nhi;
      alive[base + winner] = 0;
    }
    __syncthreads();
    if (lane == 0 && groupLive) {
      created[g] += merged;
      survivors[g] -= merged;
    }
    __syncthreads();
  }
  if (live && alive[threadIdx.x]) {
    out[grp.outIndex] = id[threadIdx.x];
    t.parent[id[threadIdx.x]] = END;  // a root until the next level links it
  }
}

/** Leaf ranks for the traversal's pruning: a dynamic leaf's position in Morton order, a static leaf's flag. */
__global__ void k_bvh_rank(DeviceWorld w, Bvh t) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < w.nBodies) t.rank[t.order[q]] = q;
}
__global__ void k_bvh_leafmax(DeviceWorld w, Bvh t) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < w.nBodies) t.maxPos[i] = w.bodies[i].size.w > 0.f ? t.rank[i] : MAXPOS_STATIC;
}
/** A refit's leaves: boxes from the current poses and maxPos from the last build's ranks. */
__global__ void k_bvh_refit_leaves(DeviceWorld w, Bvh t) {
  if (gatedOff(w)) return;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies) return;
  leafBox(w, t, i);
  t.maxPos[i] = w.bodies[i].size.w > 0.f ? t.rank[i] : MAXPOS_STATIC;
}

AVBD_D uint32_t combineMaxPos(uint32_t a, uint32_t b) {
  return umax(a & ~MAXPOS_STATIC, b & ~MAXPOS_STATIC) | ((a | b) & MAXPOS_STATIC);
}

/** Tag of a child in a pair record: its node id, or for a leaf its body index | LEAF (leaf ids are body indices). */
constexpr uint32_t LEAF = 0x80000000u;

/**
 * Refit one level, a 16-lane group per treelet (four a block): the lanes read their nodes' child ids and the boxes and
 * maxPos of children outside the treelet in parallel; the lanes then refit the nodes in LDS in rounds (a node once its
 * children are); the lanes write the nodes, their maxPos and their pair records.
 */
__global__ void __launch_bounds__(64) k_bvh_refit_treelets(Bvh t, const BvhGroup *groups, uint32_t nGroups,
                                                           uint32_t nLeaves) {
  if (t.gate && *t.gate == 0u) return;
  __shared__ float4 nodeLo[64], nodeHi[64];         // refitted box of each node slot
  __shared__ uint32_t nodeMp[64];
  __shared__ float4 childLo[2][64], childHi[2][64];  // per node slot: its two children's boxes and maxPos
  __shared__ uint32_t childMp[2][64], childId[2][64];
  __shared__ uint32_t refitted[64];                  // per node slot: refitted in an earlier round
  const uint32_t lane = threadIdx.x % WIDTH, g = threadIdx.x / WIDTH, slot = threadIdx.x;
  const uint32_t gi = blockIdx.x * 4 + g;
  const bool groupLive = gi < nGroups;
  const BvhGroup grp = groupLive ? groups[gi] : BvhGroup{0, 0, 0, 0};
  const uint32_t nodes = grp.count > 0 ? grp.count - 1 : 0;  // a treelet of k clusters holds k - 1 nodes
  const bool live = groupLive && lane < nodes;
  if (live) {
    const BvhNode n = t.nodes[grp.nodeBase + lane];
    const uint32_t ids[2] = {__float_as_uint(n.lo.w), __float_as_uint(n.hi.w)};
#pragma unroll
    for (int k = 0; k < 2; k++) {
      childId[k][slot] = ids[k];
      if (ids[k] - grp.nodeBase >= nodes) {  // outside the treelet (a leaf or a lower level's root)
        const BvhNode c = t.nodes[ids[k]];
        childLo[k][slot] = c.lo;
        childHi[k][slot] = c.hi;
        childMp[k][slot] = t.maxPos[ids[k]];
      }
    }
  }
  // Rounds, a lane per node: a node whose children are refitted (outside children always are) takes their boxes and
  // is refitted; the treelet's depth in rounds instead of its node count in sequence (min and max are exact, so the
  // order changes no bit)
  const uint32_t base = g * WIDTH;
  bool done = !live;
  refitted[slot] = 0u;
  __syncthreads();
  for (uint32_t round = 0; round < WIDTH; round++) {
    bool ready = !done;
    uint32_t in[2] = {0u, 0u};
    if (!done) {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        in[k] = childId[k][slot] - grp.nodeBase;
        if (in[k] < nodes && !refitted[base + in[k]]) ready = false;
      }
    }
    if (ready) {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (in[k] < nodes) {  // a node of the treelet (refitted in an earlier round)
          childLo[k][slot] = nodeLo[base + in[k]];
          childHi[k][slot] = nodeHi[base + in[k]];
          childMp[k][slot] = nodeMp[base + in[k]];
        }
      }
    }
    __syncthreads();  // every read of this round before any write
    if (ready) {
      nodeLo[slot] = mk4(fmin3(xyz(childLo[0][slot]), xyz(childLo[1][slot])), 0.f);
      nodeHi[slot] = mk4(fmax3(xyz(childHi[0][slot]), xyz(childHi[1][slot])), 0.f);
      nodeMp[slot] = combineMaxPos(childMp[0][slot], childMp[1][slot]);
      refitted[slot] = 1u;
      done = true;
    }
    if (!__syncthreads_or(!done)) break;  // also orders this round's writes before the next round's reads
  }
  __syncthreads();
  if (live) {
    const uint32_t nid = grp.nodeBase + lane;
    const uint32_t l = childId[0][slot], r = childId[1][slot];
    t.nodes[nid] = BvhNode{mk4(xyz(nodeLo[slot]), __uint_as_float(l)), mk4(xyz(nodeHi[slot]), __uint_as_float(r))};
    t.maxPos[nid] = nodeMp[slot];
    const uint32_t tagA = l < nLeaves ? (l | LEAF) : l, tagB = r < nLeaves ? (r | LEAF) : r;
    t.pairs[nid] = BvhPair{mk4(xyz(childLo[0][slot]), __uint_as_float(tagA)),
                           mk4(xyz(childHi[0][slot]), __uint_as_float(childMp[0][slot])),
                           mk4(xyz(childLo[1][slot]), __uint_as_float(tagB)),
                           mk4(xyz(childHi[1][slot]), __uint_as_float(childMp[1][slot]))};
  }
}

/** The quads of the internal nodes [N, nodeCount) from their pair records (after a refit wrote them). */
__global__ void k_bvh_quads(Bvh t, uint32_t nLeaves) {
  if (t.gate && *t.gate == 0u) return;
  const uint32_t id = nLeaves + blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= t.nodeCount) return;
  const BvhPair p = t.pairs[id];
  const float4 emptyLo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), emptyHi = make_float4(-INFINITY, -INFINITY,
                                                                                              -INFINITY, 0.f);
  BvhQuad qd;
  const uint32_t tagA = __float_as_uint(p.aLo.w), tagB = __float_as_uint(p.bLo.w);
  if (tagA & LEAF) {
    qd.lo[0] = p.aLo, qd.hi[0] = p.aHi, qd.lo[1] = emptyLo, qd.hi[1] = emptyHi;
  } else {
    const BvhPair c = t.pairs[tagA];
    qd.lo[0] = c.aLo, qd.hi[0] = c.aHi, qd.lo[1] = c.bLo, qd.hi[1] = c.bHi;
  }
  if (tagB & LEAF) {
    qd.lo[2] = p.bLo, qd.hi[2] = p.bHi, qd.lo[3] = emptyLo, qd.hi[3] = emptyHi;
  } else {
    const BvhPair c = t.pairs[tagB];
    qd.lo[2] = c.aLo, qd.hi[2] = c.aHi, qd.lo[3] = c.bLo, qd.hi[3] = c.bHi;
  }
  t.quads[id - nLeaves] = qd;
}

/** A query's traversal stack in LDS: entry d of lane x at stack[d * blockDim.x + x] (no bank conflicts). */
extern __shared__ uint32_t bvhStack[];

/**
 * Walk body i's world tree depth first and call visit(j) for each leaf j whose padded box overlaps i's
 * (gpuHploc.js:37-55). Each step is one 128-byte quad load holding up to four boxes two levels down: a leaf entry is
 * visited on the spot, the first internal entry descended into and the others kept on the stack. Every static leaf is
 * visited, but of two dynamic bodies only the one earlier in Morton order (rank q) visits the other: subtrees holding
 * no static leaf and no dynamic leaf later than q are skipped (maxPos), so each dynamic pair comes up once. A one-body
 * world has no pair.
 */
template <class F> AVBD_D void walkTree(const DeviceWorld &w, const Bvh &t, uint32_t i, uint32_t q, F &&visit) {
  const BvhNode self = t.nodes[i];
  const f3 qlo = xyz(self.lo), qhi = xyz(self.hi);
  const uint32_t root = t.roots[w.bodyWorld[i]];
  const uint32_t N = w.nBodies;
  if (root < N) return;
  auto hit = [&](float4 lo, float4 hi) {
    const uint32_t mp = __float_as_uint(hi.w);
    return !(hi.x < qlo.x || hi.y < qlo.y || hi.z < qlo.z || lo.x > qhi.x || lo.y > qhi.y || lo.z > qhi.z) &&
           ((mp & MAXPOS_STATIC) || (mp & ~MAXPOS_STATIC) > q);
  };
  uint32_t *stack = bvhStack + threadIdx.x;
  uint32_t at = root, sp = 0;
  // Each internal node is entered at most once; the bound only guards against a malformed tree
  for (uint32_t guard = 0; guard < t.nodeCount; guard++) {
    const BvhQuad nd = t.quads[at - N];
    uint32_t next = END;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (!hit(nd.lo[k], nd.hi[k])) continue;
      const uint32_t tag = __float_as_uint(nd.lo[k].w);
      if (tag & LEAF) visit(tag & ~LEAF);
      else if (next == END) next = tag;
      else if (sp < t.stackDepth) stack[(sp++) * blockDim.x] = tag;
    }
    if (next != END) {
      at = next;
    } else {
      if (sp == 0) break;
      at = stack[(--sp) * blockDim.x];
    }
  }
}

/** Lanes per query in the candidate pass: one per entry of a quad. */
constexpr uint32_t WALK_LANES = 4;

/** The votes of this lane's group (lanes 4g..4g+3 of the wavefront) as bits 0..3, lane 4g + k at bit k. */
AVBD_D uint32_t groupVote(bool b) {
  return uint32_t(__ballot(b) >> (__lane_id() & ~(WALK_LANES - 1u))) & ((1u << WALK_LANES) - 1u);
}

/**
 * walkTree with a quad's four entries on four lanes (the query's group: lanes 4g..4g+3 of the wavefront): each lane
 * loads its own entry (32 bytes; the group reads the quad in one 128-byte access) and tests it, so a wavefront walks
 * 16 trees and four times as many walks are in flight as with a lane per query. The group takes the same path as
 * walkTree: leaf entries visited in entry order, the first internal one descended into, the others pushed in entry
 * order onto the group's stack (in LDS, entry d at stack[d * groups + group]). visit(leaf, j, rank, count) is called
 * on all four lanes for each quad: leaf says whether this lane's entry is a leaf hit, j its body, rank the number of
 * leaf hits before it in the quad, count the quad's leaf hits (the same on every lane); the callback may use ballots.
 */
template <class F>
AVBD_D void walkTreeLanes(const DeviceWorld &w, const Bvh &t, uint32_t i, uint32_t q, F &&visit) {
  const BvhNode self = t.nodes[i];
  const f3 qlo = xyz(self.lo), qhi = xyz(self.hi);
  const uint32_t root = t.roots[w.bodyWorld[i]];
  const uint32_t N = w.nBodies;
  if (root < N) return;
  const uint32_t sub = threadIdx.x % WALK_LANES, first = __lane_id() - sub, below = (1u << sub) - 1u;
  const uint32_t groups = blockDim.x / WALK_LANES;
  uint32_t *stack = bvhStack + threadIdx.x / WALK_LANES;
  uint32_t at = root, sp = 0;
  // Each internal node is entered at most once; the bound only guards against a malformed tree
  for (uint32_t guard = 0; guard < t.nodeCount; guard++) {
    const BvhQuad *quad = t.quads + (at - N);
    const float4 lo = quad->lo[sub], hi = quad->hi[sub];
    const uint32_t mp = __float_as_uint(hi.w), tag = __float_as_uint(lo.w);
    const bool hit = !(hi.x < qlo.x || hi.y < qlo.y || hi.z < qlo.z || lo.x > qhi.x || lo.y > qhi.y || lo.z > qhi.z) &&
                     ((mp & MAXPOS_STATIC) || (mp & ~MAXPOS_STATIC) > q);
    const bool leaf = hit && (tag & LEAF) != 0, inner = hit && (tag & LEAF) == 0;
    const uint32_t leafBits = groupVote(leaf), innerBits = groupVote(inner);
    visit(leaf, tag & ~LEAF, uint32_t(__popc(leafBits & below)), uint32_t(__popc(leafBits)));
    if (innerBits != 0u) {
      const uint32_t head = uint32_t(__builtin_ctz(innerBits));
      // Internal hits after the first: pushed in entry order, as many as the stack holds
      if (inner && sub != head) {
        const uint32_t d = sp + uint32_t(__popc(innerBits & below)) - 1u;
        if (d < t.stackDepth) stack[d * groups] = tag;
      }
      sp = umin(sp + uint32_t(__popc(innerBits)) - 1u, t.stackDepth);
      at = uint32_t(__shfl(int(tag), int(first + head)));
    } else {
      if (sp == 0) break;
      __builtin_amdgcn_wave_barrier();  // the pushes above (other lanes of the group) before this read
      at = stack[(--sp) * groups];
    }
  }
}

/**
 * Pass 1: each dynamic body's query walks the tree with the box test only and keeps its first BVH_CAND candidate
 * leaves, four lanes per query (walkTreeLanes); then the same four lanes run the pair predicate on the kept
 * candidates, four at a time, writing the accepted mask (bit k of word r: candidate 32 r + k passed) and the pair
 * count. A query with more candidates than it keeps walks again, testing as it goes, and only counts. Groups take
 * bodies in Morton order (neighbouring groups walk neighbouring paths).
 */
__global__ void k_bvh_cand(DeviceWorld w, Bvh t) {
  if (gatedOff(w)) return;
  const uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) / WALK_LANES;
  if (q >= w.nBodies) return;
  const uint32_t sub = threadIdx.x % WALK_LANES;
  const uint32_t i = t.order[q];
  uint32_t *stash = t.candStash + size_t(i) * BVH_CAND;
  uint32_t n = 0;
  if (w.bodies[i].size.w > 0.f) {
    walkTreeLanes(w, t, i, q, [&](bool leaf, uint32_t j, uint32_t rank, uint32_t count) {
      if (leaf && n + rank < BVH_CAND) stash[n + rank] = j;
      n += count;
    });
  }
  uint32_t pairs = 0;
  if (n <= BVH_CAND) {
    // The group's stash writes (any of its lanes) before its reads
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    uint32_t word = 0;
    for (uint32_t base = 0; base < n; base += WALK_LANES) {
      const uint32_t slot = base + sub;
      bool accept = false;
      if (slot < n) {
        const uint32_t j = stash[slot];
        accept = acceptPair(w, umax(i, j), umin(i, j));
      }
      const uint32_t bits = groupVote(accept);
      word |= bits << (base % 32u);
      pairs += uint32_t(__popc(bits));
      const bool full = base % 32u == 32u - WALK_LANES;
      if (sub == 0 && (full || base + WALK_LANES >= n)) t.candMask[size_t(i) * BVH_CAND_WORDS + base / 32u] = word;
      if (full) word = 0;
    }
  } else {
    walkTreeLanes(w, t, i, q, [&](bool leaf, uint32_t j, uint32_t, uint32_t) {
      pairs += uint32_t(__popc(groupVote(leaf && acceptPair(w, umax(i, j), umin(i, j)))));
    });
  }
  if (sub == 0) {
    t.candCount[i] = n;
    w.pairCount[i] = pairs;
  }
}

/**
 * Pass 2 (after the scan of the counts): write each body's pairs (a > b) in candidate order; and the pair count,
 * clamped to the capacity (overflow bit 0 when it is not).
 */
__global__ void k_bvh_write(DeviceWorld w, Bvh t) {
  if (gatedOff(w)) return;
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0) {
    uint32_t total = w.pairCount[w.nBodies];
    w.counters[C_PAIR_DEMAND] = total;
    if (total > w.pairCapacity) {
      atomicOr(&w.counters[C_OVERFLOW], 1u);
      total = w.pairCapacity;
    }
    w.counters[C_PAIRS] = total;
  }
  if (q >= w.nBodies) return;
  uint32_t i = t.order[q];
  uint32_t base = w.pairCount[i];
  uint32_t count = t.candCount[i];
  uint32_t n = 0;
  if (count <= BVH_CAND) {
    for (uint32_t r = 0; r * 32 < count; r++) {
      uint32_t bits = t.candMask[size_t(i) * BVH_CAND_WORDS + r];
      while (bits) {
        uint32_t slot = r * 32 + uint32_t(__builtin_ctz(bits));
        bits &= bits - 1u;
        uint32_t j = t.candStash[size_t(i) * BVH_CAND + slot];
        if (base + n < w.pairCapacity) w.pairs[base + n] = make_uint2(umax(i, j), umin(i, j));
        n++;
      }
    }
    return;
  }
  walkTree(w, t, i, q, [&](uint32_t j) {
    uint32_t a = umax(i, j), b = umin(i, j);
    if (acceptPair(w, a, b)) {
      if (base + n < w.pairCapacity) w.pairs[base + n] = make_uint2(a, b);
      n++;
    }
  });
}

// --- launch wrappers ---------------------------------------------------------------------------------------------------

static inline uint32_t blocks(uint64_t n, uint32_t bs) {
  uint64_t g = (n + bs - 1) / bs;
  return uint32_t(g < 1 ? 1 : g);
}

__global__ void k_bvh_copy(uint32_t *dst, const uint32_t *src, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

void launch_bvh_build(const DeviceWorld &w, const Bvh &tree, const BvhPlan &plan, hipStream_t s) {
  Bvh t = tree;
  t.gate = nullptr;
  const uint32_t n = w.nBodies;
  launch_fill_u32(t.parent, END, plan.nodeCount, s);
  k_bvh_leaves<<<blocks(n, 256), 256, 0, s>>>(w, t);
  k_bvh_bounds_init<<<blocks(w.nWorlds, 256), 256, 0, s>>>(t, w.nWorlds);
  k_bvh_bounds<<<blocks(n, BOUNDS_BLOCK), BOUNDS_BLOCK, 0, s>>>(w, t);
  k_bvh_morton<<<blocks(n, 256), 256, 0, s>>>(w, t);
  // Stable LSD radix sort of (key, body)
  const uint32_t tiles = blocks(n, SORT_TILE);
  uint64_t *k0 = t.keys, *k1 = t.keysAlt;
  uint32_t *v0 = t.vals, *v1 = t.valsAlt;
  for (uint32_t pass = 0; pass < plan.sortPasses; pass++) {
    uint32_t shift = 8 * pass;
    k_sort_count<<<tiles, SORT_TILE, 0, s>>>(k0, n, shift, tiles, t.sortHist);
    launch_scan_exclusive(t.sortHist, 256 * tiles + 1, w.scanScratch, s, w.counters + C_DF_ABORT);
    k_sort_scatter<<<tiles, SORT_TILE, 0, s>>>(k0, v0, n, shift, tiles, t.sortHist, k1, v1);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  // Leaves in Morton order are the first level's clusters; the plan's levels merge them to one root per world
  const uint32_t *in = v0;
  for (size_t L = 0; L < plan.levels.size(); L++) {
    const BvhLevel &lev = plan.levels[L];
    uint32_t *out = (L % 2 == 0) ? t.clustersA : t.clustersB;
    k_bvh_level<<<blocks(lev.groupCount, 4), 64, 0, s>>>(t, t.groups + lev.groupOffset, lev.groupCount, in, out);
    in = out;
  }
  // The last level's outputs are the world roots (a one-body world's root is its leaf: its plan has a copy level)
  k_bvh_copy<<<blocks(w.nWorlds, 256), 256, 0, s>>>(t.roots, in, w.nWorlds);
  k_bvh_copy<<<blocks(n, 256), 256, 0, s>>>(t.order, v0, n);
  k_bvh_rank<<<blocks(n, 256), 256, 0, s>>>(w, t);
  k_bvh_leafmax<<<blocks(n, 256), 256, 0, s>>>(w, t);
  for (const BvhLevel &lev : plan.levels)
    k_bvh_refit_treelets<<<blocks(lev.groupCount, 4), 64, 0, s>>>(t, t.groups + lev.groupOffset, lev.groupCount, n);
  if (plan.nodeCount > n) k_bvh_quads<<<blocks(plan.nodeCount - n, 256), 256, 0, s>>>(t, n);
}

void launch_bvh_refit(const DeviceWorld &w, const Bvh &tree, const BvhPlan &plan, hipStream_t s) {
  Bvh t = tree;
  t.gate = w.gate;
  k_bvh_refit_leaves<<<blocks(w.nBodies, 256), 256, 0, s>>>(w, t);
  for (const BvhLevel &lev : plan.levels)
    k_bvh_refit_treelets<<<blocks(lev.groupCount, 4), 64, 0, s>>>(t, t.groups + lev.groupOffset, lev.groupCount,
                                                                   w.nBodies);
  if (plan.nodeCount > w.nBodies)
    k_bvh_quads<<<blocks(plan.nodeCount - w.nBodies, 256), 256, 0, s>>>(t, w.nBodies);
}

void launch_bvh_pairs(const DeviceWorld &w, const Bvh &tree, hipStream_t s) {
  Bvh t = tree;
  t.gate = w.gate;
  const size_t stackBytes = size_t(64) * t.stackDepth * sizeof(uint32_t);
  k_bvh_cand<<<blocks(uint64_t(w.nBodies) * WALK_LANES, 256), 256, stackBytes, s>>>(w, t);
  launch_scan_exclusive(w.pairCount, w.nBodies + 1, w.scanScratch, s, w.counters + C_DF_ABORT, w.gate);
  k_bvh_write<<<blocks(w.nBodies, 64), 64, stackBytes, s>>>(w, t);
}

}  // namespace avbd
