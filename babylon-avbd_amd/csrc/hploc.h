// Host/device types of the H-PLOC broadphase: node buffers, sort buffers and the level plan.
#pragma once
#include <cstdint>
#include <vector>

#include "engine_types.h"

namespace avbd {

/** A node: lo.xyz + left child (leaf: body index), hi.xyz + right child (leaf: ~0), children as u32 bits. */
struct BvhNode {
  float4 lo, hi;
};

/** An internal node as the pair search reads it: both children's boxes, tags (node id, or body | LEAF) in the lo.w
 * words and subtree maxPos in the hi.w words, so one 64-byte load decides both descents. */
struct BvhPair {
  float4 aLo, aHi, bLo, bHi;
};

/** An internal node as the traversal reads it, with its grandchildren lifted: four entries (a leaf child itself, or
 * an internal child's two children; an unused entry's box is empty), each box with its tag (node id, or body | LEAF)
 * in lo.w and its subtree maxPos in hi.w. One 128-byte load decides up to four descents, two tree levels at a time. */
struct BvhQuad {
  float4 lo[4], hi[4];
};

/** One treelet of a level: its input clusters [inStart, inStart + count), output slot, first node it creates. */
struct BvhGroup {
  uint32_t inStart, count, outIndex, nodeBase;
};

struct BvhLevel {
  uint32_t groupOffset, groupCount;
};

/** Level plan: from the body count of each world (static), the groups of every level down to one root per world. */
struct BvhPlan {
  std::vector<BvhLevel> levels;
  std::vector<BvhGroup> groups;
  uint32_t nodeCount = 0;
  uint32_t sortPasses = 4;
  uint32_t height = 1;  // bound on the tree height: a level's treelet of k clusters adds at most k - 1 levels
};

struct Bvh {
  const uint32_t *gate;  // set by the launchers from the DeviceWorld: a gated launch returns while *gate == 0
  BvhNode *nodes;      // leaves [0, N), internal nodes [N, 2N - W)
  BvhPair *pairs;      // per internal node: its children's boxes, tags and maxPos (leaves' entries unused)
  BvhQuad *quads;      // per internal node id - N: its four lifted entries (what the pair search walks)
  uint32_t *order;     // bodies in Morton order (the last build's)
  uint32_t *rank;      // per body: its position in `order`
  uint32_t *maxPos;    // per node: highest rank of a dynamic leaf below it, | MAXPOS_STATIC if a static leaf is below
  uint32_t *candStash; // per body: up to BVH_CAND candidate leaves of its query (boxes overlap), in traversal order
  uint32_t *candCount; // per body: its query's candidate count (may exceed BVH_CAND)
  uint32_t *candMask;  // per body: BVH_CAND / 32 words, bit k = candidate k passed the pair predicate
  uint32_t *parent;    // per node
  uint64_t *keys, *keysAlt;
  uint32_t *vals, *valsAlt;
  uint32_t *clustersA, *clustersB;
  uint32_t *roots;     // per world
  uint32_t *bounds;    // per world: min xyz, max xyz (order-preserving u32)
  uint32_t *sortHist;  // 256 × tiles + 1
  BvhGroup *groups;    // every level's groups
  uint32_t nodeCount;
  uint32_t stackDepth; // traversal stack entries per query (three per two levels of the plan's height bound)
};

/** Candidate leaves kept per query; a query with more is answered by a second walk that tests as it goes. */
constexpr uint32_t BVH_CAND = 128;
constexpr uint32_t BVH_CAND_WORDS = BVH_CAND / 32;

/** maxPos flag: the subtree holds a static leaf (every dynamic query must look there). */
constexpr uint32_t MAXPOS_STATIC = 0x80000000u;

/** The plan for worlds of the given body counts (bodies of world k follow those of worlds 0..k-1). */
inline BvhPlan makeBvhPlan(const std::vector<uint32_t> &counts) {
  BvhPlan plan;
  uint32_t n = 0, w = uint32_t(counts.size());
  for (uint32_t c : counts) n += c;
  plan.nodeCount = 2 * n - w;
  uint32_t worldBits = 0;
  while ((1u << worldBits) < w) worldBits++;
  plan.sortPasses = (30 + worldBits + 7) / 8;
  std::vector<uint32_t> c(counts.begin(), counts.end());
  uint32_t nodeBase = n;
  auto more = [&] {
    for (uint32_t x : c)
      if (x > 1) return true;
    return false;
  };
  while (more()) {
    BvhLevel lev{uint32_t(plan.groups.size()), 0};
    uint32_t inOff = 0, outOff = 0;
    std::vector<uint32_t> next(c.size());
    uint32_t widest = 1;
    for (size_t k = 0; k < c.size(); k++) {
      uint32_t g = (c[k] + 15) / 16;
      for (uint32_t q = 0; q < g; q++) {
        uint32_t count = std::min<uint32_t>(16, c[k] - 16 * q);
        plan.groups.push_back(BvhGroup{inOff + 16 * q, count, outOff + q, nodeBase});
        nodeBase += count - 1;
        lev.groupCount++;
        widest = std::max(widest, count);
      }
      inOff += c[k];
      outOff += g;
      next[k] = g;
    }
    plan.levels.push_back(lev);
    plan.height += widest - 1;
    c = next;
  }
  return plan;
}

void launch_bvh_build(const DeviceWorld &w, const Bvh &t, const BvhPlan &plan, hipStream_t s);
void launch_bvh_refit(const DeviceWorld &w, const Bvh &t, const BvhPlan &plan, hipStream_t s);
void launch_bvh_pairs(const DeviceWorld &w, const Bvh &t, hipStream_t s);

}  // namespace avbd
