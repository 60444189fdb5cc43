// Launch wrappers of the step's kernels (each .hip file launches its own kernels on the given stream).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_types.h"

namespace avbd {

/**
 * In-place exclusive prefix sum of n u32 values in one launch. `scratch` holds at least scan_scratch_words(n) words,
 * all zero before its first scan (each scan leaves it zero again); scans that share it run one after another on one
 * stream. A wait given up (never expected) sets *abortFlag, or scratch word 2 when abortFlag is null.
 */
void launch_scan_exclusive(uint32_t *data, uint32_t n, uint32_t *scratch, hipStream_t s,
                           uint32_t *abortFlag = nullptr, const uint32_t *gate = nullptr);
uint32_t scan_scratch_words(uint32_t n);
/** Fill n words with a value (a kernel, so it can be captured in a graph). */
void launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t s);

// Collision
void launch_begin_frame(const DeviceWorld &w, hipStream_t s);
/** Contact-reuse references (moved bodies restamped) and the last step's manifolds by pair in the hash table. */
void launch_refs_and_hash(const DeviceWorld &w, uint32_t prevCapacity, hipStream_t s);
void launch_grid(const DeviceWorld &w, hipStream_t s);
void launch_pairs(const DeviceWorld &w, hipStream_t s);
/** C_PAIRS from the scanned per-body counts (capacity-clamped, overflow bit 0). */
void launch_pairs_total(const DeviceWorld &w, hipStream_t s);
/** Contacts of every pair, then the dense manifold and point lists (and the adjacency counters cleared). */
void launch_narrowphase(const DeviceWorld &w, hipStream_t s);

// Topology and colouring
/** Per-body adjacency (after launch_narrowphase, which clears its counters). */
void launch_adjacency(const DeviceWorld &w, uint32_t manifoldCapacity, hipStream_t s);
/**
 * Colouring and the per-colour buckets; with planIterations > 0 also the dataflow solve's work list. In a world
 * without joints the last pass also warm-starts the bodies (launch_warm_start then has nothing left to do). With
 * listedRounds the Jones–Plassmann rounds run in one block over the pending bodies (one launch instead of `rounds`;
 * the same result, for steps that leave few bodies pending).
 */
void launch_coloring(const DeviceWorld &w, uint32_t rounds, uint32_t planIterations, bool listedRounds,
                     hipStream_t s);

// Solve
void launch_warm_start(const DeviceWorld &w, hipStream_t s);
void launch_primal(const DeviceWorld &w, uint32_t color, uint32_t maxColorBodies, hipStream_t s);
void launch_dual(const DeviceWorld &w, uint32_t manifoldCapacity, hipStream_t s);
void launch_velocities(const DeviceWorld &w, hipStream_t s);
/** Restitution (src/gpu/restitution.js): incoming velocities before the step; bounce rounds after it. */
void launch_restitution_before(const DeviceWorld &w, hipStream_t s);
void launch_restitution_after(const DeviceWorld &w, hipStream_t s);

/** Shape queries: every (body, query) pair traced; hits = 2 float4 per query (normal, distance or -1; index bits).
 * `best` holds 2 words per query, reset by launch_query_reset before first use and left reset by each batch. */
void launch_query_reset(uint32_t *best, uint32_t capacity, hipStream_t s);
/** Queries per castNoAlloc call: their records (32 B), best words (8 B) and hits (32 B) live in a file-scope device
 * array of queries.hip (NOALLOC_QUERY_BYTES per query), one copy per device, shared by every call on that device. */
constexpr uint32_t NOALLOC_MAX_QUERIES = 16384;
constexpr uint32_t NOALLOC_QUERY_BYTES = sizeof(QueryRec) + 2 * sizeof(uint32_t) + 2 * sizeof(float4);
/** The device address of that array on `device` (looked up once per device, with the device made current); throws
 * std::runtime_error when the lookup fails. */
void *query_scratch(int device);
void launch_queries(const DeviceWorld &w, const QueryRec *queries, uint32_t n, const QueryFilter &f, uint32_t *best,
                    float4 *hits, hipStream_t s);

/** Contact events of the step just taken (after its solve); `now` alternates with the step parity. */
void launch_contact_events(const DeviceWorld &w, const ContactWatchDev &c, uint32_t now, hipStream_t s);

/** Sleeping: state set-up and edits (outside the step), the step's passes before it, after its narrowphase (impacts,
 * then the collision refresh's preparation, gated on bodies having woken) and after it. */
void launch_sleep_init(const DeviceWorld &w, const uint8_t *eligible, uint32_t first, uint32_t count, hipStream_t s);
void launch_sleep_eligible(const DeviceWorld &w, const uint32_t *words, uint32_t n, hipStream_t s);
void launch_sleep_hold(const DeviceWorld &w, const uint32_t *indices, uint32_t n, bool islands, hipStream_t s);
void launch_sleep_before(const DeviceWorld &w, const SleepParams &p, hipStream_t s);
void launch_sleep_impact(const DeviceWorld &w, const SleepParams &p, hipStream_t s);
void launch_sleep_refresh_prepare(const DeviceWorld &w, hipStream_t s);
void launch_sleep_after(const DeviceWorld &w, const SleepParams &p, hipStream_t s);

/** Blocks of the dataflow solve that are resident at once on the current device (0 if unknown). */
uint32_t dataflow_grid();
/**
 * Every iteration's primal and dual in one launch of `grid` blocks (at most dataflow_grid()), after a launch_coloring
 * that wrote the work list; then the duals of joints without a dynamic body.
 */
void launch_dataflow(const DeviceWorld &w, uint32_t grid, uint32_t iterations, hipStream_t s);

/**
 * A batch's solve with one workgroup of `block` threads (128, 256, 512 or 1024) per world: every iteration's colours,
 * the contact duals (each fused into the pair's later colour), the joint duals and the velocities, after a
 * launch_coloring. `order` lists the worlds in launch order, `keyStart` holds (worlds + 1) × COLOR_KEYS words (written
 * here), `jointStart` (worlds + 1 words: each world's first joint slot) is null when the batch has no joint.
 */
void launch_solve_worlds(const DeviceWorld &w, const uint32_t *order, uint32_t *keyStart, const uint32_t *jointStart,
                         uint32_t iterations, uint32_t block, hipStream_t s);

// Edits between steps
/** Body commands (3 uint4 each) grouped by body: ranges[r] = (first command, count), each body's in submission
 * order; teleported bodies get marks[body] = 1. */
void launch_body_commands(Body *bodies, const uint32_t *commands, const uint2 *ranges, uint32_t n, uint32_t *marks,
                          hipStream_t s);
/** Zero the count/generation word of the last step's manifolds (counters[C_MANIFOLDS] of them) touching a marked body. */
void launch_invalidate_manifolds(Manifold *manifolds, const uint32_t *counters, uint32_t capacity,
                                 const uint32_t *marks, hipStream_t s);
void launch_clear_marks(const uint32_t *commands, const uint2 *ranges, uint32_t n, uint32_t *marks, hipStream_t s);
/** words: n (body, groups, collides-with) triples. */
void launch_set_filters(uint2 *filters, const uint32_t *words, uint32_t n, hipStream_t s);
/** Restitution coefficients of n bodies: words = (body, f32 bits) pairs. */
void launch_set_restitution(float *restitution, const uint32_t *words, uint32_t n, hipStream_t s);
void launch_set_params(WorldParams *p, const ParamEdit &e, hipStream_t s);
/** Constraint record edits: n (slot, word, value bits) triples (word 32..34: info type, body A, body B). */
void launch_edit_joints(Joint *joints, int4 *info, const uint32_t *words, uint32_t n, hipStream_t s);
/** Rest orientations of angle-locked joints and reference tangents of motors and limits, from the current poses,
 * for the n constraint slots listed. */
void launch_capture_frames(const Body *bodies, Joint *joints, const int4 *info, const uint32_t *slots, uint32_t n,
                           hipStream_t s);

}  // namespace avbd
