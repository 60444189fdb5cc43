/*
 * avbd.h — C ABI of the MI355X-native AVBD 3D rigid-body step (libavbd_hip.so).
 *
 * The entry points are what a host binding (Node N-API addon, ctypes, cgo, JNI) needs to stand in for the reference's
 * WebGPU world (`createSolverPipeline({backend})` → `scene.createSolver` → `NativeGpuSolver3D`, src/native/pipeline.js:14-93,
 * src/native/solver3D.js:104-368, reference/three-avbd/src/avbd3d/gpu/solver.ts:348-1565). Records keep the reference's
 * layouts (reference/three-avbd/src/avbd3d/gpu/layout.ts:8-55):
 *   Body     40 f32  pos.xyz friction | rot xyzw | initialPos | initialRot | size.xyz mass | moment.xyz radius |
 *                    inertialPos.xyz stamp(u32 bits) | inertialRot | vel.xyz prevUpVel | angVel.xyz shape
 *   Joint    32 f32  penLin.xyz stiffLin | penAng.xyz stiffAng | lamLin.xyz fracture | lamAng.xyz torqueArm |
 *                    C0Lin | C0Ang | rA.xyz rest | rB   (+ info i32×4: type 1 joint / 2 spring, bodyA (-1 world), bodyB, 0)
 *   Manifold 8 u32   a, b (a > b), first contact, count | generation step << 4 | normal.xyz (f32), friction (f32)
 *   Contact  16 u32  rA.xyz key | rB.xyz C0n | pen.xyz C0t1 | lam.xyz C0t2   (f32 except the key)
 * Unbounded stiffness or fracture is stored as 3e38 (the reference's BIG; layout.ts header comment).
 *
 * Ownership and threading follow the reference: a world owns its device memory; reads copy into caller buffers; a
 * handle is used by one caller at a time; avbd_world_step() enqueues on the world's HIP stream and returns without
 * waiting (reads synchronise). Errors: every call returns AVBD_OK or a negative code and records a message readable
 * with avbd_last_error() (thread-local), which a binding turns into `throw Error(msg)` as the reference does.
 * Capacity failures are not errors: they are counters plus a held step, as in the reference.
 */
#ifndef AVBD_H
#define AVBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_OK 0
#define AVBD_ERR_INVALID (-1)
#define AVBD_ERR_DEVICE (-2)
#define AVBD_ERR_DESTROYED (-3)
#define AVBD_ERR_CAPACITY (-4)

#define AVBD_BODY_FLOATS 40
#define AVBD_JOINT_FLOATS 32
#define AVBD_MANIFOLD_WORDS 8
#define AVBD_CONTACT_WORDS 16

/* Solver flags (layout.ts:74-92); AVBD_FLAGS_DEFAULT = all on, as gpuParams3D() (solver.ts:113-125). */
#define AVBD_FLAG_MATCH_NEAREST 1u
#define AVBD_FLAG_FACE_BIAS 2u
#define AVBD_FLAG_REUSE_CONTACTS 4u
#define AVBD_FLAG_START_AT_REST 8u
#define AVBD_FLAG_MASS_PENALTY 16u
#define AVBD_FLAGS_DEFAULT 31u
/* No counterpart in the reference (off by default): capsule contacts take their lever arms at the step-start rotation,
 * as sphere contacts do (wgsl-solve.ts:172-181); the reference turns capsule lever arms with the body, which lets a
 * capsule rolling hard on its side sink into what it rests on. */
#define AVBD_FLAG_CAPSULE_STEP_START_ARMS 32u

typedef struct avbd_world avbd_world;

/**
 * One world: the packed scene the reference's builder hands its GPU solver (AvbdScene3D → GpuSolver3D ctor + pack +
 * writeJoint + uploadStatics, src/native/scenes.js:652-702; solver.ts:491-580, 794-903).
 */
typedef struct AvbdSceneDesc {
  uint32_t body_count;
  const float *bodies;          /* body_count × 40 f32 */
  uint32_t joint_count;
  const float *joints;          /* joint_count × 32 f32 */
  const int32_t *joint_info;    /* joint_count × 4 i32 (type, bodyA or -1, bodyB, 0) */
  uint32_t ignore_count;
  const uint32_t *ignore_pairs; /* ignore_count × 2 (IgnoreCollision, forces.ts:300-306) */
  const uint32_t *filters;      /* body_count × 2 (groups, collides-with; setFilters) or NULL: all */
  float dt, gravity;            /* gravity: acceleration along up (negative: down) */
  float up[3];                  /* unit */
  float alpha, gamma, beta_lin, beta_ang;
  uint32_t iterations;
  uint32_t flags;               /* AVBD_FLAG_* */
  uint32_t hull_words;          /* size of the hull buffer in 16-byte words (0: no hulls) */
  const uint32_t *hulls;        /* the hull buffer (4 u32 per word; layout of avbd_hull_from_points), or NULL; a hull
                                   body's shape code (angVel.w) is AVBD_SHAPE_HULL + its record's header word
                                   (solver.ts:170-197, 813-817; wgsl-hull.ts:13-19) */
} AvbdSceneDesc;

/** Solver options (GpuSolverOptions, solver.ts:127-151, plus this library's). Zero-initialised = defaults. */
typedef struct AvbdOptions {
  uint32_t color_cap;       /* colours per iteration and colouring cap (0: 16, max 64) */
  uint32_t color_rounds;    /* Jones–Plassmann rounds per step, even (0: 16) */
  uint32_t pair_capacity;   /* initial pair/manifold capacity (0: 8 × bodies) */
  uint32_t use_graph;       /* replay each step from a captured hipGraph */
  int32_t device;           /* HIP device ordinal */
  uint32_t adaptive;        /* adapt colour cap/rounds and storage from counters read back without waiting */
  uint32_t broadphase;      /* 0 auto (H-PLOC for a world or batch of >= 10,000 bodies), 1 uniform grid, 2 H-PLOC BVH
                               (AppGpuSolver3D broadphase "auto" | "grid" | "hploc", appGpuSolver3D.js:333-402) */
  uint32_t bvh_rebuild_interval; /* H-PLOC rebuild period in steps, refit in between (0: 64; gpuHploc.js:62-80) */
  uint32_t solver;          /* 0 auto (a batch of worlds of at most 256 bodies each: 3; otherwise 1), 1 one launch
                               per colour and a dual launch per iteration, 2 dataflow: every iteration's colours and
                               duals in one launch, each body waiting only for its neighbours (its workgroups must
                               all fit on the device at once: see dataflow_share), 3 one workgroup per world: each world's
                               iterations, colours, duals and velocities in one workgroup, every world of the batch in
                               one launch (the same result from each) */
  uint32_t coloring_rounds; /* 0 auto (one block over the pending bodies while few are left pending, with `adaptive`),
                               1 one launch per Jones–Plassmann round, 2 always one block over the pending bodies
                               (the same colours either way) */
  uint32_t world_block;     /* threads per world's workgroup with solver 3: 128, 256, 512 or 1024 (0: 256) */
  uint32_t dataflow_share;  /* solver 2: the solve's grid takes 1/dataflow_share of the device's resident workgroups
                               (0 or 1: all of them). Worlds whose dataflow solves run at the same time on their own
                               streams need shares that add up to at most the device */
} AvbdOptions;

/** readCounters() (solver.ts:1429-1432; world.js:672-692). overflow bits: 1 pairs, 2 contacts, 4 manifolds. */
typedef struct AvbdCounters {
  uint32_t pairs, contacts, manifolds, overflow, clashes, colors, joints, bodies;
  uint32_t pivot_fallbacks; /* bodies whose 6x6 solve was redone in FP64 in the last step (a pivot hit its floor) */
  uint32_t solve_aborts;    /* waits given up in the last step, by the dataflow solve or a prefix scan (nonzero: that
                               step is not valid; never expected with the device to itself). A world whose dataflow
                               solve gave up runs colour launches from its next step on, once this counter has been
                               read (avbd_world_read_counters, avbd_world_adapt or the adaptive readback) */
} AvbdCounters;

/** Thread-local message of the last failed call ("" if none). */
const char *avbd_last_error(void);
/** Library version string. */
const char *avbd_version(void);
/** Number of HIP devices visible (0 without a GPU). */
int avbd_device_count(void);

/** Create a world (replaces `new NativeGpuSolver3D(device, ref, options)`, solver3D.js:104-109). */
int avbd_world_create(const AvbdSceneDesc *scene, const AvbdOptions *options, avbd_world **out);
/**
 * Create a batch: n independent worlds stepped together, one launch per phase (the reference steps each world of a
 * pipeline by its own submission, pipeline.js:49-79). Bodies and joints of world k follow those of worlds 0..k-1 in
 * every read and write. All worlds take the same iteration count.
 */
int avbd_batch_create(const AvbdSceneDesc *scenes, uint32_t n, const AvbdOptions *options, avbd_world **out);
/** Release a world (world.destroy(), solver3D.js). */
int avbd_world_destroy(avbd_world *world);

/** Enqueue n steps (world.step(), solver3D.js:350-359; GpuSolver3D.step, solver.ts:1281-1417). */
int avbd_world_step(avbd_world *world, uint32_t n);
/**
 * Enqueue n steps with HIP events around each (the reference's per-step timestamp queries, solver.ts:1389-1416) and
 * wait: ms[k] = device milliseconds of step k. For step-latency percentiles of worlds and batches.
 */
int avbd_world_step_timed(avbd_world *world, uint32_t n, float *ms);
/** avbd_world_step_timed in two halves, so several worlds' steps can be in flight at once: enqueue n event-bracketed
 * steps (at most 4096), then wait and copy their device times (count = n; capacity must be at least n). */
int avbd_world_step_timed_begin(avbd_world *world, uint32_t n);
int avbd_world_step_timed_end(avbd_world *world, float *ms, uint32_t capacity, uint32_t *count);
/** Wait for enqueued steps. */
int avbd_world_synchronize(avbd_world *world);

int avbd_world_body_count(const avbd_world *world, uint32_t *out);
int avbd_world_joint_count(const avbd_world *world, uint32_t *out);

/** readBodies() (solver.ts:1421-1423): first..first+count bodies, 40 f32 each. */
int avbd_world_read_bodies(avbd_world *world, uint32_t first, uint32_t count, float *out);
/** writeBodies (solver.ts:760-788): replace body records (poses, velocities, shapes) before the next step. */
int avbd_world_write_bodies(avbd_world *world, uint32_t first, uint32_t count, const float *in);
/**
 * readSelectedBodies(indices, {posesOnly}) (solver3D.js:223-230; bodyReadback.js:3-154): 40 f32 per body, or with
 * poses_only 8 f32 per body (pos.xyz friction, rot.xyzw).
 */
int avbd_world_read_selected(avbd_world *world, const uint32_t *indices, uint32_t n, int poses_only, float *out);
/** readJoints() (solver.ts:1425-1427): 32 f32 per joint. */
int avbd_world_read_joints(avbd_world *world, uint32_t first, uint32_t count, float *out);
int avbd_world_write_joints(avbd_world *world, uint32_t first, uint32_t count, const float *in);
/** readCounters() (solver.ts:1429-1432). */
int avbd_world_read_counters(avbd_world *world, AvbdCounters *out);
/** readPairs() (solver.ts:1471-1475): the last step's broadphase pairs (a > b), 2 u32 each, at most capacity. */
int avbd_world_read_pairs(avbd_world *world, uint32_t *out, uint32_t capacity, uint32_t *count);
/**
 * readContactList() (solver.ts:1435-1469): the last step's manifolds (8 words; `first` indexes the contact list) and
 * their contact points (16 words), densely packed. Pass NULL buffers to query the counts.
 */
int avbd_world_read_contacts(avbd_world *world, uint32_t *manifolds, uint32_t manifold_capacity, uint32_t *manifold_count,
                             uint32_t *contacts, uint32_t contact_capacity, uint32_t *contact_count);
/** Colour of each body after the last step's colouring (255: none). */
int avbd_world_read_colors(avbd_world *world, uint32_t *out);
/** Adjacency of the last step: start[bodies + 1] then the constraint ids (joints, then J + manifold index). */
int avbd_world_read_adjacency(avbd_world *world, uint32_t *start, uint32_t *list, uint32_t capacity, uint32_t *count);
/** Reference poses for contact reuse (8 f32 per body: pos.xyz 0, rot.xyzw). */
int avbd_world_read_refs(avbd_world *world, float *out);
/** Steps taken by each world of the batch. */
int avbd_world_read_steps(avbd_world *world, uint32_t *out);

/**
 * seedFrom (solver.ts:905-948): make the next step start from a given persistent state — reference poses (8 f32 per
 * body or NULL), colour state (1 u32 per body or NULL), last step's manifolds and contacts (the warm-start source, as
 * read by avbd_world_read_contacts) and each world's step counter (or NULL).
 */
int avbd_world_write_state(avbd_world *world, const float *refs, const uint32_t *colors, uint32_t manifold_count,
                           const uint32_t *manifolds, uint32_t contact_count, const uint32_t *contacts,
                           const uint32_t *steps);

/** Colour cap and Jones–Plassmann rounds (GpuSolver3D.colorCap / colorRounds). */
int avbd_world_set_coloring(avbd_world *world, uint32_t color_cap, uint32_t color_rounds);
/** The colour cap, Jones–Plassmann rounds and pair capacity in use (GpuSolver3D.colorCap / colorRounds /
 * contactCapacity, solver.ts:384-386, as adapt() leaves them); any pointer may be NULL. */
int avbd_world_get_coloring(avbd_world *world, uint32_t *color_cap, uint32_t *color_rounds, uint32_t *pair_capacity);
/** adapt(counters) (solver.ts:1484-1519): grow storage and the colour cap from the last step's counters. */
int avbd_world_adapt(avbd_world *world, int *changed);

/**
 * Run one step with HIP events around each phase (the reference's per-phase timestamp queries, solver.ts:1389-1416)
 * and around every primal and dual launch. out[13] (ms): total, collision, adjacency, colouring, warm start,
 * primal (sum over launches), dual (sum), velocities, primal launches, dual launches, broadphase (tree or grid and
 * the pair search, part of collision), narrowphase (contacts and manifold compaction, part of collision), pair search
 * (part of broadphase).
 */
int avbd_world_profile_step(avbd_world *world, float *out);

/** Device bytes held by the world. */
int avbd_world_device_bytes(const avbd_world *world, uint64_t *out);

/** Which solve the world's steps run (this library's choice of kernels; the reference has one): 1 colour launches,
 * 2 dataflow, 3 one workgroup per world. */
int avbd_world_solver_kind(const avbd_world *world, uint32_t *out);

/* ---- Edits between steps ---------------------------------------------------------------------------------------- */

/**
 * Body command record, 12 words (CommandStaging / writeCommand, src/gpu/bodyCommands2D.js:16-58; struct Command of
 * src/native/bodyCommands3D.wgsl:2): u32 body, u32 kind, u32 has_point (0 or 1), u32 0 | f32 value[4] | f32 point[4].
 *   0 linear velocity := value.xyz          1 angular velocity := value.xyz
 *   2 impulse value.xyz at point.xyz (at the centre without a point)
 *   3 force value.xyz over value.w seconds (the step's dt; applied as the impulse value.xyz × value.w)
 *   4 teleport to value.xyz, rotation := point.xyzw if has_point (unit quaternion, normalised by the caller)
 *   5 angular impulse value.xyz             6 torque value.xyz over value.w seconds
 * Impulses, forces and torques leave static bodies (mass 0) unchanged.
 */
#define AVBD_COMMAND_WORDS 12
#define AVBD_CMD_LINEAR_VELOCITY 0u
#define AVBD_CMD_ANGULAR_VELOCITY 1u
#define AVBD_CMD_IMPULSE 2u
#define AVBD_CMD_FORCE 3u
#define AVBD_CMD_TELEPORT 4u
#define AVBD_CMD_ANGULAR_IMPULSE 5u
#define AVBD_CMD_TORQUE 6u

/**
 * Queue n body commands (the commands of NativeGpuSolver3D.setLinearVelocity … teleport, setPoses and editBodies,
 * src/native/solver3D.js:130-206, src/native/runtime.js:45-175, applied by bodyCommands3D.wgsl:14-50). All n are
 * checked first (live body index, kind, finite components, nonzero rotation); on an error none is applied. They are
 * applied on the world's stream after the steps enqueued before this call and before those enqueued after it, each
 * body's commands in submission order. A teleport also resets the body's step-start and inertial poses and drops the
 * warm start of the last step's manifolds that touch it. Bodies index the whole batch.
 */
int avbd_world_enqueue_commands(avbd_world *world, const uint32_t *commands, uint32_t n);

/** Step parameters of one world of a batch (GpuParams3D, solver.ts:350, written before every step by writeParams,
 * solver.ts:1214-1254). iterations is shared by all worlds of a batch. */
typedef struct AvbdParams {
  float dt, gravity;
  float up[3];
  float alpha, gamma, beta_lin, beta_ang;
  uint32_t iterations;
  uint32_t flags; /* AVBD_FLAG_* */
} AvbdParams;

int avbd_world_get_params(avbd_world *world, uint32_t world_index, AvbdParams *out);
/** Takes effect from the next step on (a changed iteration count re-records the step graphs). */
int avbd_world_set_params(avbd_world *world, uint32_t world_index, const AvbdParams *params);
/** setFilters(indices, groups, collidesWith) (appGpuSolver3D.js:958-983): collision groups and masks of n bodies. */
int avbd_world_set_filters(avbd_world *world, const uint32_t *indices, const uint32_t *groups,
                           const uint32_t *collides_with, uint32_t n);
/** setRestitution(index, value) (appGpuSolver3D.js:455-471; src/gpu/restitution.js): restitution coefficients in
 * [0, 1] of n bodies, from the next step on. While any body's coefficient is above 0 each step ends with the
 * reference's bounce passes (incoming velocities captured before the step, 12 impulse rounds after it). */
int avbd_world_set_restitution(avbd_world *world, const uint32_t *indices, const float *values, uint32_t n);
/** The coefficients of all bodies (world->bodies floats). */
int avbd_world_get_restitution(avbd_world *world, float *out);

/* ---- Constraint edits (src/gpu/appGpuSolver3D.js:500-563, 810-837, 922-957; reference/three-avbd/src/avbd3d/gpu/
 * solver.ts:1138-1204). Constraint types of joint_info[0]: 0 released, 1 ball joint (+ angle lock), 2 spring, 3 motor,
 * 4 limit. A scene may pass motors and limits as records of their own (the layout setAngularConstraint writes: penLin
 * (1000, 1000, 0, 3e38), axes in rA.xyz / rB.xyz, speed or lower angle in rA.w, torque bound or upper angle in rB.w);
 * the world captures their reference tangents, and the rest orientation of every ball joint with a nonzero angular
 * stiffness, from the starting poses. Edits apply on the world's stream after the steps enqueued before them. */
#define AVBD_JOINT_NONE 0
#define AVBD_JOINT_BALL 1
#define AVBD_JOINT_SPRING 2
#define AVBD_JOINT_MOTOR 3
#define AVBD_JOINT_LIMIT 4

/** releaseJoints(slots) (solver.ts:1138-1162): each live slot's record zeroed and its type 0, so its bodies collide
 * again; released or out-of-range slots are ignored. Sleeping bodies of the constraints wake. */
int avbd_world_release_joints(avbd_world *world, const uint32_t *slots, uint32_t n);
/** disableConstraint(slot) (solver.ts:1200-1204): penalties, stiffnesses and multipliers zeroed (fracture and torque
 * arm kept); the constraint stops acting and its bodies collide again. */
int avbd_world_disable_constraint(avbd_world *world, uint32_t slot);
/** setWorldAnchor(slot, point) (appGpuSolver3D.js:928-941): the world point of a joint whose first body is the world. */
int avbd_world_set_world_anchor(avbd_world *world, uint32_t slot, const float *point);
/** setMotor(slot, {speed, maxTorque}) (appGpuSolver3D.js:547-563): either pointer may be NULL (kept). */
int avbd_world_set_motor(avbd_world *world, uint32_t slot, const float *speed, const float *max_torque);
/** setSpringMaterial(slot, {breakStrain, breakForce}) (appGpuSolver3D.js:810-837; springMaterial.js): positive, or
 * INFINITY for unbreakable. A spring stretched past break_strain × its rest length, or pulling with more than
 * break_force, stops acting for good. */
int avbd_world_set_spring_material(avbd_world *world, uint32_t slot, float break_strain, float break_force);
/** setAngularConstraint(slot, options) (appGpuSolver3D.js:500-526; angularConstraints.js:84-120): the slot becomes a
 * motor (type 3: lo = speed, hi = torque bound >= 0) or a limit (type 4: lo <= hi within -pi..pi) about the body-local
 * axes (normalised here), its reference tangents captured from the current poses. */
int avbd_world_set_angular_constraint(avbd_world *world, uint32_t slot, int32_t type, const float *axis_a,
                                      const float *axis_b, float lo, float hi);
/** A constraint's info word: type, body A (-1: world), body B, 0. */
int avbd_world_joint_info(avbd_world *world, uint32_t slot, int32_t *out);

/* ---- Sleeping (src/gpu/worldSleeping.js, worldSleeping.wgsl, islandSleeping.wgsl, sleepContact.wgsl) ----------------
 * A sleeping body is a fixed support: its mass (record float 19) reads 0 and its velocities 0 until it wakes, when its
 * awake mass comes back. Body commands, filter and restitution edits and body writes wake every body first, as the
 * reference's wakeAll does before its property and command flushes. */

/** enableSleeping(options) (appGpuSolver3D.js:869-881, worldSleeping.js:17-40): speed_threshold (linear speed plus
 * angular speed times bounding radius, world units/s; the reference's default 0.03) and time_threshold (seconds quiet
 * and supported; default 0.5), both positive and finite; islands -1 (auto: worlds averaging 4096 bodies or more),
 * 0 (any impact wakes the world) or 1 (connected groups sleep and wake together); start_asleep (island mode): the next
 * step puts still groups resting on still static support to sleep. Enabling again keeps the first settings. */
int avbd_world_enable_sleeping(avbd_world *world, float speed_threshold, float time_threshold, int islands,
                               int start_asleep);
/** disableSleeping() (solver3D.js:333-343): every body wakes with its mass back; sleeping stops. */
int avbd_world_disable_sleeping(avbd_world *world);
/** 1 if sleeping is enabled, 0 if not; *islands (if not NULL) its mode. */
int avbd_world_sleeping(avbd_world *world, int *islands);
/** wakeAll() (appGpuSolver3D.js:882-884): every body wakes before the next step. */
int avbd_world_wake_all(avbd_world *world);
/** wakeBody(index) (worldSleeping.js:226-235) for n bodies: held awake through the next step; in island mode their
 * islands wake at its start. */
int avbd_world_wake_bodies(avbd_world *world, const uint32_t *indices, uint32_t n);
/** setSleepEnabled(index, enabled) (solver3D.js:317-327) for n bodies (flags 0 or 1): eligibility; a body allowed to
 * sleep enables sleeping with the default thresholds; every body wakes. */
int avbd_world_set_sleep_enabled(avbd_world *world, const uint32_t *indices, const uint8_t *flags, uint32_t n);
/** readSleepStats() (appGpuSolver3D.js:885-889, worldSleeping.js:257-274) of world world_index after the last step:
 * out[0] bodies asleep, out[1] 1 if a wake-all request is pending. */
int avbd_world_sleep_stats(avbd_world *world, uint32_t world_index, uint32_t out[2]);
/** Every body's sleep state, 8 words (mass and quiet time as f32 bits, asleep, support, eligible, island parent, island
 * wake bits, flags), and 4 words per world (wake all, asleep, woken by an impact, island hit); either may be NULL. */
int avbd_world_read_sleep(avbd_world *world, uint32_t *states, uint32_t *globals);

/* ---- Shape queries (src/gpu/shapeQueries.js, shapeQueries.wgsl) --------------------------------------------------- */

typedef struct AvbdQueryOptions {
  const uint32_t *ignore; /* up to 16 body indices left out, or NULL */
  uint32_t ignore_count;
  uint32_t collides_with; /* bodies whose collision group & collides_with is 0 are left out (0xFFFFFFFF: none) */
  uint32_t membership;    /* with use_membership: bodies whose collides-with mask & membership is 0 are left out */
  int use_membership;
} AvbdQueryOptions;

/** raycastAll / sphereCastAll (appGpuSolver3D.js:306-332, shapeQueries.js:99-258): n queries (at most 65,535) of 8
 * floats — origin xyz, maximum distance, direction xyz (nonzero; normalised here), sweep radius (0: a ray) — against
 * every body's current pose; each query's nearest hit (smallest body index on a tie) into 8 floats of hits: world
 * normal xyz, distance (-1: no hit), body index (u32 bits, 0xFFFFFFFF: none), 0, 0, 0. A query starting inside or
 * touching a shape hits it at distance 0 with normal -direction. options may be NULL (no filtering). */
int avbd_world_cast(avbd_world *world, const float *queries, uint32_t n, const AvbdQueryOptions *options,
                    float *hits);

/** avbd_world_cast without allocating device memory: no counterpart in the reference; for callers that cannot allocate
 * (tight serving loops, code that must not call hipMalloc). The queries' records, nearest-hit words and hits staging
 * live in one fixed device array per GPU (AVBD_NOALLOC_MAX_QUERIES queries, 1.125 MiB), held until the process ends
 * and shared by every call on that GPU: calls on one GPU that overlap (from several threads) race. n above the
 * maximum is an AVBD_ERR_INVALID error. Same arguments and results as avbd_world_cast. */
#define AVBD_NOALLOC_MAX_QUERIES 16384u
int avbd_world_cast_noalloc(avbd_world *world, const float *queries, uint32_t n, const AvbdQueryOptions *options,
                            float *hits);

/* ---- Contact events (src/vendor/three-avbd-lib/contacts.ts ContactWatch; src/native/solver3D.js:252-311) -------- */

/** watchContacts({maxPairs, maxEvents}): event storage, made once (every body watched); after each step a pair with a
 * watched body that starts touching logs a begin event and one that stops an end event. */
int avbd_world_watch_contacts(avbd_world *world, uint32_t max_pairs, uint32_t max_events);
/** watchContacts({indices}): which bodies report their contacts (flags 0 or 1). */
int avbd_world_set_watched(avbd_world *world, const uint32_t *indices, const uint8_t *flags, uint32_t n);
/** resetContactHistory(indices) (sorted, unique): new lifetimes; their pairs begin again and do not end. */
int avbd_world_reset_contact_history(avbd_world *world, const uint32_t *indices, uint32_t n);
/** readContactEvents(): up to cap unread events, 12 words each (kind 1 begin / 2 end, a, b, step, point xyz (f32,
 * mean of the points on a), impulse (f32, |sum of normal multipliers| x dt), normal xyz (f32, b to a), 0), into out;
 * *count = copied, *written = events since the last read (above capacity some were dropped), *full = a pair set
 * overflowed. Clears the log. */
int avbd_world_read_contact_events(avbd_world *world, uint32_t *out, uint32_t cap, uint32_t *count, uint32_t *written,
                                   uint32_t *full);

/* ---- Serving loop ------------------------------------------------------------------------------------------------
 * No counterpart in the reference, whose worlds are stepped one submission at a time by their owner (pipeline.js:49-79):
 * a server takes step/rollout requests for many registered worlds as they arrive, keeps world state resident in HBM
 * under a byte budget (least recently used idle worlds are copied to host memory and restored exactly when next
 * requested), and runs admitted requests concurrently, one HIP stream per world, launching the next step of every
 * running request each turn (continuous batching at step granularity). Requests of one world run in submission order,
 * so a world's result does not depend on what else is served. Served worlds do not adapt their colour cap. */

/*
 * Convex hulls (reference/three-avbd/src/avbd3d/hull.ts: convexHull :60-66, hullFromTriangles :208-212; the body builder
 * shapes.ts:50-70; the record solver.ts:170-197). Cooking runs on the host and touches no device. The record, in
 * 16-byte words from the start of the buffer it is placed in (header at word `at`):
 *   at:   vertex start, vertex count, face start, face count
 *   at+1: edge start, edge count, index start, 0
 *   vertices (x, y, z, 0; f32 bits), faces (normal.xyz, plane offset as f32 bits; first index, index count, 0, 0),
 *   vertex indices (4 per word), edges (vertex a < vertex b, face with a→b, face with b→a),
 * all in the hull's principal frame centred on its centre of mass. A body of this hull has shape code
 * AVBD_SHAPE_HULL + at, size = `size`, mass = volume · density, moments = `moments` · density, radius = `radius`.
 * Errors (AVBD_ERR_INVALID): the points span no volume; the triangles do not close up; capacity_words too small.
 */
#define AVBD_SHAPE_HULL 3.0f
#define AVBD_HULL_MAX_VERTICES 32u
#define AVBD_HULL_MAX_WORDS 320u
typedef struct AvbdHullInfo {
  uint32_t vertex_count, face_count, edge_count;
  uint32_t words;      /* record size in 16-byte words */
  double volume;       /* per unit density */
  double moments[3];   /* principal moments per unit density */
  double size[3];      /* full widths along the principal axes */
  double radius;       /* farthest vertex from the centre of mass */
  double center[3];    /* the principal frame in the input points' frame: centre of mass and rotation (x, y, z, w) */
  double rotation[4];
} AvbdHullInfo;
/** The convex hull of n_points xyz points (at most 32 vertices kept); record may be NULL (info only). */
int avbd_hull_from_points(const double *xyz, uint32_t n_points, uint32_t at, uint32_t *record, uint32_t capacity_words,
                          AvbdHullInfo *info);
/** A hull from a closed, outward-wound triangulation of a convex point set. */
int avbd_hull_from_triangles(const double *xyz, uint32_t n_points, const uint32_t *indices, uint32_t n_indices,
                             uint32_t at, uint32_t *record, uint32_t capacity_words, AvbdHullInfo *info);

typedef struct avbd_server avbd_server;

typedef struct AvbdServerOptions {
  uint64_t hbm_budget_bytes; /* 0: 200 GiB */
  uint32_t max_active;       /* requests running at once (0: 32) */
  uint32_t lookahead;        /* steps in flight per running request (0: 2) */
  int32_t device;
  uint32_t color_cap, color_rounds, use_graph, broadphase; /* per world, as in AvbdOptions (0: defaults, graph on) */
  uint32_t small_world_bodies; /* worlds of at most this many bodies are solved by one workgroup (solver 3, workgroups
                                  of small_world_block threads); 0: every world by colour launches */
  uint32_t small_world_block;  /* 128, 256, 512 or 1024 (0: 1024) */
  uint32_t large_world_bodies; /* worlds of at least this many bodies run the dataflow solve (solver 2), each grid
                                  1/large_world_share of the device; 0: none */
  uint32_t large_world_share;  /* at least the number of such solves that can run at once (0: 4, the hardware queues) */
} AvbdServerOptions;

typedef struct AvbdServerStats {
  uint64_t submitted, completed, admissions, evictions, restores, steps, body_steps;
  double wall_seconds;                    /* first submission (since reset) to last completion */
  double step_ms_p50, step_ms_p99, step_ms_max; /* device time per step (HIP events) */
  double request_ms_p50, request_ms_p99;  /* submission to completion */
  uint64_t resident_bytes, peak_resident_bytes;
  uint32_t resident_worlds, worlds;
  uint64_t solve_aborts; /* dataflow waits given up in served steps (read when each request finishes; a world that gave
                            one up runs colour launches from then on) */
} AvbdServerStats;

int avbd_server_create(const AvbdServerOptions *options, avbd_server **out);
int avbd_server_destroy(avbd_server *server);
/** Register a world (held on the host until a request needs it); *world_id is its handle. */
int avbd_server_add_world(avbd_server *server, const AvbdSceneDesc *scene, uint32_t *world_id);
int avbd_server_world_body_count(avbd_server *server, uint32_t world_id, uint32_t *out);
/** Queue `steps` steps of a world (a step request is a rollout of length 1). */
int avbd_server_submit(avbd_server *server, uint32_t world_id, uint32_t steps, uint64_t *request_id);
/** Wait for a request; latency_ms (may be NULL) = submission to completion. */
int avbd_server_wait(avbd_server *server, uint64_t request_id, double *latency_ms);
int avbd_server_poll(avbd_server *server, uint64_t request_id, int *done);
/** The world's bodies (40 f32 each) after its queued requests have finished. */
int avbd_server_read_bodies(avbd_server *server, uint32_t world_id, float *out);
int avbd_server_stats(avbd_server *server, AvbdServerStats *out);
int avbd_server_reset_stats(avbd_server *server);

#ifdef __cplusplus
}
#endif

#endif /* AVBD_H */
