// Device records, constants and FP32 math of the AVBD 3D step for gfx950.
//
// Records keep the reference's layouts (reference/three-avbd/src/avbd3d/gpu/layout.ts:8-55) so the world API reads
// and writes them unchanged: Body 160 B (the 64 B a contact needs first), Joint 128 B, Manifold 32 B, Contact 64 B.
// Every expression is written in the reference's evaluation order (layout.ts:209-242, wgsl-*.ts) and the library is
// built with -ffp-contract=off, so a step's discrete outcomes (pairs, feature keys, colours) are reproducible bit
// for bit by a sequential FP32 restatement on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AVBD_HD __host__ __device__ __forceinline__
#define AVBD_D __device__ __forceinline__

namespace avbd {

// --- constants (layout.ts:74-106; avbd2d/gpu/layout.ts:42-100) --------------------------------------------------------
constexpr float PENALTY_MIN = 1.0f;
constexpr float PENALTY_MAX = 1e10f;
constexpr float PIVOT_FLOOR = 1e-5f;
constexpr float COLLISION_MARGIN = 0.01f;
constexpr float STICK_THRESH = 0.00001f;
constexpr float HARD = 1e30f;
constexpr float BIG = 3e38f;
constexpr float REST_SPEED = 0.1f;
constexpr float REUSE_LIN_TOL = 0.002f;
constexpr float REUSE_ANG_TOL = 0.002f;
constexpr float NEAREST_FRACTION = 0.05f;
constexpr uint32_t STICK_BIT = 0x80000000u;
constexpr uint32_t SPHERE_FEATURE = 3u << 24;
constexpr uint32_t CAPSULE_FEATURE = 4u << 24;
constexpr uint32_t MAX_COLORS = 64, NO_COLOR = 255, PENDING = 256;
// Constraint types of the joint records' info word: a released slot, a ball joint (+ angle lock), a spring, and the
// two angular constraints about a hinge axis (a motor driving the twist speed, a limit keeping the twist in a range)
constexpr int32_t T_NONE = 0, T_JOINT = 1, T_SPRING = 2, T_MOTOR = 3, T_LIMIT = 4;
constexpr float SHAPE_BOX = 0.f, SHAPE_SPHERE = 1.f, SHAPE_SAIL = 2.f, SHAPE_CAPSULE = -1.f;
/** Hulls: shape code SHAPE_HULL + the word of the body's hull record header (layout.ts:68). */
constexpr float SHAPE_HULL = 3.f;
constexpr uint32_t FLAG_MATCH_NEAREST = 1, FLAG_FACE_BIAS = 2, FLAG_REUSE_CONTACTS = 4, FLAG_START_AT_REST = 8,
                   FLAG_MASS_PENALTY = 16,
                   FLAG_CAPSULE_STEP_START_ARMS = 32;  // this build's option: capsule lever arms at x⁻ like spheres
constexpr int MAX_CONTACTS = 8;
constexpr int MAX_POLY = 8;
constexpr float SAT_AXIS_EPSILON = 1.0e-6f;
constexpr float PLANE_EPSILON = 1.0e-5f;
constexpr float CONTACT_MERGE_DIST_SQ = 1.0e-6f;

// Counter words
enum Counter : uint32_t {
  C_PAIRS = 0,
  C_CONTACTS = 1,
  C_PREV_CONTACTS = 2,
  C_OVERFLOW = 3,  // bit 0 pairs, 1 contacts, 2 manifolds (as the reference's bits 1, 2, 4)
  C_CLASHES = 4,
  C_NUM_COLORS = 5,
  C_MANIFOLDS = 6,
  C_PREV_MANIFOLDS = 7,
  C_CANDIDATES = 8,
  C_PIVOT64 = 9,  // bodies solved again in FP64 this step (an FP32 pivot took its floor)
  C_DF_ABORT = 10,  // a bounded wait gave up (dataflow solve, prefix scan; never expected: the step is then not valid)
  C_PENDING = 11,   // bodies the colouring's mark pass left pending (listed in colorPending)
  C_PAIR_DEMAND = 12,  // pairs the broadphase found, before clamping to the capacity (C_PAIRS is the clamped count)
  C_HULL_PAIRS = 13,   // pair slots with a hull this step (listed in hullPairs for the hull narrowphase)
  COUNTER_WORDS = 16
};

// --- records ------------------------------------------------------------------------------------------------------

struct Body {
  float4 pos;          // xyz, w: friction
  float4 rot;          // quaternion xyzw
  float4 ipos;         // pose at the start of the step (x-)
  float4 irot;
  float4 size;         // full widths, w: mass (0 = static)
  float4 moment;       // principal moments, w: bounding radius
  float4 inertialPos;  // inertial target, w: step the body last moved from its reference pose (u32 bits)
  float4 inertialRot;
  float4 vel;          // xyz, w: last step's velocity along up
  float4 angVel;       // xyz, w: shape code
};
static_assert(sizeof(Body) == 160, "Body record is 160 bytes");

/**
 * A constraint record (layout.ts:27-36). Ball joint: penalties, multipliers, C0 and anchors as named; the rest
 * orientation of an angle-locked joint (B relative to A when it was made) is the quaternion (c0Lin.w, c0Ang.w, rA.w,
 * rB.w) (gpuSolverKernels.js:105-130). Spring: stiffness in penLin.w, rest length in rA.w, break strain in lamLin.w and
 * break force in lamAng.w (0: unbreakable; springMaterial.js). Motor / limit (angularConstraints.js): row penalties
 * penLin.xy (penLin.w 3e38 while enabled), multipliers lamLin.xy, hinge axes rA.xyz (in A) and rB.xyz (in B),
 * reference tangents c0Lin.xyz (in A) and c0Ang.xyz (in B), warm-start values c0Lin.w / c0Ang.w, and rA.w / rB.w the
 * motor's speed and torque bound or the limit's lower and upper angle.
 */
struct Joint {
  float4 penLin;  // xyz, w: linear stiffness (>= HARD: hard); spring: its stiffness
  float4 penAng;  // xyz, w: angular stiffness
  float4 lamLin;  // xyz, w: fracture threshold (negative: breaks on pull too); spring: break strain
  float4 lamAng;  // xyz, w: torque arm; spring: break force
  float4 c0Lin;   // xyz, w: rest orientation x
  float4 c0Ang;   // xyz, w: rest orientation y
  float4 rA;      // xyz (world point when A is the world), w: spring rest length / rest orientation z
  float4 rB;      // xyz, w: rest orientation w
};
static_assert(sizeof(Joint) == 128, "Joint record is 128 bytes");

struct Manifold {
  uint4 ids;   // a, b (a > b), first contact, count | generation step << 4
  float4 geo;  // normal (B to A) xyz, friction
};
struct Contact {
  float4 rA;   // body-local anchor on A, w: key bits (feature | STICK_BIT)
  float4 rB;   // w: C0 normal
  float4 pen;  // w: C0 tangent 1
  float4 lam;  // w: C0 tangent 2
};

/** Per-world parameters (64 B). */
struct WorldParams {
  float dt, gravity, betaLin, betaAng, gamma, alpha;
  float up[3];
  uint32_t flags;
  uint32_t step;  // steps taken; the running step's number while it runs (from 1)
  uint32_t firstBody, bodyCount;
  float maxSmall;  // bodies with a larger bounding radius are this world's brute-force (large) bodies
  uint32_t pad[2];
};
static_assert(sizeof(WorldParams) == 64, "WorldParams is 64 bytes");

// --- FP32 math (explicit functions, left-to-right order, no contraction) ----------------------------------------------

struct f3 {
  float x, y, z;
};
AVBD_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
AVBD_HD f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
AVBD_HD float4 mk4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
AVBD_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
AVBD_HD void setComp(f3 &v, int i, float s) {
  if (i == 0) v.x = s;
  else if (i == 1) v.y = s;
  else v.z = s;
}
AVBD_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AVBD_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AVBD_HD f3 scale(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
AVBD_HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
AVBD_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
AVBD_HD f3 vabs(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
AVBD_HD f3 mul3(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
AVBD_HD f3 addScaled(f3 a, f3 b, float s) { return f3{a.x + b.x * s, a.y + b.y * s, a.z + b.z * s}; }
AVBD_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
AVBD_HD float length(f3 a) { return sqrtf(dot(a, a)); }
AVBD_HD f3 normalize(f3 a) { return divs(a, length(a)); }
AVBD_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
/** a < b ? a : b and a > b ? a : b — the selection the CPU restatement uses at every site. */
AVBD_HD float rmin(float a, float b) { return a < b ? a : b; }
AVBD_HD float rmax(float a, float b) { return a > b ? a : b; }
AVBD_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
AVBD_HD uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
AVBD_HD float rclamp(float x, float lo, float hi) { return rmax(lo, rmin(hi, x)); }
/** WGSL clamp(x, lo, hi) = min(max(x, lo), hi). */
AVBD_HD float wclamp(float x, float lo, float hi) { return rmin(rmax(x, lo), hi); }
AVBD_HD float rsign(float x) { return x < 0.f ? -1.f : (x > 0.f ? 1.f : 0.f); }

/** layout.ts:209-232 (qmul, qconj, qsub, qadd, qrotate). */
AVBD_HD float4 qmul(float4 a, float4 b) {
  return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                     a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
AVBD_HD float4 qconj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }
AVBD_HD float qlengthSq(float4 q) { return (q.x * q.x + q.y * q.y + q.z * q.z) + q.w * q.w; }
/** b⁻¹ = conj(b) / |b|² (componentwise divisions, as qsub forms it). */
AVBD_HD float4 qinv(float4 b) {
  float l = qlengthSq(b);
  return make_float4(-b.x / l, -b.y / l, -b.z / l, b.w / l);
}
/** qsub with b⁻¹ already formed: the same bits as qsub(a, b). */
AVBD_HD f3 qsubInv(float4 a, float4 bi) {
  float4 s = qmul(a, bi);
  return f3{s.x * 2.f, s.y * 2.f, s.z * 2.f};
}
AVBD_HD f3 qsub(float4 a, float4 b) { return qsubInv(a, qinv(b)); }
AVBD_HD float4 qadd(float4 a, f3 v) {
  float4 s = qmul(make_float4(v.x, v.y, v.z, 0.f), a);
  float4 q = make_float4(a.x + s.x * 0.5f, a.y + s.y * 0.5f, a.z + s.z * 0.5f, a.w + s.w * 0.5f);
  float l = sqrtf(qlengthSq(q));
  return make_float4(q.x / l, q.y / l, q.z / l, q.w / l);
}
AVBD_HD f3 qrotate(float4 q, f3 v) {
  float tx = (q.y * v.z - q.z * v.y) * 2.f;
  float ty = (q.z * v.x - q.x * v.z) * 2.f;
  float tz = (q.x * v.y - q.y * v.x) * 2.f;
  return f3{v.x + tx * q.w + (q.y * tz - q.z * ty), v.y + ty * q.w + (q.z * tx - q.x * tz),
            v.z + tz * q.w + (q.x * ty - q.y * tx)};
}

/**
 * atan(t) for t in [0, 1] in FP32 arithmetic only: above tan 15° the identity atan t = 30° + atan((√3 t − 1)/(√3 + t))
 * brings the argument into |u| ≤ tan 15° ≈ 0.268, where the odd series u − u³/3 + … − u¹¹/11 (Horner) is within
 * 3e-9 of atan u. The same operations in the same order run on the CPU, so both sides round alike.
 */
AVBD_HD float atanUnit(float t) {
  const float TAN15 = 0.267949192f, SQRT3 = 1.73205081f, PI6 = 0.523598776f;
  float offset = 0.f, u = t;
  if (t > TAN15) {
    u = (t * SQRT3 - 1.f) / (SQRT3 + t);
    offset = PI6;
  }
  const float u2 = u * u;
  float p = -1.f / 11.f;
  p = p * u2 + 1.f / 9.f;
  p = p * u2 - 1.f / 7.f;
  p = p * u2 + 1.f / 5.f;
  p = p * u2 - 1.f / 3.f;
  p = p * u2 + 1.f;
  return offset + u * p;
}
constexpr float PI_F = 3.14159265f, HALF_PI_F = 1.57079633f, TWO_PI_F = 6.28318531f;
/** atan2(y, x) in (−π, π] from atanUnit by octant (0 at the origin). */
AVBD_HD float atan2r(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  if (ax == 0.f && ay == 0.f) return 0.f;
  const bool steep = ay > ax;
  float a = atanUnit(steep ? ax / ay : ay / ax);
  if (steep) a = HALF_PI_F - a;
  if (x < 0.f) a = PI_F - a;
  return y < 0.f ? -a : a;
}
/** An angle wrapped into [−π, π) (the reference's atan2(sin x, cos x)). */
AVBD_HD float wrapAngle(float x) { return x - TWO_PI_F * floorf((x + PI_F) / TWO_PI_F); }

/** Contact basis rows (n, t1, t2) (layout.ts:236-242); t1 is scaled by the reciprocal of its length (one division
 * instead of three; the CPU restatement forms it the same way). */
struct Basis {
  f3 n, t1, t2;
};
/** The reciprocal of the unscaled first tangent's length (what orthonormal scales it by). */
AVBD_HD float tangentScale(f3 n) {
  const f3 t1 = fabsf(n.x) > fabsf(n.z) ? f3{-n.y, n.x, 0.f} : f3{0.f, -n.z, n.y};
  return 1.f / length(t1);
}
/** The basis from n and its tangentScale (formed once per manifold and step by the adjacency pass). */
AVBD_HD Basis orthonormalScaled(f3 n, float inv) {
  f3 t1 = fabsf(n.x) > fabsf(n.z) ? f3{-n.y, n.x, 0.f} : f3{0.f, -n.z, n.y};
  t1 = scale(t1, inv);
  return Basis{n, t1, cross(n, t1)};
}
AVBD_HD Basis orthonormal(f3 n) { return orthonormalScaled(n, tangentScale(n)); }

/** lowbias32 (avbd2d/gpu/layout.ts:106-115). */
AVBD_HD uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

AVBD_D uint32_t fbits(float f) { return __float_as_uint(f); }
AVBD_D float bitsf(uint32_t u) { return __uint_as_float(u); }

}  // namespace avbd
