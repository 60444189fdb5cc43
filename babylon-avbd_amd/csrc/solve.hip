// The AVBD iterations on gfx950: joint and body warm starts, one primal pass per colour (each body folds its joints'
// and contacts' rows into a 6x6 Newton system held in registers and solves it by LDLᵀ with the 1e-5 relative pivot
// floor), the dual/penalty ramp over joints and contact pairs, and the BDF1 velocities.
// Semantics: reference/three-avbd/src/avbd3d/gpu/wgsl-solve.ts (addRow 53-59, addJoint 90-140, evalContact 171-206,
// warmStartJoints 210-226, warmStartBodies 259-300, primal/accumulate/finishBody 314-470, dual 474-535,
// updateVelocities 537-549). A step whose contact storage overflowed or whose colouring clashed is held (every kernel
// here leaves the world untouched), as the reference's protectSolverStep does (src/gpu/gpuSolverKernels.js:72-103).
#include "engine_types.h"
#include "kernels.h"
#include "stepsetup.h"

namespace avbd {


// --- device-scope access for the dataflow solve -------------------------------------------------------------------------
// Within one launch, waves on different XCDs exchange poses and contact multipliers. Those are read and written with
// device-scope (sc1) accesses, which are coherent across the XCDs' L2 caches; a writer waits for its stores to complete
// before it publishes a stamp, and a reader reads only after it has seen the stamp.

/**
 * The dataflow solve's position in the step: iteration, colour of the running primal, the step's stamp base; and the
 * running group's item stash in LDS (slot t of the group at stash[f * DF_WAVE + t] for word f; null outside the
 * dataflow solve).
 */
struct DfCtx {
  uint32_t it, color, base;
  float *stash = nullptr;
};

/**
 * Words of a stashed contact item: the contact record (16) and the partner's share of the anchors' relative motion
 * (6). The primal stashes each item of a body's first round (its first W items); the body's fused dual then reads them
 * there instead of from memory: a pair's finisher is the only writer of its multipliers and its partner does not move
 * before the finisher's dual stamp, so the stashed values are the ones memory still holds.
 */
constexpr uint32_t STASH_WORDS = 22;

/**
 * How a primal pass runs: one launch per colour (plain memory; the dual is a pass of its own), the dataflow solve
 * (device-scope memory and stamps; each pair's dual fused into its later colour's primal), or one workgroup per world
 * (plain memory, duals fused the same way, the world's colours separated by workgroup barriers).
 */
enum : int { SOLVE_COLORS = 0, SOLVE_DATAFLOW = 1, SOLVE_WORLDS = 2 };

AVBD_D uint32_t ldAgent(const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
AVBD_D void stAgent(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// Poses and contact multipliers move as whole 16-byte records: raw buffer loads and stores with the device-scope
// cache policy (sc1), addressed from the arrays' bases (uniform, so the resource words stay scalar). The engine runs
// the dataflow solve only while both arrays span less than 4 GiB (32-bit buffer offsets).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int BUF_SC1 = 1 << 4;            // cache policy bits: sc1 (device scope)
constexpr int BUF_WORD3 = 0x00020000;      // resource word 3 of a raw (untyped) buffer on the gfx9 family
AVBD_D __amdgpu_buffer_rsrc_t bodyRsrc(const DeviceWorld &w) {
  return __builtin_amdgcn_make_buffer_rsrc(static_cast<void *>(w.bodies), 0, int(w.nBodies * uint32_t(sizeof(Body))),
                                           BUF_WORD3);
}
AVBD_D __amdgpu_buffer_rsrc_t contactRsrc(const DeviceWorld &w) {
  return __builtin_amdgcn_make_buffer_rsrc(static_cast<void *>(w.contacts), 0,
                                           int(w.pairCapacity * uint32_t(MAX_CONTACTS * sizeof(Contact))), BUF_WORD3);
}
AVBD_D float4 ldDev4(__amdgpu_buffer_rsrc_t r, uint32_t byteOffset) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, byteOffset, 0, BUF_SC1);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
AVBD_D void stDev4(__amdgpu_buffer_rsrc_t r, uint32_t byteOffset, float4 f) {
  const u32x4 v = {__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z), __float_as_uint(f.w)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, byteOffset, 0, BUF_SC1);
}
AVBD_D uint32_t bodyOff(uint32_t i, uint32_t field) { return i * uint32_t(sizeof(Body)) + field; }
AVBD_D uint32_t contactOff(uint32_t c, uint32_t field) { return c * uint32_t(sizeof(Contact)) + field; }
constexpr uint32_t OFF_POS = 0, OFF_ROT = 16, OFF_KEY = 12, OFF_PEN = 32, OFF_LAM = 48;
AVBD_D __amdgpu_buffer_rsrc_t jointRsrc(const DeviceWorld &w) {
  return __builtin_amdgcn_make_buffer_rsrc(static_cast<void *>(w.joints), 0, int(w.nJoints * uint32_t(sizeof(Joint))),
                                           BUF_WORD3);
}
/** The part of a joint record the duals write: penLin, penAng, lamLin, lamAng (its first 64 bytes). */
constexpr uint32_t JOINT_MUTABLE_RECORDS = 4;

/** This wave's earlier stores are complete before any later one is issued. */
AVBD_D void storesDone() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

/** Wait until *s >= target. Every wait ends: at w.dfSpin polls (about 65 ms by default), or at once when an earlier
 * wait gave up; a wait given up sets C_DF_ABORT, which ends every other wait. */
AVBD_D void waitFor(const DeviceWorld &w, const uint32_t *s, uint32_t target) {
  for (uint32_t spin = 0; ldAgent(s) < target; spin++) {
    if (spin >= w.dfSpin) {
      atomicOr(&w.counters[C_DF_ABORT], 1u);
      break;
    }
    if ((spin & 31u) == 31u && ldAgent(&w.counters[C_DF_ABORT]) != 0u) break;
    __builtin_amdgcn_s_sleep(1);
  }
  asm volatile("" ::: "memory");  // later loads are not moved above the wait
}

// --- joints ------------------------------------------------------------------------------------------------------------

/**
 * The two bodies of a constraint at the poses it is evaluated at. A may be the world: then rA is a world point and A's
 * rotation the identity. The colour launches read the poses from the records; the dataflow solve hands in the body's
 * own pose and reads its partner's with device-scope loads.
 */
struct JointPoses {
  float4 posA, rotA, posB, rotB;
  bool world;
};
AVBD_D JointPoses jointPoses(const DeviceWorld &w, int a, int b) {
  JointPoses P;
  P.world = a < 0;
  P.posA = P.world ? make_float4(0.f, 0.f, 0.f, 0.f) : w.bodies[a].pos;
  P.rotA = P.world ? make_float4(0.f, 0.f, 0.f, 1.f) : w.bodies[a].rot;
  P.posB = w.bodies[b].pos;
  P.rotB = w.bodies[b].rot;
  return P;
}

/** A joint's record as the solve reads it: in the dataflow solve the part the duals write with device-scope loads. */
template <int M> AVBD_D Joint loadJoint(const DeviceWorld &w, uint32_t j);
template <int M> AVBD_D void storeJoint(const DeviceWorld &w, uint32_t j, const Joint &k);
/** The poses a joint of body i is evaluated at: i's own as given, its partner's read (device scope in the dataflow). */
template <int M> AVBD_D JointPoses jointPosesFor(const DeviceWorld &w, int4 f, uint32_t i, float4 pos, float4 rot);

AVBD_D f3 anchorA(const Joint &k, const JointPoses &P) {
  if (P.world) return xyz(k.rA);
  return add(qrotate(P.rotA, xyz(k.rA)), xyz(P.posA));
}
AVBD_D f3 jointLinC(const Joint &k, const JointPoses &P) {
  return sub(anchorA(k, P), add(qrotate(P.rotB, xyz(k.rB)), xyz(P.posB)));
}
/**
 * The orientation a joint's angle rows hold B at: A's rotation times the rest orientation captured when the joint was
 * made ((c0Lin.w, c0Ang.w, rA.w, rB.w)), or A's rotation alone for a record without one (withJointRestFrames,
 * src/gpu/gpuSolverKernels.js:105-130).
 */
AVBD_D float4 restTarget(const Joint &k, const JointPoses &P) {
  const float4 t = P.rotA;
  const float4 r = make_float4(k.c0Lin.w, k.c0Ang.w, k.rA.w, k.rB.w);
  if (!(qlengthSq(r) > 0.5f)) return t;
  return make_float4((t.w * r.x + r.w * t.x) + (t.y * r.z - t.z * r.y), (t.w * r.y + r.w * t.y) + (t.z * r.x - t.x * r.z),
                     (t.w * r.z + r.w * t.z) + (t.x * r.y - t.y * r.x), t.w * r.w - ((t.x * r.x + t.y * r.y) + t.z * r.z));
}
AVBD_D f3 jointAngC(const Joint &k, const JointPoses &P) { return scale(qsub(restTarget(k, P), P.rotB), k.lamAng.w); }

// --- angular constraints: a motor or a limit about a hinge axis (src/gpu/angularConstraints.js:16-83) --------------

struct HingeFrame {
  f3 axis, ta, tb;  // A's axis, A's reference tangent, B's reference tangent, in world space
};
AVBD_D HingeFrame hingeFrame(const Joint &k, const JointPoses &P) {
  return HingeFrame{qrotate(P.rotA, xyz(k.rA)), qrotate(P.rotA, xyz(k.c0Lin)), qrotate(P.rotB, xyz(k.c0Ang))};
}
/** The twist of B about A's axis: the signed angle from A's tangent to B's. */
AVBD_D float hingeAngle(const HingeFrame &f) { return atan2r(dot(f.axis, cross(f.ta, f.tb)), dot(f.ta, f.tb)); }
/** Rows: a motor's one row (the twist this step beyond speed × dt, wrapped), a limit's lower and upper row. */
AVBD_D float2 hingeRows(const Joint &k, int32_t type, float angle, float alpha, float dt) {
  if (type == T_MOTOR) return make_float2(wrapAngle(angle - k.c0Lin.w) - k.rA.w * dt, 0.f);
  return make_float2((angle - k.rA.w) - alpha * k.c0Lin.w, (angle - k.rB.w) - alpha * k.c0Ang.w);
}
/** Forces: a motor's clamped to ± its torque bound, a limit's one-sided (lower pushes up, upper pushes down). */
AVBD_D float2 hingeForce(const Joint &k, int32_t type, float2 C) {
  const float fx = k.penLin.x * C.x + k.lamLin.x, fy = k.penLin.y * C.y + k.lamLin.y;
  if (type == T_MOTOR) return make_float2(wclamp(fx, -k.rB.w, k.rB.w), 0.f);
  return make_float2(rmin(0.f, fx), rmax(0.f, fy));
}
AVBD_D void warmHinge(const DeviceWorld &w, uint32_t j, int4 f) {
  Joint k = w.joints[j];
  if (k.penLin.w == 0.f) return;
  const WorldParams &p = w.worlds[w.bodyWorld[f.z]];
  const float angle = hingeAngle(hingeFrame(k, jointPoses(w, f.y, f.z)));
  k.c0Lin.w = f.x == T_MOTOR ? angle : rmin(0.f, angle - k.rA.w);
  k.c0Ang.w = rmax(0.f, angle - k.rB.w);
  const float decay = p.alpha * p.gamma;
  k.lamLin = mk4(scale(xyz(k.lamLin), decay), k.lamLin.w);
  k.penLin = make_float4(wclamp(k.penLin.x * p.gamma, PENALTY_MIN, PENALTY_MAX),
                         wclamp(k.penLin.y * p.gamma, PENALTY_MIN, PENALTY_MAX),
                         wclamp(k.penLin.z * p.gamma, PENALTY_MIN, PENALTY_MAX), k.penLin.w);
  w.joints[j] = k;
}
/** The dual of a motor or limit on its record; false when it has nothing to do. */
AVBD_D bool dualHinge(const WorldParams &p, Joint &k, int4 f, const JointPoses &P) {
  if (k.penLin.w == 0.f) return false;
  const float2 C = hingeRows(k, f.x, hingeAngle(hingeFrame(k, P)), p.alpha, p.dt);
  const float2 F = hingeForce(k, f.x, C);
  k.lamLin = make_float4(F.x, F.y, 0.f, k.lamLin.w);
  const float vx = f.x == T_MOTOR ? fabsf(C.x) : rmax(0.f, -C.x), vy = f.x == T_MOTOR ? fabsf(C.y) : rmax(0.f, C.y);
  k.penLin.x = rmin(k.penLin.x + vx * p.betaAng, PENALTY_MAX);
  k.penLin.y = rmin(k.penLin.y + vy * p.betaAng, PENALTY_MAX);
  return true;
}

/** Tensile failure of a spring (withSpringFracture, src/gpu/springMaterial.js:39-68): stretched past its break strain
 * (of the rest length) or pulling past its break force, it stops acting, keeping both thresholds. True when it broke. */
AVBD_D bool springFracture(Joint &k, const JointPoses &P) {
  const float strainLimit = k.lamLin.w, forceLimit = k.lamAng.w;
  if (!(k.penLin.w > 0.f && (strainLimit > 0.f || forceLimit > 0.f))) return false;
  const f3 pb = add(qrotate(P.rotB, xyz(k.rB)), xyz(P.posB));
  const float lengthNow = length(sub(anchorA(k, P), pb));
  const float extension = rmax(0.f, lengthNow - k.rA.w);
  const float strain = extension / rmax(k.rA.w, 1.0e-6f);
  const float tension = k.penLin.w * extension;
  if ((strainLimit > 0.f && k.rA.w > 1.0e-6f && strain > strainLimit) || (forceLimit > 0.f && tension > forceLimit)) {
    k.penLin = make_float4(0.f, 0.f, 0.f, 0.f);
    k.penAng = make_float4(0.f, 0.f, 0.f, 0.f);
    k.lamLin = make_float4(0.f, 0.f, 0.f, strainLimit);
    k.lamAng = make_float4(0.f, 0.f, 0.f, forceLimit);
    return true;
  }
  return false;
}
AVBD_D f3 jointLinRows(const Joint &k, const JointPoses &P, float alpha) {
  return sub(jointLinC(k, P), k.penLin.w >= HARD ? scale(xyz(k.c0Lin), alpha) : mk3(0.f, 0.f, 0.f));
}
AVBD_D f3 jointAngRows(const Joint &k, const JointPoses &P, float alpha) {
  return sub(jointAngC(k, P), k.penAng.w >= HARD ? scale(xyz(k.c0Ang), alpha) : mk3(0.f, 0.f, 0.f));
}
AVBD_D f3 penClamp(f3 pen, float gamma, float stiffness) {
  return mk3(rmin(wclamp(pen.x * gamma, PENALTY_MIN, PENALTY_MAX), stiffness),
             rmin(wclamp(pen.y * gamma, PENALTY_MIN, PENALTY_MAX), stiffness),
             rmin(wclamp(pen.z * gamma, PENALTY_MIN, PENALTY_MAX), stiffness));
}

__global__ void k_warm_joints(DeviceWorld w) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w.nJoints || held(w)) return;
  int4 f = w.info[j];
  if (f.x == T_MOTOR || f.x == T_LIMIT) {
    warmHinge(w, j, f);
    return;
  }
  if (f.x != T_JOINT) return;
  Joint k = w.joints[j];
  const WorldParams &p = w.worlds[w.bodyWorld[f.z]];
  // The w words keep the rest orientation
  const JointPoses P = jointPoses(w, f.y, f.z);
  k.c0Lin = mk4(jointLinC(k, P), k.c0Lin.w);
  k.c0Ang = mk4(jointAngC(k, P), k.c0Ang.w);
  float decay = p.alpha * p.gamma;
  k.lamLin = mk4(scale(xyz(k.lamLin), decay), k.lamLin.w);
  k.lamAng = mk4(scale(xyz(k.lamAng), decay), k.lamAng.w);
  k.penLin = mk4(penClamp(xyz(k.penLin), p.gamma, k.penLin.w), k.penLin.w);
  k.penAng = mk4(penClamp(xyz(k.penAng), p.gamma, k.penAng.w), k.penAng.w);
  w.joints[j] = k;
}

__global__ void k_warm_bodies(DeviceWorld w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies || held(w)) return;
  warmBody(w, i);
}

template <int M> AVBD_D Joint loadJoint(const DeviceWorld &w, uint32_t j) {
  Joint k = w.joints[j];
  if constexpr (M == SOLVE_DATAFLOW) {
    float4 *mut = &k.penLin;
#pragma unroll
    for (uint32_t q = 0; q < JOINT_MUTABLE_RECORDS; q++) mut[q] = ldDev4(jointRsrc(w), j * uint32_t(sizeof(Joint)) + 16u * q);
  }
  return k;
}
template <int M> AVBD_D void storeJoint(const DeviceWorld &w, uint32_t j, const Joint &k) {
  if constexpr (M == SOLVE_DATAFLOW) {
    const float4 *mut = &k.penLin;
#pragma unroll
    for (uint32_t q = 0; q < JOINT_MUTABLE_RECORDS; q++) stDev4(jointRsrc(w), j * uint32_t(sizeof(Joint)) + 16u * q, mut[q]);
  } else {
    w.joints[j] = k;
  }
}
template <int M> AVBD_D JointPoses jointPosesFor(const DeviceWorld &w, int4 f, uint32_t i, float4 pos, float4 rot) {
  auto pose = [&](int x, float4 &p, float4 &q) {
    if (uint32_t(x) == i) {
      p = pos;
      q = rot;
    } else if constexpr (M == SOLVE_DATAFLOW) {
      p = ldDev4(bodyRsrc(w), bodyOff(uint32_t(x), OFF_POS));
      q = ldDev4(bodyRsrc(w), bodyOff(uint32_t(x), OFF_ROT));
    } else {
      p = w.bodies[x].pos;
      q = w.bodies[x].rot;
    }
  };
  JointPoses P;
  P.world = f.y < 0;
  if (P.world) {
    P.posA = make_float4(0.f, 0.f, 0.f, 0.f);
    P.rotA = make_float4(0.f, 0.f, 0.f, 1.f);
  } else {
    pose(f.y, P.posA, P.rotA);
  }
  pose(f.z, P.posB, P.rotB);
  return P;
}

// --- primal --------------------------------------------------------------------------------------------------------

/** One body's Newton system: lower triangles of the linear and angular blocks, the full cross block (row = angular). */
struct Acc {
  float l00, l10, l11, l20, l21, l22;
  float a00, a10, a11, a20, a21, a22;
  float c[9];
  f3 rLin, rAng;
};

/** addRow (wgsl-solve.ts:53-59): lin += (l lᵀ) k, ang += (a aᵀ) k, cross += (a lᵀ) k, rhs += (l, a) f, each
 * accumulation one fused multiply-add (the CPU restatement uses std::fma at the same places). */
AVBD_D void addRow(Acc &s, f3 l, f3 a, float k, float f) {
  s.l00 = fmaf(l.x * l.x, k, s.l00);
  s.l10 = fmaf(l.y * l.x, k, s.l10);
  s.l11 = fmaf(l.y * l.y, k, s.l11);
  s.l20 = fmaf(l.z * l.x, k, s.l20);
  s.l21 = fmaf(l.z * l.y, k, s.l21);
  s.l22 = fmaf(l.z * l.z, k, s.l22);
  s.a00 = fmaf(a.x * a.x, k, s.a00);
  s.a10 = fmaf(a.y * a.x, k, s.a10);
  s.a11 = fmaf(a.y * a.y, k, s.a11);
  s.a20 = fmaf(a.z * a.x, k, s.a20);
  s.a21 = fmaf(a.z * a.y, k, s.a21);
  s.a22 = fmaf(a.z * a.z, k, s.a22);
  s.c[0] = fmaf(a.x * l.x, k, s.c[0]);
  s.c[1] = fmaf(a.x * l.y, k, s.c[1]);
  s.c[2] = fmaf(a.x * l.z, k, s.c[2]);
  s.c[3] = fmaf(a.y * l.x, k, s.c[3]);
  s.c[4] = fmaf(a.y * l.y, k, s.c[4]);
  s.c[5] = fmaf(a.y * l.z, k, s.c[5]);
  s.c[6] = fmaf(a.z * l.x, k, s.c[6]);
  s.c[7] = fmaf(a.z * l.y, k, s.c[7]);
  s.c[8] = fmaf(a.z * l.z, k, s.c[8]);
  s.rLin = mk3(fmaf(l.x, f, s.rLin.x), fmaf(l.y, f, s.rLin.y), fmaf(l.z, f, s.rLin.z));
  s.rAng = mk3(fmaf(a.x, f, s.rAng.x), fmaf(a.y, f, s.rAng.y), fmaf(a.z, f, s.rAng.z));
}

// --- the contact rows' arithmetic as fused multiply-adds -----------------------------------------------------------
// Every sum of products below is one chain of fused multiply-adds (one rounding per step, about half the VALU
// instructions of separate multiplies and adds); the CPU restatement forms each with std::fma in the same order.

AVBD_D float dotFused(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
AVBD_D f3 crossFused(f3 a, f3 b) {
  return mk3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
/** q v q* for a unit quaternion (layout.ts qrotate): v + 2 (w u + q⃗ × u) with u = q⃗ × v. */
AVBD_D f3 rotFused(float4 q, f3 v) {
  const f3 qv = xyz(q);
  const f3 u = crossFused(qv, v), t = crossFused(qv, u);
  return mk3(fmaf(2.f, fmaf(q.w, u.x, t.x), v.x), fmaf(2.f, fmaf(q.w, u.y, t.y), v.y),
             fmaf(2.f, fmaf(q.w, u.z, t.z), v.z));
}
/** qsub(a, b) = 2 (a b⁻¹)⃗ (layout.ts) from bi = b⁻¹, each component of the product one chain. */
AVBD_D f3 dispFused(float4 a, float4 bi) {
  return mk3(2.f * fmaf(a.w, bi.x, fmaf(a.x, bi.w, fmaf(a.y, bi.z, -(a.z * bi.y)))),
             2.f * fmaf(a.w, bi.y, fmaf(a.y, bi.w, fmaf(a.z, bi.x, -(a.x * bi.z)))),
             2.f * fmaf(a.w, bi.z, fmaf(a.z, bi.w, fmaf(a.x, bi.y, -(a.y * bi.x)))));
}

/**
 * The three rows of one contact point (normal and tangents b_q of the basis, signed by the body's side sg, lever arm r,
 * penalties k, forces f) added at once: with M = Σ k_q b_q b_qᵀ (the sign cancels) and g = sg Σ f_q b_q, the linear
 * block takes M, the cross block [r]× M, the angular block [r]× M [r]×ᵀ, the right-hand side (g, r × g). The same
 * system as three addRow calls, in about half the arithmetic (the CPU restatement forms it the same way).
 */
AVBD_D void addContactRows(Acc &s, const Basis &b, float sg, f3 r, f3 k, f3 f) {
  const f3 kn = scale(b.n, k.x), k1 = scale(b.t1, k.y), k2 = scale(b.t2, k.z);
  const float m00 = fmaf(k2.x, b.t2.x, fmaf(k1.x, b.t1.x, kn.x * b.n.x));
  const float m10 = fmaf(k2.y, b.t2.x, fmaf(k1.y, b.t1.x, kn.y * b.n.x));
  const float m11 = fmaf(k2.y, b.t2.y, fmaf(k1.y, b.t1.y, kn.y * b.n.y));
  const float m20 = fmaf(k2.z, b.t2.x, fmaf(k1.z, b.t1.x, kn.z * b.n.x));
  const float m21 = fmaf(k2.z, b.t2.y, fmaf(k1.z, b.t1.y, kn.z * b.n.y));
  const float m22 = fmaf(k2.z, b.t2.z, fmaf(k1.z, b.t1.z, kn.z * b.n.z));
  // C = [r]× M (row: angular, column: linear)
  const float c00 = fmaf(r.y, m20, -(r.z * m10)), c01 = fmaf(r.y, m21, -(r.z * m11)), c02 = fmaf(r.y, m22, -(r.z * m21));
  const float c10 = fmaf(r.z, m00, -(r.x * m20)), c11 = fmaf(r.z, m10, -(r.x * m21)), c12 = fmaf(r.z, m20, -(r.x * m22));
  const float c20 = fmaf(r.x, m10, -(r.y * m00)), c21 = fmaf(r.x, m11, -(r.y * m10)), c22 = fmaf(r.x, m21, -(r.y * m20));
  // Angular block C [r]×ᵀ, lower triangle, accumulated by the same chains
  s.a00 = fmaf(r.y, c02, fmaf(-r.z, c01, s.a00));
  s.a10 = fmaf(r.y, c12, fmaf(-r.z, c11, s.a10));
  s.a11 = fmaf(r.z, c10, fmaf(-r.x, c12, s.a11));
  s.a20 = fmaf(r.y, c22, fmaf(-r.z, c21, s.a20));
  s.a21 = fmaf(r.z, c20, fmaf(-r.x, c22, s.a21));
  s.a22 = fmaf(r.x, c21, fmaf(-r.y, c20, s.a22));
  s.l00 += m00;
  s.l10 += m10;
  s.l11 += m11;
  s.l20 += m20;
  s.l21 += m21;
  s.l22 += m22;
  s.c[0] += c00;
  s.c[1] += c01;
  s.c[2] += c02;
  s.c[3] += c10;
  s.c[4] += c11;
  s.c[5] += c12;
  s.c[6] += c20;
  s.c[7] += c21;
  s.c[8] += c22;
  const f3 g = scale(mk3(fmaf(f.z, b.t2.x, fmaf(f.y, b.t1.x, f.x * b.n.x)), fmaf(f.z, b.t2.y, fmaf(f.y, b.t1.y, f.x * b.n.y)),
                         fmaf(f.z, b.t2.z, fmaf(f.y, b.t1.z, f.x * b.n.z))),
                     sg);
  s.rLin = add(s.rLin, g);
  s.rAng = mk3(fmaf(r.y, g.z, fmaf(-r.z, g.y, s.rAng.x)), fmaf(r.z, g.x, fmaf(-r.x, g.z, s.rAng.y)),
               fmaf(r.x, g.y, fmaf(-r.y, g.x, s.rAng.z)));
}

/** A joint's rows for body i (addJoint, wgsl-solve.ts:90-140; addAngular, angularConstraints.js:36-44), from its record
 * and its bodies' poses. */
AVBD_D void addJoint(const DeviceWorld &w, Acc &acc, int4 f, const Joint &k, const JointPoses &P, float alpha,
                     uint32_t i) {
  const int a = f.y, b = f.z;
  bool isA = int(i) == a;
  float sg = isA ? 1.f : -1.f;
  if (f.x == T_MOTOR || f.x == T_LIMIT) {
    if (k.penLin.w == 0.f || (f.x == T_MOTOR && k.rB.w == 0.f)) return;
    const HingeFrame hf = hingeFrame(k, P);
    const f3 axis = scale(hf.axis, isA ? -1.f : 1.f);
    const float2 C = hingeRows(k, f.x, hingeAngle(hf), alpha, w.worlds[w.bodyWorld[b]].dt);
    const float2 F = hingeForce(k, f.x, C);
    if (f.x == T_MOTOR || F.x != 0.f) addRow(acc, mk3(0.f, 0.f, 0.f), axis, k.penLin.x, F.x);
    if (f.x == T_LIMIT && F.y != 0.f) addRow(acc, mk3(0.f, 0.f, 0.f), axis, k.penLin.y, F.y);
    return;
  }
  if (f.x == T_SPRING) {
    f3 rAW = qrotate(P.rotA, xyz(k.rA));
    f3 rBW = qrotate(P.rotB, xyz(k.rB));
    f3 d = sub(add(rAW, xyz(P.posA)), add(rBW, xyz(P.posB)));
    float len = length(d);
    if (len <= 1.0e-6f) return;
    f3 n = divs(d, len);
    float stiffness = k.penLin.w;
    f3 r = isA ? rAW : rBW;
    addRow(acc, scale(n, sg), scale(cross(r, n), sg), stiffness, stiffness * (len - k.rA.w));
    return;
  }
  f3 penLin = xyz(k.penLin), penAng = xyz(k.penAng);
  if (dot(penLin, penLin) > 0.f) {
    f3 F = add(mul3(penLin, jointLinRows(k, P, alpha)), xyz(k.lamLin));
    f3 r = isA ? qrotate(P.rotA, xyz(k.rA)) : neg(qrotate(P.rotB, xyz(k.rB)));
    f3 v = neg(r);
    addRow(acc, mk3(sg, 0.f, 0.f), mk3(0.f, -v.z, v.y), penLin.x, F.x);
    addRow(acc, mk3(0.f, sg, 0.f), mk3(v.z, 0.f, -v.x), penLin.y, F.y);
    addRow(acc, mk3(0.f, 0.f, sg), mk3(-v.y, v.x, 0.f), penLin.z, F.z);
    float rf = dot(r, F);
    acc.a00 += length(sub(scale(r, F.x), mk3(rf, 0.f, 0.f)));
    acc.a11 += length(sub(scale(r, F.y), mk3(0.f, rf, 0.f)));
    acc.a22 += length(sub(scale(r, F.z), mk3(0.f, 0.f, rf)));
  }
  if (dot(penAng, penAng) > 0.f) {
    f3 F = add(mul3(penAng, jointAngRows(k, P, alpha)), xyz(k.lamAng));
    float s = sg * k.lamAng.w;
    addRow(acc, mk3(0.f, 0.f, 0.f), mk3(s, 0.f, 0.f), penAng.x, F.x);
    addRow(acc, mk3(0.f, 0.f, 0.f), mk3(0.f, s, 0.f), penAng.y, F.y);
    addRow(acc, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, s), penAng.z, F.z);
  }
}

/**
 * The dual of a joint on its record (dual, wgsl-solve.ts:474-535 for joints; a spring's fracture; a motor's or limit's
 * ramp) at its bodies' poses; true when the record changed.
 */
AVBD_D bool dualJointRecord(const DeviceWorld &w, int4 f, Joint &k, const JointPoses &P) {
  if (f.x == T_SPRING) return springFracture(k, P);
  const WorldParams &p = w.worlds[w.bodyWorld[f.z]];
  if (f.x == T_MOTOR || f.x == T_LIMIT) return dualHinge(p, k, f, P);
  if (f.x != T_JOINT) return false;
  float alpha = p.alpha;
  f3 penLin = xyz(k.penLin), penAng = xyz(k.penAng);
  if (dot(penLin, penLin) > 0.f) {
    f3 C = jointLinRows(k, P, alpha);
    if (k.penLin.w >= HARD) k.lamLin = mk4(add(mul3(penLin, C), xyz(k.lamLin)), k.lamLin.w);
    float cap = rmin(k.penLin.w, PENALTY_MAX);
    f3 q = add(penLin, scale(vabs(C), p.betaLin));
    k.penLin = make_float4(rmin(q.x, cap), rmin(q.y, cap), rmin(q.z, cap), k.penLin.w);
  }
  if (dot(penAng, penAng) > 0.f) {
    f3 C = jointAngRows(k, P, alpha);
    if (k.penAng.w >= HARD) k.lamAng = mk4(add(mul3(penAng, C), xyz(k.lamAng)), k.lamAng.w);
    float cap = rmin(k.penAng.w, PENALTY_MAX);
    f3 q = add(penAng, scale(vabs(C), p.betaAng));
    k.penAng = make_float4(rmin(q.x, cap), rmin(q.y, cap), rmin(q.z, cap), k.penAng.w);
  }
  float frac = k.lamLin.w;
  float limit = fabsf(frac);
  f3 ll = xyz(k.lamLin), la = xyz(k.lamAng);
  bool linear = frac < 0.f && dot(ll, ll) > limit * limit;
  if (limit < BIG && (dot(la, la) > limit * limit || linear)) {
    k.penLin = make_float4(0.f, 0.f, 0.f, 0.f);
    k.penAng = make_float4(0.f, 0.f, 0.f, 0.f);
    k.lamLin = make_float4(0.f, 0.f, 0.f, frac);
    k.lamAng = make_float4(0.f, 0.f, 0.f, k.lamAng.w);
  }
  return true;
}

/** What a contact needs of one body: rotation, displacement since x- (wgsl-solve.ts:145-157), its index. */
struct PairBody {
  float4 rot;
  f3 dLin, dAng;
  uint32_t index;
};
/** Body i at pose (pos, rot): the angular displacement qsub(rot, x- rotation) from the warm start's inverse of the x-
 * rotation. */
AVBD_D PairBody pairBodyAt(const DeviceWorld &w, uint32_t i, float4 pos, float4 rot) {
  PairBody pb;
  pb.rot = rot;
  pb.dLin = sub(xyz(pos), xyz(w.bodies[i].ipos));
  pb.dAng = dispFused(rot, w.irotInv[i]);
  pb.index = i;
  return pb;
}
template <int M = SOLVE_COLORS> AVBD_D PairBody pairBody(const DeviceWorld &w, uint32_t i) {
  if constexpr (M == SOLVE_DATAFLOW)
    return pairBodyAt(w, i, ldDev4(bodyRsrc(w), bodyOff(i, OFF_POS)), ldDev4(bodyRsrc(w), bodyOff(i, OFF_ROT)));
  else
    return pairBodyAt(w, i, w.bodies[i].pos, w.bodies[i].rot);
}

struct Eval {
  f3 rSelf, C, F;  // the evaluating body's lever arm, the rows' C and forces
  float frictionScale, bounds;
};

/** Whether contact k's lever arms are taken at the step-start rotations (sphere contacts; capsule contacts with
 * FLAG_CAPSULE_STEP_START_ARMS) rather than the current ones. */
AVBD_D bool stepStartArms(const Contact &k, uint32_t flags) {
  const uint32_t feature = fbits(k.rA.w) & ~STICK_BIT;
  return feature == SPHERE_FEATURE ||
         ((flags & FLAG_CAPSULE_STEP_START_ARMS) && (feature >> 24) == (CAPSULE_FEATURE >> 24));
}

/** The partner's share of the anchors' relative motion D: its displacement, and its Δθ × r at its anchor. */
struct PartnerTerms {
  f3 lin, ang;
};
AVBD_D PartnerTerms partnerTerms(const DeviceWorld &w, const Contact &k, const PairBody &other, bool selfIsA,
                                 bool stepStart) {
  const f3 rOther = rotFused(stepStart ? w.bodies[other.index].irot : other.rot, selfIsA ? xyz(k.rB) : xyz(k.rA));
  return PartnerTerms{other.dLin, crossFused(other.dAng, rOther)};
}

/** evalContact with the partner's share of D already formed (partnerTerms). */
AVBD_D Eval evalContactTerms(const DeviceWorld &w, const Contact &k, const Basis &basis, float friction,
                             const PairBody &self, const PartnerTerms &o, bool selfIsA, bool stepStart, float alpha) {
  Eval e;
  e.rSelf = rotFused(stepStart ? w.bodies[self.index].irot : self.rot, selfIsA ? xyz(k.rA) : xyz(k.rB));
  const f3 D = add(sub(self.dLin, o.lin), sub(crossFused(self.dAng, e.rSelf), o.ang));
  const float sg = selfIsA ? 1.f : -1.f, keep = 1.f - alpha;
  e.C = mk3(fmaf(k.rB.w, keep, sg * dotFused(basis.n, D)), fmaf(k.pen.w, keep, sg * dotFused(basis.t1, D)),
            fmaf(k.lam.w, keep, sg * dotFused(basis.t2, D)));
  f3 F = mk3(fmaf(k.pen.x, e.C.x, k.lam.x), fmaf(k.pen.y, e.C.y, k.lam.y), fmaf(k.pen.z, e.C.z, k.lam.z));
  F.x = rmin(F.x, 0.f);
  e.bounds = fabsf(F.x) * friction;
  e.frictionScale = sqrtf(fmaf(F.y, F.y, F.z * F.z));
  if (e.frictionScale > e.bounds && e.frictionScale > 0.f) {
    float s = e.bounds / e.frictionScale;
    F.y = F.y * s;
    F.z = F.z * s;
  }
  e.F = F;
  return e;
}

/**
 * evalContact (wgsl-solve.ts:171-206) from either body's side: `self` is the body whose rows are wanted (selfIsA: it is
 * the contact's body A), `other` its partner. The rows' C = C0 (1 − α) + J Δx take the displacement once as the
 * relative motion of the two anchors, D = (Δx_self − Δx_other) + (Δθ_self × r_self − Δθ_other × r_other), each row as
 * C0 (1 − α) + sg (l · D) with sg = +1 for A. Exchanging the bodies negates D exactly (a − b = −(b − a) in IEEE
 * rounding), so both sides compute the same C and forces bit for bit, as the dual does from A's side. (The reference
 * sums l·Δx_A − l·Δx_B + (r_A × l)·Δθ_A − (r_B × l)·Δθ_B per row: the same value, rounded differently; the CPU
 * restatement forms it as here.)
 * Sphere contacts take their lever arms at the step-start rotation (wgsl-solve.ts:172-181): a rolling sphere would
 * otherwise see a false separation in its normal row. Capsule contacts (keys 4<<24 | j) do so only with
 * FLAG_CAPSULE_STEP_START_ARMS; by default they use the current rotation, as the reference does.
 */
AVBD_D Eval evalContact(const DeviceWorld &w, const Contact &k, const Basis &basis, float friction,
                        const PairBody &self, const PairBody &other, bool selfIsA, float alpha, uint32_t flags) {
  const bool stepStart = stepStartArms(k, flags);
  return evalContactTerms(w, k, basis, friction, self, partnerTerms(w, k, other, selfIsA, stepStart), selfIsA,
                          stepStart, alpha);
}

/**
 * Whether the dynamic body of colour ca finishes its pair with a dynamic body of colour cb in the dataflow solve: the
 * later colour runs the pair's dual right after its primal. Partners never share a colour in a solved step (a step
 * whose colouring clashed is held), so exactly one side finishes each pair.
 */
AVBD_D bool finishesPair(uint32_t ca, uint32_t cb) { return ca > cb; }

/** Stash slot value for an item that is not stashed. */
constexpr uint32_t NO_STASH = ~0u;

/** Store / load a stashed contact item (slot t of the group; word f at stash[f * DF_WAVE + t]). */
AVBD_D void stashItem(float *stash, uint32_t t, const Contact &k, const PartnerTerms &o) {
  const float v[STASH_WORDS] = {k.rA.x,  k.rA.y,  k.rA.z,  k.rA.w,  k.rB.x,  k.rB.y,  k.rB.z,  k.rB.w,
                                k.pen.x, k.pen.y, k.pen.z, k.pen.w, k.lam.x, k.lam.y, k.lam.z, k.lam.w,
                                o.lin.x, o.lin.y, o.lin.z, o.ang.x, o.ang.y, o.ang.z};
#pragma unroll
  for (uint32_t f = 0; f < STASH_WORDS; f++) stash[f * DF_WAVE + t] = v[f];
}
AVBD_D void unstashItem(const float *stash, uint32_t t, Contact &k, PartnerTerms &o) {
  float v[STASH_WORDS];
#pragma unroll
  for (uint32_t f = 0; f < STASH_WORDS; f++) v[f] = stash[f * DF_WAVE + t];
  k.rA = make_float4(v[0], v[1], v[2], v[3]);
  k.rB = make_float4(v[4], v[5], v[6], v[7]);
  k.pen = make_float4(v[8], v[9], v[10], v[11]);
  k.lam = make_float4(v[12], v[13], v[14], v[15]);
  o.lin = mk3(v[16], v[17], v[18]);
  o.ang = mk3(v[19], v[20], v[21]);
}

/**
 * One work item of body i: a joint's rows, or the three rows of contact point p of a manifold entry. In the dataflow
 * solve a contact item with a stash slot (not NO_STASH) is also stashed there for the body's fused dual.
 */
template <int M, bool J>
AVBD_D void addItem(const DeviceWorld &w, Acc &acc, uint32_t i, const PairBody &self, float4 selfPos, uint4 head,
                    uint4 tail, uint32_t p, float alpha, uint32_t flags, const DfCtx &df, uint32_t slot) {
  if constexpr (J) {  // J: the batch has joints (their code is compiled out otherwise)
    if (head.z & ENTRY_JOINT) {
      const uint32_t j = head.y;
      const int4 f = w.info[j];
      addJoint(w, acc, f, loadJoint<M>(w, j), jointPosesFor<M>(w, f, i, selfPos, self.rot), alpha, i);
      return;
    }
  }
  constexpr bool DF = M == SOLVE_DATAFLOW;
  const bool isA = (head.z & ENTRY_IS_A) != 0;
  const PairBody other = pairBody<M>(w, head.x);
  const Basis basis = orthonormalScaled(mk3(__uint_as_float(tail.x), __uint_as_float(tail.y), __uint_as_float(tail.z)),
                                        __uint_as_float(tail.w));
  Contact k = w.contacts[head.y + p];
  if constexpr (DF) {
    k.pen = ldDev4(contactRsrc(w), contactOff(head.y + p, OFF_PEN));
    k.lam = ldDev4(contactRsrc(w), contactOff(head.y + p, OFF_LAM));
  }
  const bool stepStart = stepStartArms(k, flags);
  const PartnerTerms o = partnerTerms(w, k, other, isA, stepStart);
  if constexpr (DF) {
    if (slot != NO_STASH) stashItem(df.stash, slot, k, o);
  }
  const Eval ev = evalContactTerms(w, k, basis, __uint_as_float(head.w), self, o, isA, stepStart, alpha);
  addContactRows(acc, basis, isA ? 1.f : -1.f, ev.rSelf, xyz(k.pen), ev.F);
}

/**
 * Lane `lane` of body i's system (accumulate, wgsl-solve.ts:351-413), W lanes per body (LANES, or HEAVY_LANES for a
 * body with more than LANES adjacency entries). Lane 0 holds the inertia terms. The body's entries are taken in chunks
 * of W, one entry record per lane; the chunk's work items (a joint, or one contact point) are numbered in entry order
 * by a prefix sum over the group and item t goes to lane t mod W, which finds its entry by a binary search over the
 * group's prefix sums and takes the record from the lane that loaded it. A manifold with many points is thus spread
 * over several lanes. The partial systems are then summed by a fixed tree, so the result does not depend on scheduling.
 */
template <uint32_t W, int M, bool J>
AVBD_D void accumulate(const DeviceWorld &w, uint32_t i, uint32_t first, uint32_t end, uint32_t lane, float alpha,
                       float dt, uint32_t flags, Acc &acc, const DfCtx &df, float4 &bpos, float4 &brot) {
  constexpr bool DF = M == SOLVE_DATAFLOW;
  const Body &B = w.bodies[i];
  if constexpr (DF) {
    if (df.it) waitFor(w, &w.dualDone[i], df.base + df.it);  // the body's own last iteration and its duals
  }
  // The body's pose, read once: its rows, and the update finishBody makes
  bpos = DF ? ldDev4(bodyRsrc(w), bodyOff(i, OFF_POS)) : B.pos;
  brot = DF ? ldDev4(bodyRsrc(w), bodyOff(i, OFF_ROT)) : B.rot;
  acc.l00 = acc.l10 = acc.l11 = acc.l20 = acc.l21 = acc.l22 = 0.f;
  acc.a00 = acc.a10 = acc.a11 = acc.a20 = acc.a21 = acc.a22 = 0.f;
#pragma unroll
  for (int q = 0; q < 9; q++) acc.c[q] = 0.f;
  acc.rLin = mk3(0.f, 0.f, 0.f);
  acc.rAng = mk3(0.f, 0.f, 0.f);
  if (lane == 0) {
    float dt2 = dt * dt;
    float m = B.size.w / dt2;
    f3 I = divs(xyz(B.moment), dt2);
    acc.l00 = m; acc.l11 = m; acc.l22 = m;
    acc.a00 = I.x; acc.a11 = I.y; acc.a22 = I.z;
    acc.rLin = scale(sub(xyz(bpos), xyz(B.inertialPos)), m);
    acc.rAng = mul3(I, qsub(brot, B.inertialRot));
  }
  if (first >= end) return;  // the same for every lane of the group
  const PairBody self = pairBodyAt(w, i, bpos, brot);
  for (uint32_t c0 = first; c0 < end; c0 += W) {
    const uint32_t e = c0 + lane;
    uint4 head = make_uint4(0u, 0u, 0u, 0u), tail = make_uint4(0u, 0u, 0u, 0u);
    if (e < end) {
      head = w.adjEntry[2ull * e];
      tail = w.adjEntry[2ull * e + 1];
    }
    if constexpr (DF) {
      // Each lane waits for its entry's partner before any item is evaluated: its pose of this iteration if its colour
      // comes first, else of the last one. The entry's multipliers come from the last iteration's dual, which the
      // pair's finisher ran right after its primal: this body (its own stamp was waited for) or the partner, whose
      // dual stamp then covers both.
      if (e < end && (head.z & ENTRY_DYNAMIC)) {
        const uint32_t pc = head.z >> ENTRY_COLOR_SHIFT;  // the partner's colour (annotated after the colouring)
        if (pc < df.color)
          waitFor(w, &w.bodyDone[head.x], df.base + df.it + 1u);
        else if (df.it != 0u)
          waitFor(w, finishesPair(pc, df.color) ? &w.dualDone[head.x] : &w.bodyDone[head.x], df.base + df.it);
      }
    }
    const uint32_t items = e < end ? ((head.z & ENTRY_JOINT) ? 1u : (head.z & 15u)) : 0u;
    uint32_t incl = items;
#pragma unroll
    for (uint32_t s = 1; s < W; s <<= 1) {
      uint32_t o = __shfl_up(incl, s, W);
      if (lane >= s) incl += o;
    }
    const uint32_t excl = incl - items;
    const uint32_t total = __shfl(incl, W - 1, W);
    for (uint32_t r0 = 0; r0 < total; r0 += W) {  // the same turns for every lane of the group
      const uint32_t t = r0 + lane;
      uint32_t owner = 0;  // the last lane whose first item is at or before t (prefix sums never decrease)
#pragma unroll
      for (uint32_t step = W / 2; step > 0; step >>= 1) {
        const uint32_t ex = __shfl(excl, owner + step, W);
        if (ex <= t) owner += step;
      }
      const uint4 h = make_uint4(__shfl(head.x, owner, W), __shfl(head.y, owner, W), __shfl(head.z, owner, W),
                                 __shfl(head.w, owner, W));
      const uint4 tl = make_uint4(__shfl(tail.x, owner, W), __shfl(tail.y, owner, W), __shfl(tail.z, owner, W),
                                  __shfl(tail.w, owner, W));
      const uint32_t p = t - __shfl(excl, owner, W);
      // The first round of the first chunk is stashed for the fused dual (slot t of the group)
      const uint32_t slot = DF && c0 == first && r0 == 0 ? t : NO_STASH;
      if (t < total) addItem<M, J>(w, acc, i, self, bpos, h, tl, p, alpha, flags, df, slot);
    }
  }
}

/**
 * acc += the partial system of the lane S above (lanes below S of each group of W use the result). Within a row of 16
 * lanes the value comes by DPP row shift (row_shl:S, folded into the add); S = 16 crosses rows and goes through
 * ds_bpermute. Lanes at or above S add garbage that no later level reads.
 */
template <uint32_t W, uint32_t S> AVBD_D void mergeFromAbove(Acc &acc) {
  float *a = reinterpret_cast<float *>(&acc);
  constexpr int N = sizeof(Acc) / sizeof(float);
#pragma unroll
  for (int q = 0; q < N; q++) {
    float o;
    if constexpr (S < 16)
      o = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a[q]), 0x100 + S, 0xF, 0xF, false));
    else
      o = __shfl_down(a[q], S, W);
    a[q] = a[q] + o;
  }
}

template <uint32_t W> AVBD_D void mergeTree(Acc &acc) {
  if constexpr (W >= 32) mergeFromAbove<W, 16>(acc);
  if constexpr (W >= 16) mergeFromAbove<W, 8>(acc);
  if constexpr (W >= 8) mergeFromAbove<W, 4>(acc);
  if constexpr (W >= 4) mergeFromAbove<W, 2>(acc);
  mergeFromAbove<W, 1>(acc);
}

template <class T> AVBD_D T fmaT(T a, T b, T c);
template <> AVBD_D float fmaT<float>(float a, float b, float c) { return fmaf(a, b, c); }
template <> AVBD_D double fmaT<double>(double a, double b, double c) { return fma(a, b, c); }

/**
 * LDLᵀ of the symmetric 6x6 system (lower triangle `A`, row-major by rows 1..6) with each pivot after the first floored
 * at PIVOT_FLOOR of its diagonal (wgsl-solve.ts:416-470), then x = A⁻¹ rhs. Each pivot is inverted once and its column
 * scaled by the reciprocal (6 correctly rounded divisions instead of 21), and every sum of products (D_k = A_kk −
 * Σ_j L_kj² D_j, L_ik = (A_ik − Σ_j L_ij L_kj D_j) / D_k, the forward and back substitutions) is one chain of fused
 * multiply-adds in increasing j (the CPU restatement does the same). `floored` reports whether any pivot took its
 * floor (the unfloored pivot was not above it).
 */
template <class T> AVBD_D void ldlt6(const T (&A)[21], const T (&b)[6], T (&x)[6], bool &floored) {
  const T fl = T(PIVOT_FLOOR);
  T L[6][6], D[6], r[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    T d = A[k * (k + 1) / 2 + k];
#pragma unroll
    for (int j = 0; j < k; j++) d = fmaT<T>(-(L[k][j] * L[k][j]), D[j], d);
    if (k > 0) {
      const T f = A[k * (k + 1) / 2 + k] * fl;
      floored = floored || !(d > f);
      d = d > f ? d : f;
    }
    D[k] = d;
    r[k] = T(1) / d;
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
      T a = A[i * (i + 1) / 2 + k];
#pragma unroll
      for (int j = 0; j < k; j++) a = fmaT<T>(-(L[i][j] * L[k][j]), D[j], a);
      L[i][k] = a * r[k];
    }
  }
  T y[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    y[k] = b[k];
#pragma unroll
    for (int j = 0; j < k; j++) y[k] = fmaT<T>(-L[k][j], y[j], y[k]);
  }
#pragma unroll
  for (int k = 5; k >= 0; k--) {
    x[k] = y[k] * r[k];
#pragma unroll
    for (int j = k + 1; j < 6; j++) x[k] = fmaT<T>(-L[j][k], x[j], x[k]);
  }
}

/** The FP64 solve of a system whose FP32 factorisation hit a pivot floor. */
AVBD_D void ldlt6Double(const float (&A)[21], const float (&b)[6], float (&x)[6]) {
  double Ad[21], bd[6], xd[6];
#pragma unroll
  for (int q = 0; q < 21; q++) Ad[q] = double(A[q]);
#pragma unroll
  for (int q = 0; q < 6; q++) bd[q] = double(b[q]);
  bool floored = false;
  ldlt6<double>(Ad, bd, xd, floored);
#pragma unroll
  for (int q = 0; q < 6; q++) x[q] = float(xd[q]);
}

/**
 * Solve the body's 6x6 system and move it: pos -= dx, rot = qadd(rot, -dθ) (finishBody, wgsl-solve.ts:416-470). FP32
 * first; a near-singular system (a pivot at its floor, as stiff joint chains give) is solved again in FP64 from the
 * same FP32 entries, with the same floor.
 */
template <int M>
AVBD_D void finishBody(const DeviceWorld &w, uint32_t i, const Acc &s, const DfCtx &df, float4 pos, float4 rot,
                       float4 *newPos = nullptr, float4 *newRot = nullptr) {
  const float A[21] = {s.l00, s.l10, s.l11, s.l20, s.l21, s.l22, s.c[0], s.c[1], s.c[2], s.a00, s.c[3],
                       s.c[4], s.c[5], s.a10, s.a11, s.c[6], s.c[7], s.c[8], s.a20, s.a21, s.a22};
  const float b[6] = {s.rLin.x, s.rLin.y, s.rLin.z, s.rAng.x, s.rAng.y, s.rAng.z};
  float x[6];
  bool floored = false;
  ldlt6<float>(A, b, x, floored);
  if (floored) {
    ldlt6Double(A, b, x);
    atomicAdd(&w.counters[C_PIVOT64], 1u);
  }
  Body &B = w.bodies[i];
  if constexpr (M == SOLVE_DATAFLOW) {
    *newPos = make_float4(pos.x - x[0], pos.y - x[1], pos.z - x[2], pos.w);
    *newRot = qadd(rot, mk3(-x[3], -x[4], -x[5]));
    stDev4(bodyRsrc(w), bodyOff(i, OFF_POS), *newPos);
    stDev4(bodyRsrc(w), bodyOff(i, OFF_ROT), *newRot);
    storesDone();
    stAgent(&w.bodyDone[i], df.base + df.it + 1u);
  } else {
    pos = make_float4(pos.x - x[0], pos.y - x[1], pos.z - x[2], pos.w);
    rot = qadd(rot, mk3(-x[3], -x[4], -x[5]));
    B.pos = pos;
    B.rot = rot;
    if constexpr (M == SOLVE_WORLDS) {
      *newPos = pos;
      *newRot = rot;
    }
  }
}

/**
 * The dual update of contact slot c (dual, wgsl-solve.ts:474-535) from its record kc and its evaluation at this
 * iteration's poses: the penalty ramp where the row is active, λ = the clamped force, and the stick flag.
 */
template <int M> AVBD_D void dualUpdate(const DeviceWorld &w, uint32_t c, const Contact &kc, const Eval &e,
                                         const WorldParams &p) {
  f3 pen = xyz(kc.pen);
  if (e.F.x < 0.f) pen.x = rmin(pen.x + p.betaLin * fabsf(e.C.x), PENALTY_MAX);
  uint32_t key = fbits(kc.rA.w);
  const bool ramp = e.frictionScale <= e.bounds;
  if (ramp) {
    pen.y = rmin(pen.y + p.betaLin * fabsf(e.C.y), PENALTY_MAX);
    pen.z = rmin(pen.z + p.betaLin * fabsf(e.C.z), PENALTY_MAX);
    bool stick = sqrtf(e.C.y * e.C.y + e.C.z * e.C.z) < STICK_THRESH;
    key = (key & ~STICK_BIT) | (stick ? STICK_BIT : 0u);
  }
  if constexpr (M == SOLVE_DATAFLOW) {
    if (ramp) __builtin_amdgcn_raw_buffer_store_b32(key, contactRsrc(w), contactOff(c, OFF_KEY), 0, BUF_SC1);
    stDev4(contactRsrc(w), contactOff(c, OFF_PEN), mk4(pen, kc.pen.w));
    stDev4(contactRsrc(w), contactOff(c, OFF_LAM), mk4(e.F, kc.lam.w));
  } else {
    if (ramp) w.contacts[c].rA.w = bitsf(key);
    w.contacts[c].pen = mk4(pen, kc.pen.w);
    w.contacts[c].lam = mk4(e.F, kc.lam.w);
  }
}

/**
 * The dataflow solve's dual for the pairs body i finishes (its partner is static or of an earlier colour), right after
 * its primal: each such contact point is evaluated at the body's new pose
 * (pos, rot, held by every lane of the group) and the partner's pose of this iteration (an earlier colour's was waited
 * for by the primal), and updated as the dual pass does. The items are spread over the group's lanes as in
 * accumulate. Then the body's dual stamp.
 */
template <uint32_t W, int M, bool J>
AVBD_D void fusedDual(const DeviceWorld &w, uint4 slot, uint32_t lane, float4 pos, float4 rot, const DfCtx &df) {
  constexpr bool DF = M == SOLVE_DATAFLOW;
  const uint32_t i = slot.x, first = slot.y, end = slot.z;
  const WorldParams &p = w.worlds[slot.w];
  const PairBody self = pairBodyAt(w, i, pos, rot);
  for (uint32_t c0 = first; c0 < end; c0 += W) {
    const uint32_t e = c0 + lane;
    uint4 head = make_uint4(0u, 0u, 0u, 0u), tail = make_uint4(0u, 0u, 0u, 0u);
    uint32_t items = 0;
    if (e < end) {
      head = w.adjEntry[2ull * e];
      tail = w.adjEntry[2ull * e + 1];
      const bool finishes = !(head.z & ENTRY_DYNAMIC) ||
                            finishesPair(df.color, DF ? head.z >> ENTRY_COLOR_SHIFT : w.colorA[head.x]);
      // A joint's dual is one item here in the dataflow solve (the workgroup-per-world solve runs the joints' duals
      // after each iteration's colours)
      if (finishes) items = (head.z & ENTRY_JOINT) ? (DF && J ? 1u : 0u) : (head.z & 15u);
    }
    uint32_t incl = items;
#pragma unroll
    for (uint32_t s = 1; s < W; s <<= 1) {
      uint32_t o = __shfl_up(incl, s, W);
      if (lane >= s) incl += o;
    }
    const uint32_t excl = incl - items;
    const uint32_t total = __shfl(incl, W - 1, W);
    // In the first chunk, each entry's first item number in the primal's numbering (every item, finished here or
    // not): items the primal took in its first round are in the stash
    const bool stashed = DF && c0 == first;  // the same for every lane of the group
    uint32_t exclAll = 0;
    if (stashed) {
      const uint32_t all = e < end ? ((head.z & ENTRY_JOINT) ? 1u : (head.z & 15u)) : 0u;
      uint32_t inclAll = all;
#pragma unroll
      for (uint32_t s = 1; s < W; s <<= 1) {
        uint32_t o = __shfl_up(inclAll, s, W);
        if (lane >= s) inclAll += o;
      }
      exclAll = inclAll - all;
    }
    for (uint32_t r0 = 0; r0 < total; r0 += W) {  // the same turns for every lane of the group
      const uint32_t t = r0 + lane;
      uint32_t owner = 0;
#pragma unroll
      for (uint32_t step = W / 2; step > 0; step >>= 1) {
        const uint32_t ex = __shfl(excl, owner + step, W);
        if (ex <= t) owner += step;
      }
      const uint4 h = make_uint4(__shfl(head.x, owner, W), __shfl(head.y, owner, W), __shfl(head.z, owner, W),
                                 __shfl(head.w, owner, W));
      const f3 normal = mk3(__uint_as_float(__shfl(tail.x, owner, W)), __uint_as_float(__shfl(tail.y, owner, W)),
                            __uint_as_float(__shfl(tail.z, owner, W)));
      const float tscale = __uint_as_float(__shfl(tail.w, owner, W));
      const uint32_t pt = t - __shfl(excl, owner, W);
      const uint32_t tp = stashed ? __shfl(exclAll, owner, W) + pt : W;  // the item's slot in the stash if < W
      if (J && t < total && (h.z & ENTRY_JOINT)) {
        if constexpr (J) {
          const uint32_t j = h.y;
          const int4 f = w.info[j];
          Joint k = loadJoint<M>(w, j);
          if (dualJointRecord(w, f, k, jointPosesFor<M>(w, f, i, pos, rot))) storeJoint<M>(w, j, k);
        }
      } else if (t < total) {
        // The partner is static or of an earlier colour: its pose of this iteration was waited for by the primal
        const bool isA = (h.z & ENTRY_IS_A) != 0;
        const uint32_t c = h.y + pt;
        Contact kc;
        PartnerTerms o;
        if (DF && tp < W) {
          unstashItem(df.stash, tp, kc, o);
        } else {
          kc = w.contacts[c];
          if constexpr (DF) {
            kc.pen = ldDev4(contactRsrc(w), contactOff(c, OFF_PEN));
            kc.lam = ldDev4(contactRsrc(w), contactOff(c, OFF_LAM));
          }
          o = partnerTerms(w, kc, pairBody<M>(w, h.x), isA, stepStartArms(kc, p.flags));
        }
        const Eval ev = evalContactTerms(w, kc, orthonormalScaled(normal, tscale), __uint_as_float(h.w), self, o, isA,
                                         stepStartArms(kc, p.flags), p.alpha);
        dualUpdate<M>(w, c, kc, ev, p);
      }
    }
  }
  if constexpr (DF) {
    storesDone();
    if (lane == 0) stAgent(&w.dualDone[i], df.base + df.it + 1u);
  }
}

constexpr uint32_t PRIMAL_BLOCK = 256;
constexpr uint32_t WAVE = DF_WAVE;

/** Bucket position k (of a colour half ending at `end`) solved by a group of W lanes; `lane` is the lane in the group. */
template <uint32_t W, int M = SOLVE_COLORS, bool J = true>
AVBD_D void primalGroup(const DeviceWorld &w, uint32_t k, uint32_t end, uint32_t lane, const DfCtx &df = DfCtx{}) {
  const bool live = k < end;
  const uint4 slot = live ? w.colorSlot[k] : make_uint4(0u, 0u, 0u, 0u);
  Acc acc;
  float4 bpos = make_float4(0.f, 0.f, 0.f, 0.f), brot = bpos;
  if (live) {
    const WorldParams &p = w.worlds[slot.w];
    accumulate<W, M, J>(w, slot.x, slot.y, slot.z, lane, p.alpha, p.dt, p.flags, acc, df, bpos, brot);
  } else {
    float *a = reinterpret_cast<float *>(&acc);
    for (int q = 0; q < int(sizeof(Acc) / sizeof(float)); q++) a[q] = 0.f;
  }
  mergeTree<W>(acc);
  if constexpr (M != SOLVE_COLORS) {
    float4 pos = make_float4(0.f, 0.f, 0.f, 0.f), rot = pos;
    if (live && lane == 0) finishBody<M>(w, slot.x, acc, df, bpos, brot, &pos, &rot);
    // The new pose from the group's lane 0 to all of its lanes
    pos = make_float4(__shfl(pos.x, 0, W), __shfl(pos.y, 0, W), __shfl(pos.z, 0, W), __shfl(pos.w, 0, W));
    rot = make_float4(__shfl(rot.x, 0, W), __shfl(rot.y, 0, W), __shfl(rot.z, 0, W), __shfl(rot.w, 0, W));
    if (live) fusedDual<W, M, J>(w, slot, lane, pos, rot, df);
  } else {
    if (live && lane == 0) finishBody<M>(w, slot.x, acc, df, bpos, brot);
  }
}

/**
 * One colour's primal (primal, wgsl-solve.ts:314-345). Work goes out by whole waves: first the colour's heavy bodies,
 * two to a wave (HEAVY_LANES lanes each), then its light bodies, eight to a wave (LANES lanes each), so a body with many
 * contacts does not hold back the seven bodies sharing its wave.
 */
template <bool J> __global__ void __launch_bounds__(PRIMAL_BLOCK, 4) k_primal(DeviceWorld w, uint32_t color) {
  // The hold flags and the colour's ranges are read together (neither waits for the other)
  const uint32_t overflow = w.counters[C_OVERFLOW], clashes = w.counters[C_CLASHES];
  const uint32_t heavyStart = w.colorStart[2 * color];
  const uint32_t lightStart = w.colorStart[2 * color + 1];
  const uint32_t lightEnd = w.colorStart[2 * color + 2];
  if ((overflow | clashes) != 0u) return;
  constexpr uint32_t heavyPerWave = WAVE / HEAVY_LANES, lightPerWave = WAVE / LANES;
  const uint32_t heavyWaves = (lightStart - heavyStart + heavyPerWave - 1) / heavyPerWave;
  const uint32_t waves = heavyWaves + (lightEnd - lightStart + lightPerWave - 1) / lightPerWave;
  const uint32_t lane = threadIdx.x % WAVE;
  constexpr uint32_t wavesPerBlock = PRIMAL_BLOCK / WAVE;
  // u is uniform across a wave, so every lane of a group takes the same turns (the shuffles need whole groups)
  for (uint32_t u = blockIdx.x * wavesPerBlock + threadIdx.x / WAVE; u < waves; u += gridDim.x * wavesPerBlock) {
    if (u < heavyWaves)
      primalGroup<HEAVY_LANES, SOLVE_COLORS, J>(w, heavyStart + u * heavyPerWave + lane / HEAVY_LANES, lightStart,
                                                lane % HEAVY_LANES);
    else
      primalGroup<LANES, SOLVE_COLORS, J>(w, lightStart + (u - heavyWaves) * lightPerWave + lane / LANES, lightEnd,
                                          lane % LANES);
  }
}

// --- dual --------------------------------------------------------------------------------------------------------------

/** Joint j's dual in the dual pass (plain memory, the poses in the records). */
AVBD_D void dualJoint(const DeviceWorld &w, uint32_t j) {
  const int4 f = w.info[j];
  Joint k = w.joints[j];
  if (dualJointRecord(w, f, k, jointPoses(w, f.y, f.z))) w.joints[j] = k;
}

/** Point k of manifold m: every point is a thread of its own (the points are independent of each other). */
AVBD_D void dualPoint(const DeviceWorld &w, uint32_t m, uint32_t k) {
  const Manifold mf = w.manifolds[m];
  const uint32_t c = mf.ids.z + k;
  const WorldParams &p = w.worlds[w.bodyWorld[mf.ids.x]];
  PairBody PA = pairBody<SOLVE_COLORS>(w, mf.ids.x), PB = pairBody<SOLVE_COLORS>(w, mf.ids.y);
  Basis basis = orthonormal(xyz(mf.geo));
  const Contact kc = w.contacts[c];
  dualUpdate<SOLVE_COLORS>(w, c, kc, evalContact(w, kc, basis, mf.geo.w, PA, PB, true, p.alpha, p.flags), p);
}

/** Joints first (one thread each), then one thread per contact point of the dense point list. */
__global__ void k_dual(DeviceWorld w) {
  if (held(w)) return;
  const uint32_t J = w.nJoints;
  const uint64_t total = uint64_t(J) + w.counters[C_CONTACTS];
  for (uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; id < total; id += uint64_t(gridDim.x) * blockDim.x) {
    if (id < J) {
      dualJoint(w, uint32_t(id));
    } else {
      const uint32_t mk = w.pointMap[id - J];
      dualPoint(w, mk >> 3, mk & 7u);
    }
  }
}

/** BDF1 velocities of body i from its step-start pose (updateVelocities, wgsl-solve.ts:537-549). */
AVBD_D void bodyVelocity(const DeviceWorld &w, uint32_t i) {
  const WorldParams &p = w.worlds[w.bodyWorld[i]];
  Body &B = w.bodies[i];
  f3 up = mk3(p.up[0], p.up[1], p.up[2]);
  float prevUp = dot(xyz(B.vel), up);
  if (B.size.w > 0.f) {
    B.vel = mk4(divs(sub(xyz(B.pos), xyz(B.ipos)), p.dt), prevUp);
    B.angVel = mk4(divs(qsub(B.rot, B.irot), p.dt), B.angVel.w);
  } else {
    B.vel.w = prevUp;
  }
}

__global__ void k_velocities(DeviceWorld w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nBodies || held(w)) return;
  bodyVelocity(w, i);
}

// --- dataflow solve: every iteration's colours and duals in one launch -------------------------------------------------


/**
 * Every iteration's primal (colour by colour) and dual in one launch. Unit u (see DataflowPlan) goes to wave
 * u mod (waves of the grid), and each wave takes its units in increasing order. A unit is one colour wave of one
 * iteration; each of its bodies runs the dual of the pairs it finishes right after its primal (fusedDual). A unit
 * waits only for units before it (a body for its neighbours of earlier colours in this iteration, for the primal and
 * dual of its neighbours of later colours in the last one, and for its own last iteration), so the solve computes
 * exactly what the colour-by-colour launches and the dual pass do. Every wave of the grid is resident at once (the
 * grid is sized by the occupancy), so the unit everything else waits for always runs; and every wait is bounded.
 */
template <bool J> __global__ void __launch_bounds__(PRIMAL_BLOCK, 4) k_solve_dataflow(DeviceWorld w) {
  __shared__ float stash[PRIMAL_BLOCK / WAVE][STASH_WORDS * WAVE];  // each wave's contact items, primal to dual
  if (held(w)) return;
  const DataflowPlan &plan = *w.dfPlan;
  const uint32_t lane = threadIdx.x % WAVE;
  const uint32_t waves = gridDim.x * (PRIMAL_BLOCK / WAVE);
  const uint32_t total = plan.total, perIter = plan.unitsPerIter;
  float *const waveStash = stash[threadIdx.x / WAVE];
  DfCtx df;
  df.base = plan.base;
  for (uint32_t u = blockIdx.x * (PRIMAL_BLOCK / WAVE) + threadIdx.x / WAVE; u < total; u += waves) {
    df.it = u / perIter;
    const uint32_t r = u % perIter;
    uint32_t c = 0;
    while (plan.waveStart[c + 1] <= r) c++;
    df.color = c;
    const uint32_t v = r - plan.waveStart[c];
    const uint32_t heavyStart = w.colorStart[2 * c], lightStart = w.colorStart[2 * c + 1],
                   lightEnd = w.colorStart[2 * c + 2];
    const uint32_t hw = plan.heavyWaves[c];
    if (v < hw) {
      df.stash = waveStash + (lane - lane % HEAVY_LANES);
      primalGroup<HEAVY_LANES, SOLVE_DATAFLOW, J>(w, heavyStart + v * (WAVE / HEAVY_LANES) + lane / HEAVY_LANES,
                                                  lightStart, lane % HEAVY_LANES, df);
    } else {
      df.stash = waveStash + (lane - lane % LANES);
      primalGroup<LANES, SOLVE_DATAFLOW, J>(w, lightStart + (v - hw) * (WAVE / LANES) + lane / LANES, lightEnd,
                                            lane % LANES, df);
    }
  }
}

// --- one workgroup per world: a batch's every iteration, colour and dual in one launch ------------------------------

/**
 * Where each world's bodies sit in each colour bucket: keyStart[world × COLOR_KEYS + key] is the first position of
 * bucket `key` (colour key / 2, heavy or light half) holding a body of `world` or of a later world. A bucket lists its
 * bodies in index order (the scatter ranks them by index within a chunk, chunks in order) and a world owns a contiguous
 * index range, so a world's bodies of a bucket form one run, found by a binary search on the slots' world words.
 */
__global__ void k_world_ranges(DeviceWorld w, uint32_t *keyStart) {
  const uint32_t keys = 2u * w.colorCap;
  const uint64_t id = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= uint64_t(w.nWorlds + 1u) * keys) return;
  const uint32_t world = uint32_t(id / keys), key = uint32_t(id % keys);
  uint32_t lo = w.colorStart[key], hi = w.colorStart[key + 1];
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (w.colorSlot[mid].w < world) lo = mid + 1;
    else hi = mid;
  }
  keyStart[size_t(world) * COLOR_KEYS + key] = lo;
}

/**
 * The solve of a batch of small worlds, one workgroup per world (order[blockIdx] names it; the host lists the largest
 * worlds first so they start first). The workgroup runs the world's iterations; in each, its colours in order, a
 * colour's heavy bodies two to a wave and light ones eight to a wave as k_primal does, each body running the dual of
 * the pairs it finishes right after its primal (as in the dataflow solve), then a workgroup barrier, which also makes
 * the colour's poses and multipliers visible to the next; colours without a body of the world are skipped. The joints'
 * duals follow the iteration's last colour. At the end, the world's velocities. The same arithmetic in the same order
 * as the colour launches and the dual pass, so the same bits; nothing crosses a workgroup, so there are no waits.
 */
template <uint32_t BS, bool J>
__global__ void __launch_bounds__(BS, 4)
    k_solve_worlds(DeviceWorld w, const uint32_t *order, const uint32_t *keyStart, const uint32_t *jointStart,
                   uint32_t iterations) {
  if (held(w)) return;
  const uint32_t world = order[blockIdx.x];
  const uint32_t *lo = keyStart + size_t(world) * COLOR_KEYS, *hi = lo + COLOR_KEYS;
  const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
  constexpr uint32_t waves = BS / WAVE, heavyPerWave = WAVE / HEAVY_LANES, lightPerWave = WAVE / LANES;
  const uint32_t firstJoint = jointStart ? jointStart[world] : 0u, endJoint = jointStart ? jointStart[world + 1] : 0u;
  DfCtx ctx{0u, 0u, 0u};
  for (uint32_t it = 0; it < iterations; it++) {
    for (uint32_t c = 0; c < w.colorCap; c++) {
      const uint32_t hs = lo[2 * c], he = hi[2 * c], ls = lo[2 * c + 1], le = hi[2 * c + 1];
      if (hs == he && ls == le) continue;  // the same for the whole workgroup
      ctx.color = c;
      const uint32_t hw = (he - hs + heavyPerWave - 1) / heavyPerWave;
      const uint32_t units = hw + (le - ls + lightPerWave - 1) / lightPerWave;
      for (uint32_t u = wave; u < units; u += waves) {  // uniform across a wave
        if (u < hw)
          primalGroup<HEAVY_LANES, SOLVE_WORLDS, J>(w, hs + u * heavyPerWave + lane / HEAVY_LANES, he,
                                                    lane % HEAVY_LANES, ctx);
        else
          primalGroup<LANES, SOLVE_WORLDS, J>(w, ls + (u - hw) * lightPerWave + lane / LANES, le, lane % LANES, ctx);
      }
      __syncthreads();
    }
    if (endJoint > firstJoint) {
      for (uint32_t j = firstJoint + threadIdx.x; j < endJoint; j += BS) dualJoint(w, j);
      __syncthreads();
    }
  }
  const WorldParams &p = w.worlds[world];
  for (uint32_t i = p.firstBody + threadIdx.x; i < p.firstBody + p.bodyCount; i += BS) bodyVelocity(w, i);
}

/**
 * For the dataflow solve: the iterations' duals of the joints no dynamic body lists (both ends static or asleep, or the
 * world and a static body). Their bodies do not move during the solve and nothing else reads them, so their
 * `iterations` duals run back to back, as the dual pass would run them once per iteration.
 */
__global__ void k_dual_static_joints(DeviceWorld w, uint32_t iterations) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w.nJoints || held(w)) return;
  const int4 f = w.info[j];
  const bool dynA = f.y >= 0 && w.bodies[f.y].size.w > 0.f, dynB = w.bodies[f.z].size.w > 0.f;
  if (dynA || dynB) return;
  for (uint32_t it = 0; it < iterations; it++) dualJoint(w, j);
}

// --- launch wrappers ---------------------------------------------------------------------------------------------------

static inline uint32_t blocks(uint64_t n, uint32_t bs, uint32_t cap = 1u << 20) {
  uint64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return uint32_t(g);
}

void launch_warm_start(const DeviceWorld &w, hipStream_t s) {
  if (!w.nJoints) return;  // the colouring's scatter pass warmed the bodies
  k_warm_joints<<<blocks(w.nJoints, 256), 256, 0, s>>>(w);
  k_warm_bodies<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
}
void launch_primal(const DeviceWorld &w, uint32_t color, uint32_t maxColorBodies, hipStream_t s) {
  const uint32_t grid = blocks(uint64_t(maxColorBodies) * LANES, PRIMAL_BLOCK, 4096);
  if (w.nJoints) k_primal<true><<<grid, PRIMAL_BLOCK, 0, s>>>(w, color);
  else k_primal<false><<<grid, PRIMAL_BLOCK, 0, s>>>(w, color);
}
void launch_dual(const DeviceWorld &w, uint32_t manifoldCapacity, hipStream_t s) {
  k_dual<<<blocks(uint64_t(w.nJoints) + uint64_t(manifoldCapacity) * 2, 256, 16384), 256, 0, s>>>(w);
}
uint32_t dataflow_grid() {
  int device = 0, perCu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void *>(k_solve_dataflow<true>),
                                                   PRIMAL_BLOCK, 0) != hipSuccess)
    return 0;
  return uint32_t(perCu) * uint32_t(prop.multiProcessorCount);
}
void launch_dataflow(const DeviceWorld &w, uint32_t grid, uint32_t iterations, hipStream_t s) {
  if (w.nJoints) k_solve_dataflow<true><<<grid, PRIMAL_BLOCK, 0, s>>>(w);
  else k_solve_dataflow<false><<<grid, PRIMAL_BLOCK, 0, s>>>(w);
  if (w.nJoints) k_dual_static_joints<<<blocks(w.nJoints, 256), 256, 0, s>>>(w, iterations);
}
template <bool J>
static void launch_worlds(const DeviceWorld &w, const uint32_t *order, const uint32_t *keyStart,
                          const uint32_t *jointStart, uint32_t iterations, uint32_t block, hipStream_t s) {
  switch (block) {
    case 128: k_solve_worlds<128, J><<<w.nWorlds, 128, 0, s>>>(w, order, keyStart, jointStart, iterations); break;
    case 512: k_solve_worlds<512, J><<<w.nWorlds, 512, 0, s>>>(w, order, keyStart, jointStart, iterations); break;
    case 1024: k_solve_worlds<1024, J><<<w.nWorlds, 1024, 0, s>>>(w, order, keyStart, jointStart, iterations); break;
    default: k_solve_worlds<256, J><<<w.nWorlds, 256, 0, s>>>(w, order, keyStart, jointStart, iterations); break;
  }
}
void launch_solve_worlds(const DeviceWorld &w, const uint32_t *order, uint32_t *keyStart, const uint32_t *jointStart,
                         uint32_t iterations, uint32_t block, hipStream_t s) {
  k_world_ranges<<<blocks(uint64_t(w.nWorlds + 1u) * 2u * w.colorCap, 256), 256, 0, s>>>(w, keyStart);
  if (jointStart) launch_worlds<true>(w, order, keyStart, jointStart, iterations, block, s);
  else launch_worlds<false>(w, order, keyStart, jointStart, iterations, block, s);
}
void launch_velocities(const DeviceWorld &w, hipStream_t s) {
  k_velocities<<<blocks(w.nBodies, 256), 256, 0, s>>>(w);
}

}  // namespace avbd
