// Edits between steps: ordered body commands (velocity replacements, impulses, forces, torques, teleports) with the
// warm-start invalidation a teleport needs, collision-filter writes and per-world parameter writes. Each runs on the
// world's stream, so it lands after the steps enqueued before it and before the steps enqueued after it.
#include <algorithm>

#include "kernels.h"

namespace avbd {

namespace {

/** ω change of an impulse about the centre: to the body frame, ÷ max(moment, 1e-10), back
 * (src/native/bodyCommands3D.wgsl:9-12). */
AVBD_D f3 angularImpulse(const Body &b, f3 impulse) {
  f3 local = qrotate(qconj(b.rot), impulse);
  const float floor = 1e-10f;
  f3 scaled = mk3(local.x / rmax(b.moment.x, floor), local.y / rmax(b.moment.y, floor), local.z / rmax(b.moment.z, floor));
  return qrotate(b.rot, scaled);
}

/**
 * One thread per touched body applies that body's commands in submission order (src/native/bodyCommands3D.wgsl:14-36,
 * `edit`). cmd holds 3 uint4 per command: (body, kind, has point, 0), value xyzw (f32 bits), point xyzw (f32 bits);
 * ranges[r] = (first command, count) of the r-th touched body. Teleported bodies are marked for the invalidation pass.
 */
__global__ void __launch_bounds__(64) k_body_commands(Body *bodies, const uint4 *cmd, const uint2 *ranges, uint32_t n,
                                                      uint32_t *marks) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint2 range = ranges[r];
  const uint32_t index = cmd[3 * range.x].x;
  Body b = bodies[index];
  bool teleported = false;
  for (uint32_t k = range.x; k < range.x + range.y; k++) {
    const uint4 h = cmd[3 * k];
    const uint4 vb = cmd[3 * k + 1];
    const uint4 pb = cmd[3 * k + 2];
    const float4 v = make_float4(bitsf(vb.x), bitsf(vb.y), bitsf(vb.z), bitsf(vb.w));
    const float4 pt = make_float4(bitsf(pb.x), bitsf(pb.y), bitsf(pb.z), bitsf(pb.w));
    const uint32_t kind = h.y;
    const bool hasPoint = h.z != 0u;
    if (kind == 0u) {
      b.vel = make_float4(v.x, v.y, v.z, b.vel.w);
    } else if (kind == 1u) {
      b.angVel = make_float4(v.x, v.y, v.z, b.angVel.w);
    } else if (kind == 2u || kind == 3u) {
      if (b.size.w > 0.f) {
        f3 impulse = scale(xyz(v), kind == 3u ? v.w : 1.f);
        b.vel = mk4(add(xyz(b.vel), divs(impulse, b.size.w)), b.vel.w);
        if (hasPoint) b.angVel = mk4(add(xyz(b.angVel), angularImpulse(b, cross(sub(xyz(pt), xyz(b.pos)), impulse))), b.angVel.w);
      }
    } else if (kind == 5u || kind == 6u) {
      if (b.size.w > 0.f) b.angVel = mk4(add(xyz(b.angVel), angularImpulse(b, scale(xyz(v), kind == 6u ? v.w : 1.f))), b.angVel.w);
    } else if (kind == 4u) {
      b.pos = make_float4(v.x, v.y, v.z, b.pos.w);
      if (hasPoint) b.rot = pt;
      b.ipos = make_float4(v.x, v.y, v.z, b.ipos.w);
      b.irot = b.rot;
      b.inertialPos = make_float4(v.x, v.y, v.z, b.inertialPos.w);
      b.inertialRot = b.rot;
      teleported = true;
    }
  }
  bodies[index] = b;
  if (teleported) marks[index] = 1u;
}

/** Last step's manifolds touching a teleported body lose their warm start (bodyCommands3D.wgsl:38-44). */
__global__ void k_invalidate_manifolds(Manifold *manifolds, const uint32_t *counters, uint32_t capacity,
                                       const uint32_t *marks) {
  const uint32_t count = umin(counters[C_MANIFOLDS], capacity);
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < count; m += gridDim.x * blockDim.x) {
    const uint4 ids = manifolds[m].ids;
    if ((ids.w & 15u) != 0u && (marks[ids.x] | marks[ids.y]) != 0u) manifolds[m].ids.w = 0u;
  }
}

__global__ void k_clear_marks(const uint4 *cmd, const uint2 *ranges, uint32_t n, uint32_t *marks) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) marks[cmd[3 * ranges[r].x].x] = 0u;
}

/** words: (body, groups, collides-with) triples. */
__global__ void k_set_filters(uint2 *filters, const uint32_t *words, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) filters[words[3 * k]] = make_uint2(words[3 * k + 1], words[3 * k + 2]);
}

__global__ void k_set_restitution(float *restitution, const uint32_t *words, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) restitution[words[2 * k]] = __uint_as_float(words[2 * k + 1]);
}

__global__ void k_set_params(WorldParams *p, ParamEdit e) {
  if (threadIdx.x != 0) return;
  p->dt = e.dt;
  p->gravity = e.gravity;
  p->betaLin = e.betaLin;
  p->betaAng = e.betaAng;
  p->gamma = e.gamma;
  p->alpha = e.alpha;
  p->up[0] = e.up[0];
  p->up[1] = e.up[1];
  p->up[2] = e.up[2];
  p->flags = e.flags;
}

/** Constraint record edits: (slot, word, value bits) triples, word 0..31 a word of the record, 32..34 the info
 * word's type, body A and body B. The host leaves at most one triple per (slot, word) in a launch. */
__global__ void k_edit_joints(Joint *joints, int4 *info, const uint32_t *words, uint32_t n) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t slot = words[3 * e], word = words[3 * e + 1], value = words[3 * e + 2];
  if (word < 32) reinterpret_cast<uint32_t *>(&joints[slot])[word] = value;
  else reinterpret_cast<int32_t *>(&info[slot])[word - 32] = int32_t(value);
}

/** The quaternion a ⊗ b with the reference capture's grouping (a.w b.xyz + b.w a.xyz + a.xyz × b.xyz, ...). */
AVBD_D float4 capMul(float4 a, float4 b) {
  return make_float4((a.w * b.x + b.w * a.x) + (a.y * b.z - a.z * b.y), (a.w * b.y + b.w * a.y) + (a.z * b.x - a.x * b.z),
                     (a.w * b.z + b.w * a.z) + (a.x * b.y - a.y * b.x), a.w * b.w - ((a.x * b.x + a.y * b.y) + a.z * b.z));
}
AVBD_D f3 perpendicular(f3 v) { return fabsf(v.x) > fabsf(v.z) ? mk3(-v.y, v.x, 0.f) : mk3(0.f, -v.z, v.y); }

/**
 * Frames of new constraints from the live poses: an angle-locked joint's rest orientation of B relative to A
 * (constraintFrames.wgsl:11-19), a motor's or limit's reference tangents about its axes (angularFrames.wgsl).
 */
__global__ void k_capture_frames(const Body *bodies, Joint *joints, const int4 *info, const uint32_t *slots,
                                 uint32_t n) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t j = slots[e];
  const int4 f = info[j];
  Joint k = joints[j];
  const float4 qa = f.y >= 0 ? bodies[f.y].rot : make_float4(0.f, 0.f, 0.f, 1.f);
  const float4 qb = bodies[f.z].rot;
  if (f.x == T_JOINT) {
    const float4 r = capMul(qconj(qa), qb);
    const float l = sqrtf(qlengthSq(r));
    k.c0Lin.w = r.x / l;
    k.c0Ang.w = r.y / l;
    k.rA.w = r.z / l;
    k.rB.w = r.w / l;
  } else if (f.x == T_MOTOR || f.x == T_LIMIT) {
    const f3 tangent = normalize(perpendicular(xyz(k.rA)));
    const f3 world = qrotate(qa, tangent);
    const f3 axisB = normalize(xyz(k.rB));
    f3 tangentB = qrotate(qconj(qb), world);
    tangentB = sub(tangentB, scale(axisB, dot(tangentB, axisB)));
    if (length(tangentB) < 1.0e-6f) tangentB = perpendicular(axisB);
    k.c0Lin = mk4(tangent, 0.f);
    k.c0Ang = mk4(normalize(tangentB), 0.f);
  } else {
    return;
  }
  joints[j] = k;
}

}  // namespace

void launch_edit_joints(Joint *joints, int4 *info, const uint32_t *words, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  k_edit_joints<<<(n + 255) / 256, 256, 0, s>>>(joints, info, words, n);
}

void launch_capture_frames(const Body *bodies, Joint *joints, const int4 *info, const uint32_t *slots, uint32_t n,
                           hipStream_t s) {
  if (n == 0) return;
  k_capture_frames<<<(n + 63) / 64, 64, 0, s>>>(bodies, joints, info, slots, n);
}

void launch_body_commands(Body *bodies, const uint32_t *commands, const uint2 *ranges, uint32_t n, uint32_t *marks,
                          hipStream_t s) {
  if (n == 0) return;
  k_body_commands<<<(n + 63) / 64, 64, 0, s>>>(bodies, reinterpret_cast<const uint4 *>(commands), ranges, n, marks);
}

void launch_invalidate_manifolds(Manifold *manifolds, const uint32_t *counters, uint32_t capacity,
                                 const uint32_t *marks, hipStream_t s) {
  if (capacity == 0) return;
  const uint32_t blocks = std::min<uint32_t>((capacity + 255) / 256, 2048);
  k_invalidate_manifolds<<<blocks, 256, 0, s>>>(manifolds, counters, capacity, marks);
}

void launch_clear_marks(const uint32_t *commands, const uint2 *ranges, uint32_t n, uint32_t *marks, hipStream_t s) {
  if (n == 0) return;
  k_clear_marks<<<(n + 255) / 256, 256, 0, s>>>(reinterpret_cast<const uint4 *>(commands), ranges, n, marks);
}

void launch_set_filters(uint2 *filters, const uint32_t *words, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  k_set_filters<<<(n + 255) / 256, 256, 0, s>>>(filters, words, n);
}

void launch_set_params(WorldParams *p, const ParamEdit &e, hipStream_t s) { k_set_params<<<1, 64, 0, s>>>(p, e); }

void launch_set_restitution(float *restitution, const uint32_t *words, uint32_t n, hipStream_t s) {
  k_set_restitution<<<(n + 255) / 256, 256, 0, s>>>(restitution, words, n);
}

}  // namespace avbd
