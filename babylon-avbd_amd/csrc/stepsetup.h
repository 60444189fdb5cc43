// Per-step set-up that the colouring's last passes also run: the hold flag, the body warm start (when no joint has to
// be warm-started before it) and the dataflow solve's work list.
#pragma once
#include "engine_types.h"

namespace avbd {

/**
 * A step is held (the solve leaves bodies and joints untouched) when its contact storage overflowed or its colouring
 * clashed (a body left pending after the rounds may share a colour with a neighbour), as protectSolverStep's
 * solverStepValid() (src/gpu/gpuSolverKernels.js:72-103). So a solved step always has a proper colouring.
 */
AVBD_D bool held(const DeviceWorld &w) { return (w.counters[C_OVERFLOW] | w.counters[C_CLASHES]) != 0u; }

/**
 * warmStartBodies (wgsl-solve.ts:259-300) for body i: the step-start pose, the inertial pose, and the initial guess
 * (adaptive gravity weight; a dynamic body at rest keeps its pose when the world starts bodies at rest; a static body
 * with a velocity moves along it).
 */
AVBD_D void warmBody(const DeviceWorld &w, uint32_t i) {
  const WorldParams &p = w.worlds[w.bodyWorld[i]];
  Body &B = w.bodies[i];
  float4 pos = B.pos, rot = B.rot, vel = B.vel;
  f3 angVel = xyz(B.angVel);
  float dt = p.dt, g = p.gravity;
  f3 up = mk3(p.up[0], p.up[1], p.up[2]);
  bool dyn = B.size.w > 0.f;
  f3 inertialPos = add(xyz(pos), scale(xyz(vel), dt));
  if (dyn) inertialPos = add(inertialPos, scale(up, g * (dt * dt)));
  B.inertialPos = mk4(inertialPos, B.inertialPos.w);
  B.inertialRot = qadd(rot, scale(angVel, dt));
  float wgt = 0.f;
  if (fabsf(g) > 0.f) {
    float x = (dot(xyz(vel), up) - vel.w) / dt * rsign(g) / fabsf(g);
    wgt = rmin(rmax(x, 0.f), 1.f);
  }
  bool slow = length(xyz(vel)) + length(angVel) * B.moment.w < REST_SPEED;
  bool atRest = (p.flags & FLAG_START_AT_REST) && slow;
  B.ipos = pos;
  B.irot = rot;
  w.irotInv[i] = qinv(rot);  // every contact row of the step turns displacements by it
  if (dyn && !atRest) {
    B.pos = mk4(add(add(xyz(pos), scale(xyz(vel), dt)), scale(up, g * (wgt * dt * dt))), pos.w);
    B.rot = qadd(rot, scale(angVel, dt));
  }
  bool moving = vel.x != 0.f || vel.y != 0.f || vel.z != 0.f || angVel.x != 0.f || angVel.y != 0.f || angVel.z != 0.f;
  if (!dyn && moving) {
    B.pos = mk4(add(xyz(pos), scale(xyz(vel), dt)), pos.w);
    B.rot = qadd(rot, scale(angVel, dt));
  }
}

constexpr uint32_t DF_WAVE = 64;

/**
 * The step's work list for the dataflow solve from the colour buckets' starts (COLOR_KEYS + 1 of them): wave counts
 * per colour (heavy bodies HEAVY_LANES lanes each, light ones LANES), and the stamp base past the last step's stamps.
 */
AVBD_D void writeDataflowPlan(const DeviceWorld &w, const uint32_t *starts, uint32_t iterations) {
  DataflowPlan &p = *w.dfPlan;
  p.base = p.base + p.iterations + 1u;
  p.iterations = iterations;
  uint32_t waves = 0;
  for (uint32_t c = 0; c < w.colorCap; c++) {
    const uint32_t heavy = starts[2 * c], light = starts[2 * c + 1], end = starts[2 * c + 2];
    const uint32_t hw = (light - heavy + (DF_WAVE / HEAVY_LANES) - 1) / (DF_WAVE / HEAVY_LANES);
    p.waveStart[c] = waves;
    p.heavyWaves[c] = hw;
    waves += hw + (end - light + (DF_WAVE / LANES) - 1) / (DF_WAVE / LANES);
  }
  p.waveStart[w.colorCap] = waves;
  p.unitsPerIter = waves;
  p.total = p.unitsPerIter * iterations;
}

}  // namespace avbd
