'use strict';
// The reference's Babylon-free world-API contracts (tests/backend-contract.js checkBackend and checkFloor,
// tests/packed-poses-contract.js packedPoseContract), restated for the "hip" backend of babylon-avbd_amd/js/index.js:
// the same calls and the same expectations, in 3D, the restitution part of checkContacts, and checkSleeping. What this
// backend does not build (trigger bodies) is checked to be refused with an error, or, for switching it off, accepted;
// a ray query finds the teleported box.
// Prints one JSON object with a record per contract; exits non-zero on the first failure.
const path = require('path');
const avbd = require(path.join(__dirname, '..', '..', 'babylon-avbd_amd', 'js', 'index.js'));

const check = (condition, message) => {
  if (!condition) throw Error(message);
};
const close = (a, b, tolerance = 1e-5) => check(Math.abs(a - b) <= tolerance, `${a} differs from ${b}`);
const throws = (f) => {
  try {
    f();
  } catch (e) {
    return true;
  }
  return false;
};

async function checkBackend() {
  const pipeline = await avbd.createSolverPipeline({ backend: 'hip' });
  const scene = new avbd.AvbdScene3D({ gravity: 0 });
  const zero = [0, 0, 0];
  const size = [1, 1, 1];
  const ball = scene.addBox(size, { position: zero });
  const world = scene.createSolver(pipeline, { spatialSort: false, bodyCapacity: 64 });
  try {
    const handle = world.body(ball);
    const id = handle.gpuIndex;
    check(world.backend === 'hip', 'Backend selection was lost');
    handle.teleport([2, 3, 4]).setLinearVelocity([1, 2, 3]).applyImpulse([1, 0, 0]);
    let state = await handle.readState();
    close(state.position[0], 2);
    close(state.linearVelocity[0], 2);
    // Whole-batch validation must reject without partially applying a valid edit
    check(throws(() => world.editBodies([
      { body: ball, linearVelocity: [99, 0, 0] },
      { body: ball, impulse: [NaN, 0, 0] },
    ])), 'Nonfinite batch accepted');
    state = await handle.readState();
    close(state.linearVelocity[0], 2);
    world.setVelocities([id], { linear: [0, 0, 0], angular: [0, 0, 0] });
    handle.applyImpulse([1, 0, 0], [2, 4, 4]);
    state = await handle.readState();
    close(state.linearVelocity[0], 1);
    check(state.angularVelocity[2] < 0, 'Off-centre impulse did not rotate the body');
    handle.setCollisionGroups(4, 4).setRestitution(0.6);
    check(world.getRestitution()[id] === Math.fround(0.6), 'Restitution was not set');
    // Not built here: a trigger body is refused, not ignored
    check(throws(() => handle.setTrigger(true)), 'A trigger was accepted');
    world.watchContacts({ maxPairs: 64, maxEvents: 64 });
    const noEvents = await world.readContactEvents();
    check(noEvents.events.length === 0 && noEvents.dropped === 0 && !noEvents.full, 'Events before any step');
    const ray = await world.raycast([2, 10, 4], [0, -1, 0]);
    check(ray && ray.index === id && Math.abs(ray.distance - 6.5) < 1e-5, 'Ray query missed the body');
    const selected = await world.readSelectedBodies([id], { posesOnly: true });
    check(selected.length === 8, 'Pose stride changed across backends');
    handle.setTrigger(false).setSleepEnabled(true).setSleepEnabled(false);
    world.disableSleeping();
    handle.teleport(zero).setLinearVelocity([1, 0, 0]).setAngularVelocity([0, 0, 0]);
    check(world.advance(world.params.dt, { maxFrameTime: 1 }) === 1, 'Fixed-step API differs');
    state = await handle.readState();
    close(state.position[0], world.params.dt, 0.001);
    let frame;
    let disposed;
    let after = 0;
    const binding = {
      getEngine: () => ({ getDeltaTime: () => world.params.dt * 1000 }),
      onBeforeRenderObservable: {
        add: (callback) => (frame = callback),
        remove: (observer) => {
          check(observer === frame, 'Wrong frame observer removed');
          frame = null;
        },
      },
      onDisposeObservable: {
        addOnce: (callback) => (disposed = callback),
        remove: (observer) => {
          check(observer === disposed, 'Wrong disposal observer removed');
          disposed = null;
        },
      },
    };
    world.attachToScene(binding, {
      afterStep: (solver) => {
        check(solver === world, 'afterStep received a different world');
        after++;
      },
    });
    frame();
    check(after === 1, 'Automatic scene stepping differs');
    world.detachFromScene();
    check(frame === null && disposed === null, 'Scene observers were not removed');
    const otherScene = new avbd.AvbdScene3D();
    otherScene.addBox(size);
    const other = otherScene.createSolver(pipeline, { spatialSort: false });
    check(throws(() => other.editBodies([{ body: handle, position: zero }])), 'Accepted a handle from another world');
    other.destroy();
    return { backend: 'hip', dimension: 3, passed: true };
  } finally {
    pipeline.destroy();
    check(throws(() => world.step()), 'Destroyed world can still step');
  }
}

async function checkFloor() {
  const pipeline = await avbd.createSolverPipeline({ backend: 'hip' });
  const scene = new avbd.AvbdScene3D({ iterations: 10 });
  scene.addBox([10, 1, 10], { density: 0, position: [0, -0.5, 0] });
  const ball = scene.addSphere(0.5, { position: [0, 2, 0] });
  const world = scene.createSolver(pipeline, { spatialSort: false });
  try {
    for (let i = 0; i < 180; i++) world.step();
    const state = await world.body(ball).readState();
    check(state.position.every(Number.isFinite), 'Nonfinite floor state');
    check(state.position[1] > 0.35 && state.position[1] < 0.7, `Floor failed: ${state.position[1]}`);
    return { backend: 'hip', dimension: 3, floorHeight: state.position[1], passed: true };
  } finally {
    pipeline.destroy();
  }
}

/** checkContacts (backend-contract.js), its restitution part: a sphere with restitution 1 falling at 3 m/s onto the
 * floor with gravity 0 moves up faster than 2 m/s after 12 steps. */
async function checkRestitution() {
  const pipeline = await avbd.createSolverPipeline({ backend: 'hip' });
  const scene = new avbd.AvbdScene3D({ gravity: 0 });
  scene.addBox([10, 1, 4], { density: 0, position: [0, -0.5, 0] });
  const ball = scene.addSphere(0.5, { position: [0, 0.51, 0], velocity: [0, -3, 0], restitution: 1 });
  const world = scene.createSolver(pipeline, { spatialSort: false });
  try {
    for (let i = 0; i < 12; i++) world.step();
    const state = await world.body(ball).readState();
    check(state.linearVelocity[1] > 2, `Restitution failed: ${state.linearVelocity[1]}`);
    return { backend: 'hip', dimension: 3, reboundSpeed: state.linearVelocity[1], passed: true };
  } finally {
    pipeline.destroy();
  }
}

/** checkSleeping (backend-contract.js:441-467): a box with allowSleep true resting on a floor sleeps within 180 steps;
 * an impulse (1, 2, 0) wakes it and it moves up faster than 1 m/s after one step. */
async function checkSleeping() {
  const pipeline = await avbd.createSolverPipeline({ backend: 'hip' });
  const scene = new avbd.AvbdScene3D();
  scene.addBox([10, 1, 10], { density: 0, position: [0, -0.5, 0] });
  const box = scene.addBox([1, 1, 1], { position: [0, 0.5, 0], allowSleep: true });
  const world = scene.createSolver(pipeline, { spatialSort: false });
  try {
    for (let i = 0; i < 180; i++) world.step();
    const body = world.body(box);
    const state = await body.readState();
    check(state.sleeping, 'Supported body did not sleep');
    body.applyImpulse([1, 2, 0]);
    world.step();
    const awake = await body.readState();
    check(!awake.sleeping && awake.linearVelocity[1] > 1, 'Impulse did not wake and move the body');
    return { backend: 'hip', dimension: 3, sleeping: true, waking: true, passed: true };
  } finally {
    pipeline.destroy();
  }
}

async function packedPoseContract(g, ids) {
  const expect = check;
  const near = (actual, expected, message) => check(Math.abs(actual - expected) < 1e-5, message);
  expect(typeof g.setPoses === 'function', 'Packed pose API is missing');
  const p = new Float32Array(ids.length * 3);
  const q = new Float32Array(ids.length * 4);
  for (let i = 0; i < ids.length; i++) {
    p.set([i * 2, -i, 4 + i], i * 3);
    q.set([0, 0, Math.sin(0.4) * 2, Math.cos(0.4) * 2], i * 4);
  }
  g.body(ids[0]).setLinearVelocity([1, 2, 3]).setAngularVelocity([0.2, 0.4, 0.6]);
  const before = await g.readSelectedBodies(ids);
  expect(g.setPoses(ids, p, q) === g, 'Packed poses must chain');
  const after = await g.readSelectedBodies(ids);
  for (let i = 0; i < ids.length; i++) {
    for (let k = 0; k < 3; k++) near(after[i * 40 + k], p[i * 3 + k], 'Packed position differs');
    near(after[i * 40 + 6], Math.sin(0.4), 'Quaternion was not normalized');
    near(after[i * 40 + 7], Math.cos(0.4), 'Quaternion scalar was not normalized');
    for (const k of [3, 16, 17, 18, 19, 20, 21, 22, 23, 32, 33, 34, 35, 36, 37, 38, 39])
      near(after[i * 40 + k], before[i * 40 + k], `Teleport changed non-pose field ${k}`);
  }
  g.setPoses(ids, p);
  const unchangedRotation = await g.readSelectedBodies(ids);
  for (let i = 0; i < ids.length; i++)
    for (let k = 4; k < 8; k++)
      near(unchangedRotation[i * 40 + k], after[i * 40 + k], 'Position-only batch changed rotation');
  const invalidPosition = p.slice();
  invalidPosition[invalidPosition.length - 1] = NaN;
  const zeroQ = q.slice();
  zeroQ.fill(0, zeroQ.length - 4);
  const duplicate = ids.slice();
  duplicate[duplicate.length - 1] = duplicate[0];
  for (const args of [[ids, invalidPosition], [ids, p, zeroQ], [duplicate, p], [ids, p.subarray(1)],
    [ids, Array.from(p)]]) {
    expect(throws(() => g.setPoses(...args)), 'Invalid packed poses were accepted');
    const actual = await g.readSelectedBodies(ids);
    expect(actual.every((v, i) => v === unchangedRotation[i]), 'Rejected pose batch partially changed live bodies');
  }
  g.body(ids[0]).setLinearVelocity([4, 5, 6]);
  g.setPoses(ids, p);
  g.body(ids[0]).setLinearVelocity([7, 8, 9]);
  const ordered = await g.body(ids[0]).readState();
  ordered.linearVelocity.forEach((v, i) => near(v, i + 7, 'Packed pose broke command order'));
  expect(g.setPoses(new Uint32Array(), new Float32Array()) === g, 'Empty batch failed');
  return { normalizedRotation: true, positionOnlyPreservesRotation: true, nonPoseFieldsPreserved: true,
           invalidBatchAtomic: true, commandOrder: true };
}

async function checkPackedPoses() {
  const pipeline = await avbd.createSolverPipeline({ backend: 'hip' });
  const scene = new avbd.AvbdScene3D({ gravity: 0 });
  for (let i = 0; i < 6; i++) scene.addBox([1, 1, 1], { position: [i * 3, 0, 0] });
  const world = scene.createSolver(pipeline);
  try {
    return await packedPoseContract(world, Uint32Array.from([1, 2, 4, 5]));
  } finally {
    pipeline.destroy();
  }
}

(async () => {
  const out = {};
  out.backend = await checkBackend();
  out.floor = await checkFloor();
  out.restitution = await checkRestitution();
  out.sleeping = await checkSleeping();
  out.packedPoses = await checkPackedPoses();
  process.stdout.write(JSON.stringify(out) + '\n');
})().catch((e) => {
  process.stderr.write(`${e.stack || e}\n`);
  process.exit(1);
});
