"""The reference's Babylon-free world-API contracts on the Python facade of the hip backend: checkBackend and
checkFloor of reference tests/backend-contract.js and packedPoseContract of tests/packed-poses-contract.js, the same
calls and expectations in 3D. What this backend does not build (trigger bodies) must be refused, or accepted when it
is switched off; a ray query finds the teleported box."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def close(a, b, tol=1e-5):
    assert abs(a - b) <= tol, f"{a} differs from {b}"


def test_backend_contract(avbd):
    pipeline = avbd.createSolverPipeline("hip")
    scene = avbd.AvbdScene3D(gravity=0)
    ball = scene.addBox([1, 1, 1], position=[0, 0, 0])
    world = scene.createSolver(pipeline, {"spatialSort": False})
    try:
        handle = world.body(ball)
        i = handle.gpuIndex
        handle.teleport([2, 3, 4]).setLinearVelocity([1, 2, 3]).applyImpulse([1, 0, 0])
        state = handle.readState()
        close(state["position"][0], 2)
        close(state["linearVelocity"][0], 2)
        with pytest.raises(ValueError):  # the whole batch is rejected: the valid first record does not land
            world.editBodies([{"body": ball, "linearVelocity": [99, 0, 0]}, {"body": ball, "impulse": [math.nan, 0, 0]}])
        close(handle.readState()["linearVelocity"][0], 2)
        world.setVelocities([i], linear=[0, 0, 0], angular=[0, 0, 0])
        handle.applyImpulse([1, 0, 0], [2, 4, 4])
        state = handle.readState()
        close(state["linearVelocity"][0], 1)
        assert state["angularVelocity"][2] < 0, "an off-centre impulse did not rotate the body"
        handle.setCollisionGroups(4, 4).setRestitution(0.6)
        assert world.getRestitution()[i] == np.float32(0.6)
        with pytest.raises(ValueError):
            handle.setTrigger(True)
        ray = world.raycast([2, 10, 4], [0, -1, 0])  # the teleported unit box, top face at y = 3.5
        assert ray is not None and ray["index"] == i and abs(ray["distance"] - 6.5) < 1e-5
        assert world.readSelectedBodies([i], posesOnly=True).size == 8
        handle.setTrigger(False).setSleepEnabled(True).setSleepEnabled(False)
        world.disableSleeping()
        handle.teleport([0, 0, 0]).setLinearVelocity([1, 0, 0]).setAngularVelocity([0, 0, 0])
        assert world.advance(world.params["dt"], maxFrameTime=1) == 1
        close(handle.readState()["position"][0], world.params["dt"], 0.001)
        other_scene = avbd.AvbdScene3D()
        other_scene.addBox([1, 1, 1])
        other = other_scene.createSolver(pipeline)
        with pytest.raises(ValueError):
            other.editBodies([{"body": handle, "position": [0, 0, 0]}])
        other.destroy()
    finally:
        pipeline.destroy()
    with pytest.raises(RuntimeError):
        world.step()


def test_floor_contract(avbd):
    pipeline = avbd.createSolverPipeline("hip")
    scene = avbd.AvbdScene3D(iterations=10)
    scene.addBox([10, 1, 10], density=0, position=[0, -0.5, 0])
    ball = scene.addSphere(0.5, position=[0, 2, 0])
    world = scene.createSolver(pipeline)
    try:
        for _ in range(180):
            world.step()
        y = world.body(ball).readState()["position"][1]
        assert math.isfinite(y) and 0.35 < y < 0.7, f"floor failed: {y}"
    finally:
        pipeline.destroy()


def test_packed_pose_contract(avbd):
    pipeline = avbd.createSolverPipeline("hip")
    scene = avbd.AvbdScene3D(gravity=0)
    for k in range(6):
        scene.addBox([1, 1, 1], position=[3 * k, 0, 0])
    g = scene.createSolver(pipeline)
    try:
        ids = np.array([1, 2, 4, 5], np.uint32)
        p = np.zeros(len(ids) * 3, np.float32)
        q = np.zeros(len(ids) * 4, np.float32)
        for k in range(len(ids)):
            p[3 * k:3 * k + 3] = [k * 2, -k, 4 + k]
            q[4 * k:4 * k + 4] = [0, 0, math.sin(0.4) * 2, math.cos(0.4) * 2]
        g.body(int(ids[0])).setLinearVelocity([1, 2, 3]).setAngularVelocity([0.2, 0.4, 0.6])
        before = g.readSelectedBodies(ids)
        assert g.setPoses(ids, p, q) is g
        after = g.readSelectedBodies(ids)
        for k in range(len(ids)):
            np.testing.assert_allclose(after[k, 0:3], p[3 * k:3 * k + 3], atol=1e-5)
            close(after[k, 6], math.sin(0.4))
            close(after[k, 7], math.cos(0.4))
            for f in (3, 16, 17, 18, 19, 20, 21, 22, 23, 32, 33, 34, 35, 36, 37, 38, 39):
                close(after[k, f], before[k, f])
        g.setPoses(ids, p)
        unchanged = g.readSelectedBodies(ids)
        np.testing.assert_array_equal(unchanged[:, 4:8], after[:, 4:8])
        bad_p = p.copy()
        bad_p[-1] = np.nan
        zero_q = q.copy()
        zero_q[-4:] = 0
        dup = ids.copy()
        dup[-1] = dup[0]
        for args in ((ids, bad_p), (ids, p, zero_q), (dup, p), (ids, p[1:]), (ids, p.tolist())):
            with pytest.raises((ValueError, TypeError)):
                g.setPoses(*args)
            np.testing.assert_array_equal(g.readSelectedBodies(ids), unchanged)
        g.body(int(ids[0])).setLinearVelocity([4, 5, 6])
        g.setPoses(ids, p)
        g.body(int(ids[0])).setLinearVelocity([7, 8, 9])
        np.testing.assert_allclose(g.body(int(ids[0])).readState()["linearVelocity"], [7, 8, 9], atol=1e-5)
        assert g.setPoses(np.zeros(0, np.uint32), np.zeros(0, np.float32)) is g
    finally:
        pipeline.destroy()
