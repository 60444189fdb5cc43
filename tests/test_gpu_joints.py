"""The constraint half of the world contract on the GPU against the oracle, through the C ABI: angle-locked joints
with their rest frames, hinges (two ball joints, a limit, a motor), world motors, breakable springs, a world joint,
and the edits between steps (setMotor, setWorldAnchor, disableConstraint, releaseJoints, setSpringMaterial,
setAngularConstraint). The frames each world captures at creation equal the oracle's bit for bit; then every step,
from the GPU's state on both sides, matches (pairs, keys, colours bit-exact; poses, velocities, contact and joint
multipliers within 1e-5), and a run without reseeding stays bit-identical — with the colour launches, the dataflow
solve (joint duals fused into the finishing body's unit; joints without a dynamic body dualled after it) and one
workgroup per world."""
import math

import numpy as np
import pytest

from oracle_lib import OracleWorld
from test_gpu_parity import compare_step

pytestmark = pytest.mark.gpu

OPTS = dict(colorCap=32, colorRounds=16)


def _quat(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return [*(a * math.sin(angle / 2)), math.cos(angle / 2)]


def scene(avbd):
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([40, 1, 40], density=0, friction=0.6, position=[0, -0.5, 0])
    slots = {}
    # an angle-locked chain hanging from a static post, links turned against each other
    post = s.addBox([0.4, 0.4, 0.4], density=0, position=[-4, 6, 0], rotation=_quat([0, 0, 1], 0.3))
    prev = post
    chain = []
    for k in range(3):
        link = s.addBox([0.3, 0.8, 0.3], position=[-4, 5.2 - 0.9 * k, 0], rotation=_quat([0, 1, 0], 0.4 * (k + 1)))
        chain.append(s.addJoint(prev, link, [0, -0.45 if k else -0.25, 0], [0, 0.45, 0], angularStiffness=math.inf))
        prev = link
    slots["chain"] = chain
    # a door on a hinge (about y) to a static frame, limited to [-0.6, 0.9] and driven by a motor
    frame = s.addBox([0.2, 2, 0.2], density=0, position=[0, 1.2, -3])
    door = s.addBox([1.2, 1.8, 0.1], position=[0.7, 1.2, -3])
    hinge = s.addHinge(frame, door, [0.1, 0, 0], [-0.6, 0, 0], [0, 1, 0], [0, 1, 0], span=1.6, minAngle=-0.6,
                       maxAngle=0.9, motor=dict(speed=1.5, maxTorque=20.0))
    slots["hinge"] = hinge
    # a wheel on the ground turned by a world motor about the vertical
    wheel = s.addBox([1.0, 0.3, 1.0], position=[3, 0.16, 2])
    slots["wheel_motor"] = s.addMotor(None, wheel, [0, 1, 0], [0, 1, 0], speed=3.0, maxTorque=5.0)
    # springs: a heavy box that tears its spring, a light one that holds
    top = s.addBox([0.5, 0.5, 0.5], density=0, position=[4, 7, -2])
    heavy = s.addBox([0.6, 0.6, 0.6], density=60.0, position=[4, 5.5, -2])
    slots["spring_tears"] = s.addSpring(top, heavy, [0, -0.25, 0], [0, 0.3, 0], stiffness=300, breakStrain=0.5)
    light = s.addBox([0.4, 0.4, 0.4], density=1.0, position=[4.8, 5.6, -2])
    slots["spring_holds"] = s.addSpring(top, light, [0.25, -0.25, 0], [0, 0.2, 0], stiffness=300, breakForce=500.0)
    # a box on a world joint, and a pair welded to each other resting on the ground
    hung = s.addBox([0.5, 0.5, 0.5], position=[-2, 4, 3])
    slots["world_joint"] = s.addJoint(None, hung, [-2, 4.5, 3], [0, 0.25, 0])
    a = s.addBox([0.6, 0.6, 0.6], position=[1.5, 0.31, 3], rotation=_quat([0, 1, 0], 0.2))
    b = s.addBox([0.6, 0.6, 0.6], position=[1.5, 0.92, 3], rotation=_quat([1, 0, 0], 0.3))
    slots["weld"] = s.addJoint(a, b, [0, 0.3, 0], [0, -0.3, 0], angularStiffness=math.inf)
    # a joint between two static bodies (no dynamic body lists it: its duals still run every iteration)
    slots["static_pair"] = s.addJoint(post, frame, [0, -0.2, 0], [0, 1.0, 0], linearStiffness=500)
    for x in np.linspace(-5, 5, 6):  # a few loose boxes to land on things
        s.addBox([0.5, 0.5, 0.5], position=[float(x), 3.0, 1.0])
    return s, slots


def _edit(world, oracle, step, slots):
    """The same edit on both sides at a given step."""
    if step == 20:
        world.setMotor(slots["wheel_motor"], speed=-2.0)
        oracle.set_motor(slots["wheel_motor"], speed=-2.0)
        world.setMotor(slots["hinge"]["motor"], maxTorque=40.0)
        oracle.set_motor(slots["hinge"]["motor"], max_torque=40.0)
    elif step == 30:
        world.setWorldAnchor(slots["world_joint"], [-1.6, 4.4, 3.2])
        oracle.set_world_anchor(slots["world_joint"], [-1.6, 4.4, 3.2])
    elif step == 40:
        world.disableConstraint(slots["weld"])
        oracle.disable_constraint(slots["weld"])
    elif step == 50:
        world.releaseJoints([slots["chain"][2], 10_000])
        oracle.release_joints([slots["chain"][2]])
    elif step == 60:
        world.setSpringMaterial(slots["spring_holds"], breakStrain=0.001)
        oracle.set_spring_material(slots["spring_holds"], 0.001, math.inf)
    elif step == 70:
        world.setAngularConstraint(slots["hinge"]["limit"], "limit", [0, 1, 0], [0, 1, 0], minAngle=-0.2,
                                   maxAngle=0.3)
        oracle.set_angular_constraint(slots["hinge"]["limit"], 4, [0, 1, 0], [0, 1, 0], -0.2, 0.3)


def test_frames_captured_at_creation_match_the_oracle(avbd):
    s, slots = scene(avbd)
    world = avbd.createSolverPipeline("hip").createSolver(s, OPTS)
    oracle = OracleWorld([scene(avbd)[0]], precision=32, color_cap=32, color_rounds=16)
    got = world.readJoints()
    want = oracle.joints().astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
        f"joint records differ at {np.argwhere(got.view(np.uint32) != want.view(np.uint32))[:6].tolist()}"
    for j in slots["chain"] + [slots["weld"]]:
        assert np.linalg.norm(got[j, [19, 23, 27, 31]]) > 0.99, "a rest orientation was captured"
    for j in (slots["hinge"]["limit"], slots["hinge"]["motor"], slots["wheel_motor"]):
        assert abs(np.linalg.norm(got[j, 16:19]) - 1) < 1e-6 and abs(np.linalg.norm(got[j, 20:23]) - 1) < 1e-6
    world.destroy()


SOLVERS = ["colors", "dataflow", "worlds"]


@pytest.mark.parametrize("solver", SOLVERS)
def test_constraint_steps_and_edits_match_the_oracle(avbd, solver):
    s, slots = scene(avbd)
    world = avbd.createSolverPipeline("hip").createSolver(s, dict(OPTS, solver=solver))
    assert world.solverKind() == SOLVERS.index(solver) + 1
    print(f"constraints vs oracle: {solver}", flush=True)
    oracle = OracleWorld([scene(avbd)[0]], precision=32, color_cap=32, color_rounds=16)
    for step in range(90):
        _edit(world, oracle, step, slots)
        state = world.readState()
        oracle.seed(state)
        world.step()
        oracle.step()
        compare_step(world, oracle)
    info = [world.jointInfo(j) for j in range(world.jointCount)]
    assert info[slots["chain"][2]][0] == 0 and info[slots["hinge"]["limit"]][0] == 4
    j = world.readJoints()
    assert j[slots["spring_tears"], 3] == 0, "the heavy box tore its spring"
    assert j[slots["spring_holds"], 3] == 0, "the lowered break strain tore the other"
    world.destroy()


@pytest.mark.parametrize("solver", SOLVERS)
def test_constraint_trajectory_bit_exact(avbd, solver):
    s, slots = scene(avbd)
    world = avbd.createSolverPipeline("hip").createSolver(s, dict(OPTS, solver=solver))
    print(f"constraint trajectory: {solver}", flush=True)
    oracle = OracleWorld([scene(avbd)[0]], precision=32, color_cap=32, color_rounds=16)
    for step in range(90):
        _edit(world, oracle, step, slots)
        world.step()
        oracle.step()
    got = world.readBodies()
    want = oracle.bodies().astype(np.float32)
    assert np.array_equal(got[:, 0:8], want[:, 0:8]), f"max pose difference {np.abs(got[:, :8] - want[:, :8]).max()}"
    joints = world.readJoints()
    assert np.array_equal(joints.view(np.uint32), oracle.joints().astype(np.float32).view(np.uint32))
    assert joints[slots["static_pair"], 0] == 500.0, "the static pair's duals ramped its penalty to its stiffness"
    world.destroy()


def test_constraint_edit_errors(avbd):
    s, slots = scene(avbd)
    world = avbd.createSolverPipeline("hip").createSolver(s, OPTS)
    before = world.readJoints().copy()
    bad = [lambda: world.setMotor(slots["weld"], speed=1.0),                 # not a motor
           lambda: world.setMotor(slots["wheel_motor"], maxTorque=-1.0),
           lambda: world.setWorldAnchor(slots["weld"], [0, 0, 0]),           # not a world joint
           lambda: world.setWorldAnchor(slots["world_joint"], [0, math.nan, 0]),
           lambda: world.setSpringMaterial(slots["weld"], 0.5),              # not a spring
           lambda: world.setSpringMaterial(slots["spring_holds"], -1.0),
           lambda: world.setAngularConstraint(slots["weld"], "limit", [0, 0, 0], [0, 1, 0]),
           lambda: world.setAngularConstraint(slots["weld"], "limit", minAngle=1.0, maxAngle=0.5),
           lambda: world.disableConstraint(10_000)]
    for f in bad:
        with pytest.raises((avbd.AvbdError, ValueError)):
            f()
    assert np.array_equal(world.readJoints().view(np.uint32), before.view(np.uint32)), "rejected edits change nothing"
    world.destroy()
