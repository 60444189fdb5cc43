"""Pins of the oracle's constraint features by the mechanics they must produce (CPU only):

* rest frames (constraintFrames.wgsl, gpuSolverKernels.js:105-130): an angle-locked joint holds the relative rotation
  its bodies had when it was made, not identity; the captured quaternion is normalize(q_A⁻¹ q_B);
* motors (angularConstraints.js): a world motor turns a floating body up to its speed, accelerating at no more than
  maxTorque / I while the bound holds (ω(t) = τ t / I), and reverses or stops with setMotor;
* limits: a body spun about a limited hinge axis stops at the limit;
* spring breaking (springMaterial.js): a spring stretched past breakStrain of its rest length, or pulling with more
  than breakForce, tears for good and lets go; below either it holds;
* releaseJoints / disableConstraint (solver.ts:1138-1204): the constraint stops acting and its bodies collide again;
  setWorldAnchor moves a world joint's anchor point, and the hard joint closes the gap by (1 − α) of it per step.
"""
import importlib
import math

import numpy as np
import pytest

from oracle_lib import OracleWorld

avbd = importlib.import_module("babylon-avbd_amd")


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def axis_angle(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.array([*(a * math.sin(angle / 2)), math.cos(angle / 2)])


def rel_angle(qa, qb, q0):
    """Angle of the rotation between relative orientations q_A⁻¹ q_B and q0."""
    r = qmul(qmul(qconj(qa), qb), qconj(q0))
    return 2 * math.acos(min(1.0, abs(r[3])))


def test_angle_locked_joint_holds_its_starting_relative_rotation():
    s = avbd.AvbdScene3D(gravity=-9.81)
    qa = axis_angle([0, 0, 1], 0.5)
    qb = axis_angle([0, 1, 0], 0.9)
    a = s.addBox([1, 1, 1], density=0, position=[0, 5, 0], rotation=list(qa))
    b = s.addBox([1, 0.4, 2], position=[0, 3.5, 0], rotation=list(qb))
    s.addJoint(a, b, [0, -0.75, 0], [0, 0.75, 0], linearStiffness=math.inf, angularStiffness=math.inf)
    o = OracleWorld([s], precision=32)
    rest = o.joints()[0][[19, 23, 27, 31]]
    want = qmul(qconj(qa.astype(np.float32).astype(float)), qb.astype(np.float32).astype(float))
    want /= np.linalg.norm(want)
    assert np.allclose(rest, want, atol=2e-7), (rest, want)
    q0 = qmul(qconj(o.bodies()[a, 4:8]), o.bodies()[b, 4:8])
    o.step(180)
    body = o.bodies()
    assert rel_angle(body[a, 4:8], body[b, 4:8], q0) < 2e-3, "the weld keeps the starting relative rotation"
    assert abs(body[b, 32:35]).max() < 0.05, "and holds B still against gravity"


def test_angle_locked_joint_without_a_rest_frame_would_turn():
    """The same weld with identity as its target turns B to A's orientation: the rest frame is what keeps it."""
    s = avbd.AvbdScene3D(gravity=0.0)
    a = s.addBox([1, 1, 1], density=0, position=[0, 5, 0])
    b = s.addBox([1, 1, 1], position=[0, 3.8, 0], rotation=list(axis_angle([0, 1, 0], 0.6)))
    s.addJoint(a, b, [0, -0.6, 0], [0, 0.6, 0], linearStiffness=math.inf, angularStiffness=math.inf)
    o = OracleWorld([s], precision=32)
    st = dict(bodies=o.bodies().astype(np.float32), joints=o.joints().astype(np.float32), refs=o.refs(),
              colors=o.colors(), manifolds=o.manifolds(), contacts=o.contact_words(), steps=np.array([0], np.uint32))
    st["bodies"].view(np.uint32)[:, 27] = 0
    st["joints"][0, [19, 23, 27, 31]] = 0  # no rest orientation: identity target
    o.seed(st)
    o.step(120)
    assert rel_angle(o.bodies()[a, 4:8], o.bodies()[b, 4:8], np.array([0, 0, 0, 1.0])) < 0.3  # from 0.6 rad


def _floating_box(motor=None, limit=None, spin=0.0):
    s = avbd.AvbdScene3D(gravity=0.0)
    box = s.addBox([2, 1, 1], position=[0, 0, 0], angularVelocity=[0, spin, 0])
    slot = None
    if motor is not None:
        slot = s.addMotor(None, box, [0, 1, 0], [0, 1, 0], **motor)
    if limit is not None:
        slot = s.addHinge(None, box, [0, 0, 0], [0, 0, 0], [0, 1, 0], [0, 1, 0], **limit)["limit"]
    return s, box, slot


def test_motor_reaches_its_speed():
    s, box, slot = _floating_box(motor=dict(speed=2.0, maxTorque=1e4))
    o = OracleWorld([s], precision=32)
    o.step(60)
    w = o.bodies()[box, 36:39]
    assert abs(w[1] - 2.0) < 0.02 and abs(w[0]) < 1e-3 and abs(w[2]) < 1e-3, w


def test_motor_torque_bound_limits_the_acceleration():
    s, box, slot = _floating_box(motor=dict(speed=10.0, maxTorque=0.5))
    I = s.record(box)[21]  # moment about y: m (x² + z²) / 12 = 2 (4 + 1) / 12
    o = OracleWorld([s], precision=32)
    dt = s.params["dt"]
    for k in (30, 60):
        o.step(30)
        w = o.bodies()[box, 37]
        want = 0.5 * k * dt / I
        assert abs(w - want) < 0.03 * want, (k, w, want)


def test_set_motor_changes_speed():
    s, box, slot = _floating_box(motor=dict(speed=2.0, maxTorque=1e4))
    o = OracleWorld([s], precision=32)
    o.step(30)
    o.set_motor(slot, speed=-1.0)
    o.step(30)
    assert abs(o.bodies()[box, 37] + 1.0) < 0.02
    o.set_motor(slot, max_torque=0.0)  # no torque: the motor lets the body spin freely
    o.step(30)
    assert abs(o.bodies()[box, 37] + 1.0) < 0.02


def test_limit_stops_a_spinning_body_at_its_angle():
    s, box, slot = _floating_box(limit=dict(minAngle=-0.4, maxAngle=0.7), spin=3.0)
    o = OracleWorld([s], precision=32)
    angles = []
    for _ in range(90):
        o.step(1)
        q = o.bodies()[box, 4:8]
        angles.append(2 * math.atan2(q[1], q[3]))  # rotation about y
    assert max(angles) < 0.7 + 0.03 and min(angles) > -0.4 - 0.03, (max(angles), min(angles))
    assert max(angles) > 0.6, "it reached the limit"
    assert abs(o.bodies()[box, 37]) < 0.5, "and lost most of its 3 rad/s there"


def _hanging(stiffness, mass_density, **material):
    s = avbd.AvbdScene3D(gravity=-9.81)
    top = s.addBox([1, 1, 1], density=0, position=[0, 10, 0])
    box = s.addBox([1, 1, 1], density=mass_density, position=[0, 8, 0])
    slot = s.addSpring(top, box, [0, -0.5, 0], [0, 0.5, 0], stiffness=stiffness, **material)
    return s, box, slot


@pytest.mark.parametrize("material,breaks", [(dict(breakStrain=0.2), True), (dict(breakStrain=0.8), False),
                                             (dict(breakForce=100.0), True), (dict(breakForce=300.0), False)])
def test_spring_breaks_past_its_strain_or_force(material, breaks):
    # rest 1 m, stiffness 200: a 40 kg box stretches it by 40 x 9.81 / 200 = 1.96 m at rest (strain 1.96, tension
    # 392 N), more while it first drops: it breaks at strain 0.2 or force 100; a 2 kg box stretches it by 0.098 m
    # (strain 0.098, tension 19.6 N, at most about twice that while it bounces): it holds at 0.8 or 300
    density = 40.0 if breaks else 2.0
    s, box, slot = _hanging(200.0, density, **material)
    o = OracleWorld([s], precision=32)
    o.step(240)
    k = o.joints()[slot]
    y = o.bodies()[box, 1]
    if breaks:
        assert k[3] == 0 and y < 0, "the torn spring lets the box fall"
        assert k[11] == np.float32(material.get("breakStrain", 0)) and k[15] == np.float32(material.get("breakForce", 0))
    else:
        assert k[3] == 200.0 and 7.5 < y < 8.5, (k[3], y)


def test_release_and_disable_let_go():
    for op in ("release", "disable"):
        s = avbd.AvbdScene3D(gravity=-9.81)
        top = s.addBox([1, 1, 1], density=0, position=[0, 10, 0])
        box = s.addBox([1, 1, 1], position=[0, 9.0, 0])
        slot = s.addJoint(top, box, [0, -0.5, 0], [0, 0.5, 0])
        o = OracleWorld([s], precision=32)
        o.step(30)
        assert abs(o.bodies()[box, 1] - 9.0) < 0.02
        getattr(o, "release_joints" if op == "release" else "disable_constraint")([slot] if op == "release" else slot)
        o.step(30)
        assert o.bodies()[box, 1] < 8.0, op
        # the bodies collide again: none of the pair's (box, top) contacts are ignored; the box fell away from it
        k = o.joints()[slot]
        assert k[3] == 0 and k[7] == 0


def test_set_world_anchor_moves_the_hanging_point():
    s = avbd.AvbdScene3D(gravity=-9.81)
    box = s.addBox([1, 1, 1], position=[0, 5, 0])
    slot = s.addJoint(None, box, [0, 5.5, 0], [0, 0.5, 0])
    o = OracleWorld([s], precision=32)
    o.step(20)
    o.set_world_anchor(slot, [2.0, 5.5, 0.0])
    o.step(200)
    b = o.bodies()[box]
    x, y, z, w = b[4:8]
    top = b[0:3] + np.array([2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)]) * 0.5  # pos + R (0, 0.5, 0)
    # a hard joint removes (1 - α) of the error it starts a step with (rows C − α C0): after the 2 m jump, 200 steps
    # leave 2 × 0.99^200 = 0.27 m
    err = np.linalg.norm(top - [2.0, 5.5, 0.0])
    assert abs(err - 2.0 * 0.99 ** 200) < 0.02, (top, err)


def test_oracle_atan2_against_the_library():
    from oracle_lib import lib  # noqa: F401  (the FP32 atan2 is exercised through the hinge rows above)
    rng = np.random.default_rng(3)
    # the hinge angle of a floating box turned by a known angle about the axis reads back that angle
    for angle in rng.uniform(-3.1, 3.1, 12):
        s = avbd.AvbdScene3D(gravity=0.0)
        box = s.addBox([1, 1, 1], rotation=list(axis_angle([0, 1, 0], angle)))
        s.addHinge(None, box, axisA=[0, 1, 0], axisB=[0, 1, 0], minAngle=-3.14, maxAngle=3.14)
        o = OracleWorld([s], precision=32)
        o.step(1)
        assert abs(o.bodies()[box, 37]) < 1e-4  # captured at the start: the limit is not violated, nothing moves
