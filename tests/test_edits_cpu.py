"""Edits between steps on the CPU: the oracle's body commands against closed forms, and the host staging of the
Python and JavaScript facades (src/native/runtime.js:16-175, solver3D.js:130-213, packedPoses.js:3-72).

Closed forms (independent of the quaternion routines both sides use): an impulse J at world point p changes the
linear velocity by J/m and the angular velocity by I_w⁻¹ ((p - x) × J) with I_w = R diag(I) Rᵀ built from the rotation
matrix; so the total angular momentum about the origin, I_w ω + m x × v, changes by exactly p × J. A force F or a
torque τ over dt is the impulse F·dt or τ·dt. A transposed rotation, a dropped lever arm or a wrong-signed cross
product fails these.
"""
import importlib
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import OracleWorld

ROOT = Path(__file__).resolve().parents[1]
edits = importlib.import_module("babylon-avbd_amd.edits")


def _scene(avbd):
    s = avbd.AvbdScene3D(gravity=-9.81, iterations=10, timeStep=1 / 60)
    s.addBox([20, 1, 20], density=0, friction=0.5, position=[0, -0.5, 0])
    s.addBox([0.6, 1.4, 2.2], density=1.7, position=[0.3, 0.71, -0.2], rotation=[0.1, 0.3, -0.2, 0.9])
    s.addBox([1.0, 0.5, 0.8], density=0.9, position=[3.0, 0.26, 1.0], rotation=[0.0, 0.38, 0.0, 0.92],
             velocity=[0.2, 0, -0.1], angularVelocity=[0, 0.5, 0])
    s.addBox([1, 1, 1], density=1, position=[-4, 3, 0])
    return s


def cmd(index, kind, value, point=None, w=0.0):
    r = np.zeros(12, np.uint32)
    r[0], r[1], r[2] = index, kind, 0 if point is None else 1
    f = r.view(np.float32)
    f[4:7] = value
    f[7] = w
    if point is not None:
        f[8:8 + len(point)] = point
    return r


def rotation_matrix(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def world_inertia(b):
    R = rotation_matrix(b[4:8])
    return R @ np.diag(b[20:23]) @ R.T


def angular_momentum(b):
    """I_w ω + m x × v about the origin."""
    return world_inertia(b) @ b[36:39] + b[19] * np.cross(b[0:3], b[32:35])


def oracle(avbd):
    return OracleWorld([_scene(avbd)], precision=32, color_cap=16, color_rounds=16)


def test_velocity_replacements(avbd):
    o = oracle(avbd)
    o.apply_commands(np.stack([cmd(1, 0, [1.5, -2.0, 0.25]), cmd(2, 1, [0.0, 3.0, -1.0])]))
    b = o.bodies()
    assert np.array_equal(b[1, 32:35], np.float32([1.5, -2.0, 0.25]))
    assert np.array_equal(b[2, 36:39], np.float32([0.0, 3.0, -1.0]))
    # replacements keep the w words (last step's up velocity, shape code)
    b0 = OracleWorld([_scene(avbd)]).bodies()
    assert b[1, 35] == b0[1, 35] and b[2, 39] == b0[2, 39]


@pytest.mark.parametrize("body", [1, 2])
def test_impulse_at_a_point_changes_momenta_by_the_closed_forms(avbd, body):
    o = oracle(avbd)
    before = o.bodies()[body]
    J = np.array([0.7, -1.3, 2.1])
    p = before[0:3] + np.array([0.25, 0.4, -0.3])
    o.apply_commands(cmd(body, 2, J, p)[None])
    after = o.bodies()[body]
    m = before[19]
    np.testing.assert_allclose(after[32:35] - before[32:35], J / m, rtol=1e-5, atol=1e-6)
    dw = np.linalg.solve(world_inertia(before), np.cross(p - before[0:3], J))
    np.testing.assert_allclose(after[36:39] - before[36:39], dw, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(angular_momentum(after) - angular_momentum(before), np.cross(p, J), rtol=1e-4, atol=1e-5)
    # the pose does not move
    assert np.array_equal(after[0:8], before[0:8])


def test_force_and_torque_are_impulses_over_dt(avbd):
    dt = np.float32(1 / 60)
    a, b = oracle(avbd), oracle(avbd)
    F = [3.0, 0.5, -2.0]
    T = [0.2, -0.7, 1.1]
    p = [0.5, 1.0, 0.0]
    a.apply_commands(np.stack([cmd(1, 3, F, p, w=dt), cmd(2, 6, T, w=dt)]))
    b.apply_commands(np.stack([cmd(1, 2, np.float32(F) * dt, p), cmd(2, 5, np.float32(T) * dt)]))
    assert np.array_equal(a.bodies(), b.bodies())
    before = oracle(avbd).bodies()[2]
    dw = np.linalg.solve(world_inertia(before), np.float64(T) * float(dt))
    np.testing.assert_allclose(a.bodies()[2, 36:39] - before[36:39], dw, rtol=1e-4, atol=1e-7)


def test_angular_impulse_without_a_lever_leaves_linear_velocity(avbd):
    o = oracle(avbd)
    before = o.bodies()[1]
    L = np.array([0.3, 0.9, -0.4])
    o.apply_commands(cmd(1, 5, L)[None])
    after = o.bodies()[1]
    assert np.array_equal(after[32:35], before[32:35])
    np.testing.assert_allclose(world_inertia(before) @ (after[36:39] - before[36:39]), L, rtol=1e-4, atol=1e-6)


def test_static_bodies_ignore_impulses_but_take_velocities(avbd):
    o = oracle(avbd)
    before = o.bodies()[0]
    o.apply_commands(np.stack([cmd(0, 2, [1, 2, 3], [0.5, 0, 0]), cmd(0, 3, [1, 1, 1], w=0.1),
                               cmd(0, 5, [1, 0, 0]), cmd(0, 6, [0, 1, 0], w=0.1)]))
    assert np.array_equal(o.bodies()[0], before)
    o.apply_commands(cmd(0, 0, [0, 0, 1])[None])
    assert np.array_equal(o.bodies()[0, 32:35], [0, 0, 1])


def test_teleport_sets_every_pose_word(avbd):
    o = oracle(avbd)
    before = o.bodies()[3]
    q = np.float32([0.0, 0.6, 0.0, 0.8])
    o.apply_commands(np.stack([cmd(3, 4, [1.0, 2.0, 3.0], q), cmd(2, 4, [5.0, 1.0, 5.0])]))
    b = o.bodies()
    for base in (0, 8, 24):  # pos, initial pose, inertial pose
        assert np.array_equal(b[3, base:base + 3], [1.0, 2.0, 3.0])
    for base in (4, 12, 28):
        assert np.array_equal(b[3, base:base + 4], q)
    assert np.array_equal(b[3, 32:40], before[32:40]), "velocities unchanged"
    assert b[3, 3] == before[3], "friction kept in pos.w"
    # without a rotation the rotation is kept and copied to the initial and inertial rotations
    assert np.array_equal(b[2, 12:16], b[2, 4:8]) and np.array_equal(b[2, 28:32], b[2, 4:8])


def test_commands_apply_in_submission_order(avbd):
    a, b = oracle(avbd), oracle(avbd)
    set_v = cmd(1, 0, [1.0, 0.0, 0.0])
    kick = cmd(1, 2, [0.0, 2.0, 0.0])
    a.apply_commands(np.stack([set_v, kick]))
    b.apply_commands(np.stack([kick, set_v]))
    m = np.float32(a.bodies()[1, 19])
    assert np.array_equal(a.bodies()[1, 32:35], np.float32([1.0, 0.0 + 2.0 / m, 0.0]))
    assert np.array_equal(b.bodies()[1, 32:35], [1.0, 0.0, 0.0])
    # commands of different bodies commute
    c, d = oracle(avbd), oracle(avbd)
    other = cmd(2, 2, [0.5, 0.5, 0.5], [3.0, 0.5, 1.0])
    c.apply_commands(np.stack([kick, other, set_v]))
    d.apply_commands(np.stack([other, kick, set_v]))
    assert np.array_equal(c.bodies(), d.bodies())


def test_teleport_drops_the_warm_start_of_the_bodys_manifolds(avbd):
    o = oracle(avbd)
    o.step(5)
    mf = o.manifolds()
    touching = (mf[:, 0] == 1) | (mf[:, 1] == 1)
    assert touching.any() and (~touching).any() and ((mf[:, 3] & 15) > 0).all()
    o.apply_commands(cmd(1, 4, [0.3, 5.0, -0.2])[None])
    after = o.manifolds()
    assert (after[touching, 3] == 0).all()
    assert np.array_equal(after[~touching], mf[~touching])
    # velocity-only edits keep every manifold
    o2 = oracle(avbd)
    o2.step(5)
    o2.apply_commands(cmd(1, 0, [0, 1, 0])[None])
    assert np.array_equal(o2.manifolds(), mf)


# --- host staging -------------------------------------------------------------------------------------------------


class _Staging:
    """Stands in for a world in the staging calls (no device): body count, world ranges and time steps."""

    def __init__(self, avbd, body_count=10, first=(0, 6), dts=(1 / 60, 1 / 120)):
        self.world = object.__new__(avbd.HipWorld)
        self.world.destroyed = False
        self.world.bodyCount = body_count
        self.world.world_first_body = np.array(first)
        self.world._dts = list(dts)
        self.world._pending = edits.CommandStaging()
        self.filters = []
        self.world.setFilters = lambda i, g, m: self.filters.append((list(i), list(g), list(m)))

    def words(self):
        return self.world._pending.records().copy()


EDIT_OPS = [
    ["setLinearVelocity", 1, [1, 2, 3]],
    ["applyForce", 7, [0, 10, 0], [0.5, 0.5, 0.5]],
    ["editBodies", [{"body": 2, "force": [1, 0, 0], "position": [0, 1, 0], "rotation": [0, 0, 2, 2],
                     "linearVelocity": {"x": 1, "y": 0, "z": 0}, "torque": [0, 0, 1], "worldPoint": [1, 1, 1],
                     "group": 4, "collidesWith": 5},
                    {"body": 8, "impulse": [0, 1, 0], "angularImpulse": [1, 0, 0]}]],
    ["teleport", 3, [1, 1, 1], [1, 1, 1, 1]],
    ["setPoses", [4, 5, 9], [0, 1, 0, 2, 3, 4, 5, 6, 7], [0, 0, 0, 3, 1, 1, 1, 1, 0, 0.6, 0, 0.8]],
    ["applyTorque", 9, [0, 1, 0]],
    ["setAngularVelocity", 0, [0, 0, 1]],
]


def _apply_py(world, ops):
    for method, *a in ops:
        if method == "setPoses":
            world.setPoses(np.array(a[0], np.uint32), np.array(a[1], np.float32),
                           None if len(a) < 3 else np.array(a[2], np.float32))
        else:
            getattr(world, method)(*a)


def test_edit_records_follow_the_reference_order(avbd):
    st = _Staging(avbd)
    _apply_py(st.world, EDIT_OPS)
    w = st.words()
    f = w.view(np.float32)
    kinds = [(int(r[0]), int(r[1])) for r in w]
    assert kinds == [(1, 0), (7, 3), (2, 4), (2, 0), (2, 3), (2, 6), (8, 2), (8, 5), (3, 4), (4, 4), (5, 4), (9, 4),
                     (9, 6), (0, 1)]
    # force and torque carry their world's dt in value.w; body 7 is in the second world
    assert f[1, 7] == np.float32(1 / 120) and f[4, 7] == np.float32(1 / 60) and f[12, 7] == np.float32(1 / 120)
    # a world point goes with impulses and forces only
    assert w[1, 2] == 1 and w[4, 2] == 1 and w[5, 2] == 0 and w[6, 2] == 0
    # teleport rotations are normalised in double precision, then stored as f32
    assert np.array_equal(f[2, 8:12], np.float32([0, 0, 2 / np.sqrt(8), 2 / np.sqrt(8)]))
    assert np.array_equal(f[8, 8:12], np.float32([0.5, 0.5, 0.5, 0.5]))
    # setPoses scales by 1 / hypot: the second pose's rotation (1, 1, 1, 1) becomes 0.5 each
    assert np.array_equal(f[10, 8:12], np.float32([0.5, 0.5, 0.5, 0.5]))
    assert st.filters == [([2], [4], [5])]


@pytest.mark.parametrize("op, message", [
    (["setLinearVelocity", 10, [0, 0, 0]], "A live 3D GPU body index is required"),
    (["applyImpulse", 1, [0, float("nan"), 0]], "impulse requires 3 finite components"),
    (["teleport", 1, [0, 0, 0], [0, 0, 0, 0]], "rotation must be nonzero"),
    (["editBodies", [{"body": 1, "rotation": [0, 0, 0, 1]}]], "A rotation edit requires position"),
    (["editBodies", [{"body": 1, "collidesWith": 3}]], "A collision-mask edit requires group as well"),
    (["editBodies", [{"body": 1, "linearVelocity": [1, 0, 0]}, {"body": 99, "impulse": [1, 0, 0]}]],
     "A live 3D GPU body index is required"),
    (["setPoses", [3, 2], [0, 0, 0, 1, 1, 1]], "Pose batch slots must be unique and ascending"),
    (["setPoses", [3], [0, 0, float("inf")]], "Pose positions must be finite"),
])
def test_invalid_edits_queue_nothing(avbd, op, message):
    st = _Staging(avbd)
    st.world.setLinearVelocity(0, [1, 0, 0])
    with pytest.raises(ValueError, match=message):
        _apply_py(st.world, [op])
    assert len(st.words()) == 1, "a rejected batch must leave the queue as it was"


def test_unsupported_properties_are_refused(avbd):
    """Trigger bodies are not built: their edits are refused; a restitution outside [0, 1] and a non-boolean
    allowSleep too (runtime.js:83-84); an allowSleep edit becomes setSleepEnabled."""
    st = _Staging(avbd)
    with pytest.raises(NotImplementedError):
        st.world.editBodies([{"body": 1, "isTrigger": True}])
    with pytest.raises(ValueError):
        st.world.editBodies([{"body": 1, "restitution": 1.5}])
    with pytest.raises(ValueError, match="allowSleep must be a boolean"):
        st.world.editBodies([{"body": 1, "allowSleep": 1}])
    calls = []
    st.world.setSleepEnabled = lambda i, e: calls.append((i, e))
    st.world.editBodies([{"body": 3, "allowSleep": False}])
    assert calls == [(3, False)]


NODE = shutil.which("node")


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_js_facade_stages_the_same_records(avbd, tmp_path):
    st = _Staging(avbd)
    _apply_py(st.world, EDIT_OPS)
    spec = {"bodyCount": 10, "firstBodies": [0, 6], "dts": [1 / 60, 1 / 120], "ops": EDIT_OPS}
    f = tmp_path / "ops.json"
    f.write_text(json.dumps(spec))
    r = subprocess.run([NODE, str(ROOT / "tests" / "js" / "facade_check.js"), "stage", str(f)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["error"] is None
    assert out["words"] == st.words().ravel().tolist()
    assert out["filters"] == [[[2], [4], [5]]]


def _bounce_scene(avbd, e, speed, gravity=0.0, radius=0.5, sphere=True):
    s = avbd.AvbdScene3D(gravity=gravity)
    s.addBox([10, 1, 10], density=0, position=[0, -0.5, 0])
    if sphere:
        s.addSphere(radius, position=[0, radius + 0.01, 0], velocity=[0, -speed, 0])
    else:
        s.addBox([1, 1, 1], position=[0, 0.51, 0], velocity=[0, -speed, 0])
    return s


@pytest.mark.parametrize("e", [1.0, 0.5, 0.25])
def test_restitution_rebounds_at_e_times_the_impact_speed(avbd, e):
    """A sphere hitting a static floor head on (one manifold, degree 1, no spin): the first bounce round sets the
    normal velocity to the rebound speed −e·vn exactly, so the ball leaves at e times its impact speed."""
    from oracle_lib import OracleWorld
    speed = 3.0
    w = OracleWorld([_bounce_scene(avbd, e, speed)], precision=32)
    w.set_restitution([1], [e])
    for _ in range(12):  # the first step whose velocity points up is the bounce
        w.step()
        vy = w.bodies()[1, 33]
        if vy > 0:
            break
    assert abs(vy - e * speed) < 1e-5 * speed, f"e={e}: left at {vy}, expected {e * speed}"


def test_no_restitution_no_bounce_and_slow_contacts_do_not_bounce(avbd):
    """e = 0 leaves the step untouched (bit for bit the plain step), and an approach slower than 1 m/s never bounces."""
    from oracle_lib import OracleWorld
    a = OracleWorld([_bounce_scene(avbd, 0, 3.0)], precision=32)
    b = OracleWorld([_bounce_scene(avbd, 0, 3.0)], precision=32)
    b.set_restitution([1], [0.0])
    for _ in range(10):
        a.step()
        b.step()
    np.testing.assert_array_equal(a.bodies(), b.bodies())
    slow = OracleWorld([_bounce_scene(avbd, 1.0, 0.5)], precision=32)
    slow.set_restitution([1], [1.0])
    for _ in range(12):
        slow.step()
        assert slow.bodies()[1, 33] < 0.5, "a 0.5 m/s approach bounced"


def test_reference_restitution_contract(avbd):
    """backend-contract.js checkContacts: a sphere with restitution 1 at −3 m/s onto the floor, gravity 0: after 12
    steps it moves up faster than 2 m/s. Also for a box (a four-point face manifold)."""
    from oracle_lib import OracleWorld
    for sphere in (True, False):
        w = OracleWorld([_bounce_scene(avbd, 1.0, 3.0, sphere=sphere)], precision=32)
        w.set_restitution([1], [1.0])
        w.step(12)
        assert w.bodies()[1, 33] > 2, f"restitution failed ({'sphere' if sphere else 'box'}): {w.bodies()[1, 33]}"


@pytest.mark.parametrize("precision", [32, 64])
def test_restitution_keeps_momentum_between_two_free_bodies(avbd, precision):
    """Two equal free spheres meeting head on with e = 1, no gravity and no floor: the bounce passes add equal and
    opposite velocity changes (the step without restitution, from the same state, is the reference), and they rebound
    at about their approach speed."""
    from oracle_lib import OracleWorld

    def world(e):
        s = avbd.AvbdScene3D(gravity=0)
        s.addSphere(0.5, position=[-0.505, 0, 0], velocity=[3, 0, 0])
        s.addSphere(0.5, position=[0.505, 0, 0], velocity=[-3, 0, 0])
        w = OracleWorld([s], precision=precision)
        w.set_restitution([0, 1], [e, e])
        return w
    bounced, plain = world(1.0), world(0.0)
    for _ in range(2):  # the spheres touch in the second step
        bounced.step()
        plain.step()
    change = bounced.bodies()[:, 32:35] - plain.bodies()[:, 32:35]
    assert np.abs(change[0] + change[1]).max() < (1e-5 if precision == 32 else 1e-12)
    assert bounced.bodies()[0, 32] < -2.9 and bounced.bodies()[1, 32] > 2.9
