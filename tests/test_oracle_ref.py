"""The FP64 upstream restatement (oracle/ref3d.*) against the C++ demo's golden trajectories and the reference's
known-answer tests (reference/three-avbd/tests/avbd3d-ref.test.ts:27-163). CPU only."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from oracle_lib import lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle3d")
SCENES = ["Empty", "Ground", "Dynamic Friction", "Static Friction", "Pyramid", "Rope", "Heavy Rope", "Spring",
          "Spring Ratio", "Stack", "Stack Ratio", "Soft Body", "Bridge", "Breakable"]


class Ref:
    def __init__(self, scene, iterations=0):
        self.L = lib()
        self.h = self.L.ref_create(scene, iterations)
        assert self.h

    def __del__(self):
        self.L.ref_destroy(self.h)

    def step(self, n=1):
        self.L.ref_step(self.h, n)
        return self

    def bodies(self):
        n = self.L.ref_body_count(self.h)
        out = np.zeros((n, 17))
        self.L.ref_read_bodies(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def forces(self):
        return self.L.ref_force_count(self.h)

    def joints(self):
        err = ctypes.c_double()
        n = self.L.ref_joint_stats(self.h, ctypes.byref(err))
        return n, err.value


def run(name, frames):
    return Ref(SCENES.index(name)).step(frames)


@pytest.mark.parametrize("index", range(14))
def test_matches_upstream_golden_trajectories(index):
    """avbd3d-ref.test.ts:27-49: exact force counts and max pose difference < 1e-9 at frames 1/10/60/120/180."""
    with open(os.path.join(GOLDEN, f"scene-{index}.json")) as f:
        golden = json.load(f)
    assert golden["scene"] == SCENES[index]
    r = Ref(index)
    frame = 0
    for sample in golden["samples"]:
        r.step(sample["frame"] - frame)
        frame = sample["frame"]
        assert r.forces() == sample["forces"], f"{golden['scene']} frame {frame}: force count"
        b = r.bodies()
        if len(b):
            diff = np.abs(b[:, :7] - np.asarray(sample["bodies"])).max()
            assert diff < 1e-9, f"{golden['scene']} frame {frame}: max pose diff {diff}"


def test_boxes_rest_on_the_ground_within_the_margin():
    s = run("Stack", 900).bodies()
    for i, b in enumerate(s[1:]):
        assert abs(b[2] - (i + 1)) < 0.011 * (i + 1), f"box {i} z {b[2]}"
        assert math.hypot(b[0], b[1]) < 1e-6
        assert math.hypot(*b[7:10]) < 1e-3


def test_dynamic_friction_follows_coulomb():
    s = run("Dynamic Friction", 300).bodies()
    for i, b in enumerate(s[1:]):
        slid = b[0]
        if i == 10:
            assert abs(slid - 50) < 1e-6
            continue
        mu = math.sqrt((5 - i * 0.5) * 0.5)
        expected = 100 / (2 * mu * 10)
        assert abs(slid - expected) / expected < 0.15, f"box {i}: slid {slid}, Coulomb {expected}"


def test_static_friction_on_a_30_degree_ramp():
    r = run("Static Friction", 600)
    xs = r.bodies()[2:, 0].copy()
    b = r.step(120).bodies()[2:]
    for i, body in enumerate(b):
        mu = math.sqrt((i / 10) * 0.25 + 0.25)
        if mu > math.tan(math.pi / 6) + 0.01:
            assert body[2] > 7, f"box {i} left the ramp"
            assert abs(body[0] - xs[i]) < 0.005, f"box {i} creeps"
        elif mu < math.tan(math.pi / 6) - 0.01:
            assert body[2] < 1.5, f"box {i} still on the ramp"


def test_spring_oscillates_about_its_static_equilibrium():
    r = run("Spring", 600)
    total = 0.0
    for _ in range(600):
        total += r.step().bodies()[2][2]
    assert abs(total / 600 - 9.2) < 0.05


@pytest.mark.parametrize("name,bound,frames", [("Rope", 0.02, 240), ("Heavy Rope", 0.03, 600), ("Bridge", 0.03, 240)])
def test_hard_joints_hold(name, bound, frames):
    n, err = run(name, frames).joints()
    assert n > 0 and err < bound, f"{name}: joint error {err}"


def test_breakable_joints_fracture_under_load():
    assert run("Breakable", 1).joints()[0] == 10
    n = run("Breakable", 120).joints()[0]
    assert 0 < n < 10


def test_world_anchored_drag_joint_pulls_to_target():
    r = run("Ground", 120)
    L = lib()
    D = ctypes.POINTER(ctypes.c_double)
    local = (ctypes.c_double * 3)()
    origin = (ctypes.c_double * 3)(0, -10, 1)
    direction = (ctypes.c_double * 3)(0, 1, 0)
    hit = L.ref_pick(r.h, ctypes.cast(origin, D), ctypes.cast(direction, D), ctypes.cast(local, D))
    assert hit == 1, "ray picks the box"
    target = (ctypes.c_double * 3)(3, 2, 4)
    L.ref_add_world_joint(r.h, 1, ctypes.cast(target, D), ctypes.cast(local, D), 5000.0, 0.0)
    r.step(300)
    err = L.ref_world_joint_error(r.h, 1, ctypes.cast(target, D), ctypes.cast(local, D))
    assert err < 0.05, f"anchor error {err}"


@pytest.mark.parametrize("iterations", [1, 20])
def test_every_scene_stays_finite(iterations):
    for i in range(14):
        b = Ref(i, iterations).step(40).bodies()
        assert np.isfinite(b[:, :7]).all(), f"{SCENES[i]} @ {iterations}"


def test_is_deterministic():
    a = run("Soft Body", 120).bodies()
    b = run("Soft Body", 120).bodies()
    assert np.array_equal(a[:, :7], b[:, :7])


def test_mass_properties_follow_the_box_dimensions():
    out = (ctypes.c_double * 5)()
    size = (ctypes.c_double * 3)(1, 2, 3)
    lib().ref_box_mass(ctypes.cast(size, ctypes.POINTER(ctypes.c_double)), 2.0,
                       ctypes.cast(out, ctypes.POINTER(ctypes.c_double)))
    assert out[0] == 12
    assert list(out[1:4]) == [13, 10, 5]
    assert abs(out[4] - math.hypot(0.5, 1, 1.5)) < 1e-15
