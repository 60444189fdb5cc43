"""The long-horizon statistics (babylon-avbd_amd/stats.py) against closed forms and the oracle (CPU only)."""
import math

import numpy as np

from oracle_lib import OracleWorld


def _box(avbd, rotation=(0, 0, 0, 1), pos=(0, 2, 0), size=(1, 2, 3)):
    s = avbd.AvbdScene3D()
    s.addBox(list(size), position=list(pos), rotation=list(rotation))
    return s.packed()[0]


def test_lowest_point_of_rotated_boxes_spheres_and_capsules(avbd):
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        b = _box(avbd, q, (0.3, 2.0, -1.0), (1, 2, 3))
        # brute force over the eight corners
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        corners = np.array([[sx * 0.5, sy * 1.0, sz * 1.5] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        want = (corners @ R.T + [0.3, 2.0, -1.0])[:, 1].min()
        assert abs(avbd.stats.lowest_points(b, (0, 1, 0))[0] - want) < 1e-5
    s = avbd.AvbdScene3D()
    s.addSphere(0.5, position=[0, 3, 0])
    s.addCapsule(0.2, 1.2, position=[0, 3, 0], rotation=[0, 0, math.sin(math.pi / 8), math.cos(math.pi / 8)])
    low = avbd.stats.lowest_points(s.packed()[0], (0, 1, 0))
    assert abs(low[0] - 2.5) < 1e-6
    assert abs(low[1] - (3 - 0.4 * math.cos(math.pi / 4) - 0.2)) < 1e-6  # tilted 45°: segment end lowered by cos


def test_energy_closed_form(avbd):
    s = avbd.AvbdScene3D()
    s.addBox([100, 1, 100], density=0, position=[0, -0.5, 0])
    s.addBox([1, 2, 3], position=[0, 4, 0], velocity=[1, 2, 2], angularVelocity=[0, 0, 3])
    b = s.packed()[0]
    kin, pot = avbd.stats.energy(b, (0, 1, 0), -9.81)
    m = 6.0
    Iz = m * (1 + 4) / 12
    assert abs(kin - (0.5 * m * 9 + 0.5 * Iz * 9)) < 1e-4
    assert abs(pot - m * 9.81 * 4) < 1e-3


def test_resting_box_rests_one_margin_deep(avbd):
    """A unit box settled on the ground in the oracle: its contacts sit about one 0.01 m margin deep, its lowest
    point about 0.01 m below the ground's top, and its kinetic energy vanishes."""
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
    s.addBox([1, 1, 1], position=[0, 0.55, 0])
    o = OracleWorld([s], precision=32)
    o.step(120)
    st = avbd.stats.world_statistics(o.bodies(), o.manifolds(), o.contact_words(), (0, 1, 0), -9.81)
    assert st["contacts"] == 4
    assert abs(st["penetration"]["max"] - 0.01) < 2e-3 and abs(st["penetration"]["mean"] - 0.01) < 2e-3
    assert abs(st["lowest_point"]["min"] + 0.01) < 2e-3
    assert st["kinetic_energy"] < 1e-6
