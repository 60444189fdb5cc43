"""Convex hulls and terrain on the GPU against the oracle, through the C ABI: the hull narrowphase (one wavefront per
pair: hull/hull, hull/box, sphere/hull, capsule/hull) must give the oracle's pairs, feature keys, manifold words and
colours bit for bit and its poses, velocities and multipliers within 1e-5, step by step from the GPU's state and over a
whole trajectory without reseeding; a batch of worlds with hulls (their buffers concatenated and rebased) matches the
oracle's batch; and the reference's own checks: every one of 129 hull pairs makes its one contact, including a
partial last group (tests/gpu-contact-portability.js:42-72), and a body's shape code must name a hull record."""
import math

import numpy as np
import pytest

from oracle_lib import OracleWorld
from test_gpu_parity import compare_step

pytestmark = pytest.mark.gpu

OPTS = dict(colorCap=32, colorRounds=16)


def _box_points(w, h, d):
    return [c for x in (-w / 2, w / 2) for y in (-h / 2, h / 2) for z in (-d / 2, d / 2) for c in (x, y, z)]


def _quat(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return [*(a * math.sin(angle / 2)), math.cos(angle / 2)]


def hull_scene(avbd, seed=0, n=24):
    """Hulls of several kinds, boxes, spheres and capsules dropped on a heightfield beside a ground box."""
    rng = np.random.default_rng(seed)
    s = avbd.AvbdScene3D(gravity=-9.81)
    heights = [0.3 * math.sin(0.9 * x) * math.cos(0.7 * z) for z in range(6) for x in range(6)]
    s.addHeightfield(heights=heights, columns=6, rows=6, spacing=1.5, origin=[-4, 0, -4], friction=0.7)
    s.addBox([8, 1, 8], density=0, position=[8, -0.5, 0])
    tetra = avbd.convexHull([-0.4, -0.4, -0.4, 0.4, -0.4, -0.4, 0, 0.4, -0.4, 0, 0, 0.5])
    cloud = avbd.convexHull(rng.normal(size=(30, 3)) * [0.4, 0.3, 0.35])
    cube = avbd.convexHull(_box_points(0.7, 0.7, 0.7))
    for k in range(n):
        x = float(rng.uniform(-3, 11))
        z = float(rng.uniform(-3, 3))
        pos = [x, 1.2 + 0.8 * (k % 4), z]
        q = _quat(rng.normal(size=3), float(rng.uniform(0, 3)))
        kind = k % 6
        if kind == 0:
            s.addHullShape(tetra, position=pos, rotation=q)
        elif kind == 1:
            s.addHullShape(cloud, position=pos, rotation=q)
        elif kind == 2:
            s.addHullShape(cube, position=pos, rotation=q)
        elif kind == 3:
            s.addBox([0.6, 0.5, 0.7], position=pos, rotation=q)
        elif kind == 4:
            s.addSphere(0.3, position=pos)
        else:
            s.addCapsule(0.2, 1.0, position=pos, rotation=q)
    return s


def test_hull_scene_steps_match_the_oracle(avbd):
    world = avbd.createSolverPipeline("hip").createSolver(hull_scene(avbd), OPTS)
    oracle = OracleWorld([hull_scene(avbd)], precision=32, color_cap=32, color_rounds=16)
    kinds = set()
    for step in range(120):
        print(f"hull scene step {step}", flush=True)
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        compare_step(world, oracle)
        mf, ct = world.readContactRecords()
        kinds |= {int(k) >> 24 for k in np.asarray(ct, np.uint32).reshape(-1, 16)[:, 3] & 0x7FFFFFFF}
    assert {0, 1, 3, 4} <= kinds, f"face, sphere and capsule contacts all occurred: {sorted(kinds)}"
    world.destroy()


def test_hull_scene_trajectory_bit_exact(avbd):
    world = avbd.createSolverPipeline("hip").createSolver(hull_scene(avbd, seed=1), OPTS)
    oracle = OracleWorld([hull_scene(avbd, seed=1)], precision=32, color_cap=32, color_rounds=16)
    world.step(100)
    oracle.step(100)
    got = world.readBodies()
    want = oracle.bodies().astype(np.float32)
    assert np.array_equal(got[:, 0:8], want[:, 0:8]), f"max pose difference {np.abs(got[:, :8] - want[:, :8]).max()}"
    world.destroy()


def test_hull_batch_matches_the_oracle_batch(avbd):
    scenes = lambda: [hull_scene(avbd, seed=2, n=12), hull_scene(avbd, seed=3, n=18)]
    world = avbd.createSolverPipeline("hip").createBatch(scenes(), OPTS)
    oracle = OracleWorld(scenes(), precision=32, color_cap=32, color_rounds=16)
    for step in range(40):
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        compare_step(world, oracle)
    world.destroy()


def test_129_hull_pairs_each_make_a_contact(avbd):
    s = avbd.AvbdScene3D(gravity=0, iterations=1)
    cube = _box_points(2, 2, 2)
    for i in range(129):
        s.addHull(cube, density=0, position=[i * 6, 0, 0])
        s.addSphere(0.3, position=[i * 6 + 1.2, 0, 0])
    world = avbd.createSolverPipeline("hip").createSolver(s, OPTS)
    oracle = OracleWorld([s], precision=32, color_cap=32, color_rounds=16)
    for _ in range(2):
        world.step()
        oracle.step()
        c = world.readCounters()
        assert c["manifolds"] == 129 and c["contacts"] == 129 and not c["overflow"] and not c["clashes"], c
        mf, _ = world.readContactRecords()
        pairs = {(int(a), int(b)) for a, b in np.asarray(mf, np.uint32).reshape(-1, 8)[:, 0:2]}
        assert len(pairs) == 129, "every pair ran exactly once"
        compare_step(world, oracle)
    world.destroy()


def test_shape_codes_must_name_hull_records(avbd):
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([4, 1, 4], density=0, position=[0, -0.5, 0])
    h = s.addHull(_box_points(1, 1, 1), position=[0, 1, 0])
    world = avbd.createSolverPipeline("hip").createSolver(s, OPTS)
    rec = world.readBodies()
    for bad in (rec[h, 39] + 1.0, rec[h, 39] + 0.5, 3.0 + 1e6):
        r = rec.copy()
        r[h, 39] = bad
        with pytest.raises(avbd.AvbdError):
            world.writeBodies(0, r)
    world.step(5)  # the rejected writes left the world as it was
    assert world.readBodies()[h, 39] == rec[h, 39]
    world.destroy()
