"""Restitution around the step (src/gpu/restitution.js) on the GPU against the oracle: the same bits, with bodies
bouncing off a floor and off each other, in a pile, alone and in a batch, with graph replay; the reference's contract
(a sphere with restitution 1 rebounds faster than 2 m/s); coefficients from the scene builder, edits and handles."""
import numpy as np
import pytest

from oracle_lib import OracleWorld

pytestmark = pytest.mark.gpu


def bouncy_scene(avbd, seed, n=24):
    rng = np.random.default_rng(seed)
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([20, 1, 20], density=0, friction=0.5, position=[0, -0.5, 0])
    for k in range(n):
        pos = [float(rng.uniform(-3, 3)), float(1.0 + 0.9 * k), float(rng.uniform(-3, 3))]
        e = float(rng.choice([0.0, 0.3, 0.8, 1.0]))
        if k % 3 == 0:
            s.addSphere(float(rng.uniform(0.3, 0.6)), position=pos, restitution=e, velocity=[0, -4, 0])
        else:
            s.addBox(list(rng.uniform(0.5, 1.2, 3)), position=pos, rotation=list(rng.normal(size=4)),
                     restitution=e, friction=0.4)
    return s


def _same(world, oracle):
    got, want = world.readBodies(), oracle.bodies().astype(np.float32)
    for cols in (slice(0, 8), slice(32, 35), slice(36, 39)):
        assert np.array_equal(got[:, cols], want[:, cols]), f"columns {cols} differ: max {np.abs(got[:, cols] - want[:, cols]).max()}"


@pytest.mark.parametrize("graph", [False, True])
def test_bouncing_pile_bit_exact(avbd, graph):
    world = avbd.createSolverPipeline("hip").createSolver(bouncy_scene(avbd, 3),
                                                          dict(colorCap=32, colorRounds=16, graph=graph))
    oracle = OracleWorld([bouncy_scene(avbd, 3)], precision=32, color_cap=32, color_rounds=16)
    assert world.getRestitution().max() > 0
    for _ in range(8):
        world.step(15)
        oracle.step(15)
        _same(world, oracle)
    world.destroy()


def test_batch_worlds_bounce_like_their_solo_runs(avbd):
    scenes = [bouncy_scene(avbd, s, n=10) for s in (5, 6)]
    batch = avbd.createSolverPipeline("hip").createBatch(scenes, dict(colorCap=32, colorRounds=16))
    batch.step(90)
    got = batch.readBodies()
    n = scenes[0].body_count
    for k, s in enumerate((5, 6)):
        solo = avbd.createSolverPipeline("hip").createSolver(bouncy_scene(avbd, s, n=10), dict(colorCap=32, colorRounds=16))
        solo.step(90)
        assert np.array_equal(got[k * n:(k + 1) * n, 0:8], solo.readBodies()[:, 0:8]), f"world {k}"
        solo.destroy()
    batch.destroy()


def test_reference_restitution_contract(avbd):
    """backend-contract.js checkContacts: restitution 1, gravity 0, −3 m/s onto the floor: vy > 2 after 12 steps."""
    s = avbd.AvbdScene3D(gravity=0)
    s.addBox([10, 1, 4], density=0, position=[0, -0.5, 0])
    ball = s.addSphere(0.5, position=[0, 0.51, 0], velocity=[0, -3, 0], restitution=1)
    world = s.createSolver(avbd.createSolverPipeline("hip"))
    for _ in range(12):
        world.step()
    assert world.body(ball).readState()["linearVelocity"][1] > 2
    world.destroy()


def test_coefficients_from_edits_and_handles(avbd):
    """setRestitution, the handle's setRestitution and editBodies({restitution}) change the same coefficients the
    oracle gets, from the next step on; switching the last one off drops the bounce passes (the plain step again)."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
        for k in range(6):
            s.addSphere(0.4, position=[k - 2.5, 2.0 + 0.3 * k, 0])
        return s
    world = avbd.createSolverPipeline("hip").createSolver(scene(), dict(colorCap=32, colorRounds=16, graph=True))
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    world.step(10)
    oracle.step(10)
    world.setRestitution([1, 2], [0.9, 0.5])
    world.body(3).setRestitution(1.0)
    world.editBodies([{"body": 4, "restitution": 0.7, "linearVelocity": [0, -5, 0]}])
    oracle.seed(world.readState())
    oracle.set_restitution([1, 2, 3, 4], [0.9, 0.5, 1.0, 0.7])
    np.testing.assert_array_equal(world.getRestitution()[1:5], np.float32([0.9, 0.5, 1.0, 0.7]))
    for _ in range(4):
        world.step(15)
        oracle.step(15)
        _same(world, oracle)
    with pytest.raises(ValueError):
        world.setRestitution([1], [1.5])
    world.setRestitution([1, 2, 3, 4], [0, 0, 0, 0])
    oracle.set_restitution([1, 2, 3, 4], [0, 0, 0, 0])
    world.step(20)
    oracle.step(20)
    _same(world, oracle)
    world.destroy()
