"""The dataflow solve (every iteration's colours and duals in one launch, each body waiting only for its neighbours)
computes exactly what the colour-by-colour launches compute: bit-identical to the oracle (FP32 mirror) and to the
launch path, alone, in a batch and under graph replay; and it never gives up a wait."""
import numpy as np
import pytest

from oracle_lib import OracleWorld
from test_gpu_parity import _manifold_map, compare_step

pytestmark = pytest.mark.gpu


def _world(avbd, scene, solver="dataflow", **opts):
    return avbd.createSolverPipeline("hip").createSolver(scene, dict(colorCap=32, colorRounds=16, solver=solver, **opts))


def test_trajectory_bit_exact_small_pile(avbd):
    scene = lambda: avbd.configs.box_pile(counts=(4, 4, 4), seed=7)  # noqa: E731
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    world.step(100)
    oracle.step(100)
    got = world.readBodies()
    want = oracle.bodies().astype(np.float32)
    assert world.readCounters()["solveAborts"] == 0
    assert np.array_equal(got[:, 0:8], want[:, 0:8]), f"max diff {np.abs(got[:, :8] - want[:, :8]).max()}"
    world.destroy()


def test_steps_stay_in_parity_1000_pile(avbd):
    scene = avbd.configs.pile_1000
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    for s in range(40):
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        compare_step(world, oracle)
        assert world.readCounters()["solveAborts"] == 0
    world.destroy()


@pytest.mark.parametrize("graph", [False, True])
def test_same_bits_as_the_colour_launches(avbd, graph):
    scene = avbd.configs.pile_1000
    a = _world(avbd, scene(), solver="colors", graph=graph)
    b = _world(avbd, scene(), solver="dataflow", graph=graph)
    for _ in range(4):
        a.step(30)
        b.step(30)
        assert np.array_equal(a.readBodies(), b.readBodies())
        assert b.readCounters()["solveAborts"] == 0
    # The manifold list's order is not part of the result (it follows the compaction's block order): by pair
    ga, gb = _manifold_map(*a.readContactRecords()), _manifold_map(*b.readContactRecords())
    assert ga.keys() == gb.keys()
    for pair, (cw, geo, ct) in ga.items():
        cwb, geob, ctb = gb[pair]
        assert cw == cwb and geo == geob and np.array_equal(ct, ctb), pair
    a.destroy()
    b.destroy()


def test_heavy_body_same_bits_as_the_colour_launches(avbd):
    """A 12 m plate resting on a 24 x 24 carpet of small boxes: the plate's adjacency list runs to hundreds of entries
    (a heavy body, many chunks and item rounds, its fused dual reading its first round from the stash and the rest from
    memory), each carpet box has the plate and a few neighbours. 40 steps bit-identical to the colour launches, and the
    oracle's step from the GPU's state."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        s.addBox([30, 1, 30], density=0, position=[0, -0.5, 0])
        for x in range(24):
            for z in range(24):
                s.addBox([0.4, 0.4, 0.4], position=[-5.75 + 0.5 * x, 0.2, -5.75 + 0.5 * z])
        s.addBox([12, 0.4, 12], position=[0, 0.595, 0])
        return s
    a = _world(avbd, scene(), solver="colors")
    b = _world(avbd, scene(), solver="dataflow")
    assert b.solverKind() == 2
    for _ in range(4):
        a.step(10)
        b.step(10)
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
        assert b.readCounters()["solveAborts"] == 0
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    oracle.seed(b.readState())
    b.step()
    oracle.step()
    compare_step(b, oracle)
    a.destroy()
    b.destroy()


def test_batch_worlds_match_their_solo_runs(avbd):
    scenes = [avbd.configs.box_pile(counts=(4, 4, 4), seed=s) for s in (1, 2, 3)]
    batch = avbd.createSolverPipeline("hip").createBatch(scenes, dict(colorCap=32, colorRounds=16, solver="dataflow"))
    batch.step(40)
    got = batch.readBodies()
    n = scenes[0].body_count
    for k, s in enumerate((1, 2, 3)):
        solo = _world(avbd, avbd.configs.box_pile(counts=(4, 4, 4), seed=s), solver="colors")
        solo.step(40)
        assert np.array_equal(got[k * n:(k + 1) * n], solo.readBodies()), f"world {k}"
        solo.destroy()
    batch.destroy()


def test_jointed_mix_same_bits_as_colour_launches(avbd):
    """Ball joints, spheres and capsules: the dataflow solve (joint duals fused into the unit that finishes each joint)
    steps to the same bits as the colour launches and the dual pass, bodies and joint records."""
    mix = lambda: avbd.configs.jointed_mix(n=6, spheres=30, capsules=20, area=3.0, capsule_grid=(4, 1, 5))  # noqa: E731
    a = _world(avbd, mix(), solver="dataflow")
    b = _world(avbd, mix(), solver="colors")
    assert a.solverKind() == 2
    for _ in range(5):
        a.step(30)
        b.step(30)
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
        assert np.array_equal(a.readJoints().view(np.uint32), b.readJoints().view(np.uint32))
        assert a.readCounters()["solveAborts"] == 0
    a.destroy()
    b.destroy()


def test_pivot_fallback_matches_colour_launches(avbd):
    """A chain of 8 g boxes on hard joints warm-started at penalty 1e8: the FP64 redo happens for the same bodies in
    the dataflow solve as in the colour launches, with the same bits."""
    def chain():
        s = avbd.AvbdScene3D(gravity=-9.81)
        for i in range(6):
            s.addBox([0.2, 0.2, 0.2], position=[0.2 * i + 0.1, 5, 0])
        s.addJoint(None, 0, anchorA=[0, 5, 0], anchorB=[-0.1, 0, 0])
        for i in range(5):
            s.addJoint(i, i + 1, anchorA=[0.1, 0, 0], anchorB=[-0.1, 0, 0])
        for rec, _ in s._joints:
            rec[0:3] = 1e8
        return s
    a = _world(avbd, chain(), solver="dataflow")
    b = _world(avbd, chain(), solver="colors")
    total = 0
    for _ in range(10):
        a.step()
        b.step()
        fa, fb = a.readCounters()["pivotFallbacks"], b.readCounters()["pivotFallbacks"]
        assert fa == fb
        total += fa
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
        assert np.array_equal(a.readJoints().view(np.uint32), b.readJoints().view(np.uint32))
    assert total > 0
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("mode", ["listed", "launches"])
def test_colouring_rounds_listed_and_launched_agree(avbd, mode):
    """The Jones–Plassmann rounds over the pending list in one block give the colours the per-round launches give
    (checked against the oracle on a pile whose colouring keeps changing), and the oracle's bodies."""
    scene = lambda: avbd.configs.box_pile(counts=(6, 6, 6), seed=5)  # noqa: E731
    world = _world(avbd, scene(), solver="colors", coloringRounds=mode, graph=True)
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    for _ in range(6):
        world.step(20)
        oracle.step(20)
        assert np.array_equal(world.readColors(), oracle.colors()), mode
        got, want = world.readBodies(), oracle.bodies().astype(np.float32)
        assert np.array_equal(got[:, 0:8], want[:, 0:8]), mode  # poses, bit for bit
        assert np.array_equal(got[:, 32:39], want[:, 32:39]), mode  # velocities
    world.destroy()


def test_a_wait_given_up_switches_to_colour_launches(avbd, monkeypatch):
    """A dataflow wait that gives up (forced from the host: the world is made with a spin limit of 0, so the first
    unmet poll gives up) is reported by solveAborts; reading the counters switches the world to colour launches, whose
    next steps match the oracle from the GPU's state."""
    scene = avbd.configs.pile_1000
    landed = _world(avbd, scene())
    landed.step(60)
    state = landed.readState()
    landed.destroy()
    monkeypatch.setenv("AVBD_DATAFLOW_SPIN", "0")
    world = _world(avbd, scene())
    monkeypatch.delenv("AVBD_DATAFLOW_SPIN")
    assert world.solverKind() == 2
    world.writeState(state)
    world.step()
    assert world.readCounters()["solveAborts"] > 0, "the forced give-up is reported"
    assert world.solverKind() == 1, "the world falls back to colour launches"
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    for _ in range(5):
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        compare_step(world, oracle)
        assert world.readCounters()["solveAborts"] == 0
    world.destroy()
