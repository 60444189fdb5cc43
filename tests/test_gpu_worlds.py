"""The one-workgroup-per-world solve (solver "worlds": each world's iterations, colours, fused duals, joint duals and
velocities inside one workgroup, every world of a batch in one launch) computes exactly what the colour launches and
the dual pass compute: the same bits as the colour path for every workgroup size, alone, in batches of uneven worlds,
with joints, spheres and capsules, under graph replay and the adaptive policy; and step by step the oracle's (FP32
mirror) results from the GPU's state."""
import numpy as np
import pytest

from oracle_lib import OracleWorld
from test_gpu_parity import _manifold_map, compare_step

pytestmark = pytest.mark.gpu


def _solver(avbd, scene, solver, **opts):
    return avbd.createSolverPipeline("hip").createSolver(scene, dict(colorCap=32, colorRounds=16, solver=solver, **opts))


def _batch(avbd, scenes, solver, **opts):
    return avbd.createSolverPipeline("hip").createBatch(scenes, dict(dict(colorCap=32, colorRounds=16), solver=solver,
                                                                     **opts))


def _same_contacts(a, b):
    ga, gb = _manifold_map(*a.readContactRecords()), _manifold_map(*b.readContactRecords())
    assert ga.keys() == gb.keys()
    for pair, (cw, geo, ct) in ga.items():
        cwb, geob, ctb = gb[pair]
        assert cw == cwb and geo == geob and np.array_equal(ct, ctb), pair


@pytest.mark.parametrize("block", [128, 256, 512, 1024])
def test_single_world_same_bits_as_the_colour_launches(avbd, block):
    scene = avbd.configs.pile_1000
    print(f"worlds solve: 1,000-box pile, block {block}", flush=True)
    a = _solver(avbd, scene(), "colors")
    b = _solver(avbd, scene(), "worlds", worldBlock=block)
    for _ in range(3):
        a.step(30)
        b.step(30)
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
    _same_contacts(a, b)
    a.destroy()
    b.destroy()


def test_heavy_body_same_bits_as_the_colour_launches(avbd):
    """A 12 m plate on a 24 x 24 carpet of small boxes (the plate a heavy body with hundreds of entries) in one
    workgroup: 40 steps bit-identical to the colour launches."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        s.addBox([30, 1, 30], density=0, position=[0, -0.5, 0])
        for x in range(24):
            for z in range(24):
                s.addBox([0.4, 0.4, 0.4], position=[-5.75 + 0.5 * x, 0.2, -5.75 + 0.5 * z])
        s.addBox([12, 0.4, 12], position=[0, 0.595, 0])
        return s
    a = _solver(avbd, scene(), "colors")
    b = _solver(avbd, scene(), "worlds", worldBlock=1024)
    for _ in range(4):
        a.step(10)
        b.step(10)
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
    a.destroy()
    b.destroy()


def test_single_world_steps_match_the_oracle(avbd):
    scene = avbd.configs.pile_1000
    world = _solver(avbd, scene(), "worlds")
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    world.step(60)
    for s in range(30):
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        compare_step(world, oracle)
    world.destroy()


def _uneven_piles(avbd, count=48, seed=5):
    """Box piles of log-uniform body counts in [8, 600] with per-world friction, like the batched workload's."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(8), np.log(600), count)).astype(int)
    return [avbd.configs.pile_of(int(n), 300 + k, friction=float(rng.uniform(0.2, 0.9))) for k, n in enumerate(sizes)]


@pytest.mark.parametrize("graph", [False, True])
def test_uneven_batch_same_bits_as_the_colour_launches(avbd, graph):
    """48 piles of 8..600 boxes, 120 steps from colour cap 4 with a synchronous adapt() every 10 steps (the policy
    then changes the cap at the same steps on both sides): bodies and contacts bit-identical to the colour launches."""
    print(f"worlds solve: 48 uneven piles, graph {graph}", flush=True)
    a = _batch(avbd, _uneven_piles(avbd), "colors", graph=graph, colorCap=4)
    b = _batch(avbd, _uneven_piles(avbd), "worlds", graph=graph, colorCap=4)
    assert a.solverKind() == 1 and b.solverKind() == 3
    caps = set()
    for _ in range(12):
        a.step(10)
        b.step(10)
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
        a.adapt()
        b.adapt()
        assert a.getColoring() == b.getColoring()
        caps.add(b.getColoring()["colorCap"])
    assert len(caps) > 1, "the colour cap changed during the run"
    _same_contacts(a, b)
    a.destroy()
    b.destroy()


def test_auto_picks_worlds_for_batches_of_small_worlds(avbd):
    pipe = avbd.createSolverPipeline("hip")
    batch = pipe.createBatch([avbd.configs.pile_of(n, 20 + n) for n in (8, 50, 200)], dict(colorCap=32))
    solo = pipe.createSolver(avbd.configs.pile_of(100, 7), dict(colorCap=32))
    big = pipe.createBatch([avbd.configs.pile_of(300, 8), avbd.configs.pile_of(20, 9)], dict(colorCap=32))
    assert batch.solverKind() == 3, "a batch of small worlds is solved one workgroup per world"
    assert solo.solverKind() == 1 and big.solverKind() == 1
    pipe.destroy()


def test_uneven_batch_steps_match_the_oracle(avbd):
    makers = [lambda: avbd.configs.pile_of(30, 201), lambda: avbd.configs.pile_of(300, 202, friction=0.2),
              lambda: avbd.configs.pile_of(9, 203, friction=0.9), lambda: avbd.configs.box_pile(counts=(5, 4, 5), seed=204)]
    batch = _batch(avbd, [m() for m in makers], "worlds")
    oracle = OracleWorld([m() for m in makers], precision=32, color_cap=32, color_rounds=16)
    batch.step(40)
    for _ in range(20):
        oracle.seed(batch.readState())
        batch.step()
        oracle.step()
        compare_step(batch, oracle)
    batch.destroy()


def _mix_batch(avbd):
    """Two small jointed nets over static spheres with free capsules and loose balls, and two box piles, all at 20
    iterations (a batch shares its iteration count)."""
    def mix(seed):
        s = avbd.configs.jointed_mix(n=6, spheres=30, capsules=20, area=3.0, capsule_grid=(4, 1, 5), seed=seed)
        s.addSphere(0.15, position=[2.6, 0.6, -2.6])
        return s
    return [mix(11), avbd.configs.pile_of(50, 12, iterations=20), mix(13), avbd.configs.pile_of(120, 14, iterations=20)]


def test_jointed_batch_same_bits_and_oracle(avbd):
    """Joints, spheres and capsules: 150 steps bit-identical to the colour launches (bodies and joints), then steps
    from the GPU's state against the oracle."""
    a = _batch(avbd, _mix_batch(avbd), "colors")
    b = _batch(avbd, _mix_batch(avbd), "worlds")
    for _ in range(3):
        a.step(50)
        b.step(50)
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
        assert np.array_equal(a.readJoints().view(np.uint32), b.readJoints().view(np.uint32))
    a.destroy()
    oracle = OracleWorld(_mix_batch(avbd), precision=32, color_cap=32, color_rounds=16)
    features = set()
    for _ in range(20):
        oracle.seed(b.readState())
        b.step()
        oracle.step()
        features |= compare_step(b, oracle)["features"]
    assert {3, 4} <= features, f"capsule and sphere contacts expected, saw kinds {features}"
    b.destroy()


def test_pivot_fallback_counts_match(avbd):
    """Worlds whose stiff joint chains (8 g boxes on hard joints warm-started at penalty 1e8) floor FP32 pivots: the
    FP64 redo happens for the same bodies on both paths, with the same bits."""
    def chain(offset):
        s = avbd.AvbdScene3D(gravity=-9.81)
        for i in range(6):
            s.addBox([0.2, 0.2, 0.2], position=[0.2 * i + 0.1 + offset, 5, 0])
        s.addJoint(None, 0, anchorA=[offset, 5, 0], anchorB=[-0.1, 0, 0])
        for i in range(5):
            s.addJoint(i, i + 1, anchorA=[0.1, 0, 0], anchorB=[-0.1, 0, 0])
        for rec, _ in s._joints:
            rec[0:3] = 1e8
        return s
    a = _batch(avbd, [chain(0.0), chain(1.0)], "colors")
    b = _batch(avbd, [chain(0.0), chain(1.0)], "worlds")
    total = 0
    for _ in range(10):
        a.step()
        b.step()
        fa, fb = a.readCounters()["pivotFallbacks"], b.readCounters()["pivotFallbacks"]
        assert fa == fb
        total += fb
        assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
        assert np.array_equal(a.readJoints().view(np.uint32), b.readJoints().view(np.uint32))
    assert total > 0
    a.destroy()
    b.destroy()


def test_batch_hploc_equals_grid(avbd):
    """A batch stepped with the H-PLOC broadphase (a tree per world; auto for batches of 10,000 bodies or more) finds
    the grid's pairs every step and steps to the same bits."""
    scenes = lambda: _uneven_piles(avbd, count=24, seed=9)  # noqa: E731
    a = _batch(avbd, scenes(), "colors", broadphase="grid")
    b = _batch(avbd, scenes(), "colors", broadphase="hploc", bvh={"rebuildInterval": 7})
    for _ in range(40):
        a.step()
        b.step()
        pa = {tuple(p) for p in a.readPairs().tolist()}
        pb = {tuple(p) for p in b.readPairs().tolist()}
        assert pa == pb
    assert np.array_equal(a.readBodies().view(np.uint32), b.readBodies().view(np.uint32))
    a.destroy()
    b.destroy()
