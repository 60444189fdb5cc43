"""Sleeping (src/gpu/worldSleeping.js and its shaders) on the GPU against the oracle: the same bits for the bodies and
every sleep state word, step after step, in whole-world and island mode, with and without graph replay, on the grid
and on the BVH broadphase (whose pair search runs again, gated, when an impact wakes sleepers); start asleep, wake
requests, eligibility edits and disabling; two worlds in a batch sleeping and waking on their own; the reference's
sleeping contract (backend-contract.js checkSleeping) through the facade."""
import numpy as np
import pytest

from oracle_lib import OracleWorld, set_threads

pytestmark = pytest.mark.gpu

OPTS = dict(colorCap=32, colorRounds=16)


def piles(avbd, seed, stacks=4, far=True, dropper=True):
    """Short stacks of boxes settling on a floor, a far pile, and a sphere dropped onto the first stack later."""
    rng = np.random.default_rng(seed)
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([60, 1, 60], density=0, friction=0.6, position=[0, -0.5, 0])
    for k in range(stacks):
        for h in range(1 + k % 3):
            s.addBox([1, 1, 1], position=[2.5 * k + 0.02 * rng.standard_normal(), 0.5 + 1.0 * h,
                                          0.02 * rng.standard_normal()])
    if far:
        s.addBox([1.2, 0.6, 1.0], position=[25, 0.3, 0])
    if dropper:
        s.addSphere(0.45, position=[0.1, 24, 0.05])
    return s


def _same(world, oracle, label=""):
    got, want = world.readBodies(), oracle.bodies().astype(np.float32)
    for cols in (slice(0, 8), slice(16, 20), slice(32, 35), slice(36, 39)):
        assert np.array_equal(got[:, cols], want[:, cols]), \
            f"{label} columns {cols} differ: max {np.abs(got[:, cols] - want[:, cols]).max()}"
    states, _ = world.readSleep()
    ref = oracle.sleep_state()
    bad = np.nonzero((states != ref).any(axis=1))[0]
    assert not len(bad), f"{label} sleep state of bodies {bad[:8].tolist()} differs: {states[bad[0]]} vs {ref[bad[0]]}"
    for k in range(world.worldCount):
        assert world.readSleepStats(k) == oracle.sleep_stats(k), f"{label} world {k} stats"


@pytest.mark.parametrize("islands", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_piles_sleep_and_wake_bit_exact(avbd, islands, graph):
    world = avbd.createSolverPipeline("hip").createSolver(piles(avbd, 1), dict(OPTS, graph=graph))
    oracle = OracleWorld([piles(avbd, 1)], precision=32, color_cap=32, color_rounds=16)
    world.enableSleeping({"islands": islands})
    oracle.enable_sleeping(islands=islands)
    slept = woke = False
    for chunk in range(24):
        world.step(10)
        oracle.step(10)
        _same(world, oracle, f"step {10 * (chunk + 1)}")
        asleep = oracle.sleep_state()[:, 2].astype(bool)
        slept = slept or asleep[1:-1].sum() >= 4
        woke = woke or (slept and not asleep[1])
    assert slept and woke, "the run did not cover both falling asleep and an impact waking a pile"
    world.destroy()


def test_bvh_broadphase_refresh_bit_exact(avbd):
    """12k bodies (the BVH broadphase, islands by default): two-box stacks on a lattice settle and sleep; spheres
    dropped on a few of them wake those islands, whose pairs are found again in the impact step."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        s.addBox([400, 1, 400], density=0, position=[0, -0.5, 0])
        n = 0
        for i in range(78):
            for j in range(77):
                x, z = 2.2 * i - 85, 2.2 * j - 85
                s.addBox([1, 1, 1], position=[x, 0.5, z])
                s.addBox([0.8, 0.8, 0.8], position=[x, 1.4, z])
                n += 2
        for k in range(6):
            s.addSphere(0.4, position=[2.2 * (10 * k) - 85, 14 + 2 * k, -85])
        return s
    set_threads(16)
    try:
        world = avbd.createSolverPipeline("hip").createSolver(scene(), OPTS)
        oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
        assert world.bodyCount > 12000
        world.enableSleeping()
        oracle.enable_sleeping(islands=True)
        assert world.sleepIslands
        woken = False
        for chunk in range(12):
            world.step(10)
            oracle.step(10)
            _same(world, oracle, f"step {10 * (chunk + 1)}")
            asleep = oracle.sleep_state()[:, 2].astype(bool)
            woken = woken or (asleep.sum() > 10000 and not asleep[1:-6].all())
        assert woken, "no island woke from an impact"
        world.destroy()
    finally:
        set_threads(1)


def test_start_asleep_and_requests_bit_exact(avbd):
    """startAsleep (one initializing step), then wake requests, eligibility edits, wake-all and disabling, in the same
    order on both sides."""
    world = avbd.createSolverPipeline("hip").createSolver(piles(avbd, 2, dropper=False), OPTS)
    oracle = OracleWorld([piles(avbd, 2, dropper=False)], precision=32, color_cap=32, color_rounds=16)
    world.enableSleeping({"islands": True, "startAsleep": True})  # takes its initializing step
    oracle.enable_sleeping(islands=True, start_asleep=True)
    oracle.step()
    _same(world, oracle, "start")
    assert oracle.sleep_state()[1:-1, 2].any()
    ops = [("wake", [2]), ("step", 3), ("eligible", [3], False), ("step", 40), ("wakeall",), ("step", 40),
           ("eligible", [3], True), ("step", 60), ("wake", [1, 4]), ("step", 2)]
    for op in ops:
        if op[0] == "step":
            world.step(op[1])
            oracle.step(op[1])
        elif op[0] == "wake":
            world.wakeBodies(op[1])
            oracle.wake_bodies(op[1])
        elif op[0] == "eligible":
            for i in op[1]:
                world.setSleepEnabled(i, op[2])
            oracle.set_sleep_enabled(op[1], [int(op[2])] * len(op[1]))
        else:
            world.wakeAll()
            oracle.wake_all()
        _same(world, oracle, str(op))
    world.disableSleeping()
    oracle.disable_sleeping()
    np.testing.assert_array_equal(world.readBodies()[:, 19], oracle.bodies()[:, 19].astype(np.float32))
    world.step(20)
    oracle.step(20)
    got, want = world.readBodies(), oracle.bodies().astype(np.float32)
    assert np.array_equal(got[:, 0:8], want[:, 0:8])
    world.destroy()


def test_batched_worlds_sleep_and_wake_on_their_own(avbd):
    """Whole-world mode in a batch of two: an impact in world 0 wakes world 0 only."""
    scenes = [piles(avbd, 4), piles(avbd, 5, dropper=False)]
    batch = avbd.createSolverPipeline("hip").createBatch(scenes, OPTS)
    oracle = OracleWorld([piles(avbd, 4), piles(avbd, 5, dropper=False)], precision=32, color_cap=32, color_rounds=16)
    batch.enableSleeping({"islands": False})
    oracle.enable_sleeping(islands=False)
    for chunk in range(24):
        batch.step(10)
        oracle.step(10)
        _same(batch, oracle, f"step {10 * (chunk + 1)}")
    assert batch.readSleepStats(1)["sleeping"] > 0
    batch.destroy()


def test_reference_sleeping_contract(avbd):
    """checkSleeping (backend-contract.js:441-467) through the facade: allowSleep true on a body enables sleeping; the
    box sleeps within 180 steps; an impulse wakes it and it moves up faster than 1 m/s after one step."""
    s = avbd.AvbdScene3D()
    s.addBox([10, 1, 10], density=0, position=[0, -0.5, 0])
    box = s.addBox([1, 1, 1], position=[0, 0.5, 0], allowSleep=True)
    world = s.createSolver(avbd.createSolverPipeline("hip"), {"spatialSort": False})
    assert world.sleeping
    for _ in range(180):
        world.step()
    body = world.body(box)
    state = body.readState()
    assert state["sleeping"] and state["effectiveMass"] == 0 and state["mass"] == 1
    assert world.readCounters()["sleeping"] == 1
    body.applyImpulse([1, 2, 0])
    world.step()
    awake = body.readState()
    assert not awake["sleeping"] and awake["linearVelocity"][1] > 1
    world.destroy()
