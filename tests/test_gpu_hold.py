"""Held steps and the adapt policy, through the C ABI.

A step whose pair storage overflows or whose colouring clashes is held: poses, velocities and joints keep their bits,
the counters report the failure, and adapt() recovers (the reference's "H-PLOC overflow/clashes holds motion and
recovers", tests/gpu-hploc.js:234-286, on the grid, H-PLOC, dataflow and workgroup-per-world paths). Then the adapt path against the oracle:
a pile stepped with a synchronous adapt() after every step, starting from a colour cap, round count and pair capacity
too small for it, so that it clashes, overflows, grows the cap, grows storage and shrinks the cap again (recolouring the
bodies above the new cap); every step is compared with the oracle run at the same cap, rounds and capacity from the
same state.
"""
import numpy as np
import pytest

from oracle_lib import OracleWorld
from test_gpu_parity import compare_step

pytestmark = pytest.mark.gpu

# The words of a body record the reference's test requires to be held: pos + friction, rot, vel + prevUp, angVel + shape
HELD_WORDS = np.r_[0:8, 32:40]


def _reference_scene(avbd, jointed=True):
    """tests/gpu-hploc.js:236-243: an 80 x 80 x 2 ground, two unit boxes stacked on it (the lower one sliding at
    0.1 m/s, the upper one spinning at 0.1 rad/s) and a joint between them (stiffness 1000 both ways); z up. Plus a
    third box resting beside them: capacity here counts pairs (each with room for 8 points), where the reference's
    contactCapacity = 1 counts points, so the scene needs two pairs (the jointed pair never collides) to overflow one."""
    scene = avbd.AvbdScene3D(gravity=-10.0, up=(0, 0, 1))
    scene.addBox([80, 80, 2], density=0, friction=0.7, position=[0, 0, -1])
    a = scene.addBox([1, 1, 1], density=1, friction=0.7, position=[0, 0, 0.5], velocity=[0.1, 0, 0])
    b = scene.addBox([1, 1, 1], density=1, friction=0.7, position=[0, 0, 1.5], angularVelocity=[0, 0, 0.1])
    scene.addBox([1, 1, 1], density=1, friction=0.7, position=[3, 0, 0.5])
    if jointed:
        scene.addJoint(a, b, [0, 0, 0.5], [0, 0, -0.5], linearStiffness=1000, angularStiffness=1000)
    return scene


PATHS = [
    ("grid", dict(broadphase="grid"), True),
    ("hploc", dict(broadphase="hploc"), True),
    ("dataflow", dict(broadphase="hploc", solver="dataflow"), False),  # the dataflow solve takes worlds without joints
    ("worlds", dict(broadphase="grid", solver="worlds"), True),
]


@pytest.mark.parametrize("failure", ["overflow", "clashes"])
@pytest.mark.parametrize("name,opts,jointed", PATHS, ids=[p[0] for p in PATHS])
def test_failed_step_holds_motion_and_recovers(avbd, failure, name, opts, jointed):
    options = dict(colorRounds=32, **opts)
    if failure == "overflow":
        options["capacity"] = {"pairs": 1}
    world = avbd.createSolverPipeline("hip").createSolver(_reference_scene(avbd, jointed), options)
    print(f"hold test: path {name}, failure {failure}", flush=True)
    before = world.readBodies().copy()
    joints = world.readJoints().copy()
    if failure == "clashes":
        world.setColoring(1, 2)
    world.step()
    counters = world.readCounters()
    held = world.readBodies()
    assert counters[failure] > 0, "incomplete topology is reported"
    assert np.array_equal(held[:, HELD_WORDS].view(np.uint32), before[:, HELD_WORDS].view(np.uint32)), \
        "invalid step holds pose and velocity"
    assert np.array_equal(world.readJoints().view(np.uint32), joints.view(np.uint32)), "joint forces are held"
    assert world.adapt(), "adapt() changes the capacity or the colour cap"
    world.step()
    recovered = world.readCounters()
    assert not recovered["overflow"] and not recovered["clashes"], "capacity recovery restores a complete step"
    after = world.readBodies()
    assert np.isfinite(after).all(), "finite recovered state"
    assert not np.array_equal(after[:, 0:8], before[:, 0:8]), "the recovered step moves the bodies"
    world.destroy()


def _low_pile(avbd):
    """A 6 x 3 x 6 pile of unit boxes dropped from just above the ground: contacts from the first steps on."""
    return avbd.configs.box_pile(counts=(6, 3, 6), centre=(0.0, 1.75, 0.0), seed=11)


@pytest.mark.parametrize("broadphase", ["grid", "hploc"])
def test_adapt_path_matches_the_oracle(avbd, broadphase):
    """60 steps with a synchronous adapt() after each, from colour cap 2, 2 rounds and 24 pair slots; at step 20 the
    cap is raised to 24 (adapt shrinks it again after three quiet reads) and at step 40 lowered to 3 (bodies of colours
    3 and up are coloured again). Held steps keep the bodies on both sides; solved steps match the oracle (pairs,
    keys, colours bit-exact; poses, velocities, λ within 1e-5) run with the cap, rounds and capacity the world had for
    that step."""
    world = avbd.createSolverPipeline("hip").createSolver(
        _low_pile(avbd), dict(colorCap=2, colorRounds=2, capacity={"pairs": 24}, broadphase=broadphase))
    oracle = OracleWorld([_low_pile(avbd)], precision=32, color_cap=2, color_rounds=2)
    seen = dict(overflow=0, clashes=0, solved=0, cap_up=0, cap_down=0, storage_up=0)
    for s in range(60):
        if s == 20:
            world.setColoring(24, 8)
        if s == 40:
            world.setColoring(3, 8)
        cfg = world.getColoring()
        state = world.readState()
        oracle.set_coloring(cfg["colorCap"], cfg["colorRounds"])
        oracle.set_pair_capacity(cfg["pairCapacity"])
        oracle.seed(state)
        world.step()
        oracle.step()
        c, o = world.readCounters(), oracle.counts()
        assert bool(c["overflow"]) == bool(o["overflow"]), f"step {s}: overflow {c['overflow']} vs {o['overflow']}"
        assert bool(c["clashes"]) == bool(o["clashes"]), f"step {s}: clashes {c['clashes']} vs {o['clashes']}"
        if c["overflow"] or c["clashes"]:
            seen["overflow"] += bool(c["overflow"])
            seen["clashes"] += bool(c["clashes"])
            got = world.readBodies()
            assert np.array_equal(got[:, HELD_WORDS].view(np.uint32), state["bodies"][:, HELD_WORDS].view(np.uint32))
            assert np.array_equal(oracle.bodies()[:, HELD_WORDS].astype(np.float32), state["bodies"][:, HELD_WORDS])
        else:
            seen["solved"] += 1
            compare_step(world, oracle)
        world.adapt()
        new = world.getColoring()
        seen["cap_up"] += new["colorCap"] > cfg["colorCap"]
        seen["cap_down"] += new["colorCap"] < cfg["colorCap"]
        seen["storage_up"] += new["pairCapacity"] > cfg["pairCapacity"]
    print(f"adapt path ({broadphase}): {seen}", flush=True)
    for k in ("overflow", "clashes", "cap_up", "cap_down", "storage_up"):
        assert seen[k] > 0, f"the run never saw {k}: {seen}"
    assert seen["solved"] >= 40
    world.destroy()
