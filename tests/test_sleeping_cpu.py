"""The oracle's sleeping restatement (src/gpu/worldSleeping.js, worldSleeping.wgsl, islandSleeping.wgsl,
sleepContact.wgsl), pinned by what does not retype its formulas: the quiet-time rule replayed in numpy from the bodies
and manifolds the oracle reports, the islands against scipy's connected components, the reference's own sleeping
contract (tests/backend-contract.js checkSleeping), and the behaviours its comments state (sleepers are fixed supports,
impacts wake the world or only the island hit, woken neighbours keep their mutual contacts in the impact step,
startAsleep sleeps only grounded groups, a body that may not sleep never does)."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from oracle_lib import OracleWorld

SPEED, DELAY = 0.03, 0.5
COL_MASS, COL_VEL, COL_ANG, COL_RADIUS = 19, slice(32, 35), slice(36, 39), 23


def cmd(index, kind, value):
    r = np.zeros(12, np.uint32)
    r[0], r[1] = index, kind
    r.view(np.float32)[4:7] = value
    return r


def _floor(avbd, gravity=-9.81):
    s = avbd.AvbdScene3D(gravity=gravity, iterations=10, timeStep=1 / 60)
    s.addBox([40, 1, 40], density=0, position=[0, -0.5, 0])
    return s


def _asleep(w):
    return w.sleep_state()[:, 2].astype(bool)


def _supported(w, i):
    """A manifold with body i above its partner (normal from b to a; up component past 0.3 on the right side)."""
    for m in w.manifolds():
        ny = m[4:7].view(np.float32)[1]
        if (m[0] == i and ny > 0.3) or (m[1] == i and ny < -0.3):
            return True
    return False


def test_resting_box_sleeps_once_quiet_and_supported_for_the_delay(avbd):
    """Whole-world mode: replay the quiet timer in numpy (f32 sums of dt over consecutive steps in which the box's
    speed + angular speed × bounding radius ≤ 0.03 and a contact holds it from below); the oracle puts the box to sleep
    in exactly the step the replayed timer reaches 0.5 s, with its mass and velocities zeroed."""
    s = _floor(avbd)
    s.addBox([1, 1, 1], position=[0, 0.55, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY)
    quiet, dt, slept = np.float32(0), np.float32(1 / 60), None
    for step in range(200):
        w.step()
        b = w.bodies()[1].astype(np.float32)
        asleep = bool(_asleep(w)[1])
        if asleep:
            slept = step
            break
        motion = np.float32(np.linalg.norm(b[COL_VEL]) + np.linalg.norm(b[COL_ANG]) * b[COL_RADIUS])
        ready = motion <= SPEED and _supported(w, 1)
        quiet = np.float32(quiet + dt) if ready else np.float32(0)
        assert quiet < DELAY, f"the replayed timer reached the delay at step {step} but the box is awake"
    assert slept is not None, "the box never slept"
    assert np.float32(quiet + dt) >= DELAY, f"slept at step {slept} with {quiet} s of quiet"
    b = w.bodies()[1]
    assert b[COL_MASS] == 0 and not b[COL_VEL].any() and not b[COL_ANG].any()
    st = w.sleep_state()[1]
    assert st[0].view(np.float32) == np.float32(1.0), "the sleep state keeps the awake mass"
    assert w.sleep_stats() == {"sleeping": 1, "wakeRequested": False}


def test_reference_sleeping_contract(avbd):
    """checkSleeping (backend-contract.js:441-467): a box resting on a floor sleeps within 180 steps; an impulse
    (1, 2, 0) wakes it and it moves up faster than 1 m/s after one step."""
    s = _floor(avbd)
    s.addBox([1, 1, 1], position=[0, 0.5, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY)
    w.step(180)
    assert _asleep(w)[1] and w.bodies()[1, COL_MASS] == 0
    w.apply_commands(cmd(1, 2, [1, 2, 0])[None])
    w.step()
    assert not _asleep(w)[1] and w.bodies()[1, COL_MASS] == 1 and w.bodies()[1, 33] > 1


def test_sleepers_are_fixed_supports(avbd):
    """Once asleep a box keeps its pose bit for bit, and its pair with the static floor is no longer collided."""
    s = _floor(avbd)
    s.addBox([1, 1, 1], position=[0, 0.5, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY)
    w.step(120)
    assert _asleep(w)[1]
    pose = w.bodies()[1, :8].copy()
    w.step(30)
    np.testing.assert_array_equal(w.bodies()[1, :8], pose)
    assert w.counts()["manifolds"] == 0


def _two_piles(avbd):
    """Pile A: two stacked boxes (1, 2) at x = 0; pile C: one box (3) at x = 10; a sphere (4) high above pile A."""
    s = _floor(avbd)
    s.addBox([1, 1, 1], position=[0, 0.5, 0])
    s.addBox([1, 1, 1], position=[0, 1.5, 0])
    s.addBox([1, 1, 1], position=[10, 0.5, 0])
    s.addSphere(0.5, position=[0, 40, 0])
    return s


def _until_hit(w, limit=400):
    """Step until the sphere's first manifold with box 2 exists; returns the sleep flags before that step."""
    for _ in range(limit):
        before = _asleep(w).copy()
        w.step()
        if any({int(m[0]), int(m[1])} == {2, 4} for m in w.manifolds()):
            return before
    raise AssertionError("the sphere never reached the pile")


@pytest.mark.parametrize("islands", [False, True])
def test_an_impact_wakes_the_world_or_only_the_island_hit(avbd, islands):
    """The sphere lands on sleeping pile A: in the step of impact pile A wakes (masses back before the solve, so the
    sphere does not stop on a fixed body) and, woken together, its two boxes keep their mutual contact in that very
    step (the pairs are found again after the wake). Whole-world mode wakes pile C too; island mode leaves it asleep."""
    w = OracleWorld([_two_piles(avbd)], precision=32)
    w.enable_sleeping(SPEED, DELAY, islands=islands)
    before = _until_hit(w)
    assert before[1] and before[2] and before[3], "the piles were not asleep before the impact"
    after = _asleep(w)
    assert not after[1] and not after[2], "the pile hit did not wake in the impact step"
    assert after[3] == islands, "pile C: whole-world mode wakes it, island mode does not"
    pairs = {(int(m[0]), int(m[1])) for m in w.manifolds()}
    assert (2, 1) in pairs, "the woken boxes lost their mutual contact in the impact step"
    assert w.bodies()[1, COL_MASS] == 1 and w.bodies()[2, COL_MASS] == 1


def test_islands_are_the_connected_components(avbd):
    """Island mode, all awake: after a step every dynamic body's root is the smallest index of its connected component
    over this step's manifolds between dynamic bodies and the joints (scipy's components); statics are their own."""
    s = _floor(avbd)
    rng = np.random.default_rng(3)
    for k in range(6):  # six short stacks of 1–3 boxes, some pairs of stacks touching
        x = 1.02 * k if k < 3 else 10 + 3 * k
        for h in range(1 + k % 3):
            s.addBox([1, 1, 1], position=[x, 0.5 + h, 0.01 * rng.standard_normal()])
    s.addBox([0.5, 0.5, 0.5], position=[-10, 4, 0])
    s.addBox([0.5, 0.5, 0.5], position=[-10, 5, 0])
    n = s.body_count
    s.addJoint(n - 2, n - 1, anchorA=(0, 0.5, 0), anchorB=(0, -0.5, 0))
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY, islands=True)
    w.step(3)
    mass = w.bodies()[:, COL_MASS]
    edges = [(int(m[0]), int(m[1])) for m in w.manifolds() if mass[m[0]] > 0 and mass[m[1]] > 0]
    edges.append((n - 2, n - 1))
    a, b = np.array(edges).T
    g = coo_matrix((np.ones(len(a)), (a, b)), shape=(n, n))
    _, label = connected_components(g, directed=False)
    parent = w.sleep_state()[:, 5]
    for i in range(n):
        expect = i if mass[i] <= 0 else min(j for j in range(n) if label[j] == label[i] and mass[j] > 0)
        assert parent[i] == expect, f"body {i}: root {parent[i]}, component minimum {expect}"
    assert len(set(parent[mass > 0])) >= 4


def test_start_asleep_sleeps_grounded_groups_only(avbd):
    """startAsleep (island mode): after the initializing step a still stack on the floor sleeps and a still box in
    the air does not."""
    s = _floor(avbd)
    s.addBox([1, 1, 1], position=[0, 0.5, 0])
    s.addBox([1, 1, 1], position=[0, 1.5, 0])
    s.addBox([1, 1, 1], position=[6, 5, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY, islands=True, start_asleep=True)
    w.step()
    assert _asleep(w).tolist() == [False, True, True, False]
    w.step(5)
    assert _asleep(w)[1] and _asleep(w)[2] and not _asleep(w)[3]


@pytest.mark.parametrize("islands", [False, True])
def test_a_body_that_may_not_sleep_never_does(avbd, islands):
    s = _floor(avbd)
    s.addBox([1, 1, 1], position=[0, 0.5, 0])
    s.addBox([1, 1, 1], position=[5, 0.5, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY, islands=islands, eligible=[1, 0, 1])
    w.step(150)
    assert _asleep(w).tolist() == [False, False, True]
    w.set_sleep_enabled([2], [0])  # eligibility edits wake every body
    assert not _asleep(w).any() and w.bodies()[2, COL_MASS] == 1
    w.step(150)
    assert not _asleep(w).any()


def test_waking_a_body_wakes_its_island_and_holds_it_for_a_step(avbd):
    """wakeBody (island mode): the held body's island wakes at the next step's start, another island stays asleep, and
    the held body does not fall asleep again in that step."""
    w = OracleWorld([_two_piles(avbd)], precision=32)
    w.enable_sleeping(SPEED, DELAY, islands=True)
    w.step(100)
    assert _asleep(w)[1:4].all()
    w.wake_bodies([1])
    w.step()
    assert _asleep(w)[1:4].tolist() == [False, False, True]


def test_disable_restores_every_mass(avbd):
    s = _floor(avbd)
    s.addBox([1, 1, 1], density=2, position=[0, 0.5, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY)
    w.step(120)
    assert w.bodies()[1, COL_MASS] == 0
    w.disable_sleeping()
    assert w.bodies()[1, COL_MASS] == 2
    w.step(30)
    assert w.bodies()[1, COL_MASS] == 2


def test_an_impulse_on_a_sleeper_wakes_the_world_before_it_applies(avbd):
    """Body commands wake every body first (the reference's prepareBodyCommands → wakeAll), so an impulse acts with
    the awake mass: Δv = impulse / mass."""
    s = _floor(avbd, gravity=0.0)
    s.addBox([1, 1, 1], density=2, position=[0, 0.5, 0])
    s.addBox([1, 1, 1], position=[5, 0.5, 0])
    w = OracleWorld([s], precision=32)
    w.enable_sleeping(SPEED, DELAY)
    w.step(40)
    assert _asleep(w)[1:].all()
    w.apply_commands(cmd(1, 2, [4, 0, 0])[None])
    assert not _asleep(w).any()
    np.testing.assert_allclose(w.bodies()[1, COL_VEL], [2, 0, 0], rtol=1e-6)
