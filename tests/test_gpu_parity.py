"""Parity of the HIP step with the CPU oracle (FP32 mirror), through the C ABI.

Discrete outputs (pair sets, manifold pairs, feature keys, colourings) must be bit-exact. Continuous outputs (poses,
velocities, λ, penalties) must agree to 1e-5 relative with a per-quantity scale floor:
    |gpu - cpu| <= 1e-5 * max(|cpu|, floor), floors: poses 1 m / 1, velocities 1 m/s, λ and penalties 1.
"""
import numpy as np
import pytest

from oracle_lib import OracleWorld

pytestmark = pytest.mark.gpu

REL = 1e-5


def _pair_set(p):
    return {(int(a), int(b)) for a, b in np.asarray(p).reshape(-1, 2)}


def _manifold_map(mf, ct):
    """(a, b) -> (count|step word, normal+friction words, contact words of that manifold)."""
    out = {}
    for row in np.asarray(mf, np.uint32).reshape(-1, 8):
        a, b, first, cw = (int(x) for x in row[:4])
        out[(a, b)] = (cw, tuple(int(x) for x in row[4:8]), np.asarray(ct, np.uint32)[first:first + (cw & 15)])
    return out


def _close(got, want, floor, what):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    err = np.abs(got - want)
    bound = REL * np.maximum(np.abs(want), floor)
    bad = err > bound
    if bad.any():
        i = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: {int(bad.sum())} values beyond 1e-5 relative; first at {tuple(i)}: "
                             f"gpu {got[tuple(i)]!r} cpu {want[tuple(i)]!r}")


def compare_step(world, oracle, exact=True):
    """After one step of each: pairs, manifolds, contacts, colours, bodies."""
    pg, pc = _pair_set(world.readPairs()), _pair_set(oracle.pairs())
    assert pg == pc, f"pair sets differ: gpu-only {sorted(pg - pc)[:5]}, cpu-only {sorted(pc - pg)[:5]}"
    mf, ct = world.readContactRecords()
    mg = _manifold_map(mf, ct)
    mc = _manifold_map(oracle.manifolds(), oracle.contact_words())
    assert set(mg) == set(mc), "manifold pair sets differ"
    for key in mc:
        cwg, geog, ctg = mg[key]
        cwc, geoc, ctc = mc[key]
        assert cwg == cwc, f"manifold {key}: count/generation {cwg:#x} vs {cwc:#x}"
        assert np.array_equal(ctg[:, 3], ctc[:, 3]), f"manifold {key}: feature keys differ {ctg[:, 3]} vs {ctc[:, 3]}"
        if exact:
            assert geog == geoc, f"manifold {key}: normal/friction bits differ"
    colg, colc = world.readColors(), oracle.colors()
    assert np.array_equal(colg, colc), f"colours differ at {np.flatnonzero(colg != colc)[:8]}"
    bg = world.readBodies().astype(np.float64)
    bc = oracle.bodies()
    _close(bg[:, 0:3], bc[:, 0:3], 1.0, "positions")
    _close(bg[:, 4:8], bc[:, 4:8], 1.0, "rotations")
    _close(bg[:, 32:35], bc[:, 32:35], 1.0, "velocities")
    _close(bg[:, 36:39], bc[:, 36:39], 1.0, "angular velocities")
    # λ and penalties, contact by contact (keys already equal)
    for key in mc:
        ctg = mg[key][2].view(np.float32).astype(np.float64)
        ctc = mc[key][2].view(np.float32).astype(np.float64)
        _close(ctg[:, 12:15], ctc[:, 12:15], 1.0, f"lambda {key}")
        _close(ctg[:, 8:11], ctc[:, 8:11], 1.0, f"penalty {key}")
    if world.jointCount:
        jg, jc = world.readJoints().astype(np.float64), oracle.joints()
        for name, cols in (("joint penalties", np.r_[0:3, 4:7]), ("joint lambda", np.r_[8:11, 12:15])):
            _close(jg[:, cols], jc[:, cols], 1.0, name)
    assert world.readCounters()["pivotFallbacks"] == oracle.counts()["pivotFallbacks"], "FP64 pivot fallbacks differ"
    features = {int(f) >> 24 for f in np.asarray(ct, np.uint32).reshape(-1, 16)[:, 3] & 0x7FFFFFFF} if len(ct) else set()
    return dict(pairs=len(pg), manifolds=len(mg), features=features)


def _sorted_manifolds(mf, ct):
    """Manifolds sorted by (a, b) and, in that order, each one's contact words (vectorised, for large worlds)."""
    mf = np.asarray(mf, np.uint32).reshape(-1, 8)
    ct = np.asarray(ct, np.uint32).reshape(-1, 16)
    key = (mf[:, 0].astype(np.uint64) << np.uint64(32)) | mf[:, 1].astype(np.uint64)
    order = np.argsort(key, kind="stable")
    mf = mf[order]
    counts = (mf[:, 3] & 15).astype(np.int64)
    firsts = mf[:, 2].astype(np.int64)
    total = int(counts.sum())
    within = np.arange(total) - np.repeat(np.cumsum(counts) - counts, counts)
    return key[order], mf, ct[np.repeat(firsts, counts) + within]


def compare_step_arrays(world, oracle):
    """compare_step for worlds of any size: the same checks, vectorised (pairs, manifold pairs, count/generation words,
    normals and frictions, feature keys and colours bit-exact; poses, velocities, λ and penalties within 1e-5)."""
    def pair_keys(p):
        p = np.asarray(p, np.uint32).reshape(-1, 2)
        return np.sort((p[:, 0].astype(np.uint64) << np.uint64(32)) | p[:, 1].astype(np.uint64))
    pg, pc = pair_keys(world.readPairs()), pair_keys(oracle.pairs())
    assert np.array_equal(pg, pc), f"pair sets differ: {len(pg)} vs {len(pc)} pairs"
    mf, ct = world.readContactRecords()
    kg, mg, cg = _sorted_manifolds(mf, ct)
    kc, mc, cc = _sorted_manifolds(oracle.manifolds(), oracle.contact_words())
    assert np.array_equal(kg, kc), f"manifold pair sets differ: {len(kg)} vs {len(kc)}"
    bad = np.flatnonzero(mg[:, 3] != mc[:, 3])
    assert not len(bad), f"{len(bad)} manifolds differ in count/generation, first {mg[bad[0], :2]}"
    bad = np.flatnonzero((mg[:, 4:8] != mc[:, 4:8]).any(1))
    assert not len(bad), f"{len(bad)} manifolds differ in normal/friction bits, first {mg[bad[0], :2]}"
    assert np.array_equal(cg[:, 3], cc[:, 3]), f"{int((cg[:, 3] != cc[:, 3]).sum())} feature keys differ"
    colg, colc = world.readColors(), oracle.colors()
    assert np.array_equal(colg, colc), f"colours differ at {np.flatnonzero(colg != colc)[:8]}"
    bg = world.readBodies().astype(np.float64)
    bc = oracle.bodies()
    _close(bg[:, 0:3], bc[:, 0:3], 1.0, "positions")
    _close(bg[:, 4:8], bc[:, 4:8], 1.0, "rotations")
    _close(bg[:, 32:35], bc[:, 32:35], 1.0, "velocities")
    _close(bg[:, 36:39], bc[:, 36:39], 1.0, "angular velocities")
    fg, fc = cg.view(np.float32).astype(np.float64), cc.view(np.float32).astype(np.float64)
    _close(fg[:, 12:15], fc[:, 12:15], 1.0, "lambda")
    _close(fg[:, 8:11], fc[:, 8:11], 1.0, "penalty")
    assert world.readCounters()["pivotFallbacks"] == oracle.counts()["pivotFallbacks"], "FP64 pivot fallbacks differ"
    exact = bool(np.array_equal(bg[:, 0:8].astype(np.float32), bc[:, 0:8].astype(np.float32)))
    return dict(pairs=len(pg), manifolds=len(kg), contacts=len(cg), exact_poses=exact)


def _world(avbd, scene, cap=32, rounds=16, **opts):
    return avbd.createSolverPipeline("hip").createSolver(scene, dict(colorCap=cap, colorRounds=rounds, **opts))


def test_first_step_small_pile(avbd):
    scene = lambda: avbd.configs.box_pile(counts=(4, 4, 4))  # noqa: E731
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    world.step()
    oracle.step()
    stats = compare_step(world, oracle)
    assert stats["pairs"] > 0
    world.destroy()


def test_steps_stay_in_parity_1000_pile(avbd):
    """Every step of the first 60 of the 1,000-box pile, each from the same state (the GPU's) on both sides."""
    scene = avbd.configs.pile_1000
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    for s in range(60):
        state = world.readState()
        oracle.seed(state)
        world.step()
        oracle.step()
        compare_step(world, oracle)
    world.destroy()


def test_trajectory_bit_exact_small_pile(avbd):
    """Without re-seeding, 100 steps of a 64-box pile: the FP32 GPU and FP32 oracle stay bit-identical."""
    scene = lambda: avbd.configs.box_pile(counts=(4, 4, 4), seed=7)  # noqa: E731
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    world.step(100)
    oracle.step(100)
    got = world.readBodies()
    want = oracle.bodies().astype(np.float32)
    assert np.array_equal(got[:, 0:8], want[:, 0:8]), f"max diff {np.abs(got[:, :8] - want[:, :8]).max()}"
    world.destroy()


# --- H-PLOC broadphase: the same pair set as the grid (gpu-hploc.js:28-68 pattern) ------------------------------------


def _pair_sets_equal(a, b):
    return _pair_set(a.readPairs()) == _pair_set(b.readPairs())


def test_hploc_and_grid_give_the_same_pairs_and_trajectory(avbd):
    scene = lambda: avbd.configs.box_pile(counts=(12, 12, 12), seed=21)  # noqa: E731
    grid = _world(avbd, scene(), broadphase="grid")
    tree = _world(avbd, scene(), broadphase="hploc", bvh={"rebuildInterval": 16})
    for s in range(8):
        grid.step(10)
        tree.step(10)
        assert _pair_sets_equal(grid, tree), f"pair sets differ after {10 * (s + 1)} steps"
    assert np.array_equal(grid.readBodies(), tree.readBodies())
    grid.destroy()
    tree.destroy()


def test_hploc_seeded_step_matches_the_oracle(avbd):
    scene = avbd.configs.pile_1000
    world = _world(avbd, scene(), broadphase="hploc", bvh={"rebuildInterval": 7})
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    for s in range(30):
        state = world.readState()
        oracle.seed(state)
        world.step()
        oracle.step()
        compare_step(world, oracle)
    world.destroy()


@pytest.mark.parametrize("count", [1, 2, 15, 16, 17, 129])
def test_hploc_small_and_ragged_counts(avbd, count):
    """gpu-hploc.js:81-120: pair sets equal the grid's for 1..129 bodies, with a static ground and large statics."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
        s.addBox([1, 6, 1], density=0, position=[3, 3, 0])
        rng = np.random.default_rng(count)
        for k in range(count - 2 if count > 2 else count):
            s.addBox(list(rng.uniform(0.3, 1.2, 3)), position=list(rng.uniform([-2, 0.5, -2], [2, 4, 2])),
                     rotation=list(rng.normal(size=4)))
        return s
    grid = _world(avbd, scene(), broadphase="grid")
    tree = _world(avbd, scene(), broadphase="hploc", bvh={"rebuildInterval": 5})
    for s in range(12):
        grid.step(1)
        tree.step(1)
        assert _pair_sets_equal(grid, tree), f"step {s}"
    grid.destroy()
    tree.destroy()


def test_hploc_teleport_outside_bounds(avbd):
    """A body written far outside the tree's bounds between rebuilds is still found (refit covers it)."""
    scene = lambda: avbd.configs.box_pile(counts=(6, 6, 6), seed=4)  # noqa: E731
    grid = _world(avbd, scene(), broadphase="grid")
    tree = _world(avbd, scene(), broadphase="hploc", bvh={"rebuildInterval": 64})
    grid.step(5)
    tree.step(5)
    for w in (grid, tree):
        b = w.readBodies()
        b[10, 0:3] = [500.0, 3.0, 500.0]
        b[11, 0:3] = [500.3, 3.5, 500.2]   # overlapping the teleported body
        b[10, 32:35] = 0
        b[11, 32:35] = 0
        w.writeBodies(0, b)
    for s in range(6):
        grid.step(1)
        tree.step(1)
        assert _pair_sets_equal(grid, tree)
    assert (11, 10) in _pair_set(tree.readPairs())
    grid.destroy()
    tree.destroy()


def test_hploc_query_with_more_candidates_than_kept(avbd):
    """A 12 m plate resting on a 24 x 24 carpet of small boxes: its query finds several hundred leaves, more than the
    128 candidates a query keeps, so the candidate pass counts its pairs in a second walk and the write pass walks
    again; pair sets still equal the grid's step by step, and the trajectories are the same bits."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        s.addBox([30, 1, 30], density=0, position=[0, -0.5, 0])
        for x in range(24):
            for z in range(24):
                s.addBox([0.4, 0.4, 0.4], position=[-5.75 + 0.5 * x, 0.2, -5.75 + 0.5 * z])
        s.addBox([12, 0.4, 12], position=[0, 0.595, 0])
        return s
    grid = _world(avbd, scene(), broadphase="grid")
    tree = _world(avbd, scene(), broadphase="hploc", bvh={"rebuildInterval": 5})
    plate = 24 * 24 + 1
    for s in range(12):
        grid.step(1)
        tree.step(1)
        pairs = _pair_set(tree.readPairs())
        assert pairs == _pair_set(grid.readPairs()), f"step {s}"
        assert sum(1 for a, b in pairs if plate in (a, b)) > 128
    assert np.array_equal(grid.readBodies().view(np.uint32), tree.readBodies().view(np.uint32))
    grid.destroy()
    tree.destroy()


# --- batches: a world steps to the same bits alone and inside a batch ------------------------------------------------


def test_batch_worlds_match_their_solo_runs(avbd):
    makers = [lambda: avbd.configs.pile_of(40, 101, friction=0.3),
              lambda: avbd.configs.pile_of(9, 102, friction=0.8),
              lambda: avbd.configs.pile_of(300, 103, friction=0.5),
              lambda: avbd.configs.box_pile(counts=(3, 2, 4), seed=104)]
    pipe = avbd.createSolverPipeline("hip")
    batch = pipe.createBatch([m() for m in makers], dict(colorCap=32, colorRounds=16))
    solos = [pipe.createSolver(m(), dict(colorCap=32, colorRounds=16)) for m in makers]
    batch.step(60)
    for s in solos:
        s.step(60)
    got = batch.readBodies()
    for k, s in enumerate(solos):
        first = int(batch.world_first_body[k])
        want = s.readBodies()
        assert np.array_equal(got[first:first + len(want)], want), f"world {k} differs from its solo run"
    pipe.destroy()


def test_batch_matches_the_oracle_batch(avbd):
    makers = [lambda: avbd.configs.pile_of(30, 201), lambda: avbd.configs.pile_of(70, 202, friction=0.2)]
    batch = avbd.createSolverPipeline("hip").createBatch([m() for m in makers], dict(colorCap=32, colorRounds=16))
    oracle = OracleWorld([m() for m in makers], precision=32, color_cap=32, color_rounds=16)
    for _ in range(20):
        oracle.seed(batch.readState())
        batch.step()
        oracle.step()
        compare_step(batch, oracle)
    batch.destroy()


# --- capsules, spheres and joints (the jointed + contact mix at test size) -------------------------------------------


def _small_mix(avbd):
    s = avbd.configs.jointed_mix(n=6, spheres=30, capsules=20, area=3.0, capsule_grid=(4, 1, 5))
    s.addSphere(0.15, position=[2.6, 0.6, -2.6])  # a loose ball, so the run always has sphere contacts
    return s


def test_jointed_mix_steps_stay_in_parity(avbd):
    """Net of boxes on hard ball joints, static spheres and free capsules: every step of 240 from the GPU's state."""
    world = _world(avbd, _small_mix(avbd))
    oracle = OracleWorld([_small_mix(avbd)], precision=32, color_cap=32, color_rounds=16)
    features = set()
    for s in range(240):
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        features |= compare_step(world, oracle)["features"]
    assert {3, 4} <= features, f"capsule and sphere contacts expected, saw kinds {features}"
    world.destroy()


@pytest.mark.parametrize("capsule_arms", [False, True], ids=["reference-capsule-arms", "step-start-capsule-arms"])
def test_jointed_mix_trajectory_bit_exact(avbd, capsule_arms):
    """300 steps of the small mix, with the reference's capsule lever arms and with AVBD_FLAG_CAPSULE_STEP_START_ARMS."""
    def scene():
        m = _small_mix(avbd)
        if capsule_arms:
            m.params["flags"] |= avbd._native.FLAG_CAPSULE_STEP_START_ARMS
        return m
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    world.step(300)
    oracle.step(300)
    assert np.array_equal(world.readBodies()[:, 0:8], oracle.bodies()[:, 0:8].astype(np.float32))
    assert np.array_equal(world.readJoints()[:, 0:16], oracle.joints()[:, 0:16].astype(np.float32))
    world.destroy()


def test_pivot_fallback_to_fp64_matches_the_oracle(avbd):
    """A chain of 8 g boxes on hard joints warm-started at penalty 1e8: FP32 pivots hit the floor and the 6x6 solves are
    redone in FP64 on both sides, the same bodies each time."""
    def scene():
        s = avbd.AvbdScene3D(gravity=-9.81)
        for i in range(6):
            s.addBox([0.2, 0.2, 0.2], position=[0.2 * i + 0.1, 5, 0])
        s.addJoint(None, 0, anchorA=[0, 5, 0], anchorB=[-0.1, 0, 0])
        for i in range(5):
            s.addJoint(i, i + 1, anchorA=[0.1, 0, 0], anchorB=[-0.1, 0, 0])
        for rec, _ in s._joints:
            rec[0:3] = 1e8
        return s
    world = _world(avbd, scene())
    oracle = OracleWorld([scene()], precision=32, color_cap=32, color_rounds=16)
    total = 0
    for s in range(10):
        oracle.seed(world.readState())
        world.step()
        oracle.step()
        compare_step(world, oracle)
        total += world.readCounters()["pivotFallbacks"]
    assert total > 0
    world.destroy()
