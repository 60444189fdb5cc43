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
    return dict(pairs=len(pg), manifolds=len(mg))


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
