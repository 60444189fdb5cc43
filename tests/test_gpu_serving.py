"""The serving loop (libavbd_hip.so avbd_server_*): every served world ends bit-identical to the same world stepped
alone by its own solver, whatever else was served beside it and whichever solve kernel its size picks, also when the HBM
budget forces its state out to host memory and back between requests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = dict(colorCap=16, colorRounds=16, graph=True)


def _scenes(avbd):
    return [avbd.configs.pile_of(12, 301), avbd.configs.pile_of(64, 302, friction=0.3),
            avbd.configs.boxes_world(300, 303), avbd.configs.pile_of(500, 304, friction=0.8),
            avbd.configs.boxes_world(1500, 305)]


def _solo(avbd, scene, steps):
    w = avbd.createSolverPipeline("hip").createSolver(scene, OPTS)
    w.step(steps)
    out = w.readBodies()
    w.destroy()
    return out


def _serve(avbd, budget, max_active, seed, **solve):
    rng = np.random.default_rng(seed)
    server = avbd.serve.Server(hbmBudget=budget, maxActive=max_active, colorCap=16, colorRounds=16, **solve)
    ids = [server.addWorld(s) for s in _scenes(avbd)]
    totals = [0] * len(ids)
    requests = []
    for _ in range(24):
        k = int(rng.integers(0, len(ids)))
        n = int(rng.integers(1, 40))
        requests.append(server.submit(ids[k], n))
        totals[k] += n
    for r in requests:
        assert server.wait(r) > 0
    bodies = [server.readBodies(w) for w in ids]
    stats = server.stats()
    server.destroy()
    return totals, bodies, stats


# Every world by colour launches; or the two smallest by one workgroup each and the largest by the dataflow solve on a
# quarter of the device (three requests run at once)
SOLVES = [{}, dict(smallWorldBodies=100, largeWorldBodies=1000, largeWorldShare=4)]


@pytest.mark.parametrize("solve", SOLVES, ids=["colour-launches", "per-size-solves"])
def test_served_worlds_match_their_solo_runs(avbd, solve):
    totals, bodies, stats = _serve(avbd, budget=64 << 30, max_active=3, seed=1, **solve)
    assert stats["completed"] == stats["submitted"] == 24
    assert stats["evictions"] == 0
    assert stats["solve_aborts"] == 0
    assert 0 < stats["step_ms_p50"] <= stats["step_ms_p99"]
    for scene, n, got in zip(_scenes(avbd), totals, bodies):
        assert np.array_equal(got, _solo(avbd, scene, n)), "a served world differs from its solo run"


@pytest.mark.parametrize("solve", SOLVES, ids=["colour-launches", "per-size-solves"])
def test_eviction_to_host_and_back_is_exact(avbd, solve):
    # A budget of about two small worlds: most admissions evict an idle world and later restore it
    totals, bodies, stats = _serve(avbd, budget=20 << 20, max_active=2, seed=2, **solve)
    assert stats["evictions"] > 0 and stats["restores"] > 0
    for scene, n, got in zip(_scenes(avbd), totals, bodies):
        assert np.array_equal(got, _solo(avbd, scene, n)), "a world restored from host memory diverged"
