"""The benchmark workloads themselves under the oracle (FP32 mirror), through the C ABI:

* the single large world (configs.large_world: 100,000 boxes, random half-extents and orientations, initial
  velocities) with the bench's options — auto broadphase (H-PLOC at this size), dataflow solve, hipGraph replay,
  asynchronous adaptive colour cap and storage — its first steps and three steps of the settled pile at step 300, each
  step from the GPU's state on both sides, at the colour cap the adaptive policy chose for it;
* worlds served through the serving loop (avbd_server_*), among them a 1,500-box world with the large world's
  distributions, against the oracle run for the same number of steps.
"""
import os

import numpy as np
import pytest

import oracle_lib
from oracle_lib import OracleWorld
from test_gpu_parity import compare_step_arrays

pytestmark = pytest.mark.gpu

BENCH_OPTS = {"colorCap": 16, "colorRounds": 16, "graph": True, "adaptive": True, "solver": "dataflow"}


@pytest.fixture
def oracle_threads():
    oracle_lib.set_threads(min(16, len(os.sched_getaffinity(0))))
    yield
    oracle_lib.set_threads(1)


def _step_both(world, oracle):
    state = world.readState()
    world.step(1)
    cfg = world.getColoring()  # the cap the adaptive policy set before this step
    oracle.set_coloring(cfg["colorCap"], cfg["colorRounds"])
    oracle.seed(state)
    oracle.step()
    return compare_step_arrays(world, oracle)


def test_large_world_matches_the_oracle(avbd, oracle_threads):
    scene = avbd.configs.large_world
    world = avbd.createSolverPipeline("hip").createSolver(scene(), BENCH_OPTS)
    oracle = OracleWorld([scene()], precision=32, color_cap=16, color_rounds=16)
    for s in range(2):
        print(f"large world: step {s + 1}", flush=True)
        stats = _step_both(world, oracle)
        print(f"  {stats}", flush=True)
    done = 2
    while done < 300:
        world.step(50 if done + 50 <= 300 else 300 - done)
        world.synchronize()
        done = min(done + 50, 300)
    for s in range(3):
        print(f"large world: step {301 + s}", flush=True)
        stats = _step_both(world, oracle)
        print(f"  {stats}", flush=True)
        assert stats["pairs"] > 200000 and stats["contacts"] > 300000, "the settled pile is in contact"
    c = world.readCounters()
    assert c["solveAborts"] == 0 and c["overflow"] == 0
    world.destroy()


def test_served_worlds_match_the_oracle(avbd):
    scenes = lambda: [avbd.configs.boxes_world(1500, 305), avbd.configs.pile_of(200, 306, friction=0.3),  # noqa: E731
                      avbd.configs.boxes_world(300, 307)]
    rng = np.random.default_rng(9)
    server = avbd.serve.Server(hbmBudget=64 << 30, maxActive=2, colorCap=16, colorRounds=16)
    ids = [server.addWorld(s) for s in scenes()]
    totals = [0] * len(ids)
    requests = []
    for _ in range(12):
        k = int(rng.integers(0, len(ids)))
        n = int(rng.integers(1, 25))
        requests.append(server.submit(ids[k], n))
        totals[k] += n
    for r in requests:
        server.wait(r)
    served = [server.readBodies(w) for w in ids]
    server.destroy()
    for k, (scene, n, got) in enumerate(zip(scenes(), totals, served)):
        print(f"served world {k}: {scene.body_count} bodies, {n} steps", flush=True)
        oracle = OracleWorld([scene], precision=32, color_cap=16, color_rounds=16)
        oracle.step(n)
        want = oracle.bodies().astype(np.float32)
        assert np.array_equal(got[:, 0:8], want[:, 0:8]), \
            f"served world {k}: poses differ from the oracle's, max {np.abs(got[:, 0:8] - want[:, 0:8]).max()}"
        assert np.array_equal(got[:, 32:39], want[:, 32:39]), f"served world {k}: velocities differ"
