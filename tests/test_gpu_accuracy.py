"""Accuracy of the FP32 HIP step against the FP64 restatement of the same step (the oracle at precision 64, which
takes the FP32 mirror's pairs, contacts and colours and solves in double precision), through the C ABI.

* Single step: from the GPU's state at several points of the 1,000-box pile's run (falling, landing, settling,
  settled) and of a small jointed mix, one step on each side; the error of poses, velocities, contact multipliers λ
  and joint multipliers, per quantity (max and 99th percentile, relative with the scale floors of the parity bar:
  1 m, 1, 1 m/s, 1 rad/s, 1 N). The FP32 HIP step equals the FP32 restatement bit for bit (tests/test_gpu_parity.py),
  so this is the FP32 rounding of the step itself.
* Long horizon: the full runs (300 steps of the 1,000-box pile, 300 steps of a 2,000-box world with the large
  world's distributions, 240 steps of the small mix) on the GPU in FP32 and in the FP64 restatement, separately (the
  trajectories are chaotic and part within a few steps), compared by rest-height, energy and penetration statistics;
  the mean rest height is allowed the spread that FP64 runs from starts moved by 1 µm show.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib
from oracle_lib import OracleWorld
from test_gpu_parity import _sorted_manifolds

pytestmark = pytest.mark.gpu

OPTS = dict(colorCap=32, colorRounds=16)


@pytest.fixture
def oracle_threads():
    oracle_lib.set_threads(min(16, len(os.sched_getaffinity(0))))
    yield
    oracle_lib.set_threads(1)


def _rel(got, want, floor):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    if not got.size:
        return dict(max=0.0, p99=0.0)
    r = (np.abs(got - want) / np.maximum(np.abs(want), floor)).reshape(-1)
    return dict(max=float(r.max()), p99=float(np.percentile(r, 99)))


def _single_step_errors(world, scene, cap):
    state = world.readState()
    fp64 = OracleWorld([scene], precision=64, color_cap=cap, color_rounds=16)
    fp64.seed(state)
    world.step(1)
    fp64.step(1)
    bg = world.readBodies()
    bd = fp64.bodies()
    dyn = bd[:, 19] > 0
    out = {
        "positions": _rel(bg[dyn, 0:3], bd[dyn, 0:3], 1.0),
        "rotations": _rel(bg[dyn, 4:8], bd[dyn, 4:8], 1.0),
        "velocities": _rel(bg[dyn, 32:35], bd[dyn, 32:35], 1.0),
        "angular_velocities": _rel(bg[dyn, 36:39], bd[dyn, 36:39], 1.0),
    }
    mf, ct = world.readContactRecords()
    kg, _, cg = _sorted_manifolds(mf, ct)
    _, vals = fp64.contacts()
    words = fp64.contact_words()  # the FP32 mirror's records (same order as vals)
    kd, md, cd_words = _sorted_manifolds(fp64.manifolds(), words)
    assert np.array_equal(kg, kd), "the FP64 run uses the same manifolds"
    # The FP64 values in the same sorted order: sort rows of vals the way _sorted_manifolds sorted the words
    mfd = np.asarray(fp64.manifolds(), np.uint32).reshape(-1, 8)
    order = np.argsort((mfd[:, 0].astype(np.uint64) << np.uint64(32)) | mfd[:, 1].astype(np.uint64), kind="stable")
    counts = (mfd[order, 3] & 15).astype(np.int64)
    firsts = mfd[order, 2].astype(np.int64)
    within = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts)
    vd = vals[np.repeat(firsts, counts) + within]
    out["contact_lambda"] = _rel(cg.view(np.float32)[:, 12:15], vd[:, 12:15], 1.0)
    if world.jointCount:
        jg, jd = world.readJoints(), fp64.joints()
        out["joint_lambda"] = _rel(jg[:, np.r_[8:11, 12:15]], jd[:, np.r_[8:11, 12:15]], 1.0)
    return out


def _mix(avbd):
    return avbd.configs.jointed_mix(n=8, spheres=60, capsules=60, area=8.0, capsule_grid=(6, 2, 5))


def test_single_step_error_against_fp64(avbd, oracle_threads):
    report = {}
    cases = [("1,000-box pile", avbd.configs.pile_1000, (0, 59, 60, 60, 120)),  # steps 1, 61, 122, 183, 304
             ("jointed mix (8x8 net, 60 spheres, 60 capsules)", lambda: _mix(avbd), (0, 100, 100))]
    for name, scene, gaps in cases:
        world = avbd.createSolverPipeline("hip").createSolver(scene(), OPTS)
        for gap in gaps:
            world.step(gap)
            step = int(world.readSteps()[0]) + 1
            print(f"accuracy: {name}, step {step}", flush=True)
            report[f"{name}, step {step}"] = _single_step_errors(world, scene(), OPTS["colorCap"])
        world.destroy()
    print("ACCURACY " + json.dumps(report), flush=True)
    for case, errs in report.items():
        # bounds about 3x the largest error measured on an MI355X
        assert errs["positions"]["max"] <= 1e-5, (case, errs["positions"])
        assert errs["rotations"]["max"] <= 1.5e-4, (case, errs["rotations"])
        assert errs["velocities"]["max"] <= 1e-3, (case, errs["velocities"])
        assert errs["angular_velocities"]["max"] <= 3e-2, (case, errs["angular_velocities"])
        assert errs["contact_lambda"]["p99"] <= 1e-2, (case, errs["contact_lambda"])
        if "joint_lambda" in errs:
            assert errs["joint_lambda"]["max"] <= 1e-2, (case, errs["joint_lambda"])


def _stats(avbd, world_or_oracle, scene, gpu):
    if gpu:
        mf, ct = world_or_oracle.readContactRecords()
        bodies = world_or_oracle.readBodies()
    else:
        mf, ct = world_or_oracle.manifolds(), world_or_oracle.contact_words()
        bodies = world_or_oracle.bodies()
    p = scene.params
    return avbd.stats.world_statistics(bodies, mf, ct, scene.up, p["gravity"])


def _perturbed_fp64_heights(avbd, scene, steps, runs=3, size=1e-6):
    """Mean centre heights of FP64 runs whose dynamic bodies start moved by up to `size` m (seeded): how far the
    trajectory's own chaos spreads the statistic."""
    out = []
    for seed in range(runs):
        sc = scene()
        o = OracleWorld([sc], precision=64, color_cap=OPTS["colorCap"], color_rounds=16)
        b = o.bodies().astype(np.float32)
        dyn = b[:, 19] > 0
        b[dyn, 0:3] += np.random.default_rng(100 + seed).uniform(-size, size, (int(dyn.sum()), 3)).astype(np.float32)
        o.seed(dict(bodies=b, joints=o.joints().astype(np.float32), refs=o.refs(), colors=o.colors(),
                    manifolds=o.manifolds(), contacts=o.contact_words(), steps=np.array([0], np.uint32)))
        o.step(steps)
        out.append(_stats(avbd, o, sc, False)["centre_height"]["mean"])
    return out


def test_long_horizon_statistics_against_fp64(avbd, oracle_threads):
    report = {}
    cases = [("1,000-box pile, 300 steps", avbd.configs.pile_1000, 300),
             ("2,000 boxes of the large world's distributions, 300 steps",
              lambda: avbd.configs.boxes_world(2000, avbd.configs.SEED + 11), 300),
             ("jointed mix (8x8 net, 60 spheres, 60 capsules), 240 steps", lambda: _mix(avbd), 240)]
    for name, scene, steps in cases:
        print(f"long horizon: {name}", flush=True)
        world = avbd.createSolverPipeline("hip").createSolver(scene(), OPTS)
        world.step(steps)
        gpu = _stats(avbd, world, scene(), True)
        world.destroy()
        fp64 = OracleWorld([scene()], precision=64, color_cap=OPTS["colorCap"], color_rounds=16)
        fp64.step(steps)
        cpu = _stats(avbd, fp64, scene(), False)
        spread = max(abs(h - cpu["centre_height"]["mean"]) for h in _perturbed_fp64_heights(avbd, scene, steps))
        report[name] = {"fp32_gpu": gpu, "fp64_cpu": cpu, "fp64_mean_height_spread_1um": spread}
    print("LONG_HORIZON " + json.dumps(report), flush=True)
    for name, r in report.items():
        g, c = r["fp32_gpu"], r["fp64_cpu"]
        assert abs(g["fallen"] - c["fallen"]) <= 0.01 * g["bodies"] + 2, name
        # within 5 %, or within twice what moving the FP64 run's bodies by 1 µm at the start does (the mix's capsules
        # roll off the net or stay on it, a chaotic outcome that no precision fixes)
        bound = max(0.05 * max(1.0, c["centre_height"]["mean"]), 2 * r["fp64_mean_height_spread_1um"])
        assert abs(g["centre_height"]["mean"] - c["centre_height"]["mean"]) <= bound, name
        assert abs(g["potential_energy"] - c["potential_energy"]) <= 0.05 * abs(c["potential_energy"]) + 1.0, name
        # (the minimum lowest point belongs to a body sliding off the ground's edge or to a hard landing; the
        # first percentile is the resting population's)
        assert g["lowest_point"]["p1"] > -0.05 and c["lowest_point"]["p1"] > -0.05, name
        assert g["penetration"]["p99"] < 0.05 and c["penetration"]["p99"] < 0.05, name
        assert abs(g["penetration"]["mean"] - c["penetration"]["mean"]) <= 0.2 * c["penetration"]["mean"], name
        assert g["kinetic_energy"] <= 2 * c["kinetic_energy"] + 0.05 * abs(c["potential_energy"]), name
