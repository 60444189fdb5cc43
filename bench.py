"""Benchmark: body-steps/s of the AVBD 3D step on the 100,000-box world (10 iterations) on one MI355X.

One "step" is one full AVBD step of the world (collision, adjacency, colouring, warm starts, 10 × (primal per colour,
dual), velocities) with the bodies resident in HBM. Prints one JSON line (the driver's contract) with a `roofline`
object for the dominant kernel (the per-colour primal) and a `cpu_baseline` object (the CPU restatement of the same
step, timed on a bounded sample on this host).
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def primal_bytes_per_iteration(world):
    """Algorithmic bytes one iteration's primal passes move: per dynamic body its 160 B record; per adjacency entry
    the 4 B id, the 32 B manifold and the partner's 64 B hot span; per contact point read by a body, 64 B."""
    bodies = world.readBodies()
    dynamic = int((bodies[:, 19] > 0).sum())
    start, lst = world.readAdjacency()
    mf, _ = world.readContactRecords()
    J = world.jointCount
    counts = (mf[:, 3] & 15).astype(np.int64)
    entries = len(lst)
    is_contact = lst >= J
    points = int(counts[lst[is_contact] - J].sum()) if is_contact.any() else 0
    joints = int((~is_contact).sum())
    return dynamic * 160 + entries * 100 + points * 64 + joints * 128, dict(
        dynamic=dynamic, entries=entries, contact_reads=points, joint_entries=joints)


def cpu_baseline(scene_fn, state, first_step, max_steps, budget_s, colour_cap):
    """The sequential CPU restatement (FP32) of the same step, started from the GPU's state at the start of the timed
    window and run for a bounded sample (whole steps until `budget_s` seconds or `max_steps`), one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleWorld  # the checker, timed here as the CPU baseline

    scene = scene_fn()
    oracle = OracleWorld([scene], precision=32, color_cap=colour_cap, color_rounds=16)
    oracle.seed(state)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        oracle.step(1)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    n = scene.body_count
    return dict(value=n * steps / dt, unit="body-steps/s", cores=1, kind="port",
                sample=f"steps {first_step + 1}-{first_step + steps} of the same {n}-body world from the GPU's state "
                       f"(the timed window's start), sequential FP32 restatement of the step (grid broadphase, "
                       f"colour-ordered Gauss-Seidel), {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--workload", default="large", choices=["large", "pile1000"])
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--broadphase", default="auto", choices=["auto", "grid", "hploc"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the AVBD step here is single-GPU; "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import torch

    avbd = importlib.import_module("babylon-avbd_amd")
    if args.workload == "large":
        scene_fn = avbd.configs.large_world
        workload = "100k-box pile: 100,000 boxes, half-extents U(0.25,1.0), jittered 50x40x50 lattice at 2.2 m, ground"
    else:
        scene_fn = avbd.configs.pile_1000
        workload = "1,000 unit boxes, 10x10x10 lattice at 1.1 m, ground"
    scene = scene_fn()
    n_bodies = scene.body_count
    cap = 16
    world = avbd.createSolverPipeline("hip").createSolver(
        scene, {"colorCap": cap, "colorRounds": 16, "graph": not args.no_graph, "adaptive": True,
                "broadphase": args.broadphase})
    print(f"# world: {n_bodies} bodies, {world.deviceBytes() / 2**20:.0f} MiB on device", file=sys.stderr, flush=True)

    # Warm-up: the pile falls and lands; the colour/capacity policy adapts from counters read back asynchronously
    done = 0
    while done < args.warmup:
        k = min(50, args.warmup - done)
        world.step(k)
        world.synchronize()
        done += k
        print(f"# warm-up {done}: {world.readCounters()}", file=sys.stderr, flush=True)
    counters = world.readCounters()
    state0 = None if args.no_cpu else world.readState()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.step(args.steps)
    world.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_bodies * args.steps / elapsed
    counters = world.readCounters()

    # Roofline of the dominant kernel: one profiled step (HIP events per launch on the engine's stream)
    prof = world.profileStep()
    bytes_iter, parts = primal_bytes_per_iteration(world)
    iters = scene.params["iterations"]
    primal_ms_iter = prof["primal"] / iters
    achieved = bytes_iter / (primal_ms_iter * 1e-3) / 1e9
    launches = prof["primalLaunches"] / iters
    print(f"# profile: {json.dumps(prof)}; primal bytes/iteration {bytes_iter} {parts}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene_fn, state0, args.warmup, args.cpu_steps, args.cpu_budget, 64)

    line = {
        "metric": "body-steps/s (100k-box pile, 10 iters)",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": workload, "bodies": n_bodies, "iterations": iters, "dt": scene.params["dt"],
                   "broadphase": {"auto": "auto (H-PLOC BVH: one world of >= 10,000 bodies)", "grid": "uniform grid",
                                  "hploc": "H-PLOC BVH"}[args.broadphase], "steps_timed": f"{args.warmup + 1}-{args.warmup + args.steps}",
                   "counters": counters},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_primal (one colour's bodies)",
            "bytes_per_iteration": bytes_iter,
            "launches_per_iteration": launches,
            "avg_launch_ms": prof["primal"] / max(prof["primalLaunches"], 1),
        },
        "phases_ms": prof,
        "cpu_baseline": cpu,
    }
    world.destroy()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
