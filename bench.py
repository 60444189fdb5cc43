"""Benchmark: body-steps/s of the AVBD 3D step on the 100,000-box world (10 iterations) on one MI355X.

One "step" is one full AVBD step of the world (collision, adjacency, colouring, warm starts, 10 × (primal per colour,
dual), velocities) with the bodies resident in HBM. Prints one JSON line (the driver's contract) with a `roofline`
object for the dominant kernel (the solve: one dataflow launch for every iteration, or with --solver colors the
per-colour primal) and a `cpu_baseline` object (the CPU restatement of the same step, timed on a bounded sample on
this host's cores and on one thread).
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


SETTLE_STEPS = 300  # the 100k-box pile lands and settles in these steps; every timed window starts after them


def traffic_profile(kernel):
    return os.path.join("profiles", f"{kernel}_100k_boxes_settled_counters.json")


def measured_counters(kernel):
    """The committed rocprofv3 counters of `kernel` on the settled 100k-box world (the state every timed window of
    this bench starts from): HBM bytes per launch from FETCH_SIZE/WRITE_SIZE calibrated against a copy kernel of known
    size, and the wave, VALU and wait counters; None when absent."""
    try:
        with open(os.path.join(ROOT, traffic_profile(kernel))) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def dual_bytes_per_iteration(world):
    """Algorithmic bytes one iteration's dual pass moves: per manifold its 32 B record and both bodies' 64 B pose
    spans; per contact point its 4 B list entry, its 64 B record read and 28 B written (force, penalty, key); per
    joint its 128 B record read and 64 B written."""
    mf, _ = world.readContactRecords()
    counts = (mf[:, 3] & 15).astype(np.int64)
    manifolds = int((counts > 0).sum())
    points = int(counts.sum())
    J = world.jointCount
    return manifolds * (32 + 128) + points * (4 + 64 + 28) + J * 192, dict(manifolds=manifolds, points=points)


def summarize_counters(directory, kernel, per_step):
    """Per-step sums and per-dispatch means of the counters in a rocprofv3 --pmc CSV under `directory`, over the
    dispatches of kernels whose name contains `kernel` (the last whole step of them when per_step > 0)."""
    import csv
    import glob
    paths = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not paths:
        raise SystemExit(f"no counter_collection.csv under {directory}")
    values = {}
    for path in paths:
        with open(path) as f:
            for r in csv.DictReader(f):
                if kernel not in r["Kernel_Name"]:
                    continue
                d = values.setdefault(int(r["Dispatch_Id"]), {})
                d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    ids = sorted(values)
    if per_step > 0:
        ids = ids[-per_step:]
    names = sorted({c for d in ids for c in values[d]})
    total = {c: sum(values[d].get(c, 0.0) for d in ids) for c in names}
    return {"kernel": kernel, "dispatches": len(ids), "sum": total,
            "mean_per_dispatch": {c: total[c] / max(len(ids), 1) for c in names}}


def host_threads():
    """Host cores this process may use: the OpenMP thread count the host sets, else the CPUs it is pinned to."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(scene_fn, state, first_step, max_steps, budget_s, colour_cap):
    """Two CPU figures for the same world, each started from the GPU's state at the start of the timed window and run
    for a bounded sample (whole steps until `budget_s` seconds or `max_steps`):

    * the baseline: the FP64 restatement of the reference's own CPU path (CpuSolver3D.step, cpuSolver3D.ts:144-280,
      with cpuBroadphase3D's hash grid and the face-biased box SAT), on every host core in its colour-parallel form
      (greedy body colours, a colour's Newton steps and every force's dual in parallel), and on one thread as the
      reference runs it (sequential Gauss-Seidel, newest body first); it starts from the GPU's poses and velocities with
      no contact history, as the reference's solver does for a world it is handed;
    * `fp32_port`: the FP32 restatement of this library's own step (the same pairs, colours and order as the HIP
      path, bit-identical to it), seeded with the GPU's whole state, on every host core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # the checker, timed here as the CPU baseline

    scene = scene_fn()
    n = scene.body_count
    cores = host_threads()

    def run(world_step):
        steps = 0
        t0 = time.perf_counter()
        while steps < max_steps:
            world_step()
            steps += 1
            if time.perf_counter() - t0 > budget_s:
                break
        return steps, time.perf_counter() - t0

    def reference_path(threads):
        world = oracle_lib.CpuPathWorld(scene, threads=threads, bodies=state["bodies"])
        steps, dt = run(world.step)
        manifolds = world.counts()["manifolds"]
        del world
        return steps, dt, manifolds

    steps_n, dt_n, mf_n = reference_path(cores)
    steps_1, dt_1, mf_1 = reference_path(1)

    oracle_lib.set_threads(cores)
    try:
        oracle = oracle_lib.OracleWorld([scene], precision=32, color_cap=colour_cap, color_rounds=16)
        oracle.seed(state)
        steps_p, dt_p = run(lambda: oracle.step(1))
        del oracle
    finally:
        oracle_lib.set_threads(1)
    span = lambda k: f"steps {first_step + 1}-{first_step + k}"
    return dict(
        value=n * steps_n / dt_n, unit="body-steps/s", cores=cores, kind="port", cpu=cpu_model(),
        sample=(f"{span(steps_n)} of the same {n}-body world from the GPU's poses and velocities: FP64 restatement of "
                f"the reference's CPU path (CpuSolver3D.step with cpuBroadphase3D and the face-biased SAT), "
                f"colour-parallel on {cores} OpenMP threads, {dt_n:.1f} s, {mf_n} manifolds at the end"),
        single_thread=dict(value=n * steps_1 / dt_1, unit="body-steps/s", cores=1,
                           sample=f"{span(steps_1)}, the same restatement on one thread (sequential Gauss-Seidel, "
                                  f"newest body first, as the reference), {dt_1:.1f} s, {mf_1} manifolds at the end"),
        fp32_port=dict(value=n * steps_p / dt_p, unit="body-steps/s", cores=cores,
                       sample=f"{span(steps_p)}, FP32 restatement of this library's step (same pairs, colours and "
                              f"order as the HIP path), seeded with the GPU's state, {cores} OpenMP threads, "
                              f"{dt_p:.1f} s"))


def batched_worlds(avbd, steps, warmup, sample, target_p99_ms=5.0, whole_steps=100):
    """The batched-worlds workload: 4,096 box piles (body counts log-uniform in [8, 2048], friction U(0.2, 0.9)), one
    launch per phase per batch from captured hipGraphs. The autotuner (babylon-avbd_amd/tune.py) first times short
    runs of the worlds split into several batches stepped a few at a time (each on its own stream), each with a solve
    kernel (colour launches, dataflow, one workgroup per world at two sizes), and keeps the fastest whose per-step
    latency (device time of one batch's step, HIP events) meets the 5 ms target at p99; the timed window then runs that
    choice. Beside it: the whole set stepped as one batch (every world in one launch per phase), its per-step latency
    p50/p99 and the HBM floor of its solve; and one world per launch: the first `sample` worlds of the same draw, each
    its own world (own stream and graph), stepped round-robin over the same step window."""
    opts = {"colorCap": 16, "colorRounds": 16, "graph": True, "adaptive": True}
    pipe = avbd.createSolverPipeline("hip")
    log = lambda m: print("# " + m, file=sys.stderr, flush=True)  # noqa: E731
    (groups, concurrency, solve), table = avbd.tune.batch_groups(pipe, avbd.configs.batched_piles, opts,
                                                                 target_p99_ms=target_p99_ms, log=log)
    scenes = avbd.configs.batched_piles()
    bodies = sum(s.body_count for s in scenes)
    worlds = [pipe.createBatch(part, dict(opts, **solve)) for part in avbd.tune.split_balanced(scenes, groups)]
    log(f"batch: {len(scenes)} worlds, {bodies} bodies in {len(worlds)} batches ({avbd.tune.describe(solve)}), "
        f"{sum(w.deviceBytes() for w in worlds) / 2**30:.1f} GiB on device")
    avbd.tune.run_groups(worlds, warmup, concurrency)
    ms, elapsed = [], 0.0
    done = 0
    while done < steps:
        k = min(50, steps - done)
        m, e = avbd.tune.run_groups(worlds, k, concurrency)
        ms += m
        elapsed += e
        done += k
        log(f"batch step {warmup + done}: {len(worlds)} batches, {concurrency} at a time, last step {m[-1]:.2f} ms")
    counters = [w.readCounters() for w in worlds]
    prof = worlds[0].profileStep()
    kind = worlds[0].solverKind()
    for w in worlds:
        w.destroy()
    ms = np.asarray(ms)
    agg = bodies * steps / elapsed

    # The whole set as one batch: every world's step lasts as long as the set's
    whole = pipe.createBatch(avbd.configs.batched_piles(), dict(opts, solver="colors"))
    whole_kind = whole.solverKind()
    avbd.tune.run_groups([whole], warmup, 1)
    wms, welapsed = avbd.tune.run_groups([whole], whole_steps, 1)
    wcount = whole.readCounters()
    wprof = whole.profileStep()
    whole.destroy()
    wms = np.asarray(wms)
    # HBM floor of its solve: per iteration each contact point's 64-B record read by both bodies' primal and by the
    # dual, and its 28 B of multipliers and key written; each dynamic body's 160-B record
    iters = 10
    floor_bytes = iters * (wcount["contacts"] * (3 * 64 + 28) + bodies * 160)
    log(f"whole set: step p50 {np.percentile(wms, 50):.2f} ms, p99 {np.percentile(wms, 99):.2f} ms, "
        f"solve HBM floor {floor_bytes / (HBM_PEAK_GBS * 1e9) * 1e3:.2f} ms")

    solos = [pipe.createSolver(s, opts) for s in avbd.configs.batched_piles(count=sample)]
    solo_bodies = sum(w.bodyCount for w in solos)
    for k in range(warmup):
        for w in solos:
            w.step(1)
    for w in solos:
        w.synchronize()
    log(f"one world per launch: {sample} worlds, {solo_bodies} bodies warmed up")
    t0 = time.perf_counter()
    for k in range(steps):
        for w in solos:
            w.step(1)
        if (k + 1) % 50 == 0:
            log(f"one world per launch: step {warmup + k + 1}")
    for w in solos:
        w.synchronize()
    solo_rate = solo_bodies * steps / (time.perf_counter() - t0)
    pipe.destroy()
    return {
        "workload": "4,096 box piles, body counts log-uniform in [8, 2048], friction U(0.2, 0.9), 10 iterations, "
                    "one launch per phase per batch (hipGraph)",
        "worlds": len(scenes), "bodies": bodies, "batches": groups, "concurrent_batches": concurrency,
        "solve": avbd.tune.describe(solve), "solver_kind": kind,
        "steps_timed": f"{warmup + 1}-{warmup + steps}",
        "body_steps_per_s": agg, "ms_per_round": elapsed * 1e3 / steps,
        "step_ms_p50": float(np.percentile(ms, 50)), "step_ms_p99": float(np.percentile(ms, 99)),
        "step_ms_max": float(ms.max()), "p99_target_ms": target_p99_ms,
        "latency_unit": "one batch's step (device time, HIP events); a world's step lasts as long as its batch's",
        "tuning": [{"batches": g, "concurrent": c, "solve": avbd.tune.describe(sv), "solver_kind": kd,
                    "body_steps_per_s": r, "step_ms_p50": p50, "step_ms_p99": p99, "steps": "61-80"}
                   for g, c, r, p50, p99, sv, kd in table],
        "whole_set": {
            "what": "all 4,096 worlds in one batch (one launch per phase; colour launches, solver kind "
                    f"{whole_kind}), steps {warmup + 1}-{warmup + whole_steps}",
            "body_steps_per_s": bodies * whole_steps / welapsed,
            "step_ms_p50": float(np.percentile(wms, 50)), "step_ms_p99": float(np.percentile(wms, 99)),
            "counters": wcount, "phases_ms": wprof,
            "solve_hbm_floor_ms": floor_bytes / (HBM_PEAK_GBS * 1e9) * 1e3,
            "gap": "the whole set's step cannot meet 5 ms on one MI355X: its solve alone moves at least "
                   f"{floor_bytes / 1e9:.1f} GB per step (10 iterations over {wcount['contacts']:,} contact points, "
                   "each record read by both bodies' primal and the dual and its multipliers written, plus every "
                   f"body's record), {floor_bytes / (HBM_PEAK_GBS * 1e9) * 1e3:.1f} ms at the 8 TB/s peak, before "
                   "collision, adjacency and colouring",
        },
        "counters_first_batch": counters[0], "phases_ms_first_batch": prof,
        "one_world_per_launch": {"worlds": sample, "bodies": solo_bodies, "body_steps_per_s": solo_rate,
                                 "sample": f"worlds 0-{sample - 1} of the same draw, each its own world, "
                                           f"steps {warmup + 1}-{warmup + steps} round-robin"},
        "speedup_vs_one_world_per_launch": agg / solo_rate,
    }


def jointed_mix(avbd, steps, warmup, solver="dataflow"):
    """The jointed + contact mix: a 64 x 64 net of 0.2 m boxes on hard ball joints dropped onto 2,000 static spheres
    and 20,000 free capsules; dt 1/120, 20 iterations, one world (H-PLOC broadphase, the dataflow solve)."""
    scene = avbd.configs.jointed_mix()
    n = scene.body_count
    world = avbd.createSolverPipeline("hip").createSolver(
        scene, {"colorCap": 16, "colorRounds": 16, "graph": True, "adaptive": True, "solver": solver})
    done = 0
    while done < warmup:
        k = min(50, warmup - done)
        world.step(k)
        world.synchronize()
        done += k
        print(f"# mix warm-up {done}: {world.readCounters()}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms = []
    while len(ms) < steps:
        ms.extend(world.stepTimed(min(50, steps - len(ms))).tolist())
    elapsed = time.perf_counter() - t0
    counters = world.readCounters()
    prof = world.profileStep()
    world.destroy()
    return {"workload": "64x64 jointed net (4,096 boxes, 8,064 ball joints) + 2,000 static spheres + 20,000 capsules, "
                        "ground, dt 1/120, 20 iterations", "bodies": n, "joints": len(scene._joints),
            "steps_timed": f"{warmup + 1}-{warmup + steps}", "body_steps_per_s": n * steps / elapsed,
            "ms_per_step": elapsed * 1e3 / steps, "step_ms_p50": float(np.percentile(ms, 50)),
            "step_ms_p99": float(np.percentile(ms, 99)), "solver": solver, "counters": counters,
            "phases_ms": prof}


def sleeping_pile(avbd, steps, warmup, solver, graph=True):
    """The 100k-box pile with sleeping on (island mode, the reference's default thresholds 0.03 and 0.5 s), timed over
    the same step range as the headline run: the step cost with the sleep passes, and how many bodies rest by then."""
    scene = avbd.configs.large_world()
    n = scene.body_count
    world = avbd.createSolverPipeline("hip").createSolver(scene, {"colorCap": 16, "colorRounds": 16, "graph": graph,
                                                                  "adaptive": True, "solver": solver})
    world.enableSleeping({"islands": True})
    done = 0
    while done < warmup:
        k = min(100, warmup - done)
        world.step(k)
        world.synchronize()
        done += k
    asleep_before = world.readSleepStats()["sleeping"]
    ms = []
    t0 = time.perf_counter()
    while len(ms) < steps:
        ms.extend(world.stepTimed(min(50, steps - len(ms))).tolist())
    elapsed = time.perf_counter() - t0
    asleep_after = world.readSleepStats()["sleeping"]
    print(f"# sleeping pile: {asleep_before} -> {asleep_after} asleep, {elapsed * 1e3 / steps:.3f} ms/step",
          file=sys.stderr, flush=True)
    world.destroy()
    return {"workload": "the 100k-box pile with sleeping on (islands, speed 0.03, 0.5 s)", "bodies": n,
            "steps_timed": f"{warmup + 1}-{warmup + steps}", "body_steps_per_s": n * steps / elapsed,
            "ms_per_step": elapsed * 1e3 / steps, "step_ms_p50": float(np.percentile(ms, 50)),
            "asleep_at_start": asleep_before, "asleep_at_end": asleep_after}


def serving(avbd, requests, rate, pool, max_active, small_world=0, large_world=0, lookahead=2):
    """The serving workload: Poisson arrivals (`rate` per second) of rollout requests (U{1..120} steps) for a pool of
    worlds of 16..50,000 boxes (log-uniform sizes, the single large world's body distributions), admitted by the
    serving loop against a 200 GB HBM budget and run concurrently (one stream per world). Worlds of at most
    `small_world` bodies are solved by one workgroup each (one launch per step's solve)."""
    sizes = avbd.configs.serving_pool(pool)
    server = avbd.serve.Server(hbmBudget=200 * 10**9, maxActive=max_active, lookahead=lookahead,
                               smallWorldBodies=small_world, largeWorldBodies=large_world)
    ids = [server.addWorld(avbd.configs.boxes_world(int(n), avbd.configs.SEED + 100 + k)) for k, n in enumerate(sizes)]
    warm = [server.submit(w, 1) for w in ids]  # admission and graph capture of every world
    for r in warm:
        server.wait(r)
    print(f"# serving: {pool} worlds ({int(sizes.sum())} bodies) resident", file=sys.stderr, flush=True)
    server.resetStats()
    t, world, steps = avbd.configs.serving_requests(requests, rate, pool)
    t0 = time.perf_counter()
    rids = []
    for k in range(requests):
        delay = t0 + t[k] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        rids.append(server.submit(ids[int(world[k])], int(steps[k])))
    print(f"# serving: {requests} requests submitted in {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    for r in rids:
        server.wait(r)
    st = server.stats()
    server.destroy()
    return {"workload": f"Poisson arrivals at {rate:g} req/s of rollouts U{{1..120}} steps for {pool} worlds of "
                        "16..50,000 boxes (log-uniform), 200 GB HBM budget, 10 iterations",
            "requests": requests, "worlds": pool, "pool_bodies": int(sizes.sum()),
            "offered_body_steps_per_s": float(rate * np.mean(steps) * sizes[world].mean()),
            "body_steps_per_s": st["body_steps"] / max(st["wall_seconds"], 1e-9),
            "requests_per_s": st["completed"] / max(st["wall_seconds"], 1e-9), "wall_s": st["wall_seconds"],
            "step_ms_p50": st["step_ms_p50"], "step_ms_p99": st["step_ms_p99"], "step_ms_max": st["step_ms_max"],
            "request_ms_p50": st["request_ms_p50"], "request_ms_p99": st["request_ms_p99"],
            "evictions": st["evictions"], "peak_resident_gb": st["peak_resident_bytes"] / 1e9,
            "max_active": max_active, "lookahead": lookahead, "small_world_bodies": small_world, "large_world_bodies": large_world,
            "solve_aborts": st["solve_aborts"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed steps after the pile has settled (the first 300 steps always run before them)")
    ap.add_argument("--settle", type=int, default=SETTLE_STEPS,
                    help="steps run before the warm-up so the timed window is the settled pile (default 300)")
    ap.add_argument("--workload", default="large", choices=["large", "pile1000"])
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--broadphase", default="auto", choices=["auto", "grid", "hploc"])
    ap.add_argument("--profile-steps", type=int, default=5,
                    help="steps timed per launch with HIP events after the timed window (the roofline's average)")
    ap.add_argument("--bvh-rebuild", type=int, default=0,
                    help="H-PLOC: rebuild the tree every this many steps, refit in between (0: the library default)")
    ap.add_argument("--solver", default="dataflow", choices=["auto", "colors", "dataflow"],
                    help="the large world's solve: one launch per colour, or every iteration in one dataflow launch")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-mix", action="store_true")
    ap.add_argument("--no-sleeping", action="store_true")
    ap.add_argument("--sleeping-only", action="store_true",
                    help="run only the sleeping-pile record (with --no-graph: without graph replay, for profilers)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-pool", type=int, default=64)
    ap.add_argument("--serve-active", type=int, default=32)
    ap.add_argument("--serve-lookahead", type=int, default=2, help="steps in flight per running request")
    ap.add_argument("--serve-small-world", type=int, default=512,
                    help="served worlds of at most this many bodies are solved one workgroup each")
    ap.add_argument("--serve-large-world", type=int, default=4096,
                    help="served worlds of at least this many bodies run the dataflow solve on a share of the device")
    ap.add_argument("--state", default=None,
                    help="settled-state file: loaded instead of the warm-up when it exists, written after it otherwise")
    ap.add_argument("--summarize-counters", default=None, metavar="DIR",
                    help="print per-step sums of the counters rocprofv3 --pmc wrote under DIR for --kernel, then exit")
    ap.add_argument("--kernel", default="k_primal")
    ap.add_argument("--launches-per-step", type=int, default=0,
                    help="with --summarize-counters: dispatches of --kernel per step (0: one step = all dispatches)")
    ap.add_argument("--mix-steps", type=int, default=200)
    ap.add_argument("--mix-warmup", type=int, default=400)
    ap.add_argument("--batch-steps", type=int, default=280)
    ap.add_argument("--batch-warmup", type=int, default=20)
    ap.add_argument("--batch-sample", type=int, default=128)
    args = ap.parse_args()
    if args.summarize_counters:
        print(json.dumps(summarize_counters(args.summarize_counters, args.kernel, args.launches_per_step)))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the AVBD step here is single-GPU; "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.sleeping_only:
        print(json.dumps(sleeping_pile(importlib.import_module("babylon-avbd_amd"), args.steps, args.settle + args.warmup,
                                       args.solver, not args.no_graph)))
        return

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
                "broadphase": args.broadphase, "solver": args.solver,
                "bvh": {"rebuildInterval": args.bvh_rebuild}})
    print(f"# world: {n_bodies} bodies, {world.deviceBytes() / 2**20:.0f} MiB on device", file=sys.stderr, flush=True)

    # Settling: the pile falls and lands (steps 1..settle, always run, whatever the warm-up count); the colour and
    # capacity policy adapts from counters read back asynchronously. Then the warm-up steps, then the timed window.
    settle_rate = None
    if args.state and os.path.exists(args.state):
        saved = dict(np.load(args.state))
        world.writeState(saved)
        world.setColoring(int(saved["colorsUsed"]) + 3, 16)
        print(f"# state loaded from {args.state}", file=sys.stderr, flush=True)
    else:
        done = 0
        t_settle = time.perf_counter()
        while done < args.settle:
            k = min(50, args.settle - done)
            world.step(k)
            world.synchronize()
            done += k
            print(f"# settling {done}: {world.readCounters()}", file=sys.stderr, flush=True)
        if args.settle:
            settle_rate = n_bodies * args.settle / (time.perf_counter() - t_settle)
        if args.state:
            np.savez(args.state, colorsUsed=world.readCounters()["colors"], **world.readState())
    done = 0
    while done < args.warmup:
        k = min(50, args.warmup - done)
        world.step(k)
        world.synchronize()
        done += k
        print(f"# warm-up {args.settle + done}: {world.readCounters()}", file=sys.stderr, flush=True)
    first = args.settle + args.warmup  # steps taken before the timed window
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

    # Roofline of the dominant kernel: profiled steps (HIP events per launch on the engine's stream), averaged
    profs = [world.profileStep() for _ in range(args.profile_steps)]
    prof = {k: sum(p[k] for p in profs) / len(profs) for k in profs[0]}
    bytes_iter, parts = primal_bytes_per_iteration(world)
    iters = scene.params["iterations"]
    dataflow = prof["primalLaunches"] == 1 and prof["dualLaunches"] == 0
    if dataflow:
        # One persistent launch runs every iteration's colours and duals: its bytes are both passes', 10 times
        dual_iter, dual_parts = dual_bytes_per_iteration(world)
        parts.update(dual_parts)
        kernel = "k_solve_dataflow"
        kernel_desc = "k_solve_dataflow (every iteration's primal colours and dual in one launch)"
        launch_bytes = iters * (bytes_iter + dual_iter)
        launch_ms = prof["primal"]
        launches = 1.0 / iters
        bytes_note = {"primal_bytes_per_iteration": bytes_iter, "dual_bytes_per_iteration": dual_iter}
    else:
        kernel = "k_primal"
        kernel_desc = "k_primal (one colour's bodies)"
        launches = prof["primalLaunches"] / iters
        launch_bytes = bytes_iter / launches
        launch_ms = prof["primal"] / max(prof["primalLaunches"], 1)
        bytes_note = {"primal_bytes_per_iteration": bytes_iter}
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    # HBM traffic and the VALU issue floor of the same kernel in the same (settled) state, from the committed counters
    traffic, traffic_note, valu = None, "not measured for this workload or window", None
    measured = measured_counters(kernel) if (args.workload == "large" and args.settle == SETTLE_STEPS
                                             and args.solver == "dataflow") else None
    if measured:
        traffic = measured["hbm_bytes_per_launch"]
        traffic_note = (f"{traffic_profile(kernel)}: rocprofv3 2 x FETCH_SIZE + WRITE_SIZE per launch (copy-kernel "
                        f"calibrated) on the settled pile, {measured['window']}")
        v = measured["valu"]
        valu = {"bound": "valu-issue", "insts_per_launch": v["insts_per_launch"], "floor_ms": v["floor_ms"],
                "frac": v["floor_ms"] / launch_ms, "wait_fraction": v["wait_fraction"],
                "waves_per_simd": v["waves_per_simd"], "source": traffic_profile(kernel)}
    print(f"# profile: {json.dumps(prof)}; {kernel} bytes/launch {launch_bytes} {parts}", file=sys.stderr, flush=True)

    # Long-horizon statistics of the world where the run ends (rest heights, energy, penetration)
    mf_end, ct_end = world.readContactRecords()
    statistics = dict(step=first + args.steps + args.profile_steps,
                      **avbd.stats.world_statistics(world.readBodies(), mf_end, ct_end, scene.up,
                                                    scene.params["gravity"]))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene_fn, state0, first, args.cpu_steps, args.cpu_budget, 64)

    world.destroy()
    mix = None if args.no_mix else jointed_mix(avbd, args.mix_steps, args.mix_warmup, args.solver)
    sleeping = None if args.no_sleeping else sleeping_pile(avbd, args.steps, first, args.solver)
    served = None if args.no_serving else serving(avbd, args.serve_requests, args.serve_rate, args.serve_pool,
                                                  args.serve_active, args.serve_small_world, args.serve_large_world,
                                                  args.serve_lookahead)
    batched = None
    if not args.no_batched:
        batched = batched_worlds(avbd, args.batch_steps, args.batch_warmup, args.batch_sample)

    line = {
        "metric": "body-steps/s (100k-box pile, 10 iters); batched-world p50/p99 step latency",
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
                   "broadphase": {"auto": "auto (H-PLOC BVH: a world or batch of >= 10,000 bodies)", "grid": "uniform grid",
                                  "hploc": "H-PLOC BVH"}[args.broadphase],
                   "steps_timed": f"{first + 1}-{first + args.steps}",
                   "window": f"the settled pile: steps 1-{args.settle} settle and {args.warmup} warm-up steps run "
                             "untimed first",
                   "settling_body_steps_per_s": settle_rate,
                   "counters": counters, "statistics": statistics},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": kernel_desc,
            "bytes_per_launch": launch_bytes,
            **bytes_note,
            "launches_per_iteration": launches,
            "avg_launch_ms": launch_ms,
            "valu": valu,
        },
        "phases_ms": prof,
        "cpu_baseline": cpu,
        "batched": batched,
        "jointed_mix": mix,
        "sleeping_pile": sleeping,
        "serving": served,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
