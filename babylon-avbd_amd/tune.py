"""Autotuning of the batched step: how many independent batches to split a set of worlds into, how many to step at
once, and which solve kernel each batch runs.

One batch steps every world in one launch per phase: the most work per launch, but a world's step then lasts as long
as the whole set's. Splitting the worlds into G batches stepped one after another keeps the aggregate rate close
(every batch still fills the device with its share) while each world's step lasts about 1/G as long. The solve
kernel is the other axis: one launch per colour (many small launches), the dataflow solve (one persistent launch whose
workgroups must all fit on the device at once, so batches stepped together each take a share of it) or one workgroup
per world (one launch, each world on one compute unit, with the workgroup size a further choice). The tuner times a few steps of each candidate from
the same starting state and keeps the fastest whose per-step latency (device time of one batch's step, HIP events)
meets the target at the 99th percentile.
"""
import time

import numpy as np


def split_balanced(scenes, groups):
    """Contiguous runs of the scenes with about equal body counts (at least one scene each)."""
    sizes = np.array([s.body_count for s in scenes], np.float64)
    groups = max(1, min(int(groups), len(scenes)))
    cuts = np.searchsorted(np.cumsum(sizes), sizes.sum() * np.arange(1, groups) / groups)
    bounds = [0] + sorted(set(int(c) + 1 for c in cuts if 0 < int(c) + 1 < len(scenes))) + [len(scenes)]
    return [scenes[a:b] for a, b in zip(bounds[:-1], bounds[1:])]


def run_groups(worlds, steps, concurrency=1):
    """Step every batch once per round for `steps` rounds, `concurrency` batches at a time (each on its own stream,
    so their steps overlap on the device): (per-step device ms list, elapsed s)."""
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        for k in range(0, len(worlds), concurrency):
            chunk = worlds[k:k + concurrency]
            for w in chunk:
                w.stepTimedBegin(1)
            for w in chunk:
                ms.extend(float(x) for x in w.stepTimedEnd())
    return ms, time.perf_counter() - t0


#: Candidate (batches, concurrency, solve options) for the batched workload: colour launches at several splits and
#: concurrencies (batches a multiple of the concurrency, so every round is full), the dataflow solve one batch at a
#: time or two at once (each grid half the device), one workgroup per world at two workgroup sizes.
CANDIDATES = (
    (16, 2, {"solver": "colors"}),
    (24, 3, {"solver": "colors"}),
    (24, 4, {"solver": "colors"}),
    (28, 4, {"solver": "colors"}),
    (32, 4, {"solver": "colors"}),
    (8, 1, {"solver": "dataflow"}),
    (16, 1, {"solver": "dataflow"}),
    (16, 2, {"solver": "dataflow", "dataflowShare": 2}),
    (32, 4, {"solver": "worlds", "worldBlock": 256}),
    (32, 4, {"solver": "worlds", "worldBlock": 1024}),
)


def describe(solve):
    """A short name for a candidate's solve options."""
    name = solve.get("solver", "auto")
    if "worldBlock" in solve:
        name += f"/{solve['worldBlock']}"
    if "dataflowShare" in solve:
        name += f" (1/{solve['dataflowShare']} of the device each)"
    return name


def batch_groups(pipeline, make_scenes, options, target_p99_ms=5.0, candidates=CANDIDATES, warmup=60, probe_steps=20,
                 margin=0.9, log=None):
    """Pick (batches, concurrency, solve options) for the worlds `make_scenes()` builds (fresh scenes for every
    candidate). A candidate is (batches, concurrency) or (batches, concurrency, solve options merged into `options`).
    Each is probed after `warmup` steps (so the worlds have landed and carry their full contact load) and must meet
    `margin` × the target at p99, which leaves room for the later steps' variation.

    Returns (best, table): best = (batches, concurrency, solve options); table = [(batches, concurrency,
    body_steps_per_s, p50_ms, p99_ms, solve options, solver kind run: 1 colours, 2 dataflow, 3 workgroup per
    world)]."""
    table = []
    for cand in candidates:
        g, c = cand[0], cand[1]
        solve = dict(cand[2]) if len(cand) > 2 else {}
        scenes = make_scenes()
        bodies = sum(s.body_count for s in scenes)
        worlds = [pipeline.createBatch(part, dict(options, **solve)) for part in split_balanced(scenes, g)]
        kind = worlds[0].solverKind()  # the dataflow solve falls back to colour launches past 32-bit offsets
        run_groups(worlds, warmup, c)
        ms, elapsed = run_groups(worlds, probe_steps, c)
        rate = bodies * probe_steps / elapsed
        row = (len(worlds), c, rate, float(np.percentile(ms, 50)), float(np.percentile(ms, 99)), solve, kind)
        table.append(row)
        if log:
            log(f"tune: {row[0]} batches, {c} at a time, {describe(solve)} (kind {kind}): {rate:.3e} body-steps/s, "
                f"step p50 {row[3]:.2f} ms, p99 {row[4]:.2f} ms")
        for w in worlds:
            w.destroy()
    meeting = [r for r in table if r[4] <= margin * target_p99_ms]
    best = max(meeting, key=lambda r: r[2]) if meeting else min(table, key=lambda r: r[4])
    return (best[0], best[1], best[5]), table
