"""Autotuning of the batched step: how many independent batches to split a set of worlds into.

One batch steps every world in one launch per phase: the most work per launch, but a world's step then lasts as long
as the whole set's. Splitting the worlds into G batches stepped one after another keeps the aggregate rate close
(every batch still fills the device with its share) while each world's step lasts about 1/G as long. The tuner times
a few steps of each candidate split (and how many batches run at once, each on its own stream) from the same starting
state and keeps the fastest whose per-step latency (device time of one batch's step, HIP events) meets the target at
the 99th percentile.
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


def batch_groups(pipeline, make_scenes, options, target_p99_ms=5.0,
                 candidates=((1, 1), (16, 1), (32, 2), (32, 4), (64, 4), (64, 8)), warmup=60, probe_steps=20,
                 margin=0.9, log=None):
    """Pick (batches, concurrency) for the worlds `make_scenes()` builds (fresh scenes for every candidate). Each
    candidate is probed after `warmup` steps (so the worlds have landed and carry their full contact load) and must
    meet `margin` × the target at p99, which leaves room for the later steps' variation.

    Returns (best, table): best = (batches, concurrency); table = [(batches, concurrency, body_steps_per_s, p50_ms,
    p99_ms)]."""
    table = []
    for g, c in candidates:
        scenes = make_scenes()
        bodies = sum(s.body_count for s in scenes)
        worlds = [pipeline.createBatch(part, options) for part in split_balanced(scenes, g)]
        run_groups(worlds, warmup, c)
        ms, elapsed = run_groups(worlds, probe_steps, c)
        rate = bodies * probe_steps / elapsed
        row = (len(worlds), c, rate, float(np.percentile(ms, 50)), float(np.percentile(ms, 99)))
        table.append(row)
        if log:
            log(f"tune: {row[0]} batches, {c} at a time: {rate:.3e} body-steps/s, step p50 {row[3]:.2f} ms, "
                f"p99 {row[4]:.2f} ms")
        for w in worlds:
            w.destroy()
    meeting = [r for r in table if r[4] <= margin * target_p99_ms]
    best = max(meeting, key=lambda r: r[2]) if meeting else min(table, key=lambda r: r[4])
    return (best[0], best[1]), table
