"""Autotuning of the batched step: how many independent batches to split a set of worlds into.

One batch steps every world in one launch per phase: the most work per launch, but a world's step then lasts as long
as the whole set's. Splitting the worlds into G batches stepped one after another keeps the aggregate rate close
(every batch still fills the device with its share) while each world's step lasts about 1/G as long. The tuner times
a few steps of each candidate split from the same starting state and keeps the fastest whose per-step latency (device
time of one batch's step, HIP events) meets the target at the 99th percentile.
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


def run_groups(worlds, steps):
    """Step every batch once per round, one after another, for `steps` rounds: (per-step ms list, elapsed s)."""
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        for w in worlds:
            ms.append(float(w.stepTimed(1)[0]))
    return ms, time.perf_counter() - t0


def batch_groups(pipeline, make_scenes, options, target_p99_ms=5.0, candidates=(1, 4, 8, 16, 32), warmup=10,
                 probe_steps=20, log=None):
    """Pick the number of batches for the worlds `make_scenes()` builds (fresh scenes for every candidate).

    Returns (best, table): best = the candidate group count; table = [(groups, body_steps_per_s, p50_ms, p99_ms)]."""
    table = []
    for g in candidates:
        scenes = make_scenes()
        bodies = sum(s.body_count for s in scenes)
        worlds = [pipeline.createBatch(part, options) for part in split_balanced(scenes, g)]
        run_groups(worlds, warmup)
        ms, elapsed = run_groups(worlds, probe_steps)
        rate = bodies * probe_steps / elapsed
        row = (len(worlds), rate, float(np.percentile(ms, 50)), float(np.percentile(ms, 99)))
        table.append(row)
        if log:
            log(f"tune: {row[0]} batches: {rate:.3e} body-steps/s, step p50 {row[2]:.2f} ms, p99 {row[3]:.2f} ms")
        for w in worlds:
            w.destroy()
    meeting = [r for r in table if r[3] <= target_p99_ms]
    best = max(meeting, key=lambda r: r[1]) if meeting else min(table, key=lambda r: r[3])
    return best[0], table
