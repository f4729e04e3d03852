"""Autotuner for the batched persistent launch.

The launch shape of ``fm_match_batch`` has these runtime knobs:
- ``waves_per_query``: the number of claims a query's root list is cut into,
  i.e. how many waves a query fans out to when it starts (deeper work then
  spreads by splitting). ``root_chunk`` is the same knob as a fixed number of
  roots per claim for every query; a configuration sets one or the other.
- ``tile_parents``: the LDS tile, parents staged per extension step with
  their requirement rows (at most 64; wide rows stage fewer).
- ``task_rows``: parents per queued task. This is the fork threshold, the
  replacement of the reference's soft bank limit (SubMaxMulSpill,
  src/environment.cuh).
- ``grid_waves``: persistent waves, as ``grid_frac`` of the engine's default
  grid. Fewer waves leave more LDS and L2 per wave; more waves hide more
  memory latency.

Each candidate configuration is run on the real workload in count mode and
timed with the kernel's HIP events. The counts of every run must agree, so
that tuning cannot trade correctness for speed. The fastest configuration
is returned.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field


@dataclass
class TuneResult:
    best: dict
    best_ms: float
    trials: list = field(default_factory=list)


DEFAULT_SPACE = {
    "waves_per_query": [16, 64, 256],
    "tile_parents": [32, 64],
    "task_rows": [1024, 4096],
    "grid_frac": [1.0, 1.5],
}

KNOBS = ("root_chunk", "waves_per_query", "tile_parents", "task_rows")


def autotune(engine, graph, queries, space: dict | None = None, table: bool = False, reps: int = 1,
             timeout_ms: int = 120000, log=None) -> TuneResult:
    space = space or DEFAULT_SPACE
    grid = engine.info().grid_waves
    ref_counts = None
    trials = []
    best, best_ms = None, float("inf")
    keys = list(space)
    for combo in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, combo))
        # grid_frac is relative to the engine's default grid (8 waves per CU);
        # the library caps it at the resident waves
        opts = {k: cfg[k] for k in KNOBS if k in cfg}
        opts["grid_waves"] = max(1, int(grid * cfg.get("grid_frac", 1.0)))
        ms = []
        for _ in range(reps):
            r = engine.match_batch(graph, queries, table=table, timeout_ms=timeout_ms, **opts)
            counts = [r.count(i) for i in range(len(queries))]
            if ref_counts is None:
                ref_counts = counts
            elif counts != ref_counts:
                raise AssertionError(f"autotune: configuration {opts} changed the result")
            ms.append(r.stats().kernel_ms)
            r.close()
        t = min(ms)
        trials.append({**opts, "kernel_ms": t})
        if log:
            log(f"autotune: {opts} {t:.2f} ms")
        if t < best_ms:
            best, best_ms = opts, t
    return TuneResult(best, best_ms, trials)
