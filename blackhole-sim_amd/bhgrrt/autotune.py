"""Autotuner for the persistent ray-queue kernels: steps-per-sync and launch shape.

The persistent kernels expose four knobs through ``bh_ctx_set_tuning``:
block_threads (waves per workgroup), blocks_per_cu (persistent grid size; 0 takes
the occupancy limit), steps_per_sync (RK2 steps between wave refill checks) and
refill_min (idle lanes that trigger a refill; batches always refill whole tiles).
``autotune`` times every candidate with HIP events on the context stream, keeps
the fastest (median of ``reps`` launches after one warm-up) and leaves it set.
Results can be cached in a JSON file keyed by device, precision and workload.
"""

from __future__ import annotations

import itertools
import json
import os
from pathlib import Path
from typing import Callable

import numpy as np

DEFAULT_SPACE = {
    "block_threads": (64, 128, 256),
    "steps_per_sync": (8, 16, 32, 48),
    "blocks_per_cu": (0, 1, 2),
    "refill_min": (64,),
}


def candidates(space: dict | None = None) -> list[dict]:
    sp = dict(DEFAULT_SPACE, **(space or {}))
    keys = list(sp)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(sp[k] for k in keys))]


def autotune(renderer, precision: int, launch: Callable[[], None], space: dict | None = None, reps: int = 3,
             cache: str | os.PathLike | None = None, key: str | None = None) -> dict:
    """Pick the fastest tuning for ``launch`` (one kernel launch on ``renderer``'s context).

    Returns {"best": tuning, "ms": best_ms, "trials": [(tuning, ms), ...]}; the best tuning stays set.
    """
    if cache is not None and key is not None:
        path = Path(cache)
        if path.exists():
            table = json.loads(path.read_text())
            if key in table:
                renderer.set_tuning(precision, **table[key]["best"])
                return table[key]
    trials = []
    for cand in candidates(space):
        renderer.set_tuning(precision, **cand)
        launch()
        renderer.synchronize()
        ms = []
        for _ in range(int(reps)):
            launch()
            ms.append(renderer.last_kernel_ms())
        trials.append((cand, float(np.median(ms))))
    best, best_ms = min(trials, key=lambda t: t[1])
    renderer.set_tuning(precision, **best)
    result = {"best": best, "ms": best_ms, "trials": trials}
    if cache is not None and key is not None:
        path = Path(cache)
        table = json.loads(path.read_text()) if path.exists() else {}
        table[key] = result
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(table, indent=1))
    return result
