"""Launch-shape autotuner.

Times candidate values of the library's launch knobs (include/mfcl.h
MFCL_TUNE_*) on a caller-supplied representative query and keeps the
fastest.  The query callable must run one full batched query on the
context's stream; timing uses the library's HIP events when enabled
(per-stage kernel time) and wall time around ctx.sync() otherwise.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Iterable

import ctypes as C

from . import load_library

GJK_GROUP = 1
EPA_THREADS = 2
MESH_BLOCKS_PER_CU = 3
MESH_STACK = 4
EPA_LDS_FACES = 5
GUESS_POLICY = 6
EPA_SPILL = 7
MLP_WAVES = 8
LANE_REFILL = 9


def set_knob(ctx, knob: int, value: int) -> None:
    ctx._check(load_library().mfcl_tune(ctx.h, C.c_int32(knob), C.c_int32(value)))


def tune(ctx, knob: int, candidates: Iterable[int], run: Callable[[], None], reps: int = 3,
         log: Callable[[str], None] = None) -> Dict[int, float]:
    """Sets `knob` to the candidate with the lowest median time of `run`."""
    times = {}
    for v in candidates:
        set_knob(ctx, knob, v)
        run()
        ctx.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            ctx.sync()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        times[v] = ts[len(ts) // 2]
        if log:
            log(f"[autotune] knob {knob} = {v}: {1e3 * times[v]:.3f} ms")
    best = min(times, key=times.get)
    set_knob(ctx, knob, best)
    return times
