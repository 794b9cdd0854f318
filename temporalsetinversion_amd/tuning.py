"""Autotuner of the batched form's launch shape (tsi_batch_tune).

The persistent set-inversion kernel and the tile-queue trace kernel have four knobs:

* waves_per_simd     resident wavefronts per SIMD of the persistent grid (the grid is
                     CUs x waves_per_simd blocks of 4 wavefronts): more hides more latency of each
                     wavefront's dependent chain, fewer leaves more L2 and instruction cache per
                     wavefront and a smaller saved-operand stack in use;
* claim_units        work units (8 frontier entries = 64 cells) a wavefront takes per atomic on the
                     work queue: the cells per wavefront per claim;
* trace_claim_tiles  8x8 pixel tiles a trace wavefront takes per atomic;
* trace_blocks_per_cu  blocks of the trace grid per CU.

`autotune` times one submit of the caller's frames for each value of one knob at a time, the
others held at the best found so far (coordinate descent, the set-inversion knobs judged by the
set-inversion launch's device time and the trace knobs by the trace launch's), and leaves the
batch at the best shape.  Each trial recomputes the same frames (the frames carry the reset
flag, or the caller resets the sessions), so trials compare equal work.
"""
from __future__ import annotations

import json
import time

KNOBS = {
    "waves_per_simd": (1, 2, 3),
    "claim_units": (1, 2, 4),
    "trace_claim_tiles": (1, 4, 16),
    "trace_blocks_per_cu": (2, 4, 8),
}
SET_KNOBS = ("waves_per_simd", "claim_units")
TRACE_KNOBS = ("trace_claim_tiles", "trace_blocks_per_cu")


def autotune(batch, frames, repeats: int = 2, knobs=None) -> dict:
    """Returns {"best": {knob: value}, "set_ms": .., "trace_ms": .., "trials": [...], "seconds": ..}."""
    knobs = dict(knobs or KNOBS)
    maxw = batch.info()["max_waves_per_simd"]
    knobs["waves_per_simd"] = tuple(v for v in knobs["waves_per_simd"] if v <= maxw) or (1,)
    info = batch.info()
    best = {"waves_per_simd": info["waves_per_simd"], "claim_units": info["claim_units"],
            "trace_claim_tiles": info["trace_claim_tiles"], "trace_blocks_per_cu": info["trace_blocks_per_cu"]}
    trials = []
    t0 = time.perf_counter()

    def measure(shape):
        batch.tune(**shape)
        s_best = t_best = float("inf")
        for _ in range(repeats):
            res, s_ms, t_ms = batch.run(frames)
            if any(r["status"] != 0 for r in res):
                raise RuntimeError("a frame failed during tuning")
            s_best, t_best = min(s_best, s_ms), min(t_best, t_ms)
        trials.append(dict(shape, set_ms=s_best, trace_ms=t_best))
        return s_best, t_best

    best_set, best_trace = measure(best)
    for knob in SET_KNOBS + TRACE_KNOBS:
        for v in knobs[knob]:
            if v == best[knob]:
                continue
            shape = dict(best, **{knob: v})
            s_ms, t_ms = measure(shape)
            if knob in SET_KNOBS and s_ms < best_set:
                best, best_set = shape, s_ms
            elif knob in TRACE_KNOBS and t_ms < best_trace:
                best, best_trace = shape, t_ms
    batch.tune(**best)
    return {"best": best, "set_ms": best_set, "trace_ms": best_trace, "trials": trials,
            "seconds": time.perf_counter() - t0}


def save(result: dict, path: str):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)["best"]
