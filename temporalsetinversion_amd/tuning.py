"""Autotuner of the batched form's launch shape (tsi_batch_tune).

The persistent set-inversion kernel has two knobs:

* waves_per_simd  resident wavefronts per SIMD (the grid is one block per CU of
                  4 x waves_per_simd wavefronts): more hide more of each wavefront's dependent
                  chain and memory latency, fewer leave more L2, MALL and instruction cache per
                  wavefront and keep less of the saved-operand stack live;
* claim_units     work units (8 frontier entries = 64 cells) a wavefront takes per atomic on the
                  work queue after the first round: the cells a wavefront takes per claim.

`autotune` times one submit of the caller's frames for each value of one knob at a time, the
other held at the best found so far (coordinate descent, judged by the set-inversion launch's
device time), and leaves the batch at the best shape.  Each trial recomputes the same frames
(the frames carry the reset flag, or the caller resets the sessions), so trials compare equal
work.  The trace launch has no knob: one block per 16x16 pixels of every frame.
"""
from __future__ import annotations

import json
import time

KNOBS = {
    "waves_per_simd": (1, 2, 3, 4),
    "claim_units": (1, 2, 4, 8, 16),
}


def autotune(batch, frames, repeats: int = 2, knobs=None) -> dict:
    """Returns {"best": {knob: value}, "set_ms": .., "trials": [...], "seconds": ..}."""
    knobs = dict(knobs or KNOBS)
    info = batch.info()
    maxw = info["max_waves_per_simd"]
    knobs["waves_per_simd"] = tuple(v for v in knobs["waves_per_simd"] if v <= maxw) or (1,)
    best = {"waves_per_simd": info["waves_per_simd"], "claim_units": info["claim_units"]}
    trials = []
    t0 = time.perf_counter()

    def measure(shape):
        batch.tune(**shape)
        s_best = float("inf")
        for _ in range(repeats):
            res, s_ms, _ = batch.run(frames)
            if any(r["status"] != 0 for r in res):
                raise RuntimeError("a frame failed during tuning")
            s_best = min(s_best, s_ms)
        trials.append(dict(shape, set_ms=s_best))
        return s_best

    best_ms = measure(best)
    for knob in ("waves_per_simd", "claim_units"):
        for v in knobs[knob]:
            if v == best[knob]:
                continue
            shape = dict(best, **{knob: v})
            ms = measure(shape)
            if ms < best_ms:
                best, best_ms = shape, ms
    batch.tune(**best)
    return {"best": best, "set_ms": best_ms, "trials": trials, "seconds": time.perf_counter() - t0}


def save(result: dict, path: str):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)["best"]
