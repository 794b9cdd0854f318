"""Host-side parts of the batched and served forms (no device): the batch module's generation and
compilation, its saved-operand stack size, the workload generators and the request schedule."""
import numpy as np
import pytest

import temporalsetinversion_amd as T
from temporalsetinversion_amd import batch, scenes, serve


def test_batch_module_compiles_with_the_largest_stack_of_its_tapes():
    texts = scenes.batch_test_sets()["mixed4"]
    tapes = [T.tape_from_ll(t) for t in texts]
    info = batch.precompile_batch(tapes)
    stacks = [T.kernel_source(t, prune=False)[1]["stack"] for t in tapes]
    assert info["tapes"] == 4
    assert info["stack_rows"] == max(stacks) + 1
    # the same tapes give the same module (a cache hit the second time)
    assert batch.precompile_batch(tapes)["cache_hit"] == 1


def test_batch_module_rejects_bad_tapes():
    good = T.tape_from_ll(scenes.spheres(0))
    long = np.zeros(4097, dtype=T.INSTR_DTYPE)
    with pytest.raises(T.TsiError) as e:
        batch.precompile_batch([good, long])
    assert e.value.code == -4          # TSI_E_TAPE_TOO_LONG


def test_batched_workload_generators():
    tapes = [T.tape_from_ll(t) for t in scenes.batch_tapes()]
    assert len(tapes) == 16
    assert all(100 <= len(t) <= 2000 for t in tapes)
    served = [T.tape_from_ll(t) for t in scenes.served_tapes()]
    assert len(served) == 32 and all(200 <= len(t) <= 2000 for t in served)
    reqs = scenes.batch_requests(256, 16)
    assert len(reqs) == 256
    L = np.array([r["localsubdiv"] for r in reqs])
    assert L.min() == 6 and L.max() == 10
    w = np.array([r["window"] for r in reqs])
    assert w.min() >= 1 / 30 - 1e-7 and w.max() <= 0.5 + 1e-7
    t1 = np.array([r["time"] for r in reqs])
    assert np.all(t1 - w >= -1e-6) and np.all(t1 <= 1.0)
    assert {(r["width"], r["height"]) for r in reqs} <= set(scenes.FRAME_SIZES)
    # seeded
    assert scenes.batch_requests(256, 16) == reqs


def test_poisson_schedule():
    reqs = serve.poisson_requests(32, 200.0, 10.0, seed=3)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and arr[-1] < 10.0
    assert abs(len(reqs) - 2000) < 4 * np.sqrt(2000)
    # a scene's requests are its consecutive frames, in arrival order
    for s in range(32):
        frames = [r.frame for r in reqs if r.scene == s]
        assert frames == list(range(len(frames)))
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 1 / 200.0) < 0.2 / 200.0
