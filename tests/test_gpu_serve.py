"""The served form (serve.Server, continuous batching over tsi_batch_*): a Poisson schedule of
requests for several scenes, served in real time; every scene's octree and last frame equal those
the plain path leaves after the same windows one request at a time (serve.PlainServer); a scene
that does not fit in the arena waits until an idle session is closed."""
import numpy as np
import pytest

import temporalsetinversion_amd as T
from helpers import assert_same_nodes
from temporalsetinversion_amd import scenes, serve
from temporalsetinversion_amd.batch import Batch

pytestmark = pytest.mark.gpu


def _scenes(n):
    view = T.diagonal_camera()
    return [serve.Scene(tape=i % 4, localsubdiv=5 + (i % 2), width=32 + 8 * (i % 3), height=24, view=view,
                        params=dict(lipschitz_corrected=(i % 2 == 0)), voxel_capacity=1 << 18) for i in range(n)]


def test_served_frames_equal_solo_frames():
    tapes = [T.tape_from_ll(t) for t in scenes.batch_test_sets()["mixed4"]]
    sc = _scenes(6)
    reqs = serve.poisson_requests(len(sc), 400.0, 0.3, seed=1)
    with Batch(tapes, arena_bytes=1 << 30) as b:
        srv = serve.Server(b, sc)
        out = srv.run(reqs)
        print({k: v for k, v in out.items() if k != "latencies"}, flush=True)
        assert out["requests"] == len(reqs) and out["failed"] == 0
        assert out["steps"] < len(reqs)             # requests shared submits
        plain = serve.PlainServer(tapes, sc)
        try:
            pl = plain.run(reqs)
            assert pl["failed"] == 0
            for s in sorted(srv.sid):
                assert_same_nodes(b.read_nodes(srv.sid[s]), plain.states[s].read_nodes(), f"scene {s}")
                rgba, depth = plain.states[s].trace_wait()
                brgba, bdepth = b.frame(srv.sid[s], sc[s].width, sc[s].height)
                assert np.array_equal(rgba.view(np.uint32), brgba.view(np.uint32)), f"scene {s}: colour"
                assert np.array_equal(depth.view(np.uint32), bdepth.view(np.uint32)), f"scene {s}: depth"
        finally:
            plain.close()


def test_served_admission_waits_for_the_budget():
    tapes = [T.tape_from_ll(t) for t in scenes.batch_test_sets()["mixed4"]]
    sc = _scenes(3)
    with Batch(tapes, arena_bytes=1 << 30) as probe:
        fp = max(probe.footprint(tape=s.tape, localsubdiv=s.localsubdiv, voxel_capacity=s.voxel_capacity,
                                 max_width=s.width, max_height=s.height) for s in sc)
    # room for two sessions: the third scene's requests wait until an idle session is closed
    reqs = [serve.Request(0, 0.00, 0), serve.Request(1, 0.00, 0), serve.Request(0, 0.01, 1),
            serve.Request(2, 0.02, 0), serve.Request(2, 0.06, 1), serve.Request(1, 0.12, 1)]
    with Batch(tapes, arena_bytes=2 * fp + 4096) as b:
        srv = serve.Server(b, sc, idle_close_s=0.02)
        out = srv.run(reqs)
        assert out["failed"] == 0 and out["requests"] == len(reqs)
        assert out["sessions_closed"] >= 1
        assert b.info()["sessions_open"] <= 2
