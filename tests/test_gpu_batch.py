"""The batched / served form (tsi_batch_*): every session's frame in a mixed batch equals the
frame the plain path (a tsi_ctx per scene, one launch sequence per frame) computes alone --
nodes, statistics, colour and depth bit for bit -- over several consecutive submits in which the
set of sessions changes; one session is also compared with the oracle; admission by arena
footprint; a session's capacity failure stays its own.
"""
import numpy as np
import pytest

import oracle
import temporalsetinversion_amd as T
from helpers import assert_same_nodes
from temporalsetinversion_amd import scenes
from temporalsetinversion_amd.batch import Batch

pytestmark = pytest.mark.gpu

TAPES = scenes.batch_test_sets()["mixed4"]
# (tape, localsubdiv, session parameters, frame size)
SESSIONS = [
    (0, 6, dict(), (64, 48)),
    (1, 5, dict(lipschitz_corrected=True), (48, 48)),
    (2, 6, dict(swept_volume=True, swept_window=0.25), (40, 32)),
    (3, 5, dict(swept_volume=True, swept_window=1.0 / 60.0), (64, 40)),
    (0, 4, dict(point_sample=False, swept_volume=True, swept_window=0.5), (32, 32)),
    (1, 6, dict(swept_volume=False, fast_multiplier=2.0), (0, 0)),
    (3, 4, dict(delta=0.02, lipschitz_corrected=True), (24, 16)),
    # cinematic path tracer (TSI_SHADE_CINEMATIC), then with gradient normals as well: traced by
    # their own launch beside the plain trace of the others
    (2, 5, dict(shading=4), (40, 32)),
    (0, 5, dict(shading=6), (32, 24)),
]
# which sessions take part in each submit, and the time of each session's k-th frame
SUBMITS = [[0, 1, 2, 3, 4, 5, 6, 7, 8], [0, 2, 3, 5, 7], [1, 2, 4, 6, 8], [0, 1, 2, 3, 4, 5, 6, 7, 8]]


def frame_time(sid, k):
    return float(np.float32((k + 2 * sid) / 60.0))


def _plain_runs(tapes_np, view):
    """Each session alone through the plain path: frames in the order the batch runs them."""
    per = {}
    for sid, (tape, L, kw, (w, h)) in enumerate(SESSIONS):
        frames = sum(1 for s in SUBMITS if sid in s)
        out = []
        with T.State(tapes_np[tape], width=max(w, 16), height=max(h, 16), localsubdiv=L, voxel_capacity=1 << 18,
                     prune=False, **kw) as st:
            st.reset()
            for k in range(frames):
                st.update(frame_time(sid, k), view)
                stats = st.set_inversion().stats
                img = st.trace() if w else None
                st.enqueue_root()
                out.append((st.read_nodes(), stats, img))
        per[sid] = out
    return per


def test_batch_sessions_match_the_plain_path():
    tapes_np = [T.tape_from_ll(t) for t in TAPES]
    view = T.diagonal_camera()
    plain = _plain_runs(tapes_np, view)
    with Batch(tapes_np, arena_bytes=1 << 30, waves_per_simd=2) as b:
        sids = [b.open(tape=tape, localsubdiv=L, voxel_capacity=1 << 18, max_width=64, max_height=48, **kw)
                for tape, L, kw, _ in SESSIONS]
        done = {s: 0 for s in range(len(SESSIONS))}
        for step, members in enumerate(SUBMITS):
            frames = []
            for m in members:
                w, h = SESSIONS[m][3]
                frames.append((sids[m], frame_time(m, done[m]), w, h, view))
            res, set_ms, trace_ms = b.run(frames)
            print(f"submit {step}: {len(frames)} sessions, set inversion {set_ms:.3f} ms, trace {trace_ms:.3f} ms",
                  flush=True)
            for m, r in zip(members, res):
                k = done[m]
                nodes, stats, img = plain[m][k]
                what = f"submit {step} session {m} frame {k}"
                assert r["status"] == 0, what
                assert_same_nodes(b.read_nodes(sids[m]), nodes, what)
                for key in ("cell_evals", "rev_evals", "fwd_ops", "rev_ops", "parents", "enqueued", "deallocs",
                            "allocated_nodes"):
                    assert r[key] == stats[key], f"{what}: {key} {r[key]} vs {stats[key]}"
                w, h = SESSIONS[m][3]
                if w:
                    rgba, depth = b.frame(sids[m], w, h)
                    assert np.array_equal(rgba.view(np.uint32), img[0].view(np.uint32)), f"{what}: colour"
                    assert np.array_equal(depth.view(np.uint32), img[1].view(np.uint32)), f"{what}: depth"
                done[m] += 1


def test_batch_session_matches_the_oracle():
    pair = scenes.batch_test_sets()["pair"]
    tape = oracle.parse_ll(pair[0])
    view = T.diagonal_camera()
    fr = oracle.Frame(tape, localsubdiv=5, swept_volume=True, swept_window=0.25, voxel_capacity=1 << 18)
    try:
        with Batch([tape, oracle.parse_ll(pair[1])], arena_bytes=1 << 28) as b:
            s = b.open(tape=0, localsubdiv=5, swept_volume=True, swept_window=0.25, voxel_capacity=1 << 18,
                       max_width=48, max_height=32)
            other = b.open(tape=1, localsubdiv=6, voxel_capacity=1 << 18, max_width=16, max_height=16)
            for k in range(3):
                t = float(np.float32(k / 60.0))
                res, _, _ = b.run([(other, t, 16, 16, view), (s, t, 48, 32, view)])
                ost = fr.set_inversion(t)
                orgba, odepth, _ = fr.trace(np.array(view, np.float32), 48, 32, t, T.camera_zdist(32))
                fr.enqueue_root()
                assert res[1]["cell_evals"] == ost.cell_evals
                assert_same_nodes(b.read_nodes(s), fr.export(), f"frame {k}")
                rgba, depth = b.frame(s, 48, 32)
                assert np.array_equal(rgba.view(np.uint32), orgba.view(np.uint32))
                assert np.array_equal(depth.view(np.uint32), odepth.view(np.uint32))
    finally:
        fr.close()


def test_admission_by_footprint():
    tapes_np = [T.tape_from_ll(t) for t in scenes.batch_test_sets()["one"]]
    with Batch(tapes_np, arena_bytes=64 << 20) as b:
        kw = dict(tape=0, localsubdiv=5, voxel_capacity=1 << 16, max_width=64, max_height=64)
        fp = b.footprint(**kw)
        assert 0 < fp < (64 << 20)
        opened = []
        while True:
            try:
                opened.append(b.open(**kw))
            except T.TsiError as e:
                assert e.code == -10          # TSI_E_BUDGET
                break
        assert len(opened) == (64 << 20) // fp
        assert b.info()["arena_used"] == len(opened) * fp
        # closing one admits one again, and the session computes normally
        b.close_session(opened[1])
        s = b.open(**kw)
        res, _, _ = b.run([(s, 0.0, 64, 64, T.diagonal_camera())])
        assert res[0]["status"] == 0 and res[0]["cell_evals"] > 0


def test_capacity_failure_stays_in_its_session():
    tapes_np = [T.tape_from_ll(t) for t in scenes.batch_test_sets()["pair"]]
    view = T.diagonal_camera()
    with Batch(tapes_np, arena_bytes=1 << 28) as b:
        small = b.open(tape=0, localsubdiv=6, voxel_capacity=256, max_width=16, max_height=16)
        good = b.open(tape=1, localsubdiv=5, voxel_capacity=1 << 18, max_width=32, max_height=32)
        res, _, _ = b.run([(small, 0.0, 16, 16, view), (good, 0.0, 32, 32, view)])
        assert res[0]["status"] == -3         # TSI_E_POOL_EXHAUSTED
        assert res[1]["status"] == 0
        with T.State(tapes_np[1], width=32, height=32, localsubdiv=5, voxel_capacity=1 << 18, prune=False) as st:
            st.reset()
            st.update(0.0, view)
            st.set_inversion()
            assert_same_nodes(b.read_nodes(good), st.read_nodes(), "beside a failing session")
