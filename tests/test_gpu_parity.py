"""GPU parity: libtsi's set inversion and trace against the oracle, bit for bit.

Every frame of a multi-frame temporal sequence is compared node by node (canonical
(depth, i, j, k) keys; f, t, time_err, fast, has_children, dx, dy, dz bitwise).
"""
import ctypes as C

import numpy as np
import pytest

import temporalsetinversion_amd as T
from helpers import assert_same_nodes, run_pair, tape_of

pytestmark = pytest.mark.gpu

SEQ = [k / 60.0 for k in range(0, 8)] + [0.5, 0.75, 1.0]

# evaluation engines: the tape compiled into straight-line kernels (default), and the
# interpreter with and without pruned per-group programs
ENGINES = {"compiled": dict(engine="compiled"),
           "compiled-pruned": dict(engine="compiled", prune=True, prune_live_max=100),
           "compiled-full": dict(engine="compiled", prune=False),
           "interp-pruned": dict(engine="interpreter", prune=True),
           "interp-full": dict(engine="interpreter", prune=False)}


@pytest.mark.parametrize("engine", sorted(ENGINES))
@pytest.mark.parametrize("scene,L", [("sphere", 5), ("spheres_linear", 6), ("spheres", 6)])
def test_frames_match_oracle(scene, L, engine):
    tape = tape_of(scene)
    for k, g, o, gs, os_, _, _ in run_pair(tape, SEQ, localsubdiv=L, **ENGINES[engine]):
        print(f"scene={scene} L={L} engine={engine} frame={k} nodes={len(o)} cells={os_.cell_evals} "
              f"pruned={gs['pruned_instr']}", flush=True)
        assert_same_nodes(g, o, f"{scene} frame {k}")


@pytest.mark.parametrize("engine", ["compiled", "compiled-pruned", "interp-pruned"])
@pytest.mark.parametrize("opts", [dict(swept_volume=True), dict(lipschitz_corrected=True),
                                  dict(point_sample=False), dict(fastenabled=False),
                                  dict(delta=0.05, fast_multiplier=0.5)])
def test_param_variants(opts, engine):
    tape = tape_of("spheres")
    for k, g, o, gs, os_, _, _ in run_pair(tape, SEQ[:6], localsubdiv=5, **opts, **ENGINES[engine]):
        print(f"opts={opts} frame={k} nodes={len(o)}", flush=True)
        assert_same_nodes(g, o, f"{opts} frame {k}")


@pytest.mark.parametrize("scene,L,kw", [("spheres", 6, {}), ("capsules", 5, dict(lipschitz_corrected=True)),
                                       ("csg", 5, dict(swept_volume=True))])
def test_speculative_levels_change_nothing(scene, L, kw):
    """The frame-start speculative pass over the octree's shallow nodes (on by default for the
    unpruned compiled engine): every frame's nodes and statistics equal those of frames evaluated
    level by level, and from the second frame on the shallow levels take most of their cells
    from it."""
    tape = tape_of(scene)
    runs = {}
    for spec in (True, False):
        out = []
        with T.State(tape, localsubdiv=L, voxel_capacity=1 << 20, prune=False, speculate=spec, **kw) as st:
            st.reset()
            for k, t in enumerate(SEQ[:7]):
                st.update(float(np.float32(t)))
                out.append((st.set_inversion().stats, st.read_nodes()))
                st.enqueue_root()
        runs[spec] = out
    taken = 0
    for k, ((s1, n1), (s0, n0)) in enumerate(zip(runs[True], runs[False])):
        print(f"{scene} frame {k}: cells {s1['cell_evals']}, from the speculative pass {s1['spec_cells']}", flush=True)
        assert_same_nodes(n1, n0, f"{scene} frame {k}")
        for key in ("cell_evals", "rev_evals", "fwd_ops", "rev_ops", "parents", "enqueued", "deallocs",
                    "allocated_nodes"):
            assert s1[key] == s0[key], f"frame {k}: {key}"
        assert s0["spec_cells"] == 0
        if k == 0:   # after the reset the octree is the root, a leaf: its 8 children as new cells
            assert s1["spec_cells"] == 8
        taken += s1["spec_cells"]
    assert taken > 0


def test_speculative_pass_across_depth_changes_resets_and_frames_in_flight():
    """The speculative pass runs on a stream of its own beside the previous frame's deep levels:
    frames queued two at a time, the octree depth changed between frames (tsi_set_params: more
    or fewer speculative levels than the previous frame ran) and a reset in the middle give every
    frame the nodes and statistics of the same frames without speculation."""
    tape = tape_of("capsules")
    plan = [(0, 6, False), (1, 6, False), (2, 7, False), (3, 5, False), (4, 5, True), (5, 7, False),
            (6, 7, False), (7, 6, False)]   # (frame, localsubdiv, reset before it)
    runs = {}
    for spec in (True, False):
        out = []
        with T.State(tape, localsubdiv=6, voxel_capacity=1 << 20, prune=False, lipschitz_corrected=True,
                     speculate=spec) as st:
            st.reset()
            pending = []
            for k, L, reset in plan:
                if reset or L != st.params.localsubdiv:
                    while pending:
                        out.append((pending.pop(0), st.set_inversion_result().stats, None))
                    if reset:
                        st.reset()
                    if L != st.params.localsubdiv:
                        st.set_params(localsubdiv=L)
                st.update(float(np.float32(k / 60.0)))
                st.set_inversion_async()
                st.enqueue_root()
                pending.append(k)
                if len(pending) == 2:
                    out.append((pending.pop(0), st.set_inversion_result().stats, None))
            while pending:
                out.append((pending.pop(0), st.set_inversion_result().stats, None))
            out.append((-1, None, st.read_nodes()))
        runs[spec] = out
    for (k1, s1, n1), (k0, s0, n0) in zip(runs[True], runs[False]):
        assert k1 == k0
        if s1 is not None:
            print(f"frame {k1}: cells {s1['cell_evals']}, from the speculative pass {s1['spec_cells']}", flush=True)
            for key in ("cell_evals", "rev_evals", "fwd_ops", "rev_ops", "parents", "enqueued", "deallocs"):
                assert s1[key] == s0[key], f"frame {k1}: {key}"
        else:
            assert_same_nodes(n1, n0, "after the sequence")
    assert sum(s["spec_cells"] for _, s, _ in runs[True] if s is not None) > 0


@pytest.mark.parametrize("engine", ["compiled-full", "interp-full"])
def test_stats_match_oracle(engine):
    tape = tape_of("spheres")
    for k, g, o, gs, os_, _, _ in run_pair(tape, SEQ[:5], localsubdiv=6, **ENGINES[engine]):
        s = gs
        assert s["cell_evals"] == os_.cell_evals
        assert s["rev_evals"] == os_.rev_evals
        assert s["fwd_ops"] == os_.fwd_ops
        assert s["rev_ops"] == os_.rev_ops
        assert s["enqueued"] == os_.enqueued
        assert s["pruned_instr"] == 0


@pytest.mark.parametrize("engine", ["compiled-pruned", "interp-pruned"])
@pytest.mark.parametrize("scene,L", [("spheres_linear", 6), ("capsules", 5)])
def test_pruning_prunes(scene, L, engine):
    """Pruned evaluation runs fewer records than the full tape and the same cells as the oracle.

    The compiled engine skips whole regions of 16 records (here whenever the parents' records
    apply); the 136-record spheres tape, whose MIN chain runs through every region, may skip
    none: only parity is required there."""
    tape = tape_of(scene)
    nfwd = len(tape) - 4
    for k, g, o, gs, os_, _, _ in run_pair(tape, SEQ[:4], localsubdiv=L, **ENGINES[engine]):
        assert gs["cell_evals"] == os_.cell_evals
        assert gs["rev_evals"] == os_.rev_evals
        assert gs["enqueued"] == os_.enqueued
        assert gs["fwd_ops"] <= os_.fwd_ops
        print(f"scene={scene} frame={k} cells={gs['cell_evals']} mean records run="
              f"{nfwd - gs['pruned_instr'] / max(gs['cell_evals'], 1):.1f} of {nfwd}", flush=True)
        if k == 0 and (engine != "compiled-pruned" or scene == "capsules"):
            assert gs["pruned_instr"] > 0
        assert_same_nodes(g, o, f"{scene} frame {k}")


@pytest.mark.parametrize("engine", sorted(ENGINES))
@pytest.mark.parametrize("scene,L", [("capsules", 4), ("csg", 4)])
def test_big_tapes_match_oracle(scene, L, engine):
    tape = tape_of(scene)
    for k, g, o, gs, os_, _, _ in run_pair(tape, [0.0, 1 / 60, 2 / 60, 0.3], localsubdiv=L, **ENGINES[engine]):
        print(f"scene={scene} engine={engine} frame={k} nodes={len(o)} pruned={gs['pruned_instr']}", flush=True)
        assert_same_nodes(g, o, f"{scene} frame {k}")


def test_trace_matches_oracle():
    tape = tape_of("spheres")
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    for k, g, o, gs, os_, gf, of in run_pair(tape, [0.0, 1 / 60, 0.25], width=64, height=48, trace_every=1,
                                                localsubdiv=6, view=view):
        grgba, gdepth = gf
        orgba, odepth = of
        assert np.array_equal(grgba.view(np.uint32), orgba.view(np.uint32)), f"colour differs in frame {k}"
        assert np.array_equal(gdepth.view(np.uint32), odepth.view(np.uint32)), f"depth differs in frame {k}"
        assert (odepth > 0).sum() > 0


def test_engines_are_the_ones_asked_for():
    """The default State compiles the tape; the interpreter runs only when asked."""
    tape = tape_of("spheres")
    with T.State(tape, localsubdiv=4, voxel_capacity=1 << 16) as st:
        info = st.tape_info()
        assert info["engine"] == "compiled" and info["n"] == len(tape)
    with T.State(tape, localsubdiv=4, voxel_capacity=1 << 16, engine="interpreter") as st:
        assert st.tape_info()["engine"] == "interpreter"


def test_overlapped_trace_matches_oracle():
    """The trace of frame k runs on the second stream (octree snapshot) while frame k+1's set
    inversion runs; each frame's image still equals the oracle's trace of that frame."""
    import oracle

    tape = tape_of("spheres")
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    W, H = 64, 48
    times = [float(np.float32(t)) for t in (0.0, 1 / 60, 2 / 60, 0.25)]
    fr = oracle.Frame(tape, localsubdiv=6, voxel_capacity=1 << 20)
    want = []
    for t in times:
        fr.set_inversion(t)
        rgba, depth, _ = fr.trace(view, W, H, t, T.camera_zdist(H))
        want.append((rgba, depth))
        fr.enqueue_root()
    fr.close()
    got = []
    with T.State(tape, width=W, height=H, localsubdiv=6, voxel_capacity=1 << 20) as st:
        st.reset()
        for k, t in enumerate(times):
            st.update(t, view)
            st.set_inversion()          # runs while the previous frame's trace may still run
            if k > 0:
                got.append(st.trace_wait())
            st.render_async()
        got.append(st.trace_wait())
        ms, visits, n = st.trace_totals()
        assert n == len(times) and visits > 0 and ms > 0
    for k, ((g_rgba, g_depth), (o_rgba, o_depth)) in enumerate(zip(got, want)):
        assert np.array_equal(g_rgba.view(np.uint32), o_rgba.view(np.uint32)), f"colour differs in frame {k}"
        assert np.array_equal(g_depth.view(np.uint32), o_depth.view(np.uint32)), f"depth differs in frame {k}"


def test_frame_dump_png(tmp_path):
    """main.rs:693-715: RGBA32F frame -> RGBA8 (clamp, x255, round half away from zero) -> PNG."""
    from test_outputs import read_png

    tape = tape_of("spheres")
    with T.State(tape, width=40, height=24, localsubdiv=5, voxel_capacity=1 << 18) as st:
        st.reset()
        st.update(0.0, T.look_at_inverse((2.2, 1.7, 2.6)))
        st.set_inversion()
        rgba, _ = st.trace()
        want = np.floor(np.clip(rgba, 0.0, 1.0) * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)
        assert np.array_equal(st.frame_rgba8(), want)
        st.save_png(tmp_path / "0.png")
        assert np.array_equal(read_png(tmp_path / "0.png"), want)


def test_frames_in_flight_match_blocking_frames():
    """tsi_set_inversion_async / _result (two frames queued, traces overlapped) give the same
    statistics, octree and images as the blocking calls; a third frame, a blocking frame or a
    reset while two are in flight are refused with TSI_E_PENDING."""
    from temporalsetinversion_amd import _lib

    tape = tape_of("spheres")
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    times = [float(np.float32(k / 60.0)) for k in range(6)]
    kw = dict(width=48, height=32, localsubdiv=6, voxel_capacity=1 << 20)

    with T.State(tape, **kw) as st:
        st.reset()
        want_stats, want_img = [], []
        for t in times:
            st.update(t, view)
            want_stats.append(st.set_inversion().stats)
            want_img.append(st.trace()[0])
            st.enqueue_root()
        want_nodes = st.read_nodes()

    with T.State(tape, **kw) as st:
        st.reset()
        got_stats, inflight = [], 0
        for k, t in enumerate(times):
            st.update(t, view)
            st.set_inversion_async()
            inflight += 1
            if inflight == 2:
                ctx = st._ctx
                assert _lib.lib().tsi_set_inversion_async(ctx) == -12           # TSI_E_PENDING
                assert _lib.lib().tsi_set_inversion(ctx, None, None) == -12
                assert _lib.lib().tsi_reset(ctx, (C.c_float * 6)(-5, -5, -5, 5, 5, 5)) == -12
            st.render_async()
            if inflight == 2:
                got_stats.append(st.set_inversion_result().stats)
                inflight -= 1
        got_stats.append(st.set_inversion_result().stats)
        img = st.trace_wait()[0]
        got_nodes = st.read_nodes()
        _, _, ntraces = st.trace_totals()
    assert ntraces == len(times)
    for a, b in zip(got_stats, want_stats):
        # (with pruning, the records run depend on how the frontier lists group parents into
        # wavefronts, which follows the order of their atomics: fwd_ops / rev_ops may differ)
        for key in ("cell_evals", "rev_evals", "enqueued", "parents", "deallocs"):
            assert a[key] == b[key], key
    assert_same_nodes(got_nodes, want_nodes, "after the last frame")
    assert np.array_equal(img.view(np.uint32), want_img[-1].view(np.uint32))


def test_overlapping_traces_keep_each_frame():
    """Two traces in flight use the two trace slots, each with its own snapshot and frame: trace k
    is issued while trace k-1 may still run and while frame k+1's set inversion rewrites the
    octree, and the image and depth read back after every second frame equal that frame's
    blocking trace bit for bit."""
    tape = tape_of("spheres")
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    times = [float(np.float32(k / 60.0)) for k in range(8)]
    kw = dict(width=64, height=48, localsubdiv=6, voxel_capacity=1 << 20)
    want = []
    with T.State(tape, **kw) as st:
        st.reset()
        for t in times:
            st.update(t, view)
            st.set_inversion()
            want.append(st.trace())
            st.enqueue_root()
    with T.State(tape, **kw) as st:
        st.reset()
        for k, t in enumerate(times):
            st.update(t, view)
            st.set_inversion()
            st.render_async()
            if k % 2 == 1:
                rgba, depth = st.trace_wait()
                assert np.array_equal(rgba.view(np.uint32), want[k][0].view(np.uint32)), k
                assert np.array_equal(depth.view(np.uint32), want[k][1].view(np.uint32)), k
        _, _, ntraces = st.trace_totals()
    assert ntraces == len(times)
    assert (want[0][1] >= 0).any() and (want[0][1] < 0).any()


def test_frames_outside_the_specialised_domain():
    """The tape kernels leave out fix where it is the identity for cells in [-5,5]^3 and
    t-windows in [-2, 64]; a frame at t = 70 leaves that domain, the kernels are rebuilt with every
    fix, and results stay bit-identical to the oracle before and after."""
    tape = tape_of("sphere")   # small: its kernels without the assumption compile in seconds
    times = [0.0, 1 / 60, 70.0, 70.0 + 1 / 60]
    with T.State(tape, localsubdiv=4, voxel_capacity=1 << 16) as st:
        assert st.tape_info()["specialized"]
    seen = []
    for k, g, o, gs, os_, _, _ in run_pair(tape, times, localsubdiv=5):
        print(f"frame={k} t={times[k]} nodes={len(o)}", flush=True)
        assert_same_nodes(g, o, f"frame {k} (t = {times[k]})")
        seen.append(k)
    assert seen == [0, 1, 2, 3]
    with T.State(tape, localsubdiv=4, voxel_capacity=1 << 16) as st:
        st.reset()
        st.update(70.0)
        st.set_inversion()
        assert not st.tape_info()["specialized"]


@pytest.mark.parametrize("scene,shading,L,size,kw", [
    ("spheres", 4, 6, (64, 48), {}),                              # cinematic, normals (1, 1, 1)
    ("capsules", 6, 6, (48, 36), dict(lipschitz_corrected=True)),  # + gradient normals
    ("csg", 5, 5, (40, 30), dict(swept_volume=True)),              # + normal_viz
])
def test_cinematic_matches_oracle(scene, shading, L, size, kw):
    """The cinematic path tracer (compute.wgsl:3499, 4440-4523, 4620-4648): colour, depth and
    node visits bit-identical to the oracle over three frames (jitter, bounces, sun shadow rays,
    sky; the seed follows the time)."""
    import oracle

    tape = tape_of(scene)
    W, H = size
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    fr = oracle.Frame(tape, localsubdiv=L, voxel_capacity=1 << 20, shading=shading, **kw)
    st = T.State(tape, width=W, height=H, localsubdiv=L, voxel_capacity=1 << 20, shading=shading, **kw)
    try:
        st.reset()
        for k, t in enumerate([0.0, 1 / 60, 0.25]):
            t = float(np.float32(t))
            st.update(t, view)
            gs = st.set_inversion().stats
            os_ = fr.set_inversion(t)
            assert gs["rev_evals"] == os_.rev_evals
            grgba, gdepth = st.trace()
            _, gvisits = st.last_trace()
            orgba, odepth, ovisits = fr.trace(np.array(st.uniforms.view[:], np.float32), W, H, t, T.camera_zdist(H))
            print(f"{scene} shading={shading} frame={k} visits={gvisits}", flush=True)
            assert_same_nodes(st.read_nodes(), fr.export(), f"frame {k}")
            assert np.array_equal(grgba.view(np.uint32), orgba.view(np.uint32)), f"colour differs in frame {k}"
            assert np.array_equal(gdepth.view(np.uint32), odepth.view(np.uint32)), f"depth differs in frame {k}"
            assert gvisits == ovisits
            st.enqueue_root()
            fr.enqueue_root()
    finally:
        st.close()
        fr.close()


@pytest.mark.parametrize("shading", [1, 3])   # TSI_SHADE_NORMALS, | TSI_SHADE_GRADIENT
def test_normal_viz_matches_oracle(shading):
    """normal_viz (compute.wgsl:3503-3504): octree and images bit-identical to the oracle, with
    the finite-difference normal of a tape scene (1) and with the reverse-pass gradients (3, every
    cell taking the reverse pass)."""
    tape = tape_of("spheres")
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    for k, g, o, gs, os_, gf, of in run_pair(tape, [0.0, 1 / 60, 0.25], width=64, height=48, trace_every=1,
                                                localsubdiv=6, view=view, shading=shading):
        print(f"shading={shading} frame={k} nodes={len(o)} rev={gs['rev_evals']}", flush=True)
        assert_same_nodes(g, o, f"frame {k}")
        assert gs["rev_evals"] == os_.rev_evals
        if shading == 3:
            assert gs["rev_evals"] == gs["cell_evals"]
        (grgba, gdepth), (orgba, odepth) = gf, of
        assert np.array_equal(grgba.view(np.uint32), orgba.view(np.uint32)), f"colour differs in frame {k}"
        assert np.array_equal(gdepth.view(np.uint32), odepth.view(np.uint32)), f"depth differs in frame {k}"
