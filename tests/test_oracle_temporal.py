"""Closed-form pins of the oracle's temporal update (oracle/frame.c, compute.wgsl:3598-3614,
3971-4045) and of its swept-window classification.

The field is a moving plane, f = s (x - c - v t): its gradient is (s, 0, 0) and df/dt = -s v
exactly, so every quantity the temporal update computes has a closed form that does not go
through the interval library:

* a cell re-evaluated over the window [t_a, t_b] adds min(|v| (t_b - t_a) * g, delta) to its
  time error, with g = 1 under Lipschitz correction (df divided by |grad f| = s) and g = s
  without (compute.wgsl:3972-3996);
* a cell whose time error reaches delta is re-subdivided and its error restarts at 0 (4015-4026);
  its children are then evaluated over the window since their own last evaluation;
* fast = (|df| >= delta * fast_multiplier) (4003-4009); the children of a fast cell get a
  degenerate window [t, t] (3603-3606), hence no derivative and no time error;
* a new cell of a swept frame is evaluated over [t - w, t] (3605): a cell contains the plane
  somewhere in that window iff its x-range meets [c + v (t - w), c + v t].

Times are the frame times t_k = float32(k / 60) the library and the oracle both use; sums are
formed in float64 and compared with a tolerance far below the increments, and the scene's
constants are chosen so that no cumulative error lands within 1e-4 of delta.
"""
import numpy as np
import pytest

import oracle

DELTA = 0.01
C0 = 1.3


def plane_ll(v: float, s: float) -> str:
    return ("define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) #0 {\n"
            f"  %5 = fmul float %3, {v:.9e}\n"
            "  %6 = fsub float %0, %5\n"
            f"  %7 = fsub float %6, {C0:.9e}\n"
            f"  %8 = fmul float %7, {s:.9e}\n"
            "  ret float %8\n}\n")


def times(n):
    return [float(np.float32(k / 60.0)) for k in range(n)]


def run(v, s, n_frames, L=4, **kw):
    tape = oracle.parse_ll(plane_ll(v, s))
    fr = oracle.Frame(tape, localsubdiv=L, point_sample=False, voxel_capacity=1 << 18, **kw)
    out = []
    for t in times(n_frames):
        fr.set_inversion(t)
        out.append({(int(n["depth"]), int(n["i"]), int(n["j"]), int(n["k"])): n for n in fr.export()})
        fr.enqueue_root()
    fr.close()
    return out


def depth1_plane_cells(frame):
    """The four depth-1 cells x in [0, 5] (i = 1), which contain the plane x = 1.3 + v t."""
    return [key for key in frame if key[0] == 1 and key[1] == 1]


def expected_errors(v, g, n_frames):
    """Time error after each frame and the frames that re-subdivide, for a cell evaluated every
    frame over [t_{k-1}, t_k] (a child of the root, which is re-enqueued every frame)."""
    ts = times(n_frames)
    err, errs, splits = 0.0, [0.0], [0]
    for k in range(1, n_frames):
        err += min(abs(v) * (ts[k] - ts[k - 1]) * g, DELTA)
        if err >= DELTA:
            err = 0.0
            splits.append(k)
        errs.append(err)
    return errs, splits


@pytest.mark.parametrize("lipschitz", [True, False])
def test_time_error_accumulates_per_frame_and_restarts_at_delta(lipschitz):
    v, s = 0.33, 2.0
    n = 8
    frames = run(v, s, n, lipschitz_corrected=lipschitz)
    errs, splits = expected_errors(v, 1.0 if lipschitz else s, n)
    # Lipschitz on: 0.0055 per frame, a split every second frame; off: 0.011 >= delta every frame
    assert splits == ([0, 2, 4, 6] if lipschitz else list(range(n)))
    ts = times(n)
    for k, fr in enumerate(frames):
        cells = depth1_plane_cells(fr)
        assert len(cells) == 4
        for key in cells:
            node = fr[key]
            assert node["f"] == 0.0 and node["has_children"] == 1
            assert node["t"] == np.float32(ts[k])
            assert abs(float(node["time_err"]) - errs[k]) < 1e-6, (k, key, float(node["time_err"]), errs[k])
            # fast: |df| >= delta (frame 0 takes no derivative)
            inc = abs(v) * (ts[k] - ts[k - 1]) * (1.0 if lipschitz else s) if k else 0.0
            assert node["fast"] == (1.0 if inc >= DELTA else 0.0)
            # the children were last evaluated at the last split
            last_split = max(j for j in splits if j <= k)
            kids = [c for c in fr if c[0] == 2 and c[1] // 2 == key[1] and c[2] // 2 == key[2] and c[3] // 2 == key[3]]
            assert len(kids) == 8
            for c in kids:
                assert fr[c]["t"] == np.float32(ts[last_split]), (k, c)


def test_children_of_a_split_are_evaluated_over_their_own_window():
    """At a split frame the children's windows start at their own last evaluation: two frames
    back here, so each adds min(2 |v| / 60, delta) and is itself split when that reaches delta."""
    v, s = 0.33, 1.0
    frames = run(v, s, 3, lipschitz_corrected=True)
    ts = times(3)
    fr = frames[2]
    inc = min(abs(v) * (ts[2] - ts[0]), DELTA)   # 0.011: capped at delta, so split
    assert inc >= DELTA
    d2 = [k for k in fr if k[0] == 2 and fr[k]["f"] == 0.0]
    assert d2
    for key in d2:
        node = fr[key]
        assert node["t"] == np.float32(ts[2])
        assert node["time_err"] == 0.0 and node["has_children"] == 1
        kids = [c for c in fr if c[0] == 3 and (c[1] // 2, c[2] // 2, c[3] // 2) == key[1:]]
        assert all(fr[c]["t"] == np.float32(ts[2]) for c in kids)


@pytest.mark.parametrize("mult,fast", [(1.0, True), (2.0, False)])
def test_fast_flag_and_degenerate_windows_below_a_fast_cell(mult, fast):
    """|df| = 0.9 / 60 = 0.015 per frame: fast when 0.015 >= delta * fast_multiplier.  Below a fast
    cell the children get [t, t]: no reverse pass, so their time error stays 0 and they are
    re-subdivided (degenerate window, compute.wgsl:4021)."""
    v = 0.9
    frames = run(v, 1.0, 3, lipschitz_corrected=True, fast_multiplier=mult)
    ts = times(3)
    for k in (1, 2):
        fr = frames[k]
        for key in depth1_plane_cells(fr):
            node = fr[key]
            assert node["fast"] == (1.0 if fast else 0.0)
            # 0.015 capped at delta: the error reaches delta every frame and restarts
            assert node["time_err"] == 0.0
            kids = [c for c in fr if c[0] == 2 and (c[1] // 2, c[2] // 2, c[3] // 2) == key[1:] and fr[c]["f"] == 0.0]
            for c in kids:
                assert fr[c]["t"] == np.float32(ts[k])
                if fast:
                    assert fr[c]["time_err"] == 0.0 and fr[c]["dx"] == 0.0   # no derivative taken
                else:
                    # window [t_{k-1}, t_k]: gradient midpoint s = 1 on x, 0 elsewhere
                    assert fr[c]["dx"] == 1.0 and fr[c]["dy"] == 0.0 and fr[c]["dz"] == 0.0


def test_lipschitz_divisor_is_the_gradient_norm():
    """Lipschitz correction divides df by |grad f| = s: with s = 4 the per-frame increment is
    |v| dt instead of 4 |v| dt; the stored gradient midpoint is (s, 0, 0)."""
    v, s = 0.1, 4.0   # 0.0017 and 0.0067 per frame: both below delta
    ts = times(2)
    on = run(v, s, 2, lipschitz_corrected=True)[1]
    off = run(v, s, 2, lipschitz_corrected=False)[1]
    for key in depth1_plane_cells(on):
        a, b = float(on[key]["time_err"]), float(off[key]["time_err"])
        assert abs(a - v * (ts[1] - ts[0])) < 1e-7
        assert abs(b - s * v * (ts[1] - ts[0])) < 1e-7
        assert on[key]["dx"] == np.float32(s) and on[key]["dy"] == 0.0


@pytest.mark.parametrize("L,w", [(5, 0.5), (6, 0.5), (5, 0.003)])
def test_swept_window_classification(L, w):
    """First swept frame at t = 0.5 with window w and v = 2: every new cell is evaluated over
    [t - w, t], so a cell is ambiguous (f = 0) iff its x-range meets [c + v (t - w), c + v t];
    cells wholly below are inside (f = -1), cells wholly above keep the new cell's f = 1 (swept:
    4030-4034).  Leaves are not re-subdivided, so their time error is the one increment
    min(|v| w, delta): capped at delta for w = 0.5, 0.006 for w = 0.003."""
    v, t = 2.0, 0.5
    tape = oracle.parse_ll(plane_ll(v, 1.0))
    fr = oracle.Frame(tape, localsubdiv=L, point_sample=False, swept_volume=True, swept_window=w,
                      voxel_capacity=1 << 20)
    fr.set_inversion(t)
    nodes = fr.export()
    fr.close()
    lo_band, hi_band = C0 + v * (t - w), C0 + v * t
    checked = 0
    for n in nodes[nodes["depth"] > 0]:
        size = 10.0 / (1 << int(n["depth"]))
        x0 = -5.0 + size * int(n["i"])
        x1 = x0 + size
        if min(abs(x0 - lo_band), abs(x0 - hi_band), abs(x1 - lo_band), abs(x1 - hi_band)) < 1e-5:
            continue   # a band edge on a cell face: left to the float32 rounding of the bound
        want = 0.0 if (x0 <= hi_band and x1 >= lo_band) else (-1.0 if x1 <= lo_band else 1.0)
        assert float(n["f"]) == want, (tuple(n)[:4], x0, x1, float(n["f"]))
        checked += 1
    assert checked > 100
    # every ambiguous cell above the leaf level was subdivided
    amb = nodes[(nodes["f"] == 0.0) & (nodes["depth"] < L) & (nodes["depth"] > 0)]
    assert np.all(amb["has_children"] == 1)
    leaves = nodes[(nodes["f"] == 0.0) & (nodes["depth"] == L)]
    assert len(leaves) > 0
    win = float(np.float32(t)) - float(np.float32(np.float32(t) - np.float32(w)))
    assert np.allclose(leaves["time_err"], min(v * win, DELTA), rtol=0, atol=1e-6)
