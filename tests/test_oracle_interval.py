"""Pins of the oracle's interval rules (oracle/interval.c, compute.wgsl:2481-2975) by what
the mathematics fixes: for the rational forms the interval result is exactly the hull of the
same rule applied to the corner points (rounding is monotone and the intervals are not rounded
outward); every result contains the point results of its members; the reverse pass equals
float64 finite differences on degenerate boxes; MIN/MAX choices and the short-circuit rule
change the counted reverse work exactly as compute.wgsl:3786-3963 says."""
import math

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle
from temporalsetinversion_amd import scenes

import llinterp

L = oracle.lib()
Iv = oracle.Iv
OPS = llinterp.OPS

finite = st.floats(min_value=-50.0, max_value=50.0, allow_nan=False, width=32)


def iv(a, b):
    return (min(a, b), max(a, b))


def fwd(op, side, l, r=(0.0, 0.0), lit=0.0):
    v = L.orc_forward_op(op, side, Iv(*l), Iv(*r), float(np.float32(lit)))
    return (v.lo, v.hi)


def point(op, side, x, y=0.0, lit=0.0):
    lo, hi = fwd(op, side, (x, x), (y, y), lit)
    assert lo == hi or (math.isnan(lo) and math.isnan(hi))
    return lo


TWO = [("ADD", 0), ("SUB", 0), ("MUL", 0), ("MIN", 0), ("MAX", 0)]
ONE_LIT = [("ADD", 1), ("SUB", 1), ("MUL", 1), ("DIV", 1), ("MIN", 1), ("MAX", 1)]


@settings(max_examples=300, deadline=None)
@given(finite, finite, finite, finite)
def test_two_register_forms_are_the_hull_of_their_corners(a, b, c, d):
    l, r = iv(a, b), iv(c, d)
    for name, side in TWO:
        got = fwd(OPS[name], side, l, r)
        corners = [point(OPS[name], side, x, y) for x in l for y in r]
        assert got == (min(corners), max(corners)), (name, l, r, got, corners)


literal = st.one_of(st.floats(min_value=2.0 ** -10, max_value=8.0, width=32), st.floats(min_value=-8.0, max_value=-(2.0 ** -10),
                                                                                   width=32), st.just(0.0))


@settings(max_examples=300, deadline=None)
@given(finite, finite, literal)
def test_register_literal_forms_are_the_hull_of_their_corners(a, b, c):
    l = iv(a, b)
    for name, side in ONE_LIT:
        if name == "DIV" and c == 0.0:
            continue
        got = fwd(OPS[name], side, l, lit=c)
        corners = [point(OPS[name], side, x, lit=c) for x in l]
        assert got == (min(corners), max(corners)), (name, l, c, got)
    # literal - register and literal / register (side 2)
    r = l
    got = fwd(OPS["SUB"], 2, (0.0, 0.0), r, lit=c)
    corners = [point(OPS["SUB"], 2, 0.0, y, lit=c) for y in r]
    assert got == (min(corners), max(corners))


@settings(max_examples=300, deadline=None)
@given(finite, finite)
def test_unary_forms(a, b):
    l = iv(a, b)
    zero_inside = l[0] <= 0.0 <= l[1]
    for name, extra in (("SQR", [0.0] if zero_inside else []), ("ABS", [0.0] if zero_inside else []),
                        ("NEG", []), ("SQRT", [])):
        got = fwd(OPS[name], 1, l)
        pts = [point(OPS[name], 1, x) for x in l] + extra
        assert got == (min(pts), max(pts)), (name, l, got, pts)
    # FLOOR
    assert fwd(OPS["FLOOR"], 1, l) == (math.floor(l[0]), math.floor(l[1]))


@settings(max_examples=200, deadline=None)
@given(st.floats(min_value=-20.0, max_value=20.0, allow_nan=False, width=32),
       st.floats(min_value=0.0, max_value=3.0, allow_nan=False, width=32))
def test_transcendental_forms_contain_their_members(a, w):
    l = (a, float(np.float32(a + w)))
    tol = 4 * 2.0 ** -23
    for name, f in (("SIN", math.sin), ("COS", math.cos), ("EXP", math.exp)):
        if name == "EXP" and l[1] > 13.0:   # exp stays below BIG (fix does not clamp lo from above)
            continue
        lo, hi = fwd(OPS[name], 1, l)
        for x in np.linspace(l[0], l[1], 33):
            v = min(max(f(float(np.float32(x))), -1.0e6), 1.0e6)   # fix (compute.wgsl:3462)
            scale = max(1.0, abs(v))
            # COS_I is SIN_I(a + pi/2) with a binary32 add (compute.wgsl:2959): the shifted
            # argument is rounded, so allow one ulp of it on top
            t = tol * scale + (float(np.spacing(np.float32(abs(x) + 2.0))) if name == "COS" else 0.0)
            assert lo - t <= v <= hi + t, (name, l, lo, hi, x, v)
        if name in ("SIN", "COS"):
            assert -1.0 - tol <= lo and hi <= 1.0 + tol


def test_fix_clamps_and_replaces_nan():
    big = 1.0e6
    # lo is clamped from below only and hi from above only (compute.wgsl:3462-3472)
    assert fwd(OPS["MUL"], 1, (2.0e5, 3.0e5), lit=10.0) == (2.0e6, big)
    assert fwd(OPS["MUL"], 0, (-1.0e4, 1.0e4), (-1.0e4, 1.0e4)) == (-big, big)
    lo, hi = fwd(OPS["DIV"], 1, (0.0, 0.0), lit=0.0)                           # 0/0 = NaN
    assert (lo, hi) == (-big, big)
    assert fwd(99, 0, (1.0, 2.0), (3.0, 4.0)) == (-big, -math.inf)            # unknown op: [-inf, -inf], lo clamped
    lo, hi = fwd(OPS["DIV"], 1, (1.0, 1.0), lit=4e-27)
    assert hi == big and lo > big


def _tape_eval(tape, p, deriv=True):
    return oracle.eval_tape(tape, [(c, c) for c in p], deriv)


@pytest.mark.parametrize("scene", ["spheres_linear", "capsules"])
def test_reverse_pass_is_the_gradient(scene):
    """On degenerate boxes the sensitivities of x, y, z, t equal float64 central differences."""
    text = scenes.SCENES[scene]()
    tape = oracle.parse_ll(text)
    prog = llinterp.Program(text)
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(40):
        p = [float(np.float32(v)) for v in (*rng.uniform(-1.0, 1.0, 3), rng.uniform(0.0, 1.0))]
        (lo, hi), vals, sens, fo, ro = _tape_eval(tape, p)
        h = 1e-6
        grads = []
        for k in range(4):
            q1, q0 = list(p), list(p)
            q1[k] += h
            q0[k] -= h
            grads.append((prog.eval(*q1)[0] - prog.eval(*q0)[0]) / (2 * h))
        # skip points next to a MIN/MAX switch (the derivative jumps there)
        q = [prog.eval(*[c + dv for c, dv in zip(p, rng.normal(0, 1e-5, 4))])[0] for _ in range(3)]
        if max(abs(x - lo) for x in q) > 1e-3:
            continue
        for k in range(4):
            s = 0.5 * (float(sens[k, 0]) + float(sens[k, 1]))
            assert abs(s - grads[k]) <= 2e-3 * max(1.0, abs(grads[k])), (scene, p, k, s, grads[k])
        checked += 1
    assert checked >= 20


def _mini_tape(body):
    text = ("define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) #0 {\n" + body + "\n}\n")
    return oracle.parse_ll(text)


def test_choices_and_short_circuit_decide_the_reverse_work():
    """f = max(2x, 5): while 2x < 5 the choice gives the product a zero sensitivity, and as x is
    used once the product's rule is short-circuited (compute.wgsl:3802-3805): its work is not
    counted and x's sensitivity is exactly [0, 0]."""
    tape = _mini_tape("  %5 = fmul float %0, 2.000000e+00\n"
                      "  %6 = tail call float @llvm.maxnum.f32(float %5, float 5.000000e+00) #9\n"
                      "  ret float %6")
    rc = [L.orc_rev_cost(OPS[n], 1) for n in ("MUL", "MAX", "RET")]
    args_low = [(0.0, 1.0), (0.0, 0.0), (0.0, 0.0), (0.0, 0.0)]
    (_, _), _, sens, _, ro = oracle.eval_tape(tape, args_low, True)
    assert ro == rc[1] + rc[2]
    assert sens[0].tolist() == [0.0, 0.0]
    args_high = [(3.0, 4.0), (0.0, 0.0), (0.0, 0.0), (0.0, 0.0)]
    (_, _), _, sens, _, ro = oracle.eval_tape(tape, args_high, True)
    assert ro == rc[0] + rc[1] + rc[2]
    assert sens[0].tolist() == [2.0, 2.0]
    # straddling: the undecided choice passes [1, 1] to both sides
    (_, _), _, sens, _, _ = oracle.eval_tape(tape, [(2.0, 3.0)] + args_low[1:], True)
    assert sens[0].tolist() == [2.0, 2.0]


def test_prune_keeps_exactly_the_chosen_operands():
    tape = _mini_tape("  %5 = fadd float %0, 1.000000e+00\n"
                      "  %6 = fadd float %1, 1.000000e+00\n"
                      "  %7 = tail call float @llvm.minnum.f32(float %5, float %6) #9\n"
                      "  ret float %7")
    (_, _), vals, _, _, _ = oracle.eval_tape(tape, [(0.0, 1.0), (5.0, 6.0), (0.0, 0.0), (0.0, 0.0)], False)
    live, choice = oracle.prune(tape, vals)
    assert live.tolist() == [1, 0, 0, 0, 1, 0, 1, 1]
    assert choice[6] == 1
    (_, _), vals, _, _, _ = oracle.eval_tape(tape, [(0.0, 1.0), (0.5, 6.0), (0.0, 0.0), (0.0, 0.0)], False)
    live, choice = oracle.prune(tape, vals)
    assert live.tolist() == [1, 1, 0, 0, 1, 1, 1, 1] and choice[6] == 3
