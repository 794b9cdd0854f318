"""The deterministic binary32 math routines shared by the oracle and the kernels
(oracle/detmath.c, temporalsetinversion_amd/csrc/tsi_interval.h), pinned against numpy
float64 and against the IEEE-754-2019 minimumNumber/maximumNumber rules that WGSL's
min/max follow."""
import math

import numpy as np
import pytest

import oracle

L = oracle.lib()
RNG = np.random.default_rng(1234)


def f32(x):
    return float(np.float32(x))


def ulp(r):
    return np.spacing(np.abs(np.asarray(r, np.float64)).astype(np.float32)).astype(np.float64)


def apply(fn, xs):
    return np.array([fn(float(x)) for x in xs], np.float64)


def test_sin_accuracy():
    xs = RNG.uniform(-10, 10, 4000).astype(np.float32)
    got, ref = apply(L.orc_det_sinf, xs), np.sin(xs.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 2 * ulp(ref) + 2 ** -26)
    xs = RNG.uniform(-1000, 1000, 4000).astype(np.float32)
    got, ref = apply(L.orc_det_sinf, xs), np.sin(xs.astype(np.float64))
    assert np.max(np.abs(got - ref)) <= 2 ** -22
    assert np.all(np.abs(got) <= 1.0)


def test_sin_special_values():
    assert L.orc_det_sinf(0.0) == 0.0
    for x in RNG.uniform(-50, 50, 200).astype(np.float32):
        assert L.orc_det_sinf(float(-x)) == -L.orc_det_sinf(float(x))      # odd
    # quadrant boundaries: sin(k pi/2) within 2^-23 of 0 / +-1
    for k in range(-8, 9):
        x = f32(k * math.pi / 2)
        assert abs(L.orc_det_sinf(x) - math.sin(x)) <= 2 ** -23


def test_exp_accuracy():
    xs = RNG.uniform(-80, 80, 4000).astype(np.float32)
    got, ref = apply(L.orc_det_expf, xs), np.exp(xs.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 2 * ulp(ref))
    assert L.orc_det_expf(0.0) == 1.0
    assert L.orc_det_expf(200.0) == math.inf
    assert L.orc_det_expf(-200.0) == 0.0


def test_log_accuracy():
    xs = np.exp(RNG.uniform(-30, 30, 4000)).astype(np.float32)
    got, ref = apply(L.orc_det_logf, xs), np.log(xs.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 2 * ulp(ref) + 2 ** -24)
    assert L.orc_det_logf(1.0) == 0.0
    assert L.orc_det_logf(0.0) == -math.inf


def test_atan2_accuracy_and_quadrants():
    ys = RNG.uniform(-5, 5, 4000).astype(np.float32)
    xs = RNG.uniform(-5, 5, 4000).astype(np.float32)
    got = np.array([L.orc_det_atan2f(float(a), float(b)) for a, b in zip(ys, xs)])
    ref = np.arctan2(ys.astype(np.float64), xs.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 3 * ulp(ref))
    for y, x, want in [(1.0, 0.0, math.pi / 2), (-1.0, 0.0, -math.pi / 2), (0.0, 1.0, 0.0), (1.0, 1.0, math.pi / 4),
                       (1.0, -1.0, 3 * math.pi / 4), (-1.0, -1.0, -3 * math.pi / 4)]:
        assert abs(L.orc_det_atan2f(y, x) - want) <= 2 ** -22


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 7, -1, -2, 12])
def test_powi(n):
    xs = RNG.uniform(0.01, 3, 500).astype(np.float32)
    got = np.array([L.orc_det_powi(float(a), n) for a in xs])
    ref = xs.astype(np.float64) ** n
    assert np.all(np.abs(got - ref) <= (abs(n) + 1) * ulp(ref))
    assert L.orc_det_powi(2.0, n) == 2.0 ** n                               # exact for powers of two


def test_pow045():
    """exp(0.45 log x): the log's absolute error is scaled by |0.45 log x| in the exp."""
    xs = np.concatenate([RNG.uniform(0.0, 3, 500), np.exp(RNG.uniform(-20, 2, 500))]).astype(np.float32)
    got = apply(L.orc_det_pow045, xs)
    ref = xs.astype(np.float64) ** 0.45
    bound = (3.0 + np.abs(np.log(np.maximum(xs.astype(np.float64), 1e-30)))) * ulp(ref)
    assert np.all(np.abs(got - ref) <= bound + 1e-30)
    assert L.orc_det_pow045(0.0) == 0.0 and L.orc_det_pow045(1.0) == 1.0


def test_min_max_are_minimum_number_maximum_number():
    nan = math.nan
    assert math.copysign(1.0, L.orc_wmin(-0.0, 0.0)) == -1.0
    assert math.copysign(1.0, L.orc_wmin(0.0, -0.0)) == -1.0
    assert math.copysign(1.0, L.orc_wmax(-0.0, 0.0)) == 1.0
    assert math.copysign(1.0, L.orc_wmax(0.0, -0.0)) == 1.0
    assert L.orc_wmin(nan, 1.0) == 1.0 and L.orc_wmin(1.0, nan) == 1.0
    assert L.orc_wmax(nan, -1.0) == -1.0 and L.orc_wmax(-1.0, nan) == -1.0
    assert math.isnan(L.orc_wmin(nan, nan))
    for a, b in RNG.normal(size=(200, 2)):
        a, b = f32(a), f32(b)
        assert L.orc_wmin(a, b) == min(a, b) and L.orc_wmax(a, b) == max(a, b)
