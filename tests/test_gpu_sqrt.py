"""The interval square root (SQRT_I) against the compiler's correctly rounded sqrtf, for every
binary32 in [-0, +inf] and every NaN: sqrt_nonneg leaves out the lowering's last step (x itself for +-0 and
+inf), which is redundant on that domain, and this pins that claim on the hardware whose square
root instruction the method corrects."""
import ctypes as C

import pytest

from temporalsetinversion_amd import _lib

TSI_E_ARG = -6


def _check(first: int, count: int):
    L = _lib.lib()
    f = L.tsi_check_sqrt
    f.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    f.restype = C.c_int
    bad, first_bad, checked = C.c_uint64(), C.c_uint32(), C.c_uint32()
    _lib.check(f(first, count, C.byref(bad), C.byref(first_bad), C.byref(checked)), "tsi_check_sqrt")
    return bad.value, first_bad.value, checked.value


@pytest.mark.gpu
def test_sqrt_nonneg_equals_sqrtf_on_every_nonnegative_binary32():
    # +0 .. +inf: patterns 0x00000000 .. 0x7F800000 (denormals, the scaled range below 2^-96,
    # normals and +inf), every one compared
    print("tsi_check_sqrt 0x00000000 +0x7F800001", flush=True)
    assert _check(0x00000000, 0x7F800001) == (0, 0xFFFFFFFF, 0x7F800001)
    # -0
    assert _check(0x80000000, 1) == (0, 0xFFFFFFFF, 1)
    # NaNs of either sign (the trace's vector lengths take sqrt_nonneg too, whose argument is a
    # sum of squares: >= +0, +inf or NaN)
    assert _check(0x7F800001, 0x007FFFFF) == (0, 0xFFFFFFFF, 0x007FFFFF)
    assert _check(0xFF800001, 0x007FFFFF) == (0, 0xFFFFFFFF, 0x007FFFFF)


def test_sqrt_check_rejects_bad_ranges():   # argument checks come before any device call
    L = _lib.lib()
    f = L.tsi_check_sqrt
    f.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    f.restype = C.c_int
    a, b, c = C.c_uint64(), C.c_uint32(), C.c_uint32()
    assert f(0, 0, C.byref(a), C.byref(b), C.byref(c)) == TSI_E_ARG
    assert f(0xFFFFFFF0, 32, C.byref(a), C.byref(b), C.byref(c)) == TSI_E_ARG
    assert f(0, 0x80000001, C.byref(a), C.byref(b), C.byref(c)) == TSI_E_ARG
