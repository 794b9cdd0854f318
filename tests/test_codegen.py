"""The tape code generator's range analysis (where fix() is provably the identity), run through
tsi_precompile_tape on the CPU (hiprtc needs no GPU).

fix (compute.wgsl:3462-3472) clamps to +-1e6; on the default domain (cells in [-5,5]^3,
t-windows in [-2, 3]) a value is left unfixed only when the interval rules, evaluated over the
whole domain, keep it inside the clamp.  Closed forms pin the count: x^(2^k) over [-5, 5] is at
most 5^(2^k) = 25, 625, 390625, 1.5e11, ... so the first three squarings are fix-free and every
record from the fourth on keeps its fix; 1/x (a division by an interval containing 0) is never
bounded, and neither is anything computed from it.
"""
import pytest

import temporalsetinversion_amd as T


def _tape(body, ret):
    return ("define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) #0 {\n" + body +
            f"\n  ret float {ret}\n}}\n")


def squarings(k):
    lines, prev = [], "%0"
    for i in range(k):
        lines.append(f"  %{5 + i} = fmul float {prev}, {prev}")
        prev = f"%{5 + i}"
    return _tape("\n".join(lines), prev)


@pytest.mark.parametrize("k,kept", [(3, 0), (4, 2), (6, 4)])
def test_repeated_squaring_keeps_fix_from_the_overflowing_record_on(k, kept):
    # k SQR records + RET; from the 4th square on the bound exceeds 1e6: k - 3 squares + RET keep fix
    info = T.precompile(squarings(k), prune=False)
    assert info["specialized"] and info["n"] == 4 + k + 1
    assert info["fix_kept"] == kept


def test_division_by_an_interval_is_never_bounded():
    text = _tape("  %5 = fdiv float 1.000000e+00, %0\n  %6 = fadd float %5, 1.000000e+00\n"
                 "  %7 = fmul float %1, 2.000000e+00", "%6")
    info = T.precompile(text, prune=False)
    # 1/x (DIV_LR over [-5, 5]), x + 1 after it and the RET keep fix; y * 2 does not
    assert info["fix_kept"] == 3


def test_bounded_scene_tapes_are_fix_free():
    from temporalsetinversion_amd import scenes
    for name in ("spheres", "capsules"):
        info = T.precompile(scenes.SCENES[name](), prune=False)
        assert info["specialized"] and info["fix_kept"] == 0, name
