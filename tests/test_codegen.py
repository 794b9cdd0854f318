"""The tape code generator's range analysis (where fix() is provably the identity), run through
tsi_precompile_tape on the CPU (hiprtc needs no GPU).

fix (compute.wgsl:3462-3472) clamps to +-1e6; on the default domain (cells in [-5,5]^3,
t-windows in [-2, 64]) a value is left unfixed only when the interval rules, evaluated over the
whole domain, keep it inside the clamp.  Closed forms pin the count: x^(2^k) over [-5, 5] is at
most 5^(2^k) = 25, 625, 390625, 1.5e11, ... so the first three squarings are fix-free and every
record from the fourth on keeps its fix; 1/x (a division by an interval containing 0) is never
bounded, and neither is anything computed from it.
"""
import re

import numpy as np
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


# ---- addressing of the generated modules (the pruned csg module once faulted the GPU at
# localsubdiv 10 with a variant of its stack loads that is no longer generated) ----

def _layout(tape, prune_group=16):
    """Record words and choice columns of the pruned layout, counted from the tape itself."""
    ops = (tape["opcode_uses"] >> 16) & 0x3FFF
    side = (tape["opcode_uses"] >> 30) & 3
    body = slice(4, len(tape))
    nfwd = len(tape) - 4
    minmax = int(np.sum(((ops[body] == 17) | (ops[body] == 18)) & (side[body] <= 1)))
    regions = (nfwd + prune_group - 1) // prune_group
    return 2 + (regions + 31) // 32 + (minmax + 15) // 16, (minmax + 15) // 16


def _functions(src):
    """{name: body} of the generated __device__ functions."""
    out = {}
    for m in re.finditer(r"__device__ __noinline__ \w+ (\w+)\(.*?\) \{\n", src):
        depth, i = 1, m.end()
        while depth:
            c = src[i]
            depth += (c == "{") - (c == "}")
            i += 1
        out.setdefault(m.group(1), []).append(src[m.end():i - 1])
    return out


@pytest.mark.parametrize("scene", ["spheres", "capsules", "csg"])
@pytest.mark.parametrize("prune", [True, False], ids=["pruned", "unpruned"])
def test_generated_modules_stay_inside_their_buffers(scene, prune):
    from temporalsetinversion_amd import scenes

    tape = T.tape_from_ll(scenes.SCENES[scene]())
    src, info = T.kernel_source(tape, prune=prune)
    rows = info["stack"] + 1                       # rows per wavefront tsi_upload_tape allocates
    used = [int(x) for x in re.findall(r"\(stk \+ (\d+)u\)\[lane\]", src)]
    assert used and all(u % 64 == 0 for u in used)
    assert max(used) // 64 < rows, (max(used) // 64, rows)
    record_words, mode_words = _layout(tape)
    words = [int(x) for x in re.findall(r"\brec\[size_t\((\d+)u\) \* cap\]", src)]
    words += [int(x) for x in re.findall(r"\bprec\[size_t\((\d+)u\) \* cap\]", src)]
    cols = [int(x) for x in re.findall(r"chbuf\[(\d+)u\]", src)]
    if prune and info["prune_class"]:
        assert words and cols
        assert max(words) < record_words and min(words) >= 2
        assert all(c % 64 == 0 for c in cols) and max(cols) // 64 < max(mode_words, 1)
    else:
        assert not words and not cols
    # every saved value / code word is loaded before it is read; in unpruned modules the loads
    # sit at function scope, so they precede every read on every path
    for name, bodies in _functions(src).items():
        for body in bodies:
            if "RevOut" not in src or not name.startswith("tape_reverse"):
                continue
            first_load = {}
            depth = 1
            for line in body.split("\n"):
                m = re.match(r"\s*(sv\d+|kr\d+) = ", line)
                if m and m.group(1) not in first_load:
                    first_load[m.group(1)] = depth
                for v in re.findall(r"\b(sv\d+)\.x|\b(kr\d+) >>", line):
                    v = v[0] or v[1]
                    assert v in first_load, f"{name}: {v} read before any load"
                depth += line.count("{") - line.count("}")
            if not prune:
                assert set(first_load.values()) == {1}, f"{name}: a load inside a region block"


# ---- the reverse pass's skipped subtrees (unpruned modules) ----

# forms whose product with a zero sensitivity can be NaN (an infinite derivative bound): a record
# of such a form that is not short-circuited (no z_J test: an operand with several uses) must
# never sit inside a subtree the reverse pass may skip, or skipping would drop that NaN
_UNSAFE_FORMS = {6, 7, 8, 9, 16, 18, 21, 22, 25, 26, 27, 32, 33, 34}


def _guarded_rules(src):
    """(record, form, guarded) of the rev_form_t rules of tape_reverse, guarded when the rule sits
    in an `if (lv<g>) {` block (not in its else part)."""
    body = _functions(src)["tape_reverse"][0]
    out, stack, depth = [], [], 0
    for line in body.split("\n"):
        s = line.strip()
        if s == "} else {" and stack and stack[-1][1] == depth:
            stack[-1] = ("else", depth)
        m = re.match(r"const Iv2 d_(\d+) = rev_form_t<true>\((\d+)u", s)
        if m:
            out.append((int(m.group(1)), int(m.group(2)), any(k == "guard" for k, _ in stack)))
        if re.match(r"if \(lv\d+\) \{$", s):
            stack.append(("guard", depth + 1))
        depth += line.count("{") - line.count("}")
        while stack and depth < stack[-1][1]:
            stack.pop()
    return out, body


@pytest.mark.parametrize("scene", ["mixed-11", "mixed-12", "mixed-13", "mixed-14", "csg", "capsules"])
def test_skipped_subtrees_hold_no_unsafe_rule(scene):
    from temporalsetinversion_amd import scenes
    if scene.startswith("mixed"):
        seed = int(scene.split("-")[1])
        text = scenes.mixed(seed, dict(scenes.mixed_scenes())[seed])
    else:
        text = scenes.SCENES[scene]()
    src, _ = T.kernel_source(T.tape_from_ll(text), prune=False)
    rules, body = _guarded_rules(src)
    assert "if (lv" in body, "no skipped subtree generated"
    assert any(g for _, _, g in rules), "no rule inside a skipped subtree"
    short = set(int(x) for x in re.findall(r"const bool z_(\d+) =", body))
    for j, form, guarded in rules:
        if guarded and form in _UNSAFE_FORMS:
            assert j in short, f"record {j} (form {form}) without a short-circuit inside a skipped subtree"


def test_mixed_scenes_hold_unsafe_rules_outside_skipped_subtrees():
    """The check above is not vacuous: the mixed scenes hold records of such forms without a
    short-circuit (a sine whose operand has two uses), all outside the skipped subtrees."""
    from temporalsetinversion_amd import scenes
    found = 0
    for seed, n in scenes.mixed_scenes():
        src, _ = T.kernel_source(T.tape_from_ll(scenes.mixed(seed, n)), prune=False)
        rules, body = _guarded_rules(src)
        short = set(int(x) for x in re.findall(r"const bool z_(\d+) =", body))
        unsafe = [(j, g) for j, form, g in rules if form in _UNSAFE_FORMS and j not in short]
        assert not any(g for _, g in unsafe)
        found += len(unsafe)
    assert found > 0
