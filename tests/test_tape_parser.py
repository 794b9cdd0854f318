"""Tape builder (src/ir.rs:190-366): the oracle's restatement and the library's
tsi_tape_from_ll, pinned by a hand-derived known-answer tape, by the reference's own scene
files through an independent interpreter, and by each other byte for byte."""
import json
import os

import numpy as np
import pytest

import oracle
import temporalsetinversion_amd as T
from temporalsetinversion_amd import scenes
from temporalsetinversion_amd._lib import TsiError

import llinterp

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "reference_scenes.json")

OP = llinterp.OPS

# A module in clang's dialect.  The expected tape follows ir.rs by hand:
#   fmul %0, %0         -> SQR (ir.rs:305-309): rarg becomes literal 0, side 1, %0 used once
#   fsub 0.0, %1        -> SUB, literal left (side 2), literal 0
#   llvm.maxnum(%5, %6) -> MAX (tail, binary: tokens 6 and 8), registers reindexed %k -> k-1
#   fadd %7, 0xBFE0..   -> ADD, side 1, literal -0.5 (double bits -> f32)
#   ret %8              -> RET, side 1, dummy use (ir.rs:353)
KAT_LL = """; ModuleID = 'kat.cpp'
define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) local_unnamed_addr #0 {
  %5 = fmul float %0, %0
  %6 = fsub float 0.000000e+00, %1
  %7 = tail call float @llvm.maxnum.f32(float %5, float %6) #9
  %8 = fadd float %7, 0xBFE0000000000000
  ret float %8
}
"""
#            index op  side uses larg rarg literal
KAT = [(0, 1, 3, 1, 0, 0, 0.0),
       (1, 2, 3, 1, 0, 0, 0.0),
       (2, 3, 3, 0, 0, 0, 0.0),
       (3, 4, 3, 0, 0, 0, 0.0),
       (4, OP["SQR"], 1, 1, 0, 0, 0.0),
       (5, OP["SUB"], 2, 1, 0, 1, 0.0),
       (6, OP["MAX"], 0, 1, 4, 5, 0.0),
       (7, OP["ADD"], 1, 1, 6, 0, -0.5),
       (8, OP["RET"], 1, 1, 7, 0, 0.0)]


def fields(tape):
    ou = tape["opcode_uses"].astype(np.int64)
    lr = tape["larg_rarg"].astype(np.int64)
    return dict(index=tape["index_"].astype(np.int64), op=(ou >> 16) & 0x3FFF, side=(ou >> 30) & 3,
                uses=ou & 0xFFFF, larg=lr >> 16, rarg=lr & 0xFFFF, lit=tape["literal"])


@pytest.mark.parametrize("builder", ["oracle", "library"])
def test_known_answer_tape(builder):
    t = 0.25
    tape = oracle.parse_ll(KAT_LL, t) if builder == "oracle" else T.tape_from_ll(KAT_LL, t)
    assert len(tape) == len(KAT)
    f = fields(tape)
    for row, exp in enumerate(KAT):
        got = tuple(int(f[k][row]) for k in ("index", "op", "side", "uses", "larg", "rarg")) + (float(f["lit"][row]),)
        assert got == exp, f"record {row}: {got} != {exp}"
    # ir.rs:343-348: initial box [-5, 5]^3 and [t, t]; body values zero; sens zero
    assert tape["val"][:3].tolist() == [[-5.0, 5.0]] * 3
    assert tape["val"][3].tolist() == [t, t]
    assert not tape["val"][4:].any() and not tape["sens"].any()


def _reference_scene(reference_dir, name):
    with open(os.path.join(reference_dir, "scenes", name + ".ll")) as fh:
        return fh.read()


def _fixture():
    with open(FIXTURE) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", ["cables", "cradle", "piano", "sponza"])
def test_reference_scenes_structure(reference_dir, name):
    """Opcode histogram, length and use counts of the reference's own scenes."""
    fx = _fixture()[name]
    text = _reference_scene(reference_dir, name)
    tape = oracle.parse_ll(text)
    f = fields(tape)
    assert len(tape) == 4 + fx["body"] + 1
    hist = {}
    for op in f["op"][4:]:
        hist[int(op)] = hist.get(int(op), 0) + 1
    assert hist == {OP[k]: v for k, v in fx["histogram"].items()}
    uses = fx["uses"]
    for ssa in range(len(tape) + 1):
        idx = ssa if ssa < 4 else ssa - 1        # %k -> k-1 for k > 3 (%4 is the entry label)
        if ssa == 4 or idx >= len(tape) - 1:
            continue
        assert int(f["uses"][idx]) == uses.get(str(ssa), 0), f"uses of %{ssa}"
    assert int(f["uses"][-1]) == 1             # RET's dummy use


@pytest.mark.parametrize("name", ["cables", "cradle", "piano", "sponza"])
def test_reference_scenes_evaluate(reference_dir, name):
    """The tape evaluated on degenerate boxes equals the independent float64 interpreter."""
    fx = _fixture()[name]
    tape = oracle.parse_ll(_reference_scene(reference_dir, name))
    compared = 0
    for p, v, m in zip(fx["points"], fx["values"], fx["margins"]):
        if m < 1e-3:
            continue
        (lo, hi), *_ = oracle.eval_tape(tape, [(c, c) for c in p], False)
        assert abs(hi - lo) <= 1e-5 * max(1.0, abs(v)), (p, lo, hi)
        assert abs(lo - v) <= 2e-3 * max(1.0, abs(v)), f"{name} at {p}: tape {lo} vs {v}"
        compared += 1
    assert compared >= 32


@pytest.mark.parametrize("name", ["cables", "cradle", "piano", "sponza"])
def test_library_parser_matches_oracle_on_reference_scenes(reference_dir, name):
    text = _reference_scene(reference_dir, name)
    for t in (0.0, 0.37):
        assert T.tape_from_ll(text, t).tobytes() == oracle.parse_ll(text, t).tobytes()


@pytest.mark.parametrize("scene", sorted(scenes.SCENES))
def test_generated_scenes(scene):
    """Generated scenes: both builders agree, and the tape evaluates to the interpreter's value."""
    text = scenes.SCENES[scene]()
    tape = oracle.parse_ll(text)
    assert T.tape_from_ll(text).tobytes() == tape.tobytes()
    prog = llinterp.Program(text)
    assert len(tape) == 4 + len(prog.ins) + 1
    rng = np.random.default_rng(7)
    for _ in range(12):
        p = [float(np.float32(x)) for x in (*rng.uniform(-1.2, 1.2, 3), rng.uniform(0, 1))]
        v, m = prog.eval(*p)
        if m < 1e-3:
            continue
        (lo, hi), *_ = oracle.eval_tape(tape, [(c, c) for c in p], False)
        assert abs(lo - v) <= 2e-3 * max(1.0, abs(v)), f"{scene} at {p}: {lo} vs {v}"


BAD = [
    "no field here",
    KAT_LL.replace("%5 = fmul float %0, %0", "%5 = fmul float %0, %zz"),
    KAT_LL.replace("0xBFE0000000000000", "0xZZ"),
    KAT_LL.replace("tail call float @llvm.maxnum.f32(float %5, float %6)",
                   "tail call float @llvm.maxnum.f32(float 1.0, float %6)"),   # ir.rs:247 keeps the ','
]


@pytest.mark.parametrize("k", range(len(BAD)))
def test_malformed_modules_are_rejected_by_both(k):
    """Where ir.rs would panic, both builders return an error."""
    with pytest.raises(ValueError):
        oracle.parse_ll(BAD[k])
    with pytest.raises(TsiError) as e:
        T.tape_from_ll(BAD[k])
    assert e.value.code == -7


def test_long_tapes_are_reported():
    """ir.rs:355-358 silently keeps the first 4096 records (a tape whose RET is cut off);
    both builders report the overflow instead."""
    body = ["  %5 = fadd float %0, 1.000000e+00"]
    for k in range(6, 4200):
        body.append(f"  %{k} = fadd float %{k - 1}, 1.000000e+00")
    text = ("define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) #0 {\n" + "\n".join(body)
            + f"\n  ret float %{4199}\n}}\n")
    with pytest.raises(OverflowError):
        oracle.parse_ll(text)
    with pytest.raises(TsiError) as e:
        T.tape_from_ll(text)
    assert e.value.code == -4
