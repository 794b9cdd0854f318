"""The C ABI of include/tsi.h: every declared function is exported by libtsi.so and bound by
the ctypes layer, and the ctypes struct layouts equal the C ones (checked by compiling a probe
against the header with gcc).  No device is needed: nothing here launches a kernel."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import temporalsetinversion_amd as T
from temporalsetinversion_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tsi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(tsi_\w+)\s*\(", text, flags=re.M))


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    for must in ("tsi_create", "tsi_destroy", "tsi_upload_tape", "tsi_reset", "tsi_set_uniforms",
                 "tsi_set_inversion", "tsi_trace", "tsi_enqueue_root", "tsi_tape_from_ll"):
        assert must in names


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    lib = _lib.lib()
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} not exported by libtsi.so"
    assert set(_lib.SIGNATURES) == names, (set(_lib.SIGNATURES) ^ names)


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in out


STRUCTS = {"tsi_gpu_instruction": _lib.GpuInstruction, "tsi_uniforms": _lib.Uniforms, "tsi_params": _lib.Params,
           "tsi_stats": _lib.Stats, "tsi_node": _lib.Node, "tsi_tape_info": _lib.TapeInfo}


def test_struct_layouts_match_the_header(tmp_path):
    src = tmp_path / "probe.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tsi.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'    printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                check=True).stdout.splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"
    assert C.sizeof(_lib.GpuInstruction) == 32 and C.sizeof(_lib.Uniforms) == 96   # ir.rs / uniforms.rs


def test_host_only_calls():
    lib = _lib.lib()
    p = _lib.Params()
    lib.tsi_default_params(C.byref(p), 8)
    assert (p.localsubdiv, p.point_sample, p.fastenabled) == (8, 1, 1)
    assert abs(p.delta - 0.01) < 1e-9 and abs(p.swept_window - 1.0 / 60.0) < 1e-9
    assert lib.tsi_strerror(-3) == b"voxel pool exhausted"
    assert lib.tsi_strerror(-11).startswith(b"the tape's evaluation kernels")
    # compute.wgsl:4592: res.y / tan(radians(90) / 2) == res.y (to f32 rounding)
    assert abs(T.camera_zdist(1080) - 1080.0) < 1e-3
    assert lib.tsi_create(None, None) == -6


def _has_gpu():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a GPU")
def test_no_silent_fallback_without_a_gpu():
    """Without a device the product raises; it never evaluates on the CPU."""
    with pytest.raises(_lib.TsiError) as e:
        T.State(_sphere_ll())
    assert e.value.code == -5


def _sphere_ll():
    from temporalsetinversion_amd import scenes

    return scenes.sphere_unit(1.0)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
