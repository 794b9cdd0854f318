"""ctypes binding of libtsi.so (include/tsi.h).

The library is the product: there is no Python or CPU fallback.  Loading fails loudly
when libtsi.so has not been built (run __graft_entry__.build() or
`python temporalsetinversion_amd/build.py`).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TSI_LIB", os.path.join(HERE, "_build", "libtsi.so"))
# compiled tape kernels are cached next to the library unless the caller chose a directory
JIT_CACHE = os.path.join(HERE, "_build", "jit-cache")
os.environ.setdefault("TSI_JIT_CACHE", JIT_CACHE)

TSI_OK = 0
ERRORS = {
    -1: "TSI_E_OOM", -2: "TSI_E_QUEUE_FULL", -3: "TSI_E_POOL_EXHAUSTED", -4: "TSI_E_TAPE_TOO_LONG",
    -5: "TSI_E_HIP", -6: "TSI_E_ARG", -7: "TSI_E_PARSE", -8: "TSI_E_NO_TAPE", -9: "TSI_E_UNSUPPORTED",
    -10: "TSI_E_BUDGET", -11: "TSI_E_JIT", -12: "TSI_E_PENDING", -13: "TSI_E_SCRATCH",
}
TSI_FLAG_NO_PRUNE = 1
TSI_FLAG_PRUNE_EXACT_ONLY = 2
TSI_FLAG_INTERPRET = 4
TSI_FLAG_NO_SPECULATE = 8
TSI_SHADE_NORMALS = 1
TSI_SHADE_GRADIENT = 2
TSI_SHADE_CINEMATIC = 4
TSI_FRAME_RESET = 1


class TsiError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, code)} ({code})")


class GpuInstruction(C.Structure):
    """ir.rs:88-97 GPUInstruction (32 B)."""
    _fields_ = [("index_", C.c_uint32), ("opcode_uses", C.c_uint32), ("larg_rarg", C.c_uint32),
                ("literal", C.c_float), ("val", C.c_float * 2), ("sens", C.c_float * 2)]


class Uniforms(C.Structure):
    """uniforms.rs:8-17 (96 B)."""
    _fields_ = [("view", C.c_float * 16), ("resolution", C.c_uint32 * 2), ("time", C.c_float),
                ("lasttime", C.c_float), ("rendermode", C.c_float), ("pad", C.c_float * 3)]


class Params(C.Structure):
    """compute.wgsl:3494-3511 PARAMETERS + pool sizing."""
    _fields_ = [("localsubdiv", C.c_int32), ("delta", C.c_float), ("point_sample", C.c_int32),
                ("swept_volume", C.c_int32), ("lipschitz_corrected", C.c_int32), ("fastenabled", C.c_int32),
                ("fast_multiplier", C.c_float), ("swept_window", C.c_float), ("voxel_capacity", C.c_uint64),
                ("queue_capacity", C.c_uint64), ("flags", C.c_uint32), ("prune_capacity", C.c_uint32),
                ("prune_live_max", C.c_uint32), ("shading", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("gpu_ms", C.c_double), ("eval_ms", C.c_double), ("cell_evals", C.c_uint64),
                ("rev_evals", C.c_uint64), ("fwd_ops", C.c_uint64), ("rev_ops", C.c_uint64),
                ("parents", C.c_uint64), ("deallocs", C.c_uint64), ("enqueued", C.c_uint64),
                ("allocated_nodes", C.c_uint64), ("eval_launches", C.c_uint64), ("pruned_instr", C.c_uint64),
                ("prune_launches", C.c_uint64), ("eval_kernel_ms", C.c_double), ("spec_cells", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TapeInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("stack", C.c_uint32), ("slots", C.c_uint32), ("prune_class", C.c_uint32),
                ("engine", C.c_uint32), ("cache_hit", C.c_uint32), ("specialized", C.c_uint32),
                ("fix_kept", C.c_uint32), ("compile_ms", C.c_double),
                ("log", C.c_char * 1024)]


class Node(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("i", C.c_uint32), ("j", C.c_uint32), ("k", C.c_uint32),
                ("f", C.c_float), ("t", C.c_float), ("time_err", C.c_float), ("fast", C.c_float),
                ("has_children", C.c_uint32), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float)]


class BatchConfig(C.Structure):
    _fields_ = [("arena_bytes", C.c_uint64), ("max_sessions", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("claim_units", C.c_uint32)]


class SessionDesc(C.Structure):
    _fields_ = [("tape", C.c_uint32), ("params", Params), ("root_box", C.c_float * 6), ("max_width", C.c_uint32),
                ("max_height", C.c_uint32)]


class FrameReq(C.Structure):
    _fields_ = [("session", C.c_uint32), ("time", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32),
                ("view", C.c_float * 16), ("flags", C.c_uint32)]


class FrameResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("pad", C.c_uint32), ("stats", Stats), ("node_visits", C.c_uint64)]


class BatchInfo(C.Structure):
    _fields_ = [("tapes", C.c_uint32), ("grid_blocks", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("stack_rows", C.c_uint32), ("sessions_open", C.c_uint32), ("cache_hit", C.c_uint32),
                ("max_waves_per_simd", C.c_uint32), ("claim_units", C.c_uint32), ("compile_ms", C.c_double), ("arena_bytes", C.c_uint64), ("arena_used", C.c_uint64),
                ("stack_bytes", C.c_uint64), ("log", C.c_char * 1024)]


# name -> (restype, argtypes); the list every test checks against include/tsi.h
SIGNATURES = {
    "tsi_default_params": (None, [C.POINTER(Params), C.c_int32]),
    "tsi_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "tsi_destroy": (None, [C.c_void_p]),
    "tsi_upload_tape": (C.c_int, [C.c_void_p, C.POINTER(GpuInstruction), C.c_uint32]),
    "tsi_reset": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "tsi_set_uniforms": (C.c_int, [C.c_void_p, C.POINTER(Uniforms)]),
    "tsi_set_inversion": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(Stats)]),
    "tsi_set_inversion_async": (C.c_int, [C.c_void_p]),
    "tsi_set_inversion_result": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(Stats)]),
    "tsi_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tsi_trace_async": (C.c_int, [C.c_void_p]),
    "tsi_trace_wait": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tsi_trace_totals": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    "tsi_frame_rgba8": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tsi_frame_rgba8_noalloc": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tsi_save_frame_png": (C.c_int, [C.c_void_p, C.c_char_p]),
    "tsi_write_png": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "tsi_frame_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "tsi_enqueue_root": (C.c_int, [C.c_void_p]),
    "tsi_read_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "tsi_read_tape": (C.c_int, [C.c_void_p, C.POINTER(GpuInstruction), C.c_uint32, C.POINTER(C.c_uint32)]),
    "tsi_tape_from_ll": (C.c_int, [C.c_char_p, C.c_size_t, C.c_float, C.POINTER(GpuInstruction), C.c_uint32,
                                   C.POINTER(C.c_uint32)]),
    "tsi_camera_zdist": (C.c_float, [C.c_uint32]),
    "tsi_tape_info_get": (C.c_int, [C.c_void_p, C.POINTER(TapeInfo)]),
    "tsi_last_trace_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "tsi_precompile_tape": (C.c_int, [C.POINTER(GpuInstruction), C.c_uint32, C.c_uint32, C.POINTER(TapeInfo)]),
    "tsi_tape_kernel_source": (C.c_int, [C.POINTER(GpuInstruction), C.c_uint32, C.c_uint32, C.c_char_p, C.c_size_t,
                                         C.POINTER(C.c_size_t), C.c_void_p]),
    "tsi_strerror": (C.c_char_p, [C.c_int]),
    "tsi_set_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "tsi_batch_create": (C.c_int, [C.POINTER(C.POINTER(GpuInstruction)), C.POINTER(C.c_uint32), C.c_uint32,
                                   C.POINTER(BatchConfig), C.POINTER(C.c_void_p)]),
    "tsi_batch_destroy": (None, [C.c_void_p]),
    "tsi_batch_precompile": (C.c_int, [C.POINTER(C.POINTER(GpuInstruction)), C.POINTER(C.c_uint32), C.c_uint32,
                                       C.POINTER(BatchInfo)]),
    "tsi_batch_info_get": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "tsi_batch_tune": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig)]),
    "tsi_batch_footprint": (C.c_int, [C.c_void_p, C.POINTER(SessionDesc), C.POINTER(C.c_uint64)]),
    "tsi_batch_open": (C.c_int, [C.c_void_p, C.POINTER(SessionDesc), C.POINTER(C.c_uint32)]),
    "tsi_batch_close": (C.c_int, [C.c_void_p, C.c_uint32]),
    "tsi_batch_reset": (C.c_int, [C.c_void_p, C.c_uint32]),
    "tsi_batch_submit": (C.c_int, [C.c_void_p, C.POINTER(FrameReq), C.c_uint32]),
    "tsi_batch_query": (C.c_int, [C.c_void_p]),
    "tsi_batch_wait": (C.c_int, [C.c_void_p, C.POINTER(FrameResult), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tsi_batch_read_nodes": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "tsi_batch_frame": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != TSI_OK:
        raise TsiError(rc, what)
