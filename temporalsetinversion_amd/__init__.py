"""MI355X temporal set inversion: host mirror of the reference's interface over libtsi.so.

The reference's host (src/main.rs, src/ir.rs) builds a tape from clang `.ll` text and
drives three compute entry points per frame through wgpu.  This package keeps that
interface -- `IR` for the tape builder, `State` with `update` / `compute` / `render`
for the frame protocol -- and calls the HIP kernels through the C ABI of
include/tsi.h.  There is no CPU fallback: every call goes to the device library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import GpuInstruction, Params, Stats, TsiError, Uniforms, check, lib

__all__ = ["IR", "State", "Params", "Stats", "TsiError", "look_at_inverse", "diagonal_camera", "tape_from_ll",
           "INSTR_DTYPE", "NODE_DTYPE", "OPCODES"]

INSTR_DTYPE = np.dtype([("index_", "<u4"), ("opcode_uses", "<u4"), ("larg_rarg", "<u4"), ("literal", "<f4"),
                        ("val", "<f4", 2), ("sens", "<f4", 2)])
NODE_DTYPE = np.dtype([("depth", "<u4"), ("i", "<u4"), ("j", "<u4"), ("k", "<u4"), ("f", "<f4"), ("t", "<f4"),
                       ("time_err", "<f4"), ("fast", "<f4"), ("has_children", "<u4"), ("dx", "<f4"),
                       ("dy", "<f4"), ("dz", "<f4")])

# src/ir.rs:8-43
OPCODES = {0: "NOP", 1: "ARGX", 2: "ARGY", 3: "ARGZ", 4: "ARGW", 5: "ADD", 6: "SUB", 7: "MUL", 8: "DIV", 9: "EXP",
           10: "LOG", 11: "SQR", 12: "SQRT", 13: "POW", 14: "SIN", 15: "COS", 16: "NEG", 17: "MIN", 18: "MAX",
           19: "ABS", 20: "CUP", 21: "CAP", 22: "BR", 23: "PHI", 24: "TRUNC", 25: "FLOOR", 26: "CEIL", 27: "SIGN",
           28: "RET", 29: "MOD", 30: "ATAN2"}


def tape_from_ll(text: str | bytes, time: float = 0.0) -> np.ndarray:
    """src/ir.rs IR::new: `.ll` text -> 32-byte GPUInstruction records (INSTR_DTYPE)."""
    b = text.encode() if isinstance(text, str) else text
    n = C.c_uint32(0)
    rc = lib().tsi_tape_from_ll(b, len(b), time, None, 0, C.byref(n))
    if rc not in (0, -4):
        check(rc, "tsi_tape_from_ll")
    cap = min(n.value, 4096)
    out = (GpuInstruction * max(cap, 1))()
    rc = lib().tsi_tape_from_ll(b, len(b), time, out, cap, C.byref(n))
    check(rc, "tsi_tape_from_ll")
    return np.frombuffer(bytes(out), dtype=INSTR_DTYPE, count=cap).copy()


def write_png(path: str, rgba8: np.ndarray):
    """Writes an RGBA8 image (H x W x 4, row 0 on top) as a PNG."""
    a = np.ascontiguousarray(rgba8, np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an H x W x 4 uint8 array")
    check(lib().tsi_write_png(str(path).encode(), a.ctypes.data, a.shape[1], a.shape[0]), "tsi_write_png")


def _rust_f64_debug(x: float) -> str:
    """Rust's `{:?}` of an f64: shortest round-trip digits, '.0' on integral values, exponent
    without padding ('1e-7', '1e16')."""
    r = repr(float(x))
    if "e" in r:
        m, e = r.split("e")
        return f"{m}e{int(e)}"
    return r


def timing_line(frame_number: int, rootfinding_ms_accumulated: float, frame_ms: float) -> str:
    """One line of the reference's timing.txt (main.rs:1071-1073)."""
    return (f"framenumber: {int(frame_number)}, rootfinding_time_accumulated: "
            f"{_rust_f64_debug(rootfinding_ms_accumulated)}, actual_frame_time: {_rust_f64_debug(frame_ms)}\n")


def precompile(tape, prune=True) -> dict:
    """Compiles a tape's evaluation kernels into the code-object cache (no GPU needed), for the
    pruning setting a State will use (True, "exact" or False)."""
    if isinstance(tape, IR):
        tape = tape.gpu_instructions
    elif isinstance(tape, (str, bytes)):
        tape = tape_from_ll(tape)
    t = np.ascontiguousarray(tape, dtype=INSTR_DTYPE)
    buf = (GpuInstruction * len(t)).from_buffer_copy(t.tobytes())
    info = _lib.TapeInfo()
    flags = _lib.TSI_FLAG_PRUNE_EXACT_ONLY if prune == "exact" else (0 if prune else _lib.TSI_FLAG_NO_PRUNE)
    rc = lib().tsi_precompile_tape(buf, len(t), flags, C.byref(info))
    if rc != 0:
        raise TsiError(rc, "tsi_precompile_tape\n" + info.log.decode(errors="replace"))
    return {"compile_ms": info.compile_ms, "cache_hit": bool(info.cache_hit), "n": info.n,
            "specialized": bool(info.specialized), "fix_kept": info.fix_kept}


def kernel_source(tape, prune=True) -> tuple[str, dict]:
    """The generated evaluation-kernel source for a tape and pruning setting, without compiling
    it, and the lowering's figures (n, stack rows, prune class)."""
    if isinstance(tape, IR):
        tape = tape.gpu_instructions
    elif isinstance(tape, (str, bytes)):
        tape = tape_from_ll(tape)
    t = np.ascontiguousarray(tape, dtype=INSTR_DTYPE)
    buf = (GpuInstruction * len(t)).from_buffer_copy(t.tobytes())
    flags = _lib.TSI_FLAG_PRUNE_EXACT_ONLY if prune == "exact" else (0 if prune else _lib.TSI_FLAG_NO_PRUNE)
    n = C.c_size_t(0)
    info = _lib.TapeInfo()
    check(lib().tsi_tape_kernel_source(buf, len(t), flags, None, 0, C.byref(n), C.byref(info)), "tsi_tape_kernel_source")
    out = C.create_string_buffer(n.value + 1)
    check(lib().tsi_tape_kernel_source(buf, len(t), flags, out, n.value + 1, C.byref(n), None), "tsi_tape_kernel_source")
    return out.value.decode(), {"n": info.n, "stack": info.stack, "prune_class": info.prune_class}


class IR:
    """Mirror of src/ir.rs `IR`: the tape of one scene."""

    def __init__(self, ll_text: str | bytes, time: float = 0.0):
        self.gpu_instructions = tape_from_ll(ll_text, time)

    def __len__(self):
        return len(self.gpu_instructions)

    def ops(self) -> np.ndarray:
        return (self.gpu_instructions["opcode_uses"] >> 16) & 0x3FFF

    def describe(self, j: int) -> str:
        g = self.gpu_instructions[j]
        op = int(g["opcode_uses"] >> 16) & 0x3FFF
        return (f"index = {int(g['index_'])}, uses = {int(g['opcode_uses']) & 0xFFFF}, op = {OPCODES.get(op, op)}, "
                f"larg = {int(g['larg_rarg']) >> 16}, rarg = {int(g['larg_rarg']) & 0xFFFF}, "
                f"rarg_literal = {float(g['literal'])}")


def look_at_inverse(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Camera-to-world matrix (camera.rs:33-35: inverse of cgmath look_at_dir), column-major
    float32[16] as uniforms.view holds it: columns (side, up, -forward, eye)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.zeros((4, 4), np.float64)
    m[0, :3], m[1, :3], m[2, :3], m[3, :3] = s, u, -f, eye
    m[3, 3] = 1.0
    return m.astype(np.float32).reshape(16)


def diagonal_camera(distance: float = 12.0) -> np.ndarray:
    """Eye on the (1,1,1) diagonal at `distance` from the origin, looking at the origin."""
    e = distance / np.sqrt(3.0)
    return look_at_inverse((e, e, e))


def default_params(localsubdiv: int = 8) -> Params:
    p = Params()
    lib().tsi_default_params(C.byref(p), localsubdiv)
    return p


@dataclass
class FrameResult:
    gpu_ms: float
    stats: dict


class State:
    """Mirror of src/main.rs `State`'s frame protocol over the HIP library.

    update(time, view)  -> State::update (uniform write, main.rs:498-504)
    compute()           -> State::compute (first-frame seeding + `main`, main.rs:724-814);
                           returns the device milliseconds, like the reference
    render()            -> State::render's rt_main (649-656) then buf_main (683-690)
    """

    def __init__(self, tape, width: int = 256, height: int = 256, localsubdiv: int = 8, delta: float = 0.01,
                 point_sample: bool = True, swept_volume: bool = False, lipschitz_corrected: bool = False,
                 fastenabled: bool = True, fast_multiplier: float = 1.0, swept_window: float = 1.0 / 60.0,
                 voxel_capacity: int = 1 << 24, queue_capacity: int = 0, prune=True, prune_capacity: int = 0,
                 engine: str = "compiled", prune_live_max: int = 0, shading: int = 0, speculate: bool = True):
        if isinstance(tape, IR):
            tape = tape.gpu_instructions
        elif isinstance(tape, (str, bytes)):
            tape = tape_from_ll(tape)
        self.tape = np.ascontiguousarray(tape, dtype=INSTR_DTYPE)
        p = default_params(localsubdiv)
        p.delta = delta
        p.point_sample = int(point_sample)
        p.swept_volume = int(swept_volume)
        p.lipschitz_corrected = int(lipschitz_corrected)
        p.fastenabled = int(fastenabled)
        p.fast_multiplier = fast_multiplier
        p.swept_window = np.float32(swept_window)
        p.voxel_capacity = voxel_capacity
        p.queue_capacity = queue_capacity
        # prune: True = pruned programs wherever the tape allows them, "exact" = only for tapes
        # whose pruned results are provably identical, False = the full tape in every cell
        if prune == "exact":
            p.flags = _lib.TSI_FLAG_PRUNE_EXACT_ONLY
        else:
            p.flags = 0 if prune else _lib.TSI_FLAG_NO_PRUNE
        # engine: "compiled" = the tape compiled into straight-line kernels at upload (pruned by
        # regions of 16 records; unpruned modules use 64-record regions), "interpreter" =
        # programs in constant memory, pruned per frontier group and record
        if engine not in ("compiled", "interpreter"):
            raise ValueError(f"engine must be 'compiled' or 'interpreter', not {engine!r}")
        if engine == "interpreter":
            p.flags |= _lib.TSI_FLAG_INTERPRET
        # speculate: the compiled, unpruned engine evaluates the shallow levels of each frame in
        # one launch at its start from the previous frame's frontier (TSI_FLAG_NO_SPECULATE)
        if not speculate:
            p.flags |= _lib.TSI_FLAG_NO_SPECULATE
        p.prune_capacity = prune_capacity
        p.prune_live_max = prune_live_max
        # shading: 0 = the reference's colouring; TSI_SHADE_NORMALS (1) = normal_viz with
        # finite_diff_normal's (1, 1, 1) for a tape; | TSI_SHADE_GRADIENT (2) = normal_viz with
        # the voxels' reverse-pass gradients (every cell takes the reverse pass);
        # TSI_SHADE_CINEMATIC (4) = the cinematic path tracer (8 samples, 2 bounces, sun and
        # sky), its normals as above
        p.shading = shading
        self.params = p
        self._ctx = C.c_void_p()
        check(lib().tsi_create(C.byref(p), C.byref(self._ctx)), "tsi_create")
        buf = (GpuInstruction * len(self.tape)).from_buffer_copy(self.tape.tobytes())
        rc = lib().tsi_upload_tape(self._ctx, buf, len(self.tape))
        if rc != 0:
            log = self.tape_info()["log"] if rc == -11 else ""
            lib().tsi_destroy(self._ctx)
            self._ctx = C.c_void_p()
            raise TsiError(rc, "tsi_upload_tape" + (f"\n{log}" if log else ""))
        self.uniforms = Uniforms()
        self.uniforms.resolution[0] = width
        self.uniforms.resolution[1] = height
        v = diagonal_camera()
        for i in range(16):
            self.uniforms.view[i] = float(v[i])
        self.first_eval = True
        self.width, self.height = width, height

    # -- frame protocol --------------------------------------------------------
    def update(self, time: float, view=None):
        u = self.uniforms
        u.lasttime = u.time
        u.time = float(np.float32(time))
        if view is not None:
            for i, x in enumerate(np.asarray(view, np.float32).reshape(16)):
                u.view[i] = float(x)
        check(lib().tsi_set_uniforms(self._ctx, C.byref(u)), "tsi_set_uniforms")

    def reset(self):
        v = self.tape["val"]
        box = (C.c_float * 6)(float(v[0, 0]), float(v[1, 0]), float(v[2, 0]), float(v[0, 1]), float(v[1, 1]),
                              float(v[2, 1]))
        check(lib().tsi_reset(self._ctx, box), "tsi_reset")

    def compute(self) -> float:
        if self.first_eval:
            self.reset()
            self.first_eval = False
        return self.set_inversion().gpu_ms

    def set_inversion(self) -> FrameResult:
        ms = C.c_double(0.0)
        st = Stats()
        check(lib().tsi_set_inversion(self._ctx, C.byref(ms), C.byref(st)), "tsi_set_inversion")
        return FrameResult(ms.value, st.as_dict())

    def set_inversion_async(self):
        """Queues the frame without waiting for it (at most two in flight)."""
        check(lib().tsi_set_inversion_async(self._ctx), "tsi_set_inversion_async")

    def set_inversion_result(self) -> FrameResult:
        """Waits for the oldest frame in flight; its statistics, or its error raised."""
        ms = C.c_double(0.0)
        st = Stats()
        check(lib().tsi_set_inversion_result(self._ctx, C.byref(ms), C.byref(st)), "tsi_set_inversion_result")
        return FrameResult(ms.value, st.as_dict())

    def trace(self, pixels: bool = True):
        rgba = np.zeros((self.height, self.width, 4), np.float32) if pixels else None
        depth = np.zeros((self.height, self.width), np.float32) if pixels else None
        check(lib().tsi_trace(self._ctx, rgba.ctypes.data if pixels else None,
                              depth.ctypes.data if pixels else None), "tsi_trace")
        return rgba, depth

    def last_trace(self) -> tuple[float, int]:
        """(device ms, octree node visits) of the last trace."""
        ms, visits = C.c_double(0.0), C.c_uint64(0)
        check(lib().tsi_last_trace_info(self._ctx, C.byref(ms), C.byref(visits)), "tsi_last_trace_info")
        return ms.value, visits.value

    def trace_async(self):
        """Starts tracing the current octree (a device snapshot) on a trace stream (two trace
        slots, each with its own stream, snapshot and frame: consecutive traces may overlap)."""
        check(lib().tsi_trace_async(self._ctx), "tsi_trace_async")

    def trace_wait(self, pixels: bool = True):
        rgba = np.zeros((self.height, self.width, 4), np.float32) if pixels else None
        depth = np.zeros((self.height, self.width), np.float32) if pixels else None
        check(lib().tsi_trace_wait(self._ctx, rgba.ctypes.data if pixels else None,
                                   depth.ctypes.data if pixels else None), "tsi_trace_wait")
        return rgba, depth

    def render_async(self):
        """rt_main of the current octree without waiting for it, then buf_main."""
        self.trace_async()
        self.enqueue_root()

    def trace_totals(self) -> tuple[float, int, int]:
        """(device ms, node visits, traces) summed over the completed traces."""
        ms, visits, n = C.c_double(0.0), C.c_uint64(0), C.c_uint64(0)
        check(lib().tsi_trace_totals(self._ctx, C.byref(ms), C.byref(visits), C.byref(n)), "tsi_trace_totals")
        return ms.value, visits.value, n.value

    def frame_rgba8(self) -> np.ndarray:
        """The last traced frame as RGBA8 (main.rs:709-711 conversion), rows top to bottom."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().tsi_frame_rgba8(self._ctx, out.ctypes.data), "tsi_frame_rgba8")
        return out

    def save_png(self, path: str):
        """main.rs:693-715 frame dump (images/{frame}.png in the reference)."""
        check(lib().tsi_save_frame_png(self._ctx, str(path).encode()), "tsi_save_frame_png")

    def enqueue_root(self):
        check(lib().tsi_enqueue_root(self._ctx), "tsi_enqueue_root")

    def render(self, pixels: bool = True):
        out = self.trace(pixels)
        self.enqueue_root()
        return out

    # -- parity / debug ----------------------------------------------------------
    def read_nodes(self) -> np.ndarray:
        n = C.c_uint64(0)
        check(lib().tsi_read_nodes(self._ctx, None, 0, C.byref(n)), "tsi_read_nodes")
        out = np.zeros(n.value, NODE_DTYPE)
        check(lib().tsi_read_nodes(self._ctx, out.ctypes.data, n.value, C.byref(n)), "tsi_read_nodes")
        return out

    def tape_info(self) -> dict:
        info = _lib.TapeInfo()
        check(lib().tsi_tape_info_get(self._ctx, C.byref(info)), "tsi_tape_info_get")
        d = {k: getattr(info, k) for k, _ in info._fields_}
        d["log"] = info.log.decode(errors="replace")
        d["engine"] = ("interpreter", "compiled")[info.engine]
        return d

    def read_tape(self) -> np.ndarray:
        n = C.c_uint32(0)
        check(lib().tsi_read_tape(self._ctx, None, 0, C.byref(n)), "tsi_read_tape")
        out = (GpuInstruction * n.value)()
        check(lib().tsi_read_tape(self._ctx, out, n.value, C.byref(n)), "tsi_read_tape")
        return np.frombuffer(bytes(out), dtype=INSTR_DTYPE).copy()

    def set_params(self, **kw):
        """Runtime parameters between frames (tsi_set_params): localsubdiv, delta, point_sample,
        swept_volume, swept_window, lipschitz_corrected, fastenabled, fast_multiplier."""
        p = self.params
        for k, v in kw.items():
            if k not in ("localsubdiv", "delta", "point_sample", "swept_volume", "swept_window",
                         "lipschitz_corrected", "fastenabled", "fast_multiplier"):
                raise ValueError(f"{k} cannot change after tsi_create")
            if k == "swept_window":
                v = float(np.float32(v))
            elif k != "delta" and k != "fast_multiplier":
                v = int(v)
            setattr(p, k, v)
        check(lib().tsi_set_params(self._ctx, C.byref(p)), "tsi_set_params")

    def close(self):
        if self._ctx:
            lib().tsi_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def camera_zdist(height: int) -> float:
    return float(lib().tsi_camera_zdist(height))
