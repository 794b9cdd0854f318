"""ctypes front end of the oracle (oracle/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product package
(temporalsetinversion_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class Instr(C.Structure):
    """ir.rs:88-97 GPUInstruction (32 B)."""
    _fields_ = [("index_", C.c_uint32), ("opcode_uses", C.c_uint32), ("larg_rarg", C.c_uint32),
                ("literal", C.c_float), ("val", C.c_float * 2), ("sens", C.c_float * 2)]


class Iv(C.Structure):
    _fields_ = [("lo", C.c_float), ("hi", C.c_float)]


class Params(C.Structure):
    _fields_ = [("localsubdiv", C.c_int32), ("delta", C.c_float), ("point_sample", C.c_int32),
                ("swept_volume", C.c_int32), ("lipschitz_corrected", C.c_int32), ("fastenabled", C.c_int32),
                ("fast_multiplier", C.c_float), ("swept_window", C.c_float), ("shading", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("cell_evals", C.c_uint64), ("rev_evals", C.c_uint64), ("fwd_ops", C.c_uint64),
                ("rev_ops", C.c_uint64), ("parents", C.c_uint64), ("deallocs", C.c_uint64),
                ("enqueued", C.c_uint64), ("allocated_voxels", C.c_uint32), ("status", C.c_uint32)]


class Node(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("i", C.c_uint32), ("j", C.c_uint32), ("k", C.c_uint32),
                ("f", C.c_float), ("t", C.c_float), ("time_err", C.c_float), ("fast", C.c_float),
                ("has_children", C.c_uint32), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float)]


NODE_DTYPE = np.dtype([("depth", "<u4"), ("i", "<u4"), ("j", "<u4"), ("k", "<u4"), ("f", "<f4"), ("t", "<f4"),
                       ("time_err", "<f4"), ("fast", "<f4"), ("has_children", "<u4"), ("dx", "<f4"),
                       ("dy", "<f4"), ("dz", "<f4")])
INSTR_DTYPE = np.dtype([("index_", "<u4"), ("opcode_uses", "<u4"), ("larg_rarg", "<u4"), ("literal", "<f4"),
                        ("val", "<f4", 2), ("sens", "<f4", 2)])

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_parse_ll.argtypes = [C.c_char_p, C.c_size_t, C.c_float, C.POINTER(Instr), C.c_uint32,
                                   C.POINTER(C.c_uint32)]
        L.orc_parse_ll.restype = C.c_int
        for name in ("orc_det_sinf", "orc_det_expf", "orc_det_logf", "orc_det_pow045", "orc_det_cosf",
                     "orc_det_acosf", "orc_linear_to_srgb"):
            getattr(L, name).argtypes = [C.c_float]
            getattr(L, name).restype = C.c_float
        for name in ("orc_det_atan2f", "orc_wmin", "orc_wmax", "orc_det_powf"):
            getattr(L, name).argtypes = [C.c_float, C.c_float]
            getattr(L, name).restype = C.c_float
        for name in ("orc_frand", "orc_frand0to1"):
            getattr(L, name).argtypes = [C.POINTER(C.c_uint32)]
            getattr(L, name).restype = C.c_float
        L.orc_sample_biased.argtypes = [C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        L.orc_sample_biased.restype = None
        L.orc_sunsky.argtypes = [C.POINTER(C.c_float), C.c_float, C.c_float, C.POINTER(C.c_float)]
        L.orc_sunsky.restype = None
        L.orc_det_powi.argtypes = [C.c_float, C.c_int32]
        L.orc_det_powi.restype = C.c_float
        L.orc_forward_op.argtypes = [C.c_uint32, C.c_uint32, Iv, Iv, C.c_float]
        L.orc_forward_op.restype = Iv
        L.orc_eval_tape.argtypes = [C.POINTER(Instr), C.c_uint32, C.POINTER(Iv), C.c_int, C.POINTER(Iv),
                                    C.POINTER(Iv), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_eval_tape.restype = Iv
        L.orc_fwd_cost.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_fwd_cost.restype = C.c_uint32
        L.orc_rev_cost.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_rev_cost.restype = C.c_uint32
        L.orc_prune.argtypes = [C.POINTER(Instr), C.c_uint32, C.POINTER(Iv), C.c_void_p, C.c_void_p]
        L.orc_prune.restype = None
        L.orc_create.argtypes = [C.POINTER(Params), C.POINTER(Instr), C.c_uint32, C.c_uint32]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_destroy.restype = None
        L.orc_reset.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.orc_reset.restype = C.c_int
        L.orc_enqueue_root.argtypes = [C.c_void_p]
        L.orc_enqueue_root.restype = C.c_int
        L.orc_set_inversion.argtypes = [C.c_void_p, C.c_float, C.POINTER(Stats)]
        L.orc_set_inversion.restype = C.c_int
        L.orc_trace.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
        L.orc_trace.restype = C.c_int
        L.orc_export.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_export.restype = C.c_uint64
        L.orc_leafset_build.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.c_double]
        L.orc_leafset_build.restype = C.c_void_p
        L.orc_leafset_free.argtypes = [C.c_void_p]
        L.orc_leafset_free.restype = None
        L.orc_leafset_march.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.orc_leafset_march.restype = None
        _lib = L
    return _lib


def parse_ll(text: str | bytes, time: float = 0.0) -> np.ndarray:
    """Restated ir.rs tape builder; returns an INSTR_DTYPE array of n records."""
    b = text.encode() if isinstance(text, str) else text
    cap = 4096
    out = (Instr * cap)()
    n = C.c_uint32(0)
    rc = lib().orc_parse_ll(b, len(b), time, out, cap, C.byref(n))
    if rc not in (0, -2):
        raise ValueError(f"orc_parse_ll failed: {rc}")
    m = min(n.value, cap)
    arr = np.frombuffer(bytes(out), dtype=INSTR_DTYPE, count=m).copy()
    if rc == -2:
        raise OverflowError(f"tape has {n.value} instructions (> 4096)")
    return arr


def _tape_ptr(tape: np.ndarray):
    t = np.ascontiguousarray(tape, dtype=INSTR_DTYPE)
    return t, t.ctypes.data_as(C.POINTER(Instr))


def eval_tape(tape: np.ndarray, args, take_derivative: bool):
    """One cell: returns (field(lo,hi), vals[n,2], sens[n,2], fwd_ops, rev_ops)."""
    t, p = _tape_ptr(tape)
    n = len(t)
    a = (Iv * 4)(*[Iv(float(lo), float(hi)) for lo, hi in args])
    vals = np.zeros((n, 2), np.float32)
    sens = np.zeros((n, 2), np.float32)
    fo, ro = C.c_uint64(0), C.c_uint64(0)
    f = lib().orc_eval_tape(p, n, a, int(take_derivative), vals.ctypes.data_as(C.POINTER(Iv)),
                            sens.ctypes.data_as(C.POINTER(Iv)), C.byref(fo), C.byref(ro))
    return (f.lo, f.hi), vals, sens, fo.value, ro.value


def prune(tape: np.ndarray, vals: np.ndarray):
    t, p = _tape_ptr(tape)
    n = len(t)
    v = np.ascontiguousarray(vals, np.float32)
    live = np.zeros(n, np.uint8)
    choice = np.zeros(n, np.uint8)
    lib().orc_prune(p, n, v.ctypes.data_as(C.POINTER(Iv)), live.ctypes.data, choice.ctypes.data)
    return live, choice


def frand_seq(seed: int, n: int, unit: bool = False):
    """n draws of frand (or frand0to1 with unit=True) from `seed`; returns (values, states after
    each draw)."""
    st = C.c_uint32(seed & 0xFFFFFFFF)
    f = lib().orc_frand0to1 if unit else lib().orc_frand
    vals = np.empty(n, np.float32)
    states = np.empty(n, np.uint32)
    for i in range(n):
        vals[i] = f(C.byref(st))
        states[i] = st.value
    return vals, states


def sample_biased(d, power: float, seed: int, n: int) -> np.ndarray:
    """n successive getSampleBiased(d, power) draws from `seed` (each takes two frand0to1)."""
    st = C.c_uint32(seed & 0xFFFFFFFF)
    dv = (C.c_float * 3)(*[float(x) for x in d])
    out = (C.c_float * 3)()
    res = np.empty((n, 3), np.float32)
    for i in range(n):
        lib().orc_sample_biased(dv, np.float32(power), C.byref(st), out)
        res[i] = out[:]
    return res


def sunsky(view, sun_pos=(-0.1, 0.7)) -> np.ndarray:
    v = (C.c_float * 3)(*[float(x) for x in view])
    out = (C.c_float * 3)()
    lib().orc_sunsky(v, np.float32(sun_pos[0]), np.float32(sun_pos[1]), out)
    return np.array(out[:], np.float32)


class Frame:
    """Whole-frame oracle: restated `main` + `buf_main` + `rt_main` over one octree."""

    def __init__(self, tape: np.ndarray, localsubdiv=6, delta=0.01, point_sample=True, swept_volume=False,
                 lipschitz_corrected=False, fastenabled=True, fast_multiplier=1.0, swept_window=1.0 / 60.0,
                 voxel_capacity=1 << 22, root_box=(-5.0, -5.0, -5.0, 5.0, 5.0, 5.0), shading=0):
        self._tape, p = _tape_ptr(tape)
        self.params = Params(localsubdiv, delta, int(point_sample), int(swept_volume), int(lipschitz_corrected),
                             int(fastenabled), fast_multiplier, np.float32(swept_window), int(shading))
        self._h = lib().orc_create(C.byref(self.params), p, len(self._tape), voxel_capacity)
        if not self._h:
            raise MemoryError("orc_create failed")
        self.reset(root_box)

    def reset(self, root_box=(-5.0, -5.0, -5.0, 5.0, 5.0, 5.0)):
        b = (C.c_float * 6)(*root_box)
        lib().orc_reset(self._h, b)

    def enqueue_root(self):
        return lib().orc_enqueue_root(self._h)

    def set_inversion(self, time: float) -> Stats:
        st = Stats()
        rc = lib().orc_set_inversion(self._h, np.float32(time), C.byref(st))
        if rc != 0:
            raise RuntimeError(f"orc_set_inversion failed: {rc}")
        return st

    def trace(self, view, width, height, time, zdist, row0=0, rows=0):
        v = (C.c_float * 16)(*[float(x) for x in np.asarray(view, np.float32).reshape(16)])
        rgba = np.zeros((height, width, 4), np.float32)
        depth = np.zeros((height, width), np.float32)
        visits = C.c_uint64(0)
        lib().orc_trace(self._h, v, width, height, np.float32(time), np.float32(zdist), rgba.ctypes.data,
                        depth.ctypes.data, row0, rows, C.byref(visits))
        return rgba, depth, visits.value

    def export(self) -> np.ndarray:
        n = lib().orc_export(self._h, None, 0)
        out = np.zeros(n, NODE_DTYPE)
        lib().orc_export(self._h, out.ctypes.data, n)
        return out

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pixel_rays(view, width: int, height: int, xs, ys):
    """float64 eye and unit directions of pixels (xs, ys): ray_direction (compute.wgsl:4590-4594,
    fov 90: z = height / tan(45 degrees) = height) turned by the view matrix (4630)."""
    m = np.asarray(view, np.float64).reshape(4, 4)   # column-major: m[c] is column c
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    d = np.stack([(xs - width / 2.0) + 0.5, -((ys - height / 2.0) + 0.5), -np.full_like(xs, float(height))], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    g = d @ m[:3, :3]
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    eye = np.broadcast_to(m[3, :3], g.shape).copy()
    return eye, g


def march_depth(nodes: np.ndarray, origins: np.ndarray, dirs: np.ndarray,
                root_box=(-5.0, -5.0, -5.0, 5.0, 5.0, 5.0), pad: float = 1e-4) -> np.ndarray:
    """float64 first-hit distance of each ray against a canonical node export (march.c), -1 on a
    miss: the independent restatement of the depth the trace reports."""
    nd = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    o = np.ascontiguousarray(origins, np.float64)
    d = np.ascontiguousarray(dirs, np.float64)
    box = (C.c_double * 6)(*root_box)
    h = lib().orc_leafset_build(nd.ctypes.data, len(nd), box, pad)
    if not h:
        raise MemoryError("orc_leafset_build failed")
    out = np.zeros(len(o), np.float64)
    try:
        lib().orc_leafset_march(h, o.ctypes.data, d.ctypes.data, len(o), out.ctypes.data)
    finally:
        lib().orc_leafset_free(h)
    return out
