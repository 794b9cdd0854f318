"""Batched and served forms over libtsi.so (include/tsi.h tsi_batch_*).

No counterpart in the reference, whose host (src/main.rs) renders one scene per frame.  A
`Batch` holds several tapes and sessions; a session is one scene's octree (what `State` holds),
resident in the batch's device arena.  `Batch.run(frames)` computes one frame of each listed
session in one persistent set-inversion launch and one trace launch; each session's
nodes and pixels are those `State` computes for the same tape, parameters and times.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FrameReq, FrameResult, GpuInstruction, check, lib

__all__ = ["Batch", "precompile_batch", "tape_pointers"]


def tape_pointers(tapes):
    """ctypes arrays of GpuInstruction for the tapes (kept alive by the caller) and their
    pointer / length arrays."""
    from . import INSTR_DTYPE

    bufs = []
    for t in tapes:
        t = np.ascontiguousarray(t, dtype=INSTR_DTYPE)
        bufs.append((GpuInstruction * len(t)).from_buffer_copy(t.tobytes()))
    ptrs = (C.POINTER(GpuInstruction) * len(bufs))(*[C.cast(b, C.POINTER(GpuInstruction)) for b in bufs])
    lens = (C.c_uint32 * len(bufs))(*[len(b) for b in bufs])
    return bufs, ptrs, lens


def _info(info: _lib.BatchInfo) -> dict:
    d = {k: getattr(info, k) for k, _ in info._fields_}
    d["log"] = info.log.decode(errors="replace")
    return d


def precompile_batch(tapes) -> dict:
    """Compiles the batch module of these tapes into the code-object cache (no device needed)."""
    bufs, ptrs, lens = tape_pointers(tapes)
    info = _lib.BatchInfo()
    rc = lib().tsi_batch_precompile(ptrs, lens, len(bufs), C.byref(info))
    if rc != 0:
        raise _lib.TsiError(rc, "tsi_batch_precompile\n" + info.log.decode(errors="replace"))
    return _info(info)


class Batch:
    """Sessions of several tapes in one device arena, one launch per list of frames."""

    def __init__(self, tapes, arena_bytes: int = 0, max_sessions: int = 0, waves_per_simd: int = 0,
                 claim_units: int = 0):
        self._bufs, ptrs, lens = tape_pointers(tapes)
        cfg = _lib.BatchConfig(arena_bytes, max_sessions, waves_per_simd, claim_units)
        self._b = C.c_void_p()
        check(lib().tsi_batch_create(ptrs, lens, len(self._bufs), C.byref(cfg), C.byref(self._b)), "tsi_batch_create")
        self.sessions = {}   # id -> (max_width, max_height)
        self._pending = None

    @staticmethod
    def desc(tape: int, localsubdiv: int = 8, delta: float = 0.01, point_sample: bool = True,
             swept_volume: bool = False, swept_window: float = 1.0 / 60.0, lipschitz_corrected: bool = False,
             fastenabled: bool = True, fast_multiplier: float = 1.0, shading: int = 0, voxel_capacity: int = 0,
             queue_capacity: int = 0, root_box=(-5.0, -5.0, -5.0, 5.0, 5.0, 5.0), max_width: int = 0,
             max_height: int = 0) -> _lib.SessionDesc:
        from . import default_params

        p = default_params(localsubdiv)
        p.delta = delta
        p.point_sample = int(point_sample)
        p.swept_volume = int(swept_volume)
        p.swept_window = float(np.float32(swept_window))
        p.lipschitz_corrected = int(lipschitz_corrected)
        p.fastenabled = int(fastenabled)
        p.fast_multiplier = fast_multiplier
        p.shading = shading
        p.voxel_capacity = voxel_capacity
        p.queue_capacity = queue_capacity
        d = _lib.SessionDesc()
        d.tape = tape
        d.params = p
        for i, x in enumerate(root_box):
            d.root_box[i] = float(x)
        d.max_width, d.max_height = max_width, max_height
        return d

    def footprint(self, **kw) -> int:
        n = C.c_uint64(0)
        check(lib().tsi_batch_footprint(self._b, C.byref(self.desc(**kw)), C.byref(n)), "tsi_batch_footprint")
        return n.value

    def open(self, **kw) -> int:
        """Admits a session (TsiError TSI_E_BUDGET when it does not fit) with a fresh octree."""
        sid = C.c_uint32(0)
        d = self.desc(**kw)
        check(lib().tsi_batch_open(self._b, C.byref(d), C.byref(sid)), "tsi_batch_open")
        self.sessions[sid.value] = (d.max_width or 1920, d.max_height or 1080)
        return sid.value

    def close_session(self, sid: int):
        check(lib().tsi_batch_close(self._b, sid), "tsi_batch_close")
        self.sessions.pop(sid, None)

    def reset(self, sid: int):
        check(lib().tsi_batch_reset(self._b, sid), "tsi_batch_reset")

    def submit(self, frames):
        """frames: (session, time, width, height, view[, reset]) tuples; width = height = 0: no
        trace; reset: start the session from an empty octree inside the launch."""
        reqs = (FrameReq * len(frames))()
        for i, f in enumerate(frames):
            sid, t, w, h, view = f[:5]
            reqs[i].flags = _lib.TSI_FRAME_RESET if len(f) > 5 and f[5] else 0
            reqs[i].session = sid
            reqs[i].time = float(np.float32(t))
            reqs[i].width, reqs[i].height = w, h
            v = np.asarray(view if view is not None else np.eye(4), np.float32).reshape(16)
            for k in range(16):
                reqs[i].view[k] = float(v[k])
        check(lib().tsi_batch_submit(self._b, reqs, len(frames)), "tsi_batch_submit")
        self._pending = list(frames)

    def done(self) -> bool:
        """Whether the submitted frames have finished (does not wait)."""
        rc = lib().tsi_batch_query(self._b)
        if rc == -12:
            return False
        check(rc, "tsi_batch_query")
        return True

    def wait(self):
        """(per-frame results as dicts, set-inversion device ms, trace device ms)."""
        n = len(self._pending or [])
        res = (FrameResult * max(n, 1))()
        sms, tms = C.c_double(0.0), C.c_double(0.0)
        check(lib().tsi_batch_wait(self._b, res, C.byref(sms), C.byref(tms)), "tsi_batch_wait")
        out = []
        for i in range(n):
            d = res[i].stats.as_dict()
            d["status"] = res[i].status
            d["node_visits"] = res[i].node_visits
            out.append(d)
        self._pending = None
        return out, sms.value, tms.value

    def run(self, frames):
        self.submit(frames)
        return self.wait()

    def read_nodes(self, sid: int) -> np.ndarray:
        from . import NODE_DTYPE

        n = C.c_uint64(0)
        check(lib().tsi_batch_read_nodes(self._b, sid, None, 0, C.byref(n)), "tsi_batch_read_nodes")
        out = np.zeros(n.value, NODE_DTYPE)
        check(lib().tsi_batch_read_nodes(self._b, sid, out.ctypes.data, n.value, C.byref(n)), "tsi_batch_read_nodes")
        return out

    def frame(self, sid: int, width: int, height: int):
        rgba = np.zeros((height, width, 4), np.float32)
        depth = np.zeros((height, width), np.float32)
        check(lib().tsi_batch_frame(self._b, sid, rgba.ctypes.data, depth.ctypes.data), "tsi_batch_frame")
        return rgba, depth

    def tune(self, waves_per_simd: int = 0, claim_units: int = 0):
        """Launch shape for the next submits (0 keeps a knob as it is)."""
        cfg = _lib.BatchConfig(0, 0, waves_per_simd, claim_units)
        check(lib().tsi_batch_tune(self._b, C.byref(cfg)), "tsi_batch_tune")

    def info(self) -> dict:
        info = _lib.BatchInfo()
        check(lib().tsi_batch_info_get(self._b, C.byref(info)), "tsi_batch_info_get")
        return _info(info)

    def close(self):
        if self._b:
            lib().tsi_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
