"""Shared helpers of the parity tests (oracle = checker, libtsi = product)."""
from __future__ import annotations

import numpy as np

import oracle
import temporalsetinversion_amd as T
from temporalsetinversion_amd import scenes

FIELDS = ("f", "t", "time_err", "fast", "has_children", "dx", "dy", "dz")


def canon(nodes: np.ndarray) -> np.ndarray:
    order = np.lexsort((nodes["k"], nodes["j"], nodes["i"], nodes["depth"]))
    return nodes[order]


def assert_same_nodes(gpu: np.ndarray, orc: np.ndarray, what: str = ""):
    g, o = canon(gpu), canon(orc)
    assert len(g) == len(o), f"{what}: {len(g)} GPU nodes vs {len(o)} oracle nodes"
    for key in ("depth", "i", "j", "k"):
        assert np.array_equal(g[key], o[key]), f"{what}: node keys differ ({key})"
    for fld in FIELDS:
        a, b = g[fld], o[fld]
        if a.dtype.kind == "f":
            same = a.view(np.uint32) == b.view(np.uint32)
        else:
            same = a == b
        if not np.all(same):
            bad = np.nonzero(~same)[0][:5]
            rows = [(int(g["depth"][i]), int(g["i"][i]), int(g["j"][i]), int(g["k"][i]), a[i], b[i]) for i in bad]
            raise AssertionError(f"{what}: field {fld} differs in {int((~same).sum())} nodes, e.g. {rows}")


def run_pair(tape, times, width=0, height=0, trace_every=0, voxel_capacity=1 << 20, view=None, **params):
    """Runs the same frame sequence on the GPU library and on the oracle; yields per frame
    (k, gpu_nodes, oracle_nodes, gpu_stats, oracle_stats, gpu_frame, oracle_frame)."""
    fr = oracle.Frame(tape, voxel_capacity=voxel_capacity,
                      **{k: v for k, v in params.items() if k not in ("prune", "engine", "prune_live_max")})
    st = T.State(tape, width=max(width, 16), height=max(height, 16), voxel_capacity=voxel_capacity, **params)
    try:
        st.reset()
        for k, t in enumerate(times):
            t = float(np.float32(t))
            st.update(t, view)
            gstats = st.set_inversion().stats
            ost = fr.set_inversion(t)
            gframe = oframe = None
            if trace_every and k % trace_every == 0:
                gframe = st.trace()
                view = np.array(st.uniforms.view[:], np.float32)
                rgba, depth, _ = fr.trace(view, width, height, t, T.camera_zdist(height))
                oframe = (rgba, depth)
            yield k, st.read_nodes(), fr.export(), gstats, ost, gframe, oframe
            st.enqueue_root()
            fr.enqueue_root()
    finally:
        st.close()
        fr.close()


def tape_of(name: str, **kw):
    if name == "sphere":
        return oracle.parse_ll(scenes.sphere_unit(1.0))
    return oracle.parse_ll(scenes.SCENES[name](**kw))
