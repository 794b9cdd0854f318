"""tsi_frame_rgba8_noalloc: the RGBA8 frame dump with its staging buffer in the library's static
device memory (no hipMalloc).  Called through the C ABI and compared byte for byte with
tsi_frame_rgba8 (which stages through a buffer it allocates), on two calls in a row, the second
with a smaller frame, so the static buffer still holds the first call's bytes beyond the second
frame; frames above the buffer's size are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import temporalsetinversion_amd as T
from helpers import tape_of

pytestmark = pytest.mark.gpu


def _both(st, w, h):
    st.uniforms.resolution[0] = w
    st.uniforms.resolution[1] = h
    st.width, st.height = w, h
    st.update(0.0)
    rgba, _ = st.trace()
    ref = np.zeros((h, w, 4), np.uint8)
    got = np.full((h, w, 4), 7, np.uint8)
    T.check(T.lib().tsi_frame_rgba8(st._ctx, ref.ctypes.data), "tsi_frame_rgba8")
    T.check(T.lib().tsi_frame_rgba8_noalloc(st._ctx, got.ctypes.data), "tsi_frame_rgba8_noalloc")
    return rgba, ref, got


def test_noalloc_frame_dump_matches_the_allocating_one():
    with T.State(tape_of("spheres"), width=96, height=64, localsubdiv=6, voxel_capacity=1 << 18) as st:
        st.reset()
        st.update(0.0)
        st.set_inversion()
        rgba, ref, got = _both(st, 96, 64)
        print(f"96x64: {int((rgba[..., 0] < 1.0).sum())} shaded pixels", flush=True)
        assert np.array_equal(got, ref)
        assert len(np.unique(ref.reshape(-1, 4), axis=0)) > 4          # a picture, not a constant
        # second call, smaller frame (the static buffer keeps the larger frame's tail)
        rgba2, ref2, got2 = _both(st, 40, 24)
        assert np.array_equal(got2, ref2)
        assert got2.shape == (24, 40, 4)
        # a frame above the static buffer is refused before any device work
        st.uniforms.resolution[0] = 4096
        st.uniforms.resolution[1] = 4096
        T.check(T.lib().tsi_set_uniforms(st._ctx, C.byref(st.uniforms)), "tsi_set_uniforms")
        big = np.zeros(16, np.uint8)
        assert T.lib().tsi_frame_rgba8_noalloc(st._ctx, big.ctypes.data) == -13   # TSI_E_SCRATCH


def test_noalloc_frame_dump_across_contexts():
    """Two contexts on one GPU share the static buffer; calls that do not overlap each see their
    own frame."""
    with T.State(tape_of("spheres"), width=64, height=48, localsubdiv=5, voxel_capacity=1 << 18) as a, \
         T.State(tape_of("capsules"), width=32, height=32, localsubdiv=5, voxel_capacity=1 << 18) as b:
        for st in (a, b):
            st.reset()
            st.update(0.0)
            st.set_inversion()
        _, ra, ga = _both(a, 64, 48)
        _, rb, gb = _both(b, 32, 32)
        assert np.array_equal(ga, ra) and np.array_equal(gb, rb)
        ga2 = np.zeros_like(ra)
        T.check(T.lib().tsi_frame_rgba8_noalloc(a._ctx, ga2.ctypes.data), "tsi_frame_rgba8_noalloc")
        assert np.array_equal(ga2, ra)
