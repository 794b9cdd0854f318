"""The trace's two forms (one 16x16-pixel block per tile, and the persistent form whose
wavefronts claim 8x8-pixel quadrants on a few blocks per CU, chosen per frame from the measured
times unless $TSI_TRACE_BLOCKS fixes it) shade every pixel with the same code: frames must be
identical, bit for bit, in both forms, for the plain, normal and cinematic shadings, and the
plain frames equal the oracle's."""
import os

import numpy as np
import pytest

import temporalsetinversion_amd as T
from helpers import run_pair, tape_of
from temporalsetinversion_amd import _lib

pytestmark = pytest.mark.gpu


def _frames(blocks, shading, tape_text, L, size, times):
    old = os.environ.get("TSI_TRACE_BLOCKS")
    os.environ["TSI_TRACE_BLOCKS"] = str(blocks)   # read when the context is created
    try:
        st = T.State(T.tape_from_ll(tape_text), width=size[0], height=size[1], localsubdiv=L, shading=shading)
    finally:
        if old is None:
            del os.environ["TSI_TRACE_BLOCKS"]
        else:
            os.environ["TSI_TRACE_BLOCKS"] = old
    out = []
    try:
        st.reset()
        view = T.look_at_inverse((2.2, 1.7, 2.6))
        for t in times:
            st.update(float(np.float32(t)), view)
            st.set_inversion()
            out.append(st.trace())
            st.enqueue_root()
    finally:
        st.close()
    return out


@pytest.mark.parametrize("shading", [0, _lib.TSI_SHADE_NORMALS, _lib.TSI_SHADE_CINEMATIC])
def test_persistent_trace_equals_full_launch(shading):
    from temporalsetinversion_amd import scenes
    text = scenes.SCENES["spheres"]()
    times = [0.0, 1 / 60, 0.25]
    size = (72, 40) if shading == _lib.TSI_SHADE_CINEMATIC else (200, 120)   # not multiples of 16
    print(f"trace forms shading={shading} size={size}", flush=True)
    full = _frames(0, shading, text, 6, size, times)
    pers = _frames(1, shading, text, 6, size, times)
    for k, ((fr, fd), (pr, pd)) in enumerate(zip(full, pers)):
        assert np.array_equal(fr.view(np.uint32), pr.view(np.uint32)), f"colour differs in frame {k}"
        assert np.array_equal(fd.view(np.uint32), pd.view(np.uint32)), f"depth differs in frame {k}"
        assert (fd > 0).sum() > 0


def test_persistent_trace_matches_oracle():
    tape = tape_of("spheres")
    view = T.look_at_inverse((2.2, 1.7, 2.6))
    old = os.environ.get("TSI_TRACE_BLOCKS")
    os.environ["TSI_TRACE_BLOCKS"] = "2"
    try:
        for k, g, o, gs, os_, gf, of in run_pair(tape, [0.0, 1 / 60, 0.25], width=64, height=48, trace_every=1,
                                                    localsubdiv=6, view=view):
            assert np.array_equal(gf[0].view(np.uint32), of[0].view(np.uint32)), f"colour differs in frame {k}"
            assert np.array_equal(gf[1].view(np.uint32), of[1].view(np.uint32)), f"depth differs in frame {k}"
    finally:
        if old is None:
            del os.environ["TSI_TRACE_BLOCKS"]
        else:
            os.environ["TSI_TRACE_BLOCKS"] = old
