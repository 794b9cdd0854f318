"""Failure handling forced from the host: capacities too small for the scene.

The reference drops work silently when its queue is full (compute.wgsl:3256-3258) and
bump-allocates voxels without a bound (3437-3441).  The library bounds both in the kernels
(frontier appends stop at the list capacity, blocks past the pool are not handed out) and
reports the frame as TSI_E_QUEUE_FULL or TSI_E_POOL_EXHAUSTED; every kernel still runs to its
end.  After a new tape and tsi_reset the context computes frames identical to the oracle again.

Sizes (oracle counts at localsubdiv 6, t = 0): the capsules scene needs 4017 nodes (502 blocks of
8) and 344 parents at depth 5; the spheres scene 689 nodes and at most 46 parents per level.
"""
import numpy as np
import pytest

import temporalsetinversion_amd as T
from helpers import assert_same_nodes, tape_of
import oracle

pytestmark = pytest.mark.gpu

TIMES = [0.0, 1 / 60, 2 / 60]


def _recovers(st):
    """Upload the spheres tape, reset, and compare three frames with the oracle."""
    tape = tape_of("spheres")
    buf = (T._lib.GpuInstruction * len(tape)).from_buffer_copy(np.ascontiguousarray(tape).tobytes())
    T.check(T.lib().tsi_upload_tape(st._ctx, buf, len(tape)), "tsi_upload_tape")
    st.tape = tape
    st.reset()
    fr = oracle.Frame(tape, localsubdiv=6, voxel_capacity=1 << 20)
    try:
        for k, t in enumerate(TIMES):
            t = float(np.float32(t))
            st.update(t)
            st.set_inversion()
            fr.set_inversion(t)
            assert_same_nodes(st.read_nodes(), fr.export(), f"frame {k} after the reset")
            st.enqueue_root()
            fr.enqueue_root()
    finally:
        fr.close()


def test_pool_exhausted_is_reported_and_a_reset_recovers():
    with T.State(tape_of("capsules"), localsubdiv=6, voxel_capacity=1024, prune=False) as st:
        st.reset()
        st.update(0.0)
        with pytest.raises(T.TsiError) as e:
            st.set_inversion()
        assert e.value.code == -3                      # TSI_E_POOL_EXHAUSTED
        nodes = st.read_nodes()                        # the octree stays readable, inside the pool
        assert 0 < len(nodes) <= 1024
        # the asynchronous form reports the same frame error when it is collected
        st.enqueue_root()
        st.update(1 / 60)
        st.set_inversion_async()
        with pytest.raises(T.TsiError) as e:
            st.set_inversion_result()
        assert e.value.code == -3
        _recovers(st)


def test_queue_full_is_reported_and_a_reset_recovers():
    with T.State(tape_of("capsules"), localsubdiv=6, voxel_capacity=1 << 16, queue_capacity=64, prune=False) as st:
        st.reset()
        st.update(0.0)
        with pytest.raises(T.TsiError) as e:
            st.set_inversion()
        assert e.value.code == -2                      # TSI_E_QUEUE_FULL
        assert len(st.read_nodes()) > 0
        _recovers(st)


def test_queue_full_with_pruning_buffers():
    """The compiled engine's pruned path sizes its liveness records by the list capacity too."""
    with T.State(tape_of("capsules"), localsubdiv=6, voxel_capacity=1 << 16, queue_capacity=64, prune=True,
                 prune_live_max=100) as st:
        st.reset()
        for t in TIMES:
            st.update(float(np.float32(t)))
            with pytest.raises(T.TsiError) as e:
                st.set_inversion()
            assert e.value.code == -2
            st.enqueue_root()
        _recovers(st)
