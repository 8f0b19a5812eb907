"""GPU: a full pass of every kernel captured in one HIP graph.  Graph
replays, with eager passes and synchronisations in between, must leave the
same bits as the same number of eager passes on an identical second set of
buffers (every kernel here is deterministic)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = {"n_stream": 100_003, "n_scan": 70_001, "n_sort": 5_000, "gemm": (96, 80, 64), "jacobi": (130, 3),
         "mass": (50, 4), "diffusion": (20, 3)}


@pytest.mark.parametrize("sizes", [SMALL, dict(SMALL, mass=(33, 7), diffusion=(9, 5), jacobi=(66, 1))])
def test_graph_pass_replays_match_eager(rpx, sizes):
    from rajaperf_amd.suite_pass import SuitePass
    st = rpx.Stream()
    g_pass, e_pass = SuitePass(sizes), SuitePass(sizes)
    with rpx.Graph(st) as graph:
        g_pass.run(st)
    st.synchronize()
    for _ in range(3):
        graph.launch()
        e_pass.run(st)
        st.synchronize()
        got, want = g_pass.outputs(), e_pass.outputs()
        for k in want:
            assert np.array_equal(got[k], want[k]), k


def test_null_stream_cannot_be_captured(rpx):
    with pytest.raises(rpx.RpxError):
        rpx.lib().rpx_graph_begin_capture(None)
