"""GPU: a full pass of every kernel captured in one HIP graph.  Graph
replays, with eager passes and synchronisations in between, must leave the
same bits as the same number of eager passes on an identical second set of
buffers -- except SCAN.  A tile's decoupled look-back stops at the first
predecessor that has published its inclusive prefix, and which one that is
depends on timing, so two runs of the same scan may associate the FP64 sums
differently.  SCAN outputs are held to that reassociation's bound instead:
at most (tiles + 64) roundings of partial sums no larger than the prefix of
|x| (integer-valued inputs, where association cannot matter, are checked
bit for bit in the scan tests)."""

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
            if k == "scan":
                x = np.abs(g_pass.sx.to_array())
                tiles = -(-x.size // 1024)  # at least the number of scan tiles
                bound = (tiles + 64) * np.finfo(np.float64).eps * np.concatenate([[0.0], np.cumsum(x)[:-1]])
                assert np.all(np.abs(got[k] - want[k]) <= bound), k
            else:
                assert np.array_equal(got[k], want[k]), k


def test_null_stream_cannot_be_captured(rpx):
    with pytest.raises(rpx.RpxError):
        rpx.lib().rpx_graph_begin_capture(None)
