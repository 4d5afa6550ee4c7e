"""Layout tuner: the reference's threshold sweep (CPU) and an end-to-end run (GPU)."""
import numpy as np
import pytest

from warpkernel_amd import fe
from warpkernel_amd.tuning import thresholds


def test_threshold_sweep_matches_reference_driver():
    # wpk2.cu:156-157: inc = (max-min)*0.1 if positive else 1; th = min..max step inc
    assert thresholds(24, 81) == list(range(24, 82, 5))
    assert thresholds(9, 9) == [9]
    assert thresholds(3, 10) == list(range(3, 11, 1))
    assert thresholds(1, 400) == list(range(1, 401, 39))


@pytest.mark.gpu
def test_tune_layout_returns_a_working_engine():
    import torch
    from oracle import pyoracle as O
    from warpkernel_amd.tuning import tune_layout
    m = fe.hex_laplace_3d(24)
    x = torch.from_numpy(fe.uniform_x(m.ncols, seed=8)).cuda()
    y = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    best = tune_layout(m.rowptr, m.colind, m.vals, x, y, reps=3, threshold_list=[4, 7, 14])
    assert best.variant in (1, 2) and best.ms > 0 and len(best.table) == 4
    _check_choice(best, m, x, y)
    # group widths 64 and 32, lanes-per-row caps 32 and 64 (64 only at width 64)
    best = tune_layout(m.rowptr, m.colind, m.vals, x, y, reps=2, threshold_list=[4, 14],
                       group_width=(64, 32), max_tpr=(32, 64))
    assert len(best.table) == 1 + 2 + 2 + 1 + 2
    assert {(t[3], t[4]) for t in best.table if t[0] == 2} == {(64, 32), (64, 64), (32, 32)}
    _check_choice(best, m, x, y)


def _check_choice(best, m, x, y):
    import torch
    from oracle import pyoracle as O
    y.fill_(float("nan"))
    best.engine.launch(x, y)
    torch.cuda.synchronize()
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x.cpu().numpy())
    if best.variant == 1:
        assert np.array_equal(y.cpu().numpy(), ref)
    else:
        L = O.scan2(m.rowptr, best.threshold, best.group_width, best.max_tpr)
        pv, pc = O.process(L, m.vals, m.colind)
        assert np.array_equal(y.cpu().numpy(), O.wpk_spmv(L, pv, pc, x.cpu().numpy()))
    best.engine.close()


@pytest.mark.gpu
def test_stream_read_helper():
    """The streaming-read measurement helper runs on aligned buffers, leaves the
    sink alone unless the data's XOR hits its fixed value, and refuses sizes
    that are not 16-byte multiples before launching."""
    import torch
    from warpkernel_amd._capi import WK_ERR_ARG, WK_OK, lib
    print("[gpu] stream read helper", flush=True)
    buf = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    sink = torch.full((4,), 7, dtype=torch.int32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    assert lib().wk_stream_read(buf.data_ptr(), buf.numel() * 8, sink.data_ptr(), st) == WK_OK
    assert lib().wk_stream_read(buf.data_ptr(), 0, None, st) == WK_OK
    torch.cuda.synchronize()
    assert sink.tolist() == [7, 7, 7, 7]  # XOR of zeros is 0, not the fixed value
    assert lib().wk_stream_read(buf.data_ptr(), 24, None, st) == WK_ERR_ARG
