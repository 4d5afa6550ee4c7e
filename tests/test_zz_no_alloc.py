"""wk_engine_set_colinds_noalloc against wk_engine_set_colinds (device
columns): the same columns land, the same x-window state follows, and the
launches after the swap give the same bits, on two calls in a row, the second
on a smaller engine; an out-of-range column is refused with nothing changed."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import _capi as capi, engine, fe, structure
from warpkernel_amd._capi import check, lib

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def noalloc(eng, pcols_dev):
    f = lib().wk_engine_set_colinds_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return f(eng._h, pcols_dev.data_ptr(), None)


def swap_and_run(m, s, eng, pcols, x, use_noalloc):
    """Swap in pcols (device), run and run_x; returns (columns, windows, y)."""
    d = dev(pcols)
    if use_noalloc:
        check(noalloc(eng, d))
    else:
        check(lib().wk_engine_set_colinds(eng._h, d.data_ptr(), 1, None))
    y = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(dev(x), y)
    torch.cuda.synchronize()
    return eng.device_colinds, eng.x_window_entries(), y.cpu().numpy()


@pytest.mark.parametrize("n", [24])
def test_noalloc_matches_set_colinds_two_calls(n):
    cases = [fe.hex_elastic_3d(n), fe.hex_elastic_3d(n // 2)]  # the second call is smaller
    for call, m in enumerate(cases):
        say(f"[gpu] set_colinds_noalloc call {call}: {m.name}, {m.nnz} nnz")
        s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
        x = fe.uniform_x(m.ncols, seed=9100 + call)
        ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
        twins = [engine(s, m.vals, m.colind), engine(s, m.vals, m.colind)]
        orig = twins[0].device_colinds
        built = twins[0].x_window_entries()
        assert built > 0
        kr = s.reorder_columns_coalesced(m.colind)
        xr = s.reorder_x(x)
        # Kr columns in, then the original columns back, through both entry points
        a = swap_and_run(m, s, twins[0], kr, xr, False)
        b = swap_and_run(m, s, twins[1], kr, xr, True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(b[0], kr)
        assert a[1] == b[1] == 0
        assert np.array_equal(a[2], b[2])
        assert np.array_equal(b[2], ref)  # WPK1_rx: Kr columns, x reordered, y in row order
        a = swap_and_run(m, s, twins[0], orig, x, False)
        b = swap_and_run(m, s, twins[1], orig, x, True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(b[0], orig)
        assert a[1] == b[1] == built
        assert np.array_equal(a[2], b[2]) and np.array_equal(b[2], ref)
        # out of range: refused, nothing changed
        bad = orig.copy()
        bad[len(bad) - 1] = m.ncols
        rc = noalloc(twins[1], dev(bad))
        assert rc == capi.WK_ERR_ARG, rc
        assert np.array_equal(twins[1].device_colinds, orig)
        assert twins[1].x_window_entries() == built
        for e in twins:
            e.close()
