"""The Kr row-sort variant (reorder_columns_rowsort, warpkernel.hpp:414-444;
driver pattern optimize.cu:645-704) on the GPU through the C ABI.

The row sort changes only the order of each row's entries, so the bar is the
oracle's summation over the row-sorted CSR: bit-exact for WPK1 (CSR fma chain)
and WPK2 (lane emulation on the row-sorted padded arrays), run and run_x.
Also: the reference's own construction (engine over vals[reorder_A_rows] and
the permuted columns, then run_x) gives the same bits as WK_ENGINE_ROWSORT;
update_values with CSR-order values lands in the sorted slots; out-of-range
padded columns are refused before they reach the device."""
import sys

import numpy as np
import pytest

from conftest import random_csr
from oracle import pyoracle as O
from warpkernel_amd import WKError, engine, fe, structure, structure2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def out_vec(n):
    return torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def check_engine(rp, cols, vals, x, variant, th, W):
    n = rp.size - 1
    nnz = int(rp[-1])
    if variant == 1:
        s = structure(W).scan(nnz, n, vals, rp, cols)
        L = O.scan1(rp, W)
    else:
        s = structure2(W, 32).scan(nnz, n, vals, rp, cols, th)
        L = O.scan2(rp, th, W, 32)
    kcols, src = O.reorder_columns_rowsort(L, rp, cols)
    cols2, vals2 = cols[src], vals[src]
    if variant == 1:
        ref = O.csr_spmv(rp, cols2, vals2, x)
    else:
        pv, pc = O.process(L, vals2, cols2)
        ref = O.wpk_spmv(L, pv, pc, x)
    eng = engine(s, vals, cols, kr=True, rowsort=True)
    y = out_vec(n)
    eng.launch(dev(x), y)
    assert np.array_equal(host(y), ref)
    xr = O.reorder_x(L, x)
    yr = out_vec(n)
    eng.launch_x(dev(xr), yr)
    assert np.array_equal(host(yr), ref[L.row_map])
    return s, L, eng, kcols, src, ref, xr


@pytest.mark.parametrize("name", ["hex_laplace_3d_48", "tet_poisson_3d_96_perm", "q1_laplace_2d_128"])
@pytest.mark.parametrize("W", [64, 32])
def test_fe_rowsort_run_and_run_x(name, W):
    say(f"[gpu] rowsort {name} W={W}")
    m = fe.CONFIGS[name]()
    x = fe.uniform_x(m.ncols, seed=2400)
    s, L, eng, kcols, src, ref, xr = check_engine(m.rowptr, m.colind, m.vals, x, 1, 0, W)
    eng.close()
    # the reference's construction: permuted values and columns, plain engine, run_x
    kc = s.reorder_columns_rowsort(m.colind)
    assert np.array_equal(kc, kcols) and np.array_equal(s.reorder_A_rows, src)
    e2 = engine(s, m.vals[s.reorder_A_rows], kc)
    yr = out_vec(m.nrows)
    e2.launch_x(dev(xr), yr)
    assert np.array_equal(host(yr), ref[L.row_map])
    e2.close()
    _, _, e3, _, _, _, _ = check_engine(m.rowptr, m.colind, m.vals, x, 2, 5, W)
    e3.close()


@pytest.mark.parametrize("seed", range(3))
def test_random_rowsort_with_duplicates(seed):
    say(f"[gpu] rowsort random seed={seed}")
    rng = np.random.default_rng(1200 + seed)
    n = 3000
    rp, cols, vals = random_csr(rng, n, n, seed)
    for r in range(0, n, 7):  # duplicate columns inside rows: ties keep CSR order
        if rp[r + 1] - rp[r] >= 2:
            cols[rp[r] + 1] = cols[rp[r]]
    x = rng.uniform(-1, 1, n)
    for variant, th in ((1, 0), (2, 2), (2, 11)):
        _, _, eng, _, _, _, _ = check_engine(rp, cols, vals, x, variant, th, 64)
        eng.close()


def test_update_values_follow_the_sort():
    say("[gpu] rowsort update_values")
    m = fe.CONFIGS["q1_laplace_2d_128"]()
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, m.ncols)
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, rowsort=True)
    v2 = rng.uniform(-1, 1, m.nnz)
    eng.update_values(dev(v2))
    fresh = engine(s, v2, m.colind, rowsort=True)
    y1, y2 = out_vec(m.nrows), out_vec(m.nrows)
    eng.launch(dev(x), y1)
    fresh.launch(dev(x), y2)
    assert np.array_equal(host(y1), host(y2))
    eng.close()
    fresh.close()


def test_out_of_range_columns_refused():
    say("[gpu] colinds range check")
    m = fe.CONFIGS["q1_laplace_2d_128"]()
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    good = eng.device_colinds
    bad = good.copy()
    bad[5] = m.ncols
    with pytest.raises(WKError):
        eng.device_colinds = bad
    from warpkernel_amd import _capi as capi
    from warpkernel_amd._capi import lib
    dbad = dev(bad)
    rc = lib().wk_engine_set_colinds(eng._h, dbad.data_ptr(), 1, None)
    assert rc == capi.WK_ERR_ARG
    assert np.array_equal(eng.device_colinds, good)
    bad_csr = m.colind.copy()
    bad_csr[3] = -1
    with pytest.raises(WKError):
        engine(s, m.vals, bad_csr)
    eng.close()
