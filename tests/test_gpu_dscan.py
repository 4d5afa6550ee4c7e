"""The scan/process step on the GPU (wk_engine_create_device) against the host
scan: every layout array, the padded values / columns / Kr columns and the
CSR-to-slot map bit for bit, then the SpMV of the resulting engine bit-exact
against the oracle.  Malformed CSR input is refused by the device-side check."""
import sys

import numpy as np
import pytest

from conftest import random_csr
from oracle import pyoracle as O
from warpkernel_amd import WKError, engine, fe
from warpkernel_amd import _capi as capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def compare(rp, cols, vals, variant, threshold=0, dtype="f64", kr=True):
    """Device-scanned engine vs the oracle's scan + process; returns both."""
    nrows = rp.size - 1
    L = O.scan1(rp, 64) if variant == 1 else O.scan2(rp, threshold, 64, 32)
    e = engine.from_device(dev(rp), dev(cols), dev(vals), variant=variant, threshold=threshold,
                           dtype=dtype, kr=kr)
    k = e.kernel
    assert (k.nrows, k.ncols, k.nz) == (nrows, nrows, int(rp[-1]))
    for f in ("nznrows", "allocate_nz", "nwarps"):
        assert getattr(k, f) == getattr(L, f), f
    if variant == 2:
        assert (k.min_nz, k.max_nz) == (L.min_nz, L.max_nz)
    for f in ("row_map", "row_map_inv", "reorder_rows", "warp_offsets", "max_nz_rows"):
        assert np.array_equal(getattr(k, f), getattr(L, f)), f
    if variant == 2:
        assert np.array_equal(k.threads_per_row, L.threads_per_row)
        assert np.array_equal(k.row_offset_warp, L.row_offset_warp)
    pv, pc = O.process(L, vals, cols)
    got_v = e.download(capi.WK_ENGINE_VALUES)
    want_v = pv if dtype == "f64" else pv.astype(np.float32)
    assert np.array_equal(got_v.view(np.uint8), want_v.view(np.uint8))
    assert np.array_equal(e.download(capi.WK_ENGINE_COLS), pc)
    if kr:
        assert np.array_equal(e.download(capi.WK_ENGINE_COLS_KR), O.reorder_columns_coalesced(L, cols))
    if nrows:
        check_windows_match_host(e, rp, cols, vals, variant, threshold, dtype, kr)
    return L, e, pv, pc


def check_windows_match_host(e, rp, cols, vals, variant, threshold, dtype, kr):
    """The device-built x windows equal the host-built ones of the same layout
    (the device builder leaves out layouts with a group of over 8192 slots)."""
    from warpkernel_amd import structure, structure2
    n = rp.size - 1
    s = (structure(64).scan(int(rp[-1]), n, vals, rp, cols) if variant == 1 else
         structure2(64, 32).scan(int(rp[-1]), n, vals, rp, cols, threshold))
    h = engine(s, vals, cols, dtype=dtype, kr=kr)
    arrays = [capi.WK_ENGINE_WIN_IDX, capi.WK_ENGINE_WIN_OFF, capi.WK_ENGINE_WIN_COLS]
    if kr:
        arrays.append(capi.WK_ENGINE_WIN_COLS_KR)
    host_has = h.download(capi.WK_ENGINE_WIN_OFF).size > 0
    dev_has = e.download(capi.WK_ENGINE_WIN_OFF).size > 0
    if host_has and int(np.max(s.max_nz_rows)) * 64 <= 8192:
        assert dev_has
    if host_has and dev_has:
        for w in arrays:
            assert np.array_equal(e.download(w), h.download(w)), w
    h.close()


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("nrows", [1, 65, 3001])
def test_random_layouts_and_run(mode, nrows):
    say(f"[gpu] dscan random mode={mode} nrows={nrows}")
    rng = np.random.default_rng(100 * mode + nrows)
    rp, cols, vals = random_csr(rng, nrows, nrows, mode)
    x = rng.uniform(-1, 1, nrows)
    ref = O.csr_spmv(rp, cols, vals, x)
    L, e, _, _ = compare(rp, cols, vals, 1)
    dy = torch.full((nrows,), float("nan"), dtype=torch.float64, device="cuda")
    e.launch(dev(x), dy)
    assert np.array_equal(host(dy), ref)
    dyr = torch.full((nrows,), float("nan"), dtype=torch.float64, device="cuda")
    e.launch_x(dev(O.reorder_x(L, x)), dyr)
    assert np.array_equal(host(dyr), ref[L.row_map])
    e.close()
    for th in (1, 7, 60):
        L2, e2, pv, pc = compare(rp, cols, vals, 2, th)
        dy2 = torch.full((nrows,), float("nan"), dtype=torch.float64, device="cuda")
        e2.launch(dev(x), dy2)
        assert np.array_equal(host(dy2), O.wpk_spmv(L2, pv, pc, x))
        e2.close()


@pytest.mark.parametrize("name,thresholds", [
    ("hex_elastic_3d_32", (27,)),
    ("tet_poisson_3d_96", (5,)),
    ("tet_poisson_3d_96_perm", ()),
    ("q1_laplace_2d_128", (4,)),
])
def test_fe_layouts(name, thresholds):
    say(f"[gpu] dscan {name}")
    m = fe.CONFIGS[name]()
    x = fe.uniform_x(m.ncols, seed=2003)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    _, e, _, _ = compare(m.rowptr, m.colind, m.vals, 1)
    dy = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    e.launch(dev(x), dy)
    assert np.array_equal(host(dy), ref)
    e.close()
    for th in thresholds:
        _, e2, _, _ = compare(m.rowptr, m.colind, m.vals, 2, th)
        e2.close()


def test_f32_values_and_update():
    say("[gpu] dscan f32 + update_values")
    rng = np.random.default_rng(5)
    rp, cols, vals = random_csr(rng, 700, 700, 1)
    _, e, _, _ = compare(rp, cols, vals, 1, dtype="f32", kr=False)
    e.close()
    L, e, _, _ = compare(rp, cols, vals, 1)
    vals2 = rng.uniform(-1, 1, vals.size)
    e.update_values(dev(vals2))
    torch.cuda.synchronize()
    pv, _ = O.process(L, vals2, cols)
    assert np.array_equal(e.download(capi.WK_ENGINE_VALUES), pv)
    e.close()


def test_empty_and_all_empty_rows():
    say("[gpu] dscan empty")
    rp = np.zeros(1, np.int32)
    e = engine.from_device(dev(rp), dev(np.zeros(0, np.int32)), dev(np.zeros(0)), ncols=0)
    assert e.kernel.nrows == 0 and e.kernel.allocate_nz == 0
    e.close()
    rp = np.zeros(130, np.int32)
    _, e, _, _ = compare(rp, np.zeros(0, np.int32), np.zeros(0), 1)
    dy = torch.full((129,), float("nan"), dtype=torch.float64, device="cuda")
    e.launch(dev(np.ones(129)), dy)
    assert np.array_equal(host(dy), np.zeros(129))
    e.close()


def test_malformed_csr_refused():
    say("[gpu] dscan malformed")
    rp = np.array([0, 2, 1, 3], np.int32)  # decreasing
    cols = np.array([0, 1, 2], np.int32)
    with pytest.raises(WKError) as ex:
        engine.from_device(dev(rp), dev(cols), dev(np.ones(3)))
    assert ex.value.code == capi.WK_ERR_CSR
    rp = np.array([0, 1, 2, 3], np.int32)
    with pytest.raises(WKError) as ex:
        engine.from_device(dev(rp), dev(np.array([0, 3, 1], np.int32)), dev(np.ones(3)))
    assert ex.value.code == capi.WK_ERR_CSR
    with pytest.raises(WKError) as ex:  # rowptr ends short of nnz
        engine.from_device(dev(np.array([0, 1, 1, 2], np.int32)), dev(cols), dev(np.ones(3)))
    assert ex.value.code == capi.WK_ERR_CSR
    rp = np.array([0, 9000], np.int32)  # longer than the device histogram holds
    with pytest.raises(WKError) as ex:
        engine.from_device(dev(rp), dev(np.zeros(9000, np.int32)), dev(np.ones(9000)), ncols=1)
    assert ex.value.code == capi.WK_ERR_LIMIT
