"""x windows (stage_lds = 3) on the GPU through the C ABI.

Each group's distinct columns are gathered once into LDS and every slot
carries a 16-bit position in that list; the products and their order do not
change, so the bars are those of the other kernels: FP64 y bit-exact against
the oracle (CSR fma chain for WPK1, lane emulation for WPK2), run, run_x and
the residual; FP32 engines bit-identical to the staged kernel without windows.
Structure: every slot's window entry reproduces its padded column, in the
plain and in the Kr numbering (pads read column 0 of either).  Layouts whose
groups exceed the window limit, and engines whose columns were swapped, run
the staged kernel instead."""
import sys

import numpy as np
import pytest

from conftest import random_csr
from oracle import pyoracle as O
from warpkernel_amd import engine, fe, structure, structure2
from warpkernel_amd import _capi as capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

WIN = dict(stage_lds=3, nontemporal=1, rows_per_lane=1, k_tile=1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def nan_vec(n):
    return torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def check_windows(eng, L, cols):
    """Every slot's window entry is its padded column (plain and Kr)."""
    idx = eng.download(capi.WK_ENGINE_WIN_IDX).astype(np.int64)
    off = eng.download(capi.WK_ENGINE_WIN_OFF).astype(np.int64)
    wc = eng.download(capi.WK_ENGINE_WIN_COLS)
    wk = eng.download(capi.WK_ENGINE_WIN_COLS_KR)
    assert idx.size == L.allocate_nz and off.size == L.nwarps + 1
    group = np.repeat(np.arange(L.nwarps), L.max_nz_rows.astype(np.int64) * 64)
    _, pc = O.process(L, np.zeros(cols.size), cols)
    assert np.array_equal(wc[off[group] + idx], pc)
    assert np.array_equal(wk[off[group] + idx], O.reorder_columns_coalesced(L, cols))
    sizes = np.diff(off)
    assert sizes.max() <= 2048
    for g in range(0, L.nwarps, max(1, L.nwarps // 50)):  # ascending, distinct
        w = wc[off[g]:off[g + 1]]
        assert np.all(np.diff(w) >= 0)
    return int(sizes.max())


@pytest.mark.parametrize("name,th", [("hex_elastic_3d_32", 27), ("tet_poisson_3d_96", 5),
                                      ("tet_poisson_3d_96_perm", 5), ("q1_laplace_2d_128", 4)])
@pytest.mark.parametrize("u,wpb,st", [(8, 4, 3), (16, 1, 3), (8, 2, 3), (16, 1, 5), (8, 4, 5)])
def test_fe_window_run_run_x_residual(name, th, u, wpb, st):
    """stage 5: positions packed 10 bits each where every window fits 1024
    entries (all four matrices), else the stage 3 kernel."""
    say(f"[gpu] window {name} u={u} wpb={wpb} stage={st}")
    m = fe.CONFIGS[name]()
    rng = np.random.default_rng(21)
    x = rng.uniform(-1, 1, m.ncols)
    bvec = rng.uniform(-1, 1, m.nrows)
    cfg = dict(WIN, unroll=u, waves_per_block=wpb, stage_lds=st)
    for variant in (1, 2):
        if variant == 1:
            s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
            L = O.scan1(m.rowptr, 64)
            ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
        else:
            s = structure2(64, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, th)
            L = O.scan2(m.rowptr, th, 64, 32)
            pv, pc = O.process(L, m.vals, m.colind)
            ref = O.wpk_spmv(L, pv, pc, x)
        eng = engine(s, m.vals, m.colind, kr=True)
        if u == 8 and wpb == 4:
            check_windows(eng, L, m.colind)
        y = nan_vec(m.nrows)
        eng.launch(dev(x), y, cfg=cfg)
        assert np.array_equal(host(y), ref)
        yx = nan_vec(m.nrows)
        eng.launch_x(dev(O.reorder_x(L, x)), yx, cfg=cfg)
        assert np.array_equal(host(yx), ref[L.row_map])
        r = nan_vec(m.nrows)
        eng.residual(dev(bvec), dev(x), r, cfg=cfg)
        assert np.array_equal(host(r), bvec - ref)
        eng.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_random_window_or_fallback(mode):
    """Row-length mixes up to 5000 entries: some layouts exceed the window limit
    and run the staged kernel; all are bit-exact."""
    say(f"[gpu] window random mode={mode}")
    rng = np.random.default_rng(300 + mode)
    n = 6000
    rp, cols, vals = random_csr(rng, n, n, mode)
    x = rng.uniform(-1, 1, n)
    ref = O.csr_spmv(rp, cols, vals, x)
    s = structure(64).scan(int(rp[-1]), n, vals, rp, cols)
    eng = engine(s, vals, cols, kr=True)
    for st in (3, 5):  # packed positions only where every window fits 1024 entries
        y = nan_vec(n)
        eng.launch(dev(x), y, cfg=dict(WIN, unroll=8, waves_per_block=4, stage_lds=st))
        assert np.array_equal(host(y), ref), st
    if eng.download(capi.WK_ENGINE_WIN_OFF).size:
        check_windows(eng, O.scan1(rp, 64), cols)
    eng.close()


def test_fp32_window_matches_staged():
    say("[gpu] window fp32")
    m = fe.CONFIGS["hex_elastic_3d_32"]()
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, dtype="f32")
    x = dev(fe.uniform_x(m.ncols, seed=5).astype(np.float32))
    for acc in (0, 1):
        y2, y3 = nan_vec(m.nrows), nan_vec(m.nrows)
        eng.launch(x, y2, cfg=dict(WIN, stage_lds=2, unroll=8, waves_per_block=4, accumulate=acc))
        eng.launch(x, y3, cfg=dict(WIN, unroll=8, waves_per_block=4, accumulate=acc))
        assert torch.equal(y2, y3)
        y3.fill_(float("nan"))
        eng.launch(x, y3, cfg=dict(WIN, stage_lds=5, unroll=16, waves_per_block=1, accumulate=acc))
        assert torch.equal(y2, y3)
    eng.close()


@pytest.mark.parametrize("where", ["host", "device"])
def test_swapped_columns_drop_and_restore_the_windows(where):
    """The drivers' device_colinds swap (wpk1.cu:175) replaces the columns the
    windows were built from: launches run the staged kernel while they differ,
    and the windows return when the original columns are restored
    (wpk1.cu:196), from a host array or a device tensor."""
    say(f"[gpu] window after device_colinds swap ({where})")
    m = fe.CONFIGS["hex_elastic_3d_32"]()
    x = fe.uniform_x(m.ncols, seed=9)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    built = eng.x_window_entries()
    assert built > 0
    put = (lambda a: a) if where == "host" else dev
    restore = eng.device_colinds
    reorder_cols = s.reorder_columns_coalesced(m.colind)
    for st in (3, 5, 7):
        eng.device_colinds = put(reorder_cols)
        assert eng.x_window_entries() == 0
        yx = nan_vec(m.nrows)
        eng.launch_x(dev(s.reorder_x(x)), yx, cfg=dict(WIN, unroll=8, waves_per_block=4, stage_lds=st))
        assert np.array_equal(host(yx), ref[s.row_map])
        eng.device_colinds = put(restore)
        assert eng.x_window_entries() == built
        y = nan_vec(m.nrows)
        eng.launch(dev(x), y, cfg=dict(WIN, unroll=16, waves_per_block=1, stage_lds=st))
        assert np.array_equal(host(y), ref)
    # columns that differ in one slot only keep the windows off
    one = restore.copy()
    one[len(one) // 2] = (one[len(one) // 2] + 1) % m.ncols
    eng.device_colinds = put(one)
    assert eng.x_window_entries() == 0
    eng.device_colinds = put(restore)
    assert eng.x_window_entries() == built
    # out-of-range columns are refused and change nothing
    bad = restore.copy()
    bad[7] = m.ncols
    with pytest.raises(capi.WKError):
        eng.device_colinds = put(bad)
    assert np.array_equal(eng.device_colinds, restore) and eng.x_window_entries() == built
    eng.close()


def test_adjoint_engine_refuses_column_swap():
    say("[gpu] adjoint engine refuses device_colinds")
    m = fe.hex_laplace_3d(8)
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, adjoint=True)
    cols = eng.device_colinds
    with pytest.raises(capi.WKError) as e:
        eng.device_colinds = cols
    assert e.value.code == capi.WK_ERR_STATE
    eng.close()


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6, 7, 8, 16, 32])
def test_multi_rhs_all_paths_bitexact(k):
    """k right-hand sides through the default choice, the direct kernel at every
    k-tile and the windowed kernel at k-tiles 2, 4 and 8 (wider than k: one
    masked tile): FP64 bit-exact against the oracle (WPK1 CSR order, WPK2 lane
    order), run and run_x."""
    say(f"[gpu] multi-RHS k={k}")
    m = fe.CONFIGS["hex_elastic_3d_32"]() if k in (4, 8) else fe.tet_poisson_3d(40, drop=0.05)
    x = fe.uniform_x(m.ncols, k, seed=4000 + k).reshape(m.ncols, k)
    cfgs = [None] + [dict(k_tile=t, stage_lds=0) for t in (1, 2, 4, 8, 16, 32) if t <= k] + \
        [dict(WIN, k_tile=t, unroll=8, waves_per_block=w, stage_lds=st) for t in (2, 4, 8) if t < 2 * k
         for w in (1, 2, 4) for st in (3, 4)] + \
        [dict(WIN, k_tile=t, unroll=16, waves_per_block=w, stage_lds=4) for t in (4, 8) if t < 2 * k
         for w in (1, 2)] + \
        [dict(WIN, k_tile=8, unroll=8, waves_per_block=w, stage_lds=6) for w in (1, 2, 4)]
    for variant in (1, 2):
        if variant == 1:
            s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
            L = O.scan1(m.rowptr, 64)
            ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
        else:
            s = structure2(64, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 5)
            L = O.scan2(m.rowptr, 5, 64, 32)
            pv, pc = O.process(L, m.vals, m.colind)
            ref = O.wpk_spmv(L, pv, pc, x)
        eng = engine(s, m.vals, m.colind, kr=True)
        xr = np.empty_like(x)
        xr[L.row_map_inv] = x  # reorder_x, row by row
        xd, xrd = dev(x), dev(xr)
        for cfg in cfgs:
            y = torch.full((m.nrows, k), float("nan"), dtype=torch.float64, device="cuda")
            eng.launch(xd, y, cfg=cfg)
            assert np.array_equal(host(y), ref), (variant, cfg)
            eng.launch_x(xrd, y, cfg=cfg)
            assert np.array_equal(host(y), ref[L.row_map]), (variant, cfg, "run_x")
        eng.close()


@pytest.mark.parametrize("k", [2, 4, 8])
def test_multi_rhs_fp32_windowed_matches_direct(k):
    say(f"[gpu] multi-RHS fp32 k={k}")
    """Every window shape, WPK1 and WPK2, run and run_x: with values in
    registers (stage 4) up to one wave per right-hand side shares a window."""
    m = fe.CONFIGS["hex_elastic_3d_32"]()
    xh = fe.uniform_x(m.ncols, k, seed=4100 + k).reshape(m.ncols, k).astype(np.float32)
    x = dev(xh)
    for s in (structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind),
              structure2(64, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 27)):
        eng = engine(s, m.vals, m.colind, dtype="f32", kr=True)
        xr = np.empty_like(xh)
        xr[np.asarray(s.row_map_inv)] = xh  # reorder_x, row by row
        xrd = dev(xr)
        for acc in (0, 1):
            for launch, xin in ((eng.launch, x), (eng.launch_x, xrd)):
                y0 = torch.full((m.nrows, k), float("nan"), dtype=torch.float64, device="cuda")
                y1 = torch.full_like(y0, float("nan"))
                launch(xin, y0, cfg=dict(k_tile=k, stage_lds=0, accumulate=acc))
                for wpb in (1, 2, 4, 8):
                    for st, u in ((3, 8), (4, 8), (4, 16), (6, 8)):
                        y1.fill_(float("nan"))
                        launch(xin, y1, cfg=dict(WIN, k_tile=k, unroll=u, waves_per_block=wpb,
                                                 accumulate=acc, stage_lds=st))
                        assert torch.equal(y0, y1), (s.variant, acc, wpb, st, u, launch.__name__)
        eng.close()


@pytest.mark.parametrize("k", [3, 8, 12])
def test_multi_rhs_residual_and_strided_windows(k):
    """r = b - A x with k right-hand sides through the windowed k-tiles (the
    default and explicit k-tiles 2 and 4), and x, y, b as column slices of
    wider arrays (leading dimension k + 3): bit-exact against the oracle."""
    say(f"[gpu] multi-RHS residual k={k}")
    m = fe.CONFIGS["hex_elastic_3d_32"]()
    ld = k + 3
    xw = fe.uniform_x(m.ncols, ld, seed=4200 + k).reshape(m.ncols, ld)
    bw = fe.uniform_x(m.nrows, ld, seed=4300 + k).reshape(m.nrows, ld)
    x, b = xw[:, :k], bw[:, :k]
    ax = O.csr_spmv(m.rowptr, m.colind, m.vals, np.ascontiguousarray(x))
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    xd, bd = dev(xw), dev(bw)
    for cfg in (None, dict(WIN, k_tile=2, unroll=8, waves_per_block=1, xcd_remap=1),
                dict(WIN, k_tile=4, unroll=8, waves_per_block=2, xcd_remap=0),
                dict(WIN, k_tile=8, unroll=8, waves_per_block=2, xcd_remap=1, stage_lds=4)):
        yw = torch.full((m.nrows, ld), float("nan"), dtype=torch.float64, device="cuda")
        eng.launch(xd[:, :k], yw[:, :k], cfg=cfg)
        got = host(yw)
        assert np.array_equal(got[:, :k], ax), cfg
        assert np.isnan(got[:, k:]).all()  # columns beyond k untouched
        rw = torch.full_like(yw, float("nan"))
        eng.residual(bd[:, :k], xd[:, :k], rw[:, :k], cfg=cfg)
        assert np.array_equal(host(rw)[:, :k], b - ax), cfg
    eng.close()


@pytest.mark.parametrize("k", [16, 32])
def test_multi_rhs_columns_match_single_rhs_at_full_size(k):
    """Hex elasticity 64^3 (the headline matrix): every right-hand side of a
    k-wide launch (default shape: windows of k-tile values, or the values held
    in registers across passes) has the same per-row order as a one-column
    launch, so each column equals the k = 1 result bit for bit."""
    say(f"[gpu] multi-RHS columns vs k=1, hex_elastic_3d_64, k={k}")
    m = fe.CONFIGS["hex_elastic_3d_64"]()
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    x = fe.uniform_x(m.ncols, k, seed=4500 + k).reshape(m.ncols, k)
    y = torch.full((m.nrows, k), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(dev(x), y)
    yk = host(y)
    y1 = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    for j in (0, 3, k // 2, k - 1):
        eng.launch(dev(np.ascontiguousarray(x[:, j])), y1)
        assert np.array_equal(yk[:, j], host(y1)), (j, eng.get_launch(k=k).as_dict())
    eng.close()


@pytest.mark.parametrize("k", [3, 5, 6, 7, 9, 11, 13, 15])
def test_masked_last_tile_on_elasticity(k):
    """The default launch at k not a power of two on hex elasticity: a masked
    last tile for a rest of 3 or more columns, alone or after whole tiles in
    one launch (x of the missing columns read as 0, their y untouched), node
    windows for a rest of one.  run and the
    residual bit-exact against the oracle, and y past column k never written."""
    say(f"[gpu] masked last k-tile, hex elasticity 32^3, k={k}")
    m = fe.CONFIGS["hex_elastic_3d_32"]()
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    x = fe.uniform_x(m.ncols, k, seed=4100 + k).reshape(m.ncols, k)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    # y and b as column blocks of wider arrays: a write past column k shows
    ybig = torch.full((m.nrows, k + 3), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(dev(x), ybig[:, :k])
    yh = host(ybig)
    assert np.array_equal(yh[:, :k], ref)
    assert np.isnan(yh[:, k:]).all()
    b = fe.uniform_x(m.nrows, k, seed=4200 + k).reshape(m.nrows, k)
    rbig = torch.full((m.nrows, k + 3), float("nan"), dtype=torch.float64, device="cuda")
    eng.residual(dev(b), dev(x), rbig[:, :k])
    rh = host(rbig)
    assert np.array_equal(rh[:, :k], b - ref)
    assert np.isnan(rh[:, k:]).all()
    eng.close()
