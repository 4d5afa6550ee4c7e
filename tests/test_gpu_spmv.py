"""GPU parity of the WPK SpMV against the oracle (MI355X, through the C ABI).

Bars: layouts bit-exact (CPU tests); FP64 y bit-exact against the oracle's
summation order -- the CSR fma chain for WPK1, the lane-by-lane emulation for
WPK2 -- and within 1e-14 normwise of the CSR product for WPK2."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import WKError, engine, fe, structure, structure2

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


def normwise(y, ref):
    return float(np.linalg.norm(y - ref) / max(np.linalg.norm(ref), 1e-300))


_CACHE = {}


def matrix(name):
    if name not in _CACHE:
        _CACHE.clear()
        _CACHE[name] = fe.CONFIGS[name]()
    return _CACHE[name]


def run_wpk1(m, W, x):
    s = structure(W).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, kr=True)
    dx = dev(x)
    dy = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(dx, dy)
    y = host(dy)
    dxr = dev(s.reorder_x(x))
    dyr = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch_x(dxr, dyr)
    return s, eng, y, host(dyr)


@pytest.mark.parametrize("name", ["q1_laplace_2d_128", "hex_elastic_3d_64", "tet_poisson_3d_96",
                                  "tet_poisson_3d_96_perm", "hex_laplace_3d_48"])
@pytest.mark.parametrize("W", [64, 32])
def test_wpk1_run_and_run_x_bitexact(name, W):
    say(f"[gpu] wpk1 {name} W={W}")
    m = matrix(name)
    x = fe.uniform_x(m.ncols, seed=2001)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    s, eng, y, yr = run_wpk1(m, W, x)
    assert np.array_equal(y, ref)
    assert np.array_equal(yr, ref[s.row_map])
    assert eng.verify(ref, y) and eng.verify_x(ref, yr)


@pytest.mark.parametrize("name,thresholds", [
    ("q1_laplace_2d_128", (1, 4, 9)),
    ("hex_elastic_3d_64", (10, 27, 81)),
    ("tet_poisson_3d_96", (2, 5, 15)),
    ("tet_poisson_3d_96_perm", (3,)),
])
@pytest.mark.parametrize("W", [64, 32])
def test_wpk2_bitexact_vs_emulation(name, thresholds, W):
    m = matrix(name)
    x = fe.uniform_x(m.ncols, seed=2002)
    csr = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    for T in thresholds:
        for mt in ((32, 64) if W == 64 else (32,)):
            say(f"[gpu] wpk2 {name} W={W} T={T} max_tpr={mt}")
            s = structure2(W, mt).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, T)
            L = O.scan2(m.rowptr, T, W, mt)
            pv, pc = O.process(L, m.vals, m.colind)
            ref = O.wpk_spmv(L, pv, pc, x)
            eng = engine(s, m.vals, m.colind, kr=True)
            dy = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
            eng.launch(dev(x), dy)
            y = host(dy)
            assert np.array_equal(y, ref)
            assert normwise(y, csr) <= 1e-14
            kr = O.reorder_columns_coalesced(L, m.colind)
            xr = O.reorder_x(L, x)
            refx = O.wpk_spmv(L, pv, kr, xr, kr=True)
            dyr = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
            eng.launch_x(dev(xr), dyr)
            assert np.array_equal(host(dyr), refx)
            del eng


def test_dropin_driver_flow_wpk1():
    """The reference driver's sequence (wpk1.cu:114-242): run, swap in the Kr
    columns, run_x on reordered x, restore, run with reordered x and row map."""
    say("[gpu] drop-in driver flow")
    m = matrix("hex_elastic_3d_32")
    x = fe.uniform_x(m.ncols, seed=2003)
    yref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    k1 = structure()
    k1.scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    reorder_cols = k1.reorder_columns_coalesced(m.colind)
    dreordered_x = dev(k1.reorder_x(x))
    eng = engine(k1, m.vals, m.colind)
    dx = dev(x)
    for wpb in (1, 2, 4, 8):
        blocksize = wpb * 64
        nblocks = (k1.nwarps + wpb - 1) // wpb
        dy = torch.zeros(m.nrows, dtype=torch.float64, device="cuda")
        t = eng.run(nblocks, blocksize, dx, dy)
        assert t > 0 and eng.verify(yref, host(dy))
        restore_col = eng.device_colinds
        dy = torch.zeros(m.nrows, dtype=torch.float64, device="cuda")
        eng.device_colinds = reorder_cols
        eng.run_x(nblocks, blocksize, dreordered_x, dy)
        assert eng.verify_x(yref, host(dy))
        dy = torch.zeros(m.nrows, dtype=torch.float64, device="cuda")
        eng.run(nblocks, blocksize, dreordered_x, dy)  # WPK1_rx
        assert np.array_equal(host(dy), yref)
        eng.device_colinds = restore_col
        assert np.array_equal(eng.device_colinds, restore_col)


@pytest.mark.parametrize("cfg", [
    dict(waves_per_block=1), dict(waves_per_block=8), dict(unroll=4), dict(unroll=16),
    dict(xcd_remap=0), dict(nontemporal=0), dict(nontemporal=2), dict(nontemporal=2, waves_per_block=4),
    dict(nontemporal=2, rows_per_lane=2), dict(nontemporal=2, unroll=16),
    dict(nontemporal=2, rows_per_lane=2, unroll=4),
    dict(rows_per_lane=2), dict(rows_per_lane=2, unroll=4),
    dict(rows_per_lane=4, unroll=4), dict(rows_per_lane=4), dict(blocks_per_cu=3),
    dict(rows_per_lane=2, waves_per_block=1, blocks_per_cu=6),
    dict(stage_lds=1), dict(stage_lds=1, unroll=16), dict(stage_lds=1, waves_per_block=1, unroll=16),
    dict(stage_lds=2), dict(stage_lds=2, unroll=16), dict(stage_lds=2, waves_per_block=2, unroll=16),
])
def test_launch_options_do_not_change_results(cfg):
    say(f"[gpu] launch options {cfg}")
    m = fe.tet_poisson_3d(40, drop=0.05)  # mixed row lengths, empty rows, odd group count
    x = fe.uniform_x(m.ncols, seed=2004)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, kr=True)
    dy = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(dev(x), dy, cfg=cfg)
    assert np.array_equal(host(dy), ref)
    dyr = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch_x(dev(s.reorder_x(x)), dyr, cfg=cfg)
    assert np.array_equal(host(dyr), ref[s.row_map])
    # WPK2: split groups next to unsplit ones in one wave
    L = O.scan2(m.rowptr, 4, 64, 32)
    pv, pc = O.process(L, m.vals, m.colind)
    s2 = structure2().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 4)
    e2 = engine(s2, m.vals, m.colind)
    dy2 = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    e2.launch(dev(x), dy2, cfg=cfg)
    assert np.array_equal(host(dy2), O.wpk_spmv(L, pv, pc, x))


def test_empty_rows_written_as_zero():
    say("[gpu] empty rows")
    m = fe.tet_poisson_3d(20, drop=0.6)
    empty = m.row_lengths() == 0
    assert empty.any()
    x = fe.uniform_x(m.ncols, seed=2005)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    L2 = O.scan2(m.rowptr, 3, 64, 32)
    pv, pc = O.process(L2, m.vals, m.colind)
    ref2 = O.wpk_spmv(L2, pv, pc, x)  # tree order of the split rows
    for cls, args, want in ((structure, (), ref), (structure2, (3,), ref2)):
        s = cls().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, *args)
        eng = engine(s, m.vals, m.colind, kr=True)
        dy = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        eng.launch(dev(x), dy)
        y = host(dy)
        assert np.all(y[empty] == 0) and np.array_equal(y, want)


def test_tiny_and_empty_matrices():
    say("[gpu] tiny matrices")
    for rp in ([0, 2, 7, 7, 10, 15, 16], [0], [0, 0, 0], [0, 1]):
        rp = np.array(rp, np.int32)
        n = rp.size - 1
        nnz = int(rp[-1])
        rng = np.random.default_rng(n)
        ci = rng.integers(0, max(n, 1), nnz).astype(np.int32)
        for r in range(n):
            ci[rp[r]:rp[r + 1]].sort()
        va = rng.uniform(-1, 1, nnz)
        x = rng.uniform(-1, 1, max(n, 1))[:n]
        ref = O.csr_spmv(rp, ci, va, x) if n else np.zeros(0)
        for W in (32, 64):
            s = structure(W).scan(nnz, n, va, rp, ci)
            eng = engine(s, va, ci)
            dy = torch.full((max(n, 1),), float("nan"), dtype=torch.float64, device="cuda")
            eng.launch(dev(x) if n else dy, dy)
            assert np.array_equal(host(dy)[:n], ref)


def test_update_values_on_device():
    say("[gpu] GPU-side process")
    m = matrix("hex_laplace_3d_48")
    x = fe.uniform_x(m.ncols, seed=2006)
    s = structure2().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 9)
    eng = engine(s, m.vals, m.colind, adjoint=True)
    v2 = np.random.default_rng(9).uniform(-2, 2, m.nnz)
    eng.update_values(dev(v2))
    dy = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    eng.launch(dev(x), dy)
    L = O.scan2(m.rowptr, 9, 64, 32)
    pv, pc = O.process(L, v2, m.colind)
    assert np.array_equal(host(dy), O.wpk_spmv(L, pv, pc, x))
    # the transposed layout was updated too
    dx = torch.empty(m.ncols, dtype=torch.float64, device="cuda")
    eng.run_adjoint(dev(x), dx)
    cp, ri, perm = O.transpose(m.nrows, m.ncols, m.rowptr, m.colind)
    LT = O.scan2(cp, 9, 64, 32)
    pvt, pct = O.process(LT, v2[perm], ri)
    assert np.array_equal(host(dx), O.wpk_spmv(LT, pvt, pct, x))


def test_residual_fp64():
    say("[gpu] residual")
    m = matrix("hex_laplace_3d_48")
    x = fe.uniform_x(m.ncols, seed=2007)
    b = fe.uniform_x(m.nrows, seed=3007)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    r = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    eng.residual(dev(b), dev(x), r)
    assert np.array_equal(host(r), b - O.csr_spmv(m.rowptr, m.colind, m.vals, x))


def test_errors_cross_the_abi():
    say("[gpu] error paths")
    m = fe.q1_laplace_2d(8)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, dtype="f32")
    r = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    x = torch.zeros(m.ncols, dtype=torch.float64, device="cuda")
    with pytest.raises(WKError):
        eng.residual(x, x, r)
    with pytest.raises(WKError):
        eng.run_adjoint(x, r)  # no transposed layout
    with pytest.raises(WKError):
        eng.launch(x, r, cfg=dict(waves_per_block=3))


def test_autotuner_keeps_results_and_stores_choice():
    say("[gpu] autotuner")
    m = fe.tet_poisson_3d(40, drop=0.05)
    x = fe.uniform_x(m.ncols, seed=2008)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, kr=True, adjoint=True)
    dx, dy = dev(x), torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    best, ms = eng.tune(dx, dy, reps=3)
    assert ms > 0 and eng.get_launch().as_dict() == best
    dy.fill_(float("nan"))
    eng.launch(dx, dy)
    assert np.array_equal(host(dy), ref)
    best_adj, _ = eng.tune(dy, torch.empty(m.ncols, dtype=torch.float64, device="cuda"), op=2, reps=3)
    assert eng.get_launch(op=2).as_dict() == best_adj
    k = 8
    xk = fe.uniform_x(m.ncols, k, seed=2009).reshape(m.ncols, k)
    yk = torch.empty((m.nrows, k), dtype=torch.float64, device="cuda")
    bk, _ = eng.tune(dev(xk), yk, reps=3)
    assert bk["k_tile"] in (1, 2, 4, 8)
    yk.fill_(float("nan"))
    eng.launch(dev(xk), yk)
    assert np.array_equal(host(yk), O.csr_spmv(m.rowptr, m.colind, m.vals, xk))


@pytest.mark.parametrize("cfg", [None, dict(rows_per_lane=1, unroll=8), dict(rows_per_lane=2, unroll=4),
                                 dict(rows_per_lane=4, unroll=8, waves_per_block=1, xcd_remap=0)])
def test_gather_ceiling_sums_x_over_slots(cfg):
    """The gather-ceiling kernel is the direct kernel with every slot's value
    1.0: the layout SpMV with unit values (pads read column 0)."""
    say(f"[gpu] gather ceiling {cfg}")
    m = fe.tet_poisson_3d(40, drop=0.05)
    x = fe.uniform_x(m.ncols, seed=2006)
    for variant in (1, 2):
        if variant == 1:
            L = O.scan1(m.rowptr, 64)
            s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
        else:
            L = O.scan2(m.rowptr, 4, 64, 32)
            s = structure2().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 4)
        _, pc = O.process(L, m.vals, m.colind)
        ref = O.wpk_spmv(L, np.ones(len(pc)), pc, x)
        eng = engine(s, m.vals, m.colind)
        dy = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        eng.gather_ceiling(dev(x), dy, cfg=cfg)
        assert np.array_equal(host(dy), ref), variant


@pytest.mark.parametrize("name", ["hex_laplace_3d_48", "tet_poisson_3d_96"])
@pytest.mark.parametrize("k", [8, 16, 32])
def test_multi_rhs_defaults_on_short_rows_bitexact(name, k):
    """The layout defaults at several right-hand sides on short rows (27-entry hex
    diffusion rows take the values-in-registers kernel from k = 8; 15-entry tet
    rows keep the register-fed windows): bit-exact against the CSR oracle
    column by column, run and run_x."""
    say(f"[gpu] multi-RHS default {name} k={k}")
    m = matrix(name)
    x = fe.uniform_x(m.ncols, k, seed=2400 + k).reshape(m.ncols, k)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, kr=True)
    y = torch.full((m.nrows, k), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(dev(x), y)
    assert np.array_equal(host(y), ref), eng.get_launch(k=k).as_dict()
    xr = np.empty_like(x)
    xr[np.asarray(s.row_map_inv)] = x
    y.fill_(float("nan"))
    eng.launch_x(dev(xr), y)
    assert np.array_equal(host(y), ref[np.asarray(s.row_map)])
