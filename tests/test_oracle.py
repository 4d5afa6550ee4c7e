"""The oracle itself, pinned against the hand-worked vectors, closed forms and
brute-force restatements (CPU only)."""
import numpy as np
import pytest

from conftest import random_csr, random_rowptr
from oracle import pyoracle as O


def _rowptr_of(case):
    if "rowptr" in case:
        return np.array(case["rowptr"], np.int32)
    spec = case["rowptr_spec"]
    lens = np.zeros(spec["nrows"], np.int64)
    for k, v in spec["lengths"].items():
        lens[int(k)] = v
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)


def _scan(case, rp):
    if case["variant"] == 1:
        return O.scan1(rp, case["W"])
    return O.scan2(rp, case["threshold"], case["W"], case["max_tpr"])


@pytest.mark.parametrize("idx", range(4))
def test_oracle_matches_golden(golden, idx):
    case = golden["cases"][idx]
    rp = _rowptr_of(case)
    L = _scan(case, rp)
    e = case["expect"]
    for key, want in e.items():
        if key == "pad_slots":
            used = np.zeros(L.allocate_nz, bool)
            used[L.reorder_rows] = True
            assert np.array_equal(np.flatnonzero(~used), want)
        elif key == "max_nz_rows_all":
            assert np.all(L.max_nz_rows == want)
        elif key == "threads_per_row_all":
            assert np.all(L.threads_per_row == want)
        elif key == "reorder_rows_rule":
            tid, i = np.meshgrid(np.arange(25), np.arange(4), indexing="ij")
            assert np.array_equal(L.reorder_rows[(i + 4 * tid).ravel()], (tid + 32 * i).ravel())
        elif isinstance(want, list):
            assert np.array_equal(getattr(L, key), want), key
        else:
            assert getattr(L, key) == want, key


def test_oracle_closed_form_hex_elasticity(golden):
    """allocate_nz of the 64^3 hex elasticity pattern from its row-length census."""
    cf = golden["closed_forms"]["hex_elasticity_64"]
    lens = np.concatenate([np.full(750141, 81), np.full(71442, 54), np.full(2268, 36), np.full(24, 24)])
    # the census is order-free for the scan; interleave to defeat a sort-free shortcut
    lens = np.random.default_rng(7).permutation(lens)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    assert rp[-1] == cf["nnz"]
    for W in (32, 64):
        L = O.scan1(rp, W)
        assert L.nwarps == cf[f"W{W}"]["nwarps"]
        assert L.allocate_nz == cf[f"W{W}"]["allocate_nz"]


def _decode_check(L, rp):
    """Brute-force: every CSR entry lands in a distinct slot, and the slot
    decodes (group, lane, step) back to the row and chunk position the layout
    promises."""
    W = L.W
    rr = L.reorder_rows.astype(np.int64)
    assert rr.size == rp[-1]
    assert np.unique(rr).size == rr.size
    assert rr.size == 0 or rr.max() < L.allocate_nz
    g = np.searchsorted(L.warp_offsets, rr, side="right") - 1
    off = rr - L.warp_offsets[g]
    lane, step = off % W, off // W
    assert np.all(step < L.max_nz_rows[g])
    rows = np.repeat(np.arange(L.nrows), np.diff(rp))
    pos_in_row = np.arange(rr.size) - rp[rows]
    if L.variant == 1:
        assert np.array_equal(L.row_map[g * W + lane], rows)
        assert np.array_equal(step, pos_in_row)
    else:
        tpr = L.threads_per_row[g].astype(np.int64)
        A = L.max_nz_rows[g].astype(np.int64)
        sorted_row = L.row_offset_warp[g].astype(np.int64) + lane // tpr
        assert np.array_equal(L.row_map[sorted_row], rows)
        assert np.array_equal((lane % tpr) * A + step, pos_in_row)
    # row_map sorts by length descending, then row ascending
    lens = np.diff(rp)[L.row_map]
    key = -lens.astype(np.int64) * (L.nrows + 1) + L.row_map
    assert np.all(np.diff(key) > 0)
    assert np.array_equal(L.row_map_inv[L.row_map], np.arange(L.nrows))


@pytest.mark.parametrize("seed", range(12))
def test_oracle_scan_decodes(seed):
    rng = np.random.default_rng(seed)
    for mode in range(4):
        rp = random_rowptr(rng, int(rng.integers(1, 300)), mode)
        for W in (32, 64):
            _decode_check(O.scan1(rp, W), rp)
            for T in (0, 1, 3, 17):
                for mt in (2, 8, 32):
                    _decode_check(O.scan2(rp, T, W, mt), rp)


def test_oracle_empty_matrix():
    rp = np.zeros(5, np.int32)
    L = O.scan1(rp, 64)
    assert (L.nznrows, L.nwarps, L.allocate_nz) == (0, 0, 0)
    L2 = O.scan2(rp, 3, 64, 32)
    assert list(L2.row_offset_warp) == [4]
    assert (L2.min_nz, L2.max_nz) == (4, 0)


def test_process_and_kr_reconstruct_csr():
    rng = np.random.default_rng(3)
    rp, ci, va = random_csr(rng, 200, 200, 1)
    L = O.scan1(rp, 64)
    pv, pc = O.process(L, va, ci)
    assert np.array_equal(pv[L.reorder_rows], va)
    assert np.array_equal(pc[L.reorder_rows], ci)
    pads = np.ones(L.allocate_nz, bool)
    pads[L.reorder_rows] = False
    assert np.all(pv[pads] == 0) and np.all(pc[pads] == 0)
    kr = O.reorder_columns_coalesced(L, ci)
    assert np.array_equal(kr[L.reorder_rows], L.row_map_inv[ci])
    assert np.all(kr[pads] == 0)
    x = rng.uniform(-1, 1, 200)
    xr = O.reorder_x(L, x)
    assert np.array_equal(xr[L.row_map_inv], x)


@pytest.mark.parametrize("mode", range(4))
def test_csr_spmv_against_dense(mode):
    rng = np.random.default_rng(10 + mode)
    rp, ci, va = random_csr(rng, 150, 120, mode)
    D = np.zeros((150, 120))
    for r in range(150):
        for t in range(rp[r], rp[r + 1]):
            D[r, ci[t]] += va[t]
    x = rng.uniform(-1, 1, (120, 3))
    for fused in (True, False):
        y = O.csr_spmv(rp, ci, va, x, fused=fused)
        scale = np.abs(D) @ np.abs(x) + 1e-300
        assert np.all(np.abs(y - D @ x) <= 1e-13 * scale)
    # fused differs from unfused somewhere on random data (fma is really used)
    yf = O.csr_spmv(rp, ci, va, x[:, 0], fused=True)
    yu = O.csr_spmv(rp, ci, va, x[:, 0], fused=False)
    if rp[-1] > 200:
        assert not np.array_equal(yf, yu)


def _fma(a, b, c):
    """Correctly rounded a*b + c (float() of a Fraction rounds to nearest even)."""
    from fractions import Fraction as F
    return float(F(a) * F(b) + F(c))


def test_csr_spmv_first_term_then_fma():
    """Row of three: ((a0*x0) fused with a1*x1) fused with a2*x2."""
    rp = np.array([0, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    va = np.array([0.1, 0.7, 1.3])
    x = np.array([0.3, -0.9, 0.45])
    s = 0.1 * 0.3
    s = _fma(0.7, -0.9, s)
    s = _fma(1.3, 0.45, s)
    assert O.csr_spmv(rp, ci, va, x)[0] == s
    # and the unfused chain rounds the products separately
    u = 0.1 * 0.3
    u = u + 0.7 * -0.9
    u = u + 1.3 * 0.45
    assert O.csr_spmv(rp, ci, va, x, fused=False)[0] == u


@pytest.mark.parametrize("seed", range(6))
def test_wpk_emulation_variant1_bitexact_vs_csr(seed):
    rng = np.random.default_rng(100 + seed)
    rp, ci, va = random_csr(rng, 300, 300, seed % 4)
    x = rng.uniform(-1, 1, (300, 2))
    ref = O.csr_spmv(rp, ci, va, x)
    for W in (32, 64):
        L = O.scan1(rp, W)
        pv, pc = O.process(L, va, ci)
        assert np.array_equal(O.wpk_spmv(L, pv, pc, x), ref)
        # Kr form: y'[i] = y[row_map[i]]
        kr = O.reorder_columns_coalesced(L, ci)
        xr = np.empty_like(x)
        xr[L.row_map_inv] = x
        assert np.array_equal(O.wpk_spmv(L, pv, kr, xr, kr=True), ref[L.row_map])


def _chunk_tree_reference(rp, ci, va, x, L):
    """Independent WPK2 restatement: chunks of A (= max_nz_rows of the row's
    group) per lane, pads as +0*x[0], then a pairwise tree over tpr lanes."""
    y = np.zeros(L.nrows)
    W = L.W
    for g in range(L.nwarps):
        tpr, A = int(L.threads_per_row[g]), int(L.max_nz_rows[g])
        for r in range(W // tpr):
            rs = int(L.row_offset_warp[g]) + r
            if rs >= L.nznrows:
                continue
            row = int(L.row_map[rs])
            lanes = []
            for t in range(tpr):
                ents = [(va[q], x[ci[q]]) for q in range(rp[row] + t * A, min(rp[row] + (t + 1) * A, rp[row + 1]))]
                ents += [(0.0, x[0])] * (A - len(ents))
                s = ents[0][0] * ents[0][1]
                for a, b in ents[1:]:
                    s = _fma(a, b, s)
                lanes.append(s)
            while len(lanes) > 1:
                lanes = [lanes[i] + lanes[i + 1] for i in range(0, len(lanes), 2)]
            y[row] = lanes[0]
    return y


@pytest.mark.parametrize("seed", range(4))
def test_wpk_emulation_variant2_vs_chunk_tree(seed):
    rng = np.random.default_rng(200 + seed)
    rp, ci, va = random_csr(rng, 120, 150, 1)
    x = rng.uniform(-1, 1, 150)
    L = O.scan2(rp, 7, 64, 32)
    pv, pc = O.process(L, va, ci)
    assert np.array_equal(O.wpk_spmv(L, pv, pc, x), _chunk_tree_reference(rp, ci, va, x, L))


@pytest.mark.parametrize("seed", range(4))
def test_wpk_emulation_variant2_close_to_csr(seed):
    rng = np.random.default_rng(300 + seed)
    rp, ci, va = random_csr(rng, 200, 220, seed % 4)
    x = rng.uniform(-1, 1, 220)
    ref = O.csr_spmv(rp, ci, va, x)
    for T in (1, 4, 19):
        L = O.scan2(rp, T, 64, 32)
        pv, pc = O.process(L, va, ci)
        y = O.wpk_spmv(L, pv, pc, x)
        scale = np.abs(O.csr_spmv(rp, ci, np.abs(va), np.abs(x))) + 1e-300
        assert np.all(np.abs(y - ref) <= 64 * 2.2e-16 * scale)


def test_transpose_and_dvals():
    rng = np.random.default_rng(5)
    rp, ci, va = random_csr(rng, 90, 70, 0)
    cp, ri, perm = O.transpose(90, 70, rp, ci)
    D = np.zeros((90, 70))
    D[np.repeat(np.arange(90), np.diff(rp)), ci] = va
    T = np.zeros((70, 90))
    T[np.repeat(np.arange(70), np.diff(cp)), ri] = va[perm]
    assert np.array_equal(T, D.T)
    # rows of each column ascending
    for c in range(70):
        assert np.all(np.diff(ri[cp[c]:cp[c + 1]]) > 0)
    # dvals = d(0.5*||A x - b||^2)/dvals by central differences (quadratic: exact up to rounding)
    x = rng.uniform(-1, 1, 70)
    b = rng.uniform(-1, 1, 90)
    dy = O.csr_spmv(rp, ci, va, x) - b
    g = O.dvals(rp, ci, dy, x)
    h = 1e-4
    for j in rng.choice(va.size, 20, replace=False):
        vp, vm = va.copy(), va.copy()
        vp[j] += h
        vm[j] -= h
        fp = 0.5 * np.sum((O.csr_spmv(rp, ci, vp, x) - b) ** 2)
        fm = 0.5 * np.sum((O.csr_spmv(rp, ci, vm, x) - b) ** 2)
        assert abs((fp - fm) / (2 * h) - g[j]) <= 1e-7 * max(1.0, abs(g[j]))


@pytest.mark.parametrize("idx", range(2))
def test_oracle_rowsort_golden(golden, idx):
    case = golden["rowsort_cases"][idx]
    rp = np.array(case["rowptr"], np.int32)
    cols = np.array(case["colind"], np.int32)
    L = O.scan1(rp, case["W"])
    out, src = O.reorder_columns_rowsort(L, rp, cols)
    assert out.tolist() == case["expect"]["cols_out"]
    assert src.tolist() == case["expect"]["reorder_A_rows"]
    vals = np.arange(1.0, cols.size + 1.0)
    pv = O.reorder_values_rowsort(L, vals, src)
    assert np.array_equal(pv[L.reorder_rows], vals[src])
    assert np.count_nonzero(pv) == cols.size


@pytest.mark.parametrize("seed", range(3))
def test_oracle_rowsort_properties(seed):
    """Each row of the output is its own entries (a permutation that stays in the
    row), their permuted columns ascending, ties in CSR order."""
    rng = np.random.default_rng(900 + seed)
    n = 700
    rp, cols, vals = random_csr(rng, n, n, seed)
    cols[rp[1]:rp[1] + 3] = cols[rp[1]] if rp[2] - rp[1] >= 3 else cols[rp[1]:rp[1] + 3]
    L = O.scan1(rp, 64)
    out, src = O.reorder_columns_rowsort(L, rp, cols)
    assert np.array_equal(np.sort(src), np.arange(cols.size))
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.array_equal(rows[src], rows)
    assert np.array_equal(out, L.row_map_inv[cols[src]])
    for r in range(n):
        k = out[rp[r]:rp[r + 1]]
        s = src[rp[r]:rp[r + 1]]
        assert np.all(np.diff(k) >= 0)
        ties = np.diff(k) == 0
        assert np.all(np.diff(s)[ties] > 0)
