"""MatrixMarket reading/writing and layout files (CPU, through the C ABI).

MatrixMarket: the hand-worked fixtures in tests/golden/mm (each header
variant, unsorted and duplicate entries, blank lines, CRLF), a write -> read
round trip of FE matrices that must return the same bits, and malformed files
as errors.  CUSP's reader is not in the reference tree; these fixtures pin the
restated behaviour, not CUSP's own code ("parity unpinned" against CUSP).

Layout files: save -> load returns every array and scalar unchanged (WPK1,
WPK2, with and without the row sort); damaged or truncated files are refused."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, random_csr
from oracle import pyoracle as O
from warpkernel_amd import WKError, fe, structure, structure2
from warpkernel_amd.io import load_structure, read_matrix_market, save_structure, write_matrix_market

MM = os.path.join(ROOT, "tests", "golden", "mm")


def _expected():
    with open(os.path.join(MM, "expected.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["general.mtx", "symmetric.mtx", "skew.mtx", "pattern.mtx",
                                  "integer_crlf.mtx"])
def test_matrix_market_fixtures(name):
    want = _expected()[name]
    m = read_matrix_market(os.path.join(MM, name))
    assert (m.nrows, m.ncols) == (want["nrows"], want["ncols"])
    assert m.rowptr.tolist() == want["rowptr"]
    assert m.colind.tolist() == want["colind"]
    assert m.vals.tolist() == want["vals"]


@pytest.mark.parametrize("name", ["q1_laplace_2d_128", "hex_laplace_3d_16"])
def test_matrix_market_round_trip_is_bit_exact(tmp_path, name):
    m = fe.CONFIGS[name]() if name in fe.CONFIGS else fe.hex_laplace_3d(16)
    p = tmp_path / "m.mtx"
    write_matrix_market(p, m)
    r = read_matrix_market(p)
    assert (r.nrows, r.ncols) == (m.nrows, m.ncols)
    assert np.array_equal(r.rowptr, m.rowptr)
    assert np.array_equal(r.colind, m.colind)
    assert np.array_equal(r.vals.view(np.uint64), m.vals.view(np.uint64))


def test_matrix_market_round_trip_extreme_values(tmp_path):
    vals = np.array([5e-324, -1.7976931348623157e308, 0.1, -0.0, 1 / 3, 2.0 ** -1022, 123456789.125])
    n = vals.size
    m = fe.CSR("x", n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32)[::-1].copy(), vals)
    write_matrix_market(tmp_path / "x.mtx", m)
    r = read_matrix_market(tmp_path / "x.mtx")
    assert np.array_equal(r.vals.view(np.uint64), vals.view(np.uint64))
    assert np.array_equal(r.colind, m.colind)


@pytest.mark.parametrize("text", [
    "3 3 1\n1 1 1\n",                                                     # no banner
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",        # dense format
    "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",  # complex
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",      # truncated
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",      # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n",      # column 0
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n",        # missing value
])
def test_matrix_market_malformed(tmp_path, text):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(WKError):
        read_matrix_market(p)


def _same_layout(a, b):
    for f in ("nrows", "ncols", "nznrows", "nz", "allocate_nz", "nwarps"):
        assert getattr(a, f) == getattr(b, f), f
    for f in ("row_map", "row_map_inv", "reorder_rows", "warp_offsets", "max_nz_rows"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("W", [64, 32])
def test_layout_file_round_trip(tmp_path, variant, W):
    rng = np.random.default_rng(77)
    rp, cols, vals = random_csr(rng, 4000, 4000, 1)
    nnz = int(rp[-1])
    s = (structure(W).scan(nnz, 4000, vals, rp, cols) if variant == 1 else
         structure2(W, 32).scan(nnz, 4000, vals, rp, cols, 7))
    s.reorder_columns_rowsort(cols)
    p = tmp_path / "layout.wk"
    s.save(p, "wpk_driver")
    t, name = load_structure(p)
    assert name == "wpk_driver"
    assert type(t) is type(s)
    _same_layout(s, t)
    assert np.array_equal(t.reorder_A_rows, s.reorder_A_rows)
    assert np.array_equal(t._rowptr, rp)
    if variant == 2:
        assert (t.min_nz, t.max_nz, t.nmax_per_thread) == (s.min_nz, s.max_nz, 7)
        assert np.array_equal(t.threads_per_row, s.threads_per_row)
        assert np.array_equal(t.row_offset_warp, s.row_offset_warp)
    # the loaded layout processes like the scanned one, and like the oracle's
    pv, pc = t.process_host(vals, cols)
    L = O.scan1(rp, W) if variant == 1 else O.scan2(rp, 7, W, 32)
    pv2, pc2 = O.process(L, vals, cols)
    assert np.array_equal(pv, pv2) and np.array_equal(pc, pc2)


def test_layout_file_damage_is_refused(tmp_path):
    m = fe.CONFIGS["q1_laplace_2d_128"]()
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    p = tmp_path / "a.wk"
    save_structure(s, p)
    raw = bytearray(p.read_bytes())
    for at in (0, 40, len(raw) // 2, len(raw) - 3):
        bad = bytearray(raw)
        bad[at] ^= 0x5A
        q = tmp_path / f"bad{at}.wk"
        q.write_bytes(bytes(bad))
        with pytest.raises(WKError):
            load_structure(q)
    q = tmp_path / "short.wk"
    q.write_bytes(bytes(raw[: len(raw) // 3]))
    with pytest.raises(WKError):
        load_structure(q)
    with pytest.raises(WKError):
        load_structure(tmp_path / "missing.wk")
