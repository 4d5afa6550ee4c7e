"""run_assembly (row sums of the padded values, warpkernel.hpp:78-100, 898-929)
on the GPU through the C ABI.

Checks, all bit for bit:
  - WPK1: the sequential sum of each row's entries in CSR order plus its
    trailing pads, computed densely in numpy in the engine's value type
    (float64, or float32 for an FP32 engine) -- the kernel's definition, not
    the oracle's code;
  - WPK1 and WPK2: the oracle's lane emulation of the SpMV at x = 1
    (fma(v, 1, s) is the rounded s + v, so the two sums agree exactly);
  - rows without entries are 0."""
import sys

import numpy as np
import pytest

from conftest import random_csr
from oracle import pyoracle as O
from warpkernel_amd import engine, fe, structure, structure2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sequential_rowsums(rp, vals, dtype):
    """s = -0; s = s + v for every entry of the row in CSR order (pads add +0,
    which changes nothing unless the sum is -0: the generators never give that)."""
    lens = np.diff(rp)
    n = lens.size
    width = int(lens.max()) if n else 0
    dense = np.zeros((n, width), dtype)
    rows = np.repeat(np.arange(n), lens)
    pos = np.arange(int(rp[-1])) - np.repeat(rp[:-1], lens)
    dense[rows, pos] = vals.astype(dtype)
    s = np.full(n, -0.0, dtype)
    for i in range(width):
        s = s + dense[:, i]
    return s.astype(np.float64)


def assembly(eng, n):
    y = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch_assembly(y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


CASES = [("q1_laplace_2d_128", (4,)), ("hex_elastic_3d_32", (27,)), ("tet_poisson_3d_96", (5,))]


@pytest.mark.parametrize("name,thresholds", CASES)
@pytest.mark.parametrize("W", [64, 32])
def test_fe_rowsums(name, thresholds, W):
    say(f"[gpu] assembly {name} W={W}")
    m = fe.CONFIGS[name]()
    ones = np.ones(m.ncols)
    want = sequential_rowsums(m.rowptr, m.vals, np.float64)
    assert np.array_equal(want, O.csr_spmv(m.rowptr, m.colind, m.vals, ones))
    s = structure(W).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind)
    assert np.array_equal(assembly(eng, m.nrows), want)
    eng.close()
    for th in thresholds:
        L = O.scan2(m.rowptr, th, W, 32)
        pv, pc = O.process(L, m.vals, m.colind)
        s2 = structure2(W, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, th)
        e2 = engine(s2, m.vals, m.colind)
        assert np.array_equal(assembly(e2, m.nrows), O.wpk_spmv(L, pv, pc, ones))
        e2.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("W", [64, 32])
def test_random_rowsums_with_empty_rows(mode, W):
    say(f"[gpu] assembly random mode={mode} W={W}")
    rng = np.random.default_rng(40 + mode)
    rp, cols, vals = random_csr(rng, 2500, 2500, mode)
    ones = np.ones(2500)
    s = structure(W).scan(int(rp[-1]), 2500, vals, rp, cols)
    eng = engine(s, vals, cols)
    y = assembly(eng, 2500)
    assert np.array_equal(y, sequential_rowsums(rp, vals, np.float64))
    assert np.all(y[np.diff(rp) == 0] == 0)
    eng.close()
    for th in (1, 9):
        L = O.scan2(rp, th, W, 32)
        pv, pc = O.process(L, vals, cols)
        s2 = structure2(W, 32).scan(int(rp[-1]), 2500, vals, rp, cols, th)
        e2 = engine(s2, vals, cols)
        assert np.array_equal(assembly(e2, 2500), O.wpk_spmv(L, pv, pc, ones))
        e2.close()


def test_f32_engine_sums_in_float():
    say("[gpu] assembly f32")
    m = fe.CONFIGS["hex_laplace_3d_48"]()
    s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, dtype="f32")
    assert np.array_equal(assembly(eng, m.nrows), sequential_rowsums(m.rowptr, m.vals, np.float32))
    eng.close()


def test_timed_form_and_device_scanned_engine():
    say("[gpu] assembly timed + device-scanned engine")
    m = fe.CONFIGS["q1_laplace_2d_128"]()
    want = sequential_rowsums(m.rowptr, m.vals, np.float64)
    d = [torch.from_numpy(a).cuda() for a in (m.rowptr, m.colind, m.vals)]
    eng = engine.from_device(*d)
    y = torch.full((m.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    sec = eng.run_assembly(0, 256, y)
    assert sec > 0
    assert np.array_equal(y.cpu().numpy(), want)
    eng.close()
