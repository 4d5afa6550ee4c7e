"""Engines built from files (GPU, through the C ABI): a MatrixMarket matrix
scanned, its layout saved and loaded again (serialize.cu:130-160), then every
engine form -- run, run_x, adjoint, dvals, row sort -- from the loaded layout
against the same engine from the scanned one, bit for bit, and against the
oracle."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import engine, fe, structure2
from warpkernel_amd.io import load_structure, read_matrix_market, write_matrix_market

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_engines_from_a_loaded_layout(tmp_path):
    print("[gpu] io: mtx -> scan -> save -> load -> engines", flush=True)
    sys.stdout.flush()
    m0 = fe.CONFIGS["tet_poisson_3d_96"]()
    write_matrix_market(tmp_path / "tet.mtx", m0)
    m = read_matrix_market(tmp_path / "tet.mtx")
    assert np.array_equal(m.vals, m0.vals) and np.array_equal(m.colind, m0.colind)
    s = structure2(64, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 5)
    s.save(tmp_path / "tet.k2", "io-test")
    t, name = load_structure(tmp_path / "tet.k2")
    assert name == "io-test"
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, m.ncols)
    dy = rng.uniform(-1, 1, m.nrows)
    outs = []
    for lay in (s, t):
        e = engine(lay, m.vals, m.colind, kr=True, adjoint=True)
        y, yx = (torch.empty(m.nrows, dtype=torch.float64, device="cuda") for _ in range(2))
        dxv = torch.empty(m.ncols, dtype=torch.float64, device="cuda")
        dv = torch.empty(m.nnz, dtype=torch.float64, device="cuda")
        e.launch(dev(x), y)
        e.launch_x(dev(lay.reorder_x(x)), yx)
        e.run_adjoint(dev(dy), dxv)
        e.dvals(dev(dy), dev(x), dv)
        outs.append([host(a) for a in (y, yx, dxv, dv)])
        e.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    L = O.scan2(m.rowptr, 5, 64, 32)
    pv, pc = O.process(L, m.vals, m.colind)
    assert np.array_equal(outs[1][0], O.wpk_spmv(L, pv, pc, x))
    assert np.array_equal(outs[1][3], O.dvals(m.rowptr, m.colind, dy, x))
