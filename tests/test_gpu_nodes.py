"""Node windows (stage_lds = 7) on the GPU through the C ABI.

Where every triple of a lane's steps reads columns 3n, 3n+1, 3n+2 (3D
elasticity: three unknowns per node, each row's entries in whole nodes) or
one window position three times (pads), the engine also holds each group's
window as node entries of three x values, and per lane one 10-bit entry index
per three steps.  Structure: every slot, decoded through the node window,
names its padded column (pads read column 0).  Products: the node kernel sums
the same slots in the same order as the packed-window kernel (stage_lds 5),
so y is bit-identical to it in every precision and, in FP64, to the oracle.
run_x reads entries of three Kr columns each (a node's columns are not
adjacent in that numbering).  Layouts without node windows and k-tiles above
1 run as stage 5."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import engine, fe, structure, structure2
from warpkernel_amd import _capi as capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

PK = dict(stage_lds=5, nontemporal=1, rows_per_lane=1, k_tile=1, unroll=16)
ND = dict(PK, stage_lds=7)

_CACHE = {}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def nan_like(n, k=1):
    return torch.full((n,) if k == 1 else (n, k), float("nan"), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def matrix(name):
    if name not in _CACHE:
        _CACHE.clear()
        _CACHE[name] = fe.CONFIGS[name]() if name in fe.CONFIGS else fe.hex_elastic_3d(int(name))
    return _CACHE[name]


class BlockCSR:
    """A matrix of dense 3x3 blocks over a random node graph: node i couples to
    itself and to up to max_deg - 1 other random nodes (none for the nodes
    listed empty, whose three rows are empty), each coupling a dense block, so
    every row is whole nodes in ascending column order."""

    def __init__(self, nnodes, max_deg, seed, empty=0.0):
        rng = np.random.default_rng(seed)
        rowptr, cols = [0], []
        for i in range(nnodes):
            if rng.random() < empty:
                nbrs = np.zeros(0, np.int64)
            else:
                nbrs = np.unique(np.concatenate([[i], rng.integers(0, nnodes, rng.integers(0, max_deg))]))
            rc = (3 * nbrs[:, None] + np.arange(3)[None, :]).ravel()
            for _ in range(3):
                cols.append(rc)
                rowptr.append(rowptr[-1] + rc.size)
        self.nrows = self.ncols = 3 * nnodes
        self.rowptr = np.array(rowptr, np.int32)
        self.colind = np.concatenate(cols).astype(np.int32)
        self.nnz = int(self.rowptr[-1])
        self.vals = rng.uniform(-1.0, 1.0, self.nnz)


def layout(m, variant, th=27):
    """The engine's scan and the oracle's layout of the same matrix."""
    if variant == 1:
        return structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind), O.scan1(m.rowptr, 64)
    return (structure2(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, th),
            O.scan2(m.rowptr, th, 64, 32))


def decode_node_columns(eng, L):
    """The column each slot reads through the node window, from the engine's
    arrays: entry index from the lane's word, then the entry's three columns."""
    off = eng.download(capi.WK_ENGINE_WIN_OFF).astype(np.int64)
    nc = eng.download(capi.WK_ENGINE_ND_COLS).astype(np.int64)
    cnt = eng.download(capi.WK_ENGINE_ND_CNT).astype(np.int64)
    noff = eng.download(capi.WK_ENGINE_ND_OFF).astype(np.int64)
    pos = eng.download(capi.WK_ENGINE_ND_POS)
    G = L.nwarps
    lens = L.max_nz_rows.astype(np.int64)
    assert cnt.size == G and noff.size == G + 1 and nc.size == off[-1]
    assert np.array_equal(np.diff(noff), 3 * ((lens + 26) // 27)) and pos.size == noff[-1] * 64
    assert np.all(cnt <= np.diff(off)) and np.all(cnt <= 1024)
    # per slot (group, step, lane), in slot order
    group = np.repeat(np.arange(G), lens * 64)
    within = np.arange(group.size, dtype=np.int64) - L.warp_offsets[group]
    step, lane = within // 64, within % 64
    trip = step // 3
    word = pos[(noff[group] + trip // 3) * 64 + lane].astype(np.int64)
    ent = (word >> (10 * (trip % 3))) & 1023
    assert np.all(ent < cnt[group])
    e = nc[off[group] + ent]
    cols = np.where(e >= 0, e + step % 3, -e - 1)
    kr = eng.download(capi.WK_ENGINE_ND_COLS_KR)
    cols_kr = None
    if kr.size:
        assert kr.size == 3 * off[-1]
        cols_kr = kr[3 * (off[group] + ent) + step % 3]
    return cols, cols_kr, cnt


@pytest.mark.parametrize("variant", [1, 2])
def test_node_windows_name_each_slots_column(variant):
    say(f"[gpu] node windows decode, variant {variant}")
    m = matrix("hex_elastic_3d_32")
    s, L = layout(m, variant)
    eng = engine(s, m.vals, m.colind, kr=True)
    cols, cols_kr, cnt = decode_node_columns(eng, L)
    _, pc = O.process(L, np.zeros(m.nnz), m.colind)  # pads read column 0
    assert np.array_equal(cols, pc)
    # Kr numbering: pads read column 0 of that numbering
    assert np.array_equal(cols_kr, O.reorder_columns_coalesced(L, m.colind))
    # one entry per node a group touches (a third of its window columns), and
    # one for pads where the group has them
    wsz = np.diff(eng.download(capi.WK_ENGINE_WIN_OFF).astype(np.int64))
    assert np.all(cnt * 3 >= wsz) and np.all(cnt * 3 <= wsz + 3)
    eng.close()


@pytest.mark.parametrize("name", ["hex_laplace_3d_48", "tet_poisson_3d_96", "q1_laplace_2d_128"])
def test_scalar_layouts_have_no_node_windows(name):
    """Scalar problems: rows are not whole nodes, so stage 7 runs as stage 5."""
    say(f"[gpu] no node windows on {name}")
    m = matrix(name)
    s, _ = layout(m, 1)
    eng = engine(s, m.vals, m.colind)
    assert eng.download(capi.WK_ENGINE_ND_CNT).size == 0
    x = dev(fe.uniform_x(m.ncols, seed=7101))
    y5, y7 = nan_like(m.nrows), nan_like(m.nrows)
    eng.launch(x, y5, cfg=PK)
    eng.launch(x, y7, cfg=ND)
    assert torch.equal(y5, y7)
    eng.close()


@pytest.mark.parametrize("name", ["hex_elastic_3d_32", "hex_elastic_3d_64"])
@pytest.mark.parametrize("variant", [1, 2])
def test_node_kernel_bitexact(name, variant):
    say(f"[gpu] node kernel {name} variant {variant}")
    m = matrix(name)
    s, L = layout(m, variant)
    eng = engine(s, m.vals, m.colind, kr=True)
    assert eng.download(capi.WK_ENGINE_ND_CNT).size == L.nwarps
    x = fe.uniform_x(m.ncols, seed=7102)
    dxr = dev(s.reorder_x(x))
    if variant == 1:
        ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    else:
        pv, pc = O.process(L, m.vals, m.colind)
        ref = O.wpk_spmv(L, pv, pc, x)
    b = fe.uniform_x(m.nrows, seed=7103)
    # unroll 16: 27-step batches, 8: 18-step
    for wpb, remap, u in ((1, 1, 16), (2, 0, 8), (4, 1, 16), (1, 1, 8)):
        say(f"[gpu]   waves_per_block {wpb} xcd_remap {remap} unroll {u}")
        cfg = dict(ND, waves_per_block=wpb, xcd_remap=remap, unroll=u)
        y = nan_like(m.nrows)
        eng.launch(dev(x), y, cfg=cfg)
        assert np.array_equal(host(y), ref)
        # the residual through the same kernel
        r, r5 = nan_like(m.nrows), nan_like(m.nrows)
        eng.residual(dev(b), dev(x), r, cfg=cfg)
        eng.residual(dev(b), dev(x), r5, cfg=dict(cfg, stage_lds=5))
        assert torch.equal(r, r5) and np.array_equal(host(r), b - ref)
        # run_x: entries of three Kr columns
        yx, yx5 = nan_like(m.nrows), nan_like(m.nrows)
        eng.launch_x(dxr, yx, cfg=cfg)
        eng.launch_x(dxr, yx5, cfg=dict(cfg, stage_lds=5))
        assert torch.equal(yx, yx5)
        if variant == 1:
            assert np.array_equal(host(yx), ref[s.row_map])
    eng.close()


@pytest.mark.parametrize("k", [2, 3])
def test_node_kernel_strided_columns(k):
    """k right-hand sides at stage 7: one launch per column of x (row stride
    k), each equal to the single-column product."""
    say(f"[gpu] node kernel k={k}")
    m = matrix("hex_elastic_3d_32")
    s, _ = layout(m, 1)
    eng = engine(s, m.vals, m.colind)
    X = fe.uniform_x(m.ncols, k, seed=7104)
    Y = nan_like(m.nrows, k)
    eng.launch(dev(X), Y, cfg=dict(ND, unroll=8 if k == 3 else 16))
    Yh = host(Y)
    for j in range(k):
        assert np.array_equal(Yh[:, j], O.csr_spmv(m.rowptr, m.colind, m.vals, X[:, j]))
    eng.close()


@pytest.mark.parametrize("acc", [0, 1])
def test_node_kernel_fp32_matches_packed(acc):
    say(f"[gpu] node kernel fp32 accumulate {acc}")
    m = matrix("hex_elastic_3d_32")
    s, _ = layout(m, 1)
    eng = engine(s, m.vals, m.colind, dtype="f32")
    assert eng.download(capi.WK_ENGINE_ND_CNT).size > 0
    x = dev(fe.uniform_x(m.ncols, seed=7105).astype(np.float32))
    y5, y7 = nan_like(m.nrows), nan_like(m.nrows)
    for wpb, u in ((1, 16), (4, 8), (2, 8)):
        eng.launch(x, y5, cfg=dict(PK, waves_per_block=wpb, accumulate=acc, unroll=u))
        eng.launch(x, y7, cfg=dict(ND, waves_per_block=wpb, accumulate=acc, unroll=u))
        assert torch.equal(y5, y7)
    eng.close()


def test_node_windows_from_the_device_scan_and_run_x():
    """Engines scanned on the device get the same node windows, in both
    numberings."""
    say("[gpu] node windows, device scan and run_x")
    m = matrix("hex_elastic_3d_32")
    s, L = layout(m, 1)
    he = engine(s, m.vals, m.colind, kr=True)
    de = engine.from_device(dev(m.rowptr), dev(m.colind), dev(m.vals), ncols=m.ncols, kr=True)
    off = he.download(capi.WK_ENGINE_WIN_OFF).astype(np.int64)
    cnt = he.download(capi.WK_ENGINE_ND_CNT).astype(np.int64)
    keep = np.zeros(off[-1], bool)  # only each group's first ND_CNT entries are written
    for g in range(L.nwarps):
        keep[off[g]:off[g] + cnt[g]] = True
    for w in (capi.WK_ENGINE_ND_COLS, capi.WK_ENGINE_ND_COLS_KR, capi.WK_ENGINE_ND_CNT,
              capi.WK_ENGINE_ND_OFF, capi.WK_ENGINE_ND_POS):
        a, b = he.download(w), de.download(w)
        assert a.size == b.size > 0, w
        if w == capi.WK_ENGINE_ND_COLS:
            a, b = a[keep], b[keep]
        if w == capi.WK_ENGINE_ND_COLS_KR:
            a, b = a.reshape(-1, 3)[keep], b.reshape(-1, 3)[keep]
        assert np.array_equal(a, b), w
    x = fe.uniform_x(m.ncols, seed=7106)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    y = nan_like(m.nrows)
    de.launch(dev(x), y, cfg=ND)
    assert np.array_equal(host(y), ref)
    yx = nan_like(m.nrows)
    he.launch_x(dev(s.reorder_x(x)), yx, cfg=ND)
    assert np.array_equal(host(yx), ref[s.row_map])
    he.close()
    de.close()


def test_adjoint_through_node_windows():
    """The transposed layout of an elasticity matrix has node windows too: dx =
    A^T dy at stage 7 equals stage 5 and the oracle's transposed fma chain."""
    say("[gpu] adjoint through node windows")
    m = matrix("hex_elastic_3d_32")
    s, _ = layout(m, 1)
    eng = engine(s, m.vals, m.colind, adjoint=True)
    dy = fe.uniform_x(m.nrows, seed=7107)
    cp, ri, perm = O.transpose(m.nrows, m.ncols, m.rowptr, m.colind)
    ref = O.csr_spmv(cp, ri, m.vals[perm], dy)
    assert eng.get_launch(op=capi.WK_OP_ADJOINT).as_dict()["stage_lds"] == 7
    dx7, dx5 = nan_like(m.ncols), nan_like(m.ncols)
    eng.run_adjoint(dev(dy), dx7, cfg=ND)
    eng.run_adjoint(dev(dy), dx5, cfg=PK)
    assert torch.equal(dx7, dx5)
    assert np.array_equal(host(dx7), ref)
    eng.close()


@pytest.mark.parametrize("max_deg,empty", [(9, 0.0), (9, 0.02), (12, 0.0), (14, 0.02)])
def test_node_windows_on_random_block_graphs(max_deg, empty):
    """Rows of mixed length in whole nodes (pads in most groups, some empty
    rows): up to 27 entries per row the kernel takes each group's values
    straight into registers, beyond that in 18- and 27-step batches with a
    partial last batch.  run, run_x and the residual, FP64 against the oracle;
    FP32 against stage 5."""
    say(f"[gpu] node windows on a random block graph, max_deg {max_deg}, empty {empty}")
    m = BlockCSR(20000, max_deg, seed=7110 + max_deg, empty=empty)
    s, L = layout(m, 1)
    eng = engine(s, m.vals, m.colind, kr=True)
    cols, cols_kr, _ = decode_node_columns(eng, L)
    _, pc = O.process(L, np.zeros(m.nnz), m.colind)
    assert np.array_equal(cols, pc)
    assert np.array_equal(cols_kr, O.reorder_columns_coalesced(L, m.colind))
    x = fe.uniform_x(m.ncols, seed=7111)
    ref = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    b = fe.uniform_x(m.nrows, seed=7112)
    for u, wpb in ((16, 1), (8, 2), (16, 4)):
        cfg = dict(ND, unroll=u, waves_per_block=wpb)
        y, yx, r = nan_like(m.nrows), nan_like(m.nrows), nan_like(m.nrows)
        eng.launch(dev(x), y, cfg=cfg)
        eng.launch_x(dev(s.reorder_x(x)), yx, cfg=cfg)
        eng.residual(dev(b), dev(x), r, cfg=cfg)
        assert np.array_equal(host(y), ref)
        assert np.array_equal(host(yx), ref[s.row_map])
        assert np.array_equal(host(r), b - ref)
    eng.close()
    e32 = engine(s, m.vals, m.colind, dtype="f32")
    x32 = dev(x.astype(np.float32))
    for acc in (0, 1):
        y5, y7 = nan_like(m.nrows), nan_like(m.nrows)
        e32.launch(x32, y5, cfg=dict(PK, accumulate=acc))
        e32.launch(x32, y7, cfg=dict(ND, accumulate=acc))
        assert torch.equal(y5, y7)
    e32.close()


def test_split_layout_with_long_groups():
    """WPK2 at threshold 81 on hex elasticity keeps every row whole (81-step
    groups through the split kernel's 27- and 18-step batches); at threshold
    27 every group has at most 27 steps (values straight into registers)."""
    say("[gpu] node windows, WPK2 thresholds 81 and 27")
    m = matrix("hex_elastic_3d_32")
    x = fe.uniform_x(m.ncols, seed=7113)
    for th in (81, 27):
        s, L = layout(m, 2, th)
        eng = engine(s, m.vals, m.colind)
        assert eng.download(capi.WK_ENGINE_ND_CNT).size == L.nwarps
        pv, pc = O.process(L, m.vals, m.colind)
        ref = O.wpk_spmv(L, pv, pc, x)
        for u in (16, 8):
            y = nan_like(m.nrows)
            eng.launch(dev(x), y, cfg=dict(ND, unroll=u))
            assert np.array_equal(host(y), ref)
        eng.close()
