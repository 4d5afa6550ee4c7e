"""One launch over many uneven FE matrices (one wave per group, or persistent
waves over equal-cost ranges), and HIP graph replay."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import batch, engine, fe, graph, structure, structure2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


_MIX = {}


def mix():
    if not _MIX:
        _MIX["m"] = fe.batch_mix(64)
    return _MIX["m"]


def expected(m, i, x):
    if i % 3 == 2:  # WPK2 engines: lane-order oracle
        L = O.scan2(m.rowptr, 8, 64, 32)
        pv, pc = O.process(L, m.vals, m.colind)
        return O.wpk_spmv(L, pv, pc, x)
    return O.csr_spmv(m.rowptr, m.colind, m.vals, x)


def make_engines(ms, dtype="f64"):
    engs = []
    for i, m in enumerate(ms):
        if i % 3 == 2:
            s = structure2().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, 8)
        else:
            s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
        vals = m.vals if dtype == "f64" else m.vals.astype(np.float32).astype(np.float64)
        engs.append(engine(s, vals, m.colind, dtype=dtype))
    return engs


@pytest.mark.parametrize("k,stage,wpc", [(1, 0, 0), (1, 1, 0), (1, 2, 0), (1, 3, 0), (1, 5, 0),
                                         (4, 3, 0), (4, 5, 0), (8, 5, 0), (12, 5, 0), (3, 0, 0),
                                         (1, 3, 8), (1, 5, 8), (8, 5, 16), (1, 0, 64), (4, 2, 16)])
def test_batch_of_64_matrices_bitexact(k, stage, wpc):
    ms = mix()
    say(f"[gpu] batch of {len(ms)} matrices, {sum(m.nnz for m in ms)} nnz, k={k}, stage={stage}, "
        f"waves_per_cu={wpc}")
    engs = make_engines(ms)
    xs_h = [fe.uniform_x(m.ncols, k, seed=7000 + i).reshape((m.ncols,) if k == 1 else (m.ncols, k))
            for i, m in enumerate(ms)]
    xs = [dev(x) for x in xs_h]
    ys = [torch.full((m.nrows,) if k == 1 else (m.nrows, k), float("nan"), dtype=torch.float64,
                     device="cuda") for m in ms]
    b = batch(engs, waves_per_cu=wpc, stage_lds=stage)
    b.bind(xs, ys)
    b.run()
    torch.cuda.synchronize()
    for i, m in enumerate(ms):
        assert np.array_equal(ys[i].cpu().numpy(), expected(m, i, xs_h[i])), (i, m.name)


@pytest.mark.parametrize("k", [1, 8])
def test_batch_fp32_compensated(k):
    ms = mix()[:16]
    say(f"[gpu] batch FP32 k={k}")
    engs = make_engines(ms, "f32")
    xs_h = [fe.uniform_x(m.ncols, k, seed=7100 + i).astype(np.float32) for i, m in enumerate(ms)]
    shape = (lambda m: (m.nrows,)) if k == 1 else (lambda m: (m.nrows, k))
    for wpc, stage in ((4, 3), (0, 5)):
        ys = [torch.full(shape(m), float("nan"), dtype=torch.float64, device="cuda") for m in ms]
        b = batch(engs, waves_per_cu=wpc, stage_lds=stage)
        b.bind([dev(x) for x in xs_h], ys)
        b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(ms):
            vq = m.vals.astype(np.float32).astype(np.float64)
            ref = O.csr_spmv(m.rowptr, m.colind, vq, xs_h[i].astype(np.float64))
            y = ys[i].cpu().numpy()
            assert np.linalg.norm(y - ref) <= 1e-10 * np.linalg.norm(ref), (stage, i)
        b.close()


def test_graph_replays_match_eager():
    say("[gpu] graph capture and replay")
    ms = mix()[:8]
    engs = make_engines(ms)
    xs_h = [fe.uniform_x(m.ncols, seed=7200 + i) for i, m in enumerate(ms)]
    xs = [dev(x) for x in xs_h]
    ys = [torch.empty(m.nrows, dtype=torch.float64, device="cuda") for m in ms]
    b = batch(engs)
    b.bind(xs, ys)
    # a single engine with value updates, forward and adjoint in the same graph
    m = fe.hex_laplace_3d(20)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    e = engine(s, m.vals, m.colind, adjoint=True)
    v = dev(m.vals.copy())
    x = dev(fe.uniform_x(m.ncols, seed=7300))
    y = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    dx = torch.empty(m.ncols, dtype=torch.float64, device="cuda")
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    with graph.capture(st) as g:
        e.update_values(v, stream=st)
        e.launch(x, y, stream=st)
        e.run_adjoint(y, dx, stream=st)
        b.run(stream=st)
    cp, ri, perm = O.transpose(m.nrows, m.ncols, m.rowptr, m.colind)
    for rep in range(4):
        # eager work between replays: new values and new inputs
        vals_h = np.random.default_rng(rep).uniform(-1, 1, m.nnz)
        x_h = fe.uniform_x(m.ncols, seed=7400 + rep)
        v.copy_(torch.from_numpy(vals_h))
        x.copy_(torch.from_numpy(x_h))
        for i in range(len(ms)):
            xs_h[i] = fe.uniform_x(ms[i].ncols, seed=7500 + 10 * rep + i)
            xs[i].copy_(torch.from_numpy(xs_h[i]))
        torch.cuda.synchronize()
        g.launch(st)
        st.synchronize()
        yref = O.csr_spmv(m.rowptr, m.colind, vals_h, x_h)
        assert np.array_equal(y.cpu().numpy(), yref), rep
        assert np.array_equal(dx.cpu().numpy(), O.csr_spmv(cp, ri, vals_h[perm], yref)), rep
        for i, mi in enumerate(ms):
            assert np.array_equal(ys[i].cpu().numpy(), expected(mi, i, xs_h[i])), (rep, i)
    g.close()


@pytest.mark.parametrize("k", [4, 8])
def test_batch_window_fit_per_matrix(k):
    """With several right-hand sides each matrix gets a window only where its
    k-tile window fits the block's LDS: a banded matrix of about 1,250 window
    entries per group (3.1 slots per entry) is windowed at k-tile 4 (40 KB)
    and reads x directly at k-tile 8 (80 KB), in the same launch as FE
    matrices with windows; all bit-exact."""
    say(f"[gpu] batch window fit k={k}")
    n = 4096
    rp = np.arange(n + 1, dtype=np.int32) * 60
    cols = ((np.arange(n)[:, None] + 20 * np.arange(60)[None, :]) % n)
    cols = np.sort(cols, axis=1).astype(np.int32).ravel()
    rng = np.random.default_rng(11)
    ms = list(mix()[:6])
    mats = [(m.rowptr, m.colind, m.vals, m.nrows, m.ncols) for m in ms]
    mats.append((rp, cols, rng.uniform(-1, 1, cols.size), n, n))
    engs = [engine(structure().scan(int(r[-1]), nr, v, r, c), v, c) for r, c, v, nr, nc in mats]
    xs_h = [fe.uniform_x(nc, k, seed=7700 + i).reshape(nc, k) for i, (_, _, _, _, nc) in enumerate(mats)]
    ys = [torch.full((nr, k), float("nan"), dtype=torch.float64, device="cuda") for _, _, _, nr, _ in mats]
    b = batch(engs)
    b.bind([dev(x) for x in xs_h], ys)
    b.run()
    torch.cuda.synchronize()
    for i, (r, c, v, nr, nc) in enumerate(mats):
        assert np.array_equal(ys[i].cpu().numpy(), O.csr_spmv(r, c, v, xs_h[i])), i
    b.close()


@pytest.mark.parametrize("k", [1, 8])
def test_graph_replay_after_rebind_same_k(k):
    """A graph that captured wk_batch_run keeps working after re-binds with the
    same k: the batch's device arrays are allocated once and rewritten in
    place, so each replay reads the vectors of the latest bind."""
    say(f"[gpu] graph replay after re-bind k={k}")
    ms = mix()[:10]
    engs = make_engines(ms)
    shape = (lambda m: (m.ncols,)) if k == 1 else (lambda m: (m.ncols, k))
    yshape = (lambda m: (m.nrows,)) if k == 1 else (lambda m: (m.nrows, k))
    b = batch(engs)
    xs_h = [fe.uniform_x(m.ncols, k, seed=7800 + i).reshape(shape(m)) for i, m in enumerate(ms)]
    ys = [torch.full(yshape(m), float("nan"), dtype=torch.float64, device="cuda") for m in ms]
    b.bind([dev(x) for x in xs_h], ys)
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    with graph.capture(st) as g:
        b.run(stream=st)
    for rep in range(3):
        # new vectors in new allocations, bound eagerly between replays
        xs_h = [fe.uniform_x(m.ncols, k, seed=7900 + 100 * rep + i).reshape(shape(m))
                for i, m in enumerate(ms)]
        xs = [dev(x) for x in xs_h]
        ys = [torch.full(yshape(m), float("nan"), dtype=torch.float64, device="cuda") for m in ms]
        b.bind(xs, ys)
        torch.cuda.synchronize()
        g.launch(st)
        st.synchronize()
        for i, m in enumerate(ms):
            assert np.array_equal(ys[i].cpu().numpy(), expected(m, i, xs_h[i])), (rep, i)
    g.close()
    b.close()
