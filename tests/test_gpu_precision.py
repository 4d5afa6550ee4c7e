"""Precision and multi-RHS parity on the batching/precision matrix: 3D trilinear
hex linear elasticity on a 32^3 element grid (107,811 rows, 8,214,057 nnz), k
in {1, 4, 8, 16, 32} right-hand sides, WPK1 and WPK2 (threshold 27), run and
run_x, each with the engine's default launch and with the launch the autotuner
picks.  Bars, all against the oracle:

  - FP64 values: bit-exact (WPK1: the CSR fma chain in row order; WPK2: the
    lane emulation with warpKernel2's pairwise tree, warpkernel.hpp:154-162).
  - FP32 values, WK_ACC_F64: a float*float product is exact in FP64, so the
    FP64 fma chain over the quantised inputs is bit-exact against the oracle
    run on those inputs (converted to double).
  - FP32 values, WK_ACC_TWOSUM (TwoProd + TwoSum, FP64 result): normwise
    relative error <= 1e-10 against the FP64 CSR oracle on the same quantised
    inputs.  The quantisation error (against the FP64 oracle on the unquantised
    inputs) is printed beside it; it is a property of FP32 storage, not of the
    accumulation, and is not gated.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from warpkernel_amd import engine, fe, structure, structure2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KS = (1, 4, 8, 16, 32)
TH = 27
TWOSUM_BOUND = 1e-10
WIN = dict(nontemporal=1, rows_per_lane=1)

# explicit FP32 shapes beyond the defaults: window k-tiles shared by two and
# four waves, register-fed and staged, 16-step batches, and the direct kernel
# at the widest k-tiles
FP32_SHAPES = {
    1: [dict(WIN, k_tile=1, unroll=16, waves_per_block=1, stage_lds=5),
        dict(WIN, k_tile=1, unroll=8, waves_per_block=4, stage_lds=3),
        dict(k_tile=1, stage_lds=0, unroll=8, waves_per_block=2)],
    4: [dict(WIN, k_tile=4, unroll=8, waves_per_block=4, stage_lds=4),
        dict(WIN, k_tile=4, unroll=16, waves_per_block=1, stage_lds=4)],
    8: [dict(WIN, k_tile=8, unroll=16, waves_per_block=2, stage_lds=4),
        dict(WIN, k_tile=8, unroll=8, waves_per_block=8, stage_lds=4),
        dict(WIN, k_tile=8, unroll=8, waves_per_block=2, stage_lds=3)],
    16: [dict(WIN, k_tile=8, unroll=16, waves_per_block=2, stage_lds=4),
         dict(WIN, k_tile=8, unroll=8, waves_per_block=4, stage_lds=4),
         dict(k_tile=16, stage_lds=0, unroll=8, waves_per_block=2)],
    32: [dict(WIN, k_tile=8, unroll=16, waves_per_block=2, stage_lds=4),
         dict(WIN, k_tile=8, unroll=8, waves_per_block=8, stage_lds=4),
         dict(k_tile=32, stage_lds=0, unroll=8, waves_per_block=2)],
}
# values held in registers across 32-byte k-tiles (8 FP32 / 4 FP64 right-hand
# sides per pass), one to four waves per block
for _k in (8, 16, 32):
    FP32_SHAPES[_k] += [dict(WIN, k_tile=8, unroll=8, waves_per_block=w, stage_lds=6) for w in (1, 4)]
FP64_SHAPES = {k: [dict(WIN, k_tile=min(k, 8), unroll=8, waves_per_block=w, stage_lds=6, xcd_remap=r)
                   for w, r in ((1, 1), (2, 0), (4, 1))] for k in KS}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def normwise(y, ref):
    return float(np.linalg.norm(y - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="module")
def hex32():
    m = fe.CONFIGS["hex_elastic_3d_32"]()
    vq = m.vals.astype(np.float32).astype(np.float64)  # the FP32 engines' values, exactly
    out = {"m": m, "vq": vq}
    for variant in (1, 2):
        if variant == 1:
            s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
            L = O.scan1(m.rowptr, 64)
        else:
            s = structure2(64, 32).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, TH)
            L = O.scan2(m.rowptr, TH, 64, 32)
        assert np.array_equal(np.asarray(s.reorder_rows), L.reorder_rows)
        e64 = engine(s, m.vals, m.colind, kr=True)
        e32 = engine(s, vq, m.colind, dtype="f32", kr=True)
        # the layout defaults, read before any tuning replaces them
        out[("default", variant)] = {(e, op, k): eng.get_launch(op, k).as_dict()
                                     for e, eng in (("f64", e64), ("f32", e32))
                                     for op in (0, 1) for k in KS}
        out[variant] = (s, L, e64, e32)
    yield out
    for variant in (1, 2):
        out[variant][2].close()
        out[variant][3].close()


def reference(m, L, variant, vals, x):
    """The oracle's y for this variant: CSR row order (WPK1) or the WPK2 lane
    emulation on the padded arrays."""
    if variant == 1:
        return O.csr_spmv(m.rowptr, m.colind, vals, x)
    pv, pc = O.process(L, vals, m.colind)
    return O.wpk_spmv(L, pv, pc, x)


def run_both(eng, xd, xrd, nrows, k, cfg):
    """(run, run_x) results of one launch configuration."""
    y = torch.full((nrows, k), float("nan"), dtype=torch.float64, device="cuda")
    eng.launch(xd, y, cfg=cfg)
    yr = host(y).copy()
    y.fill_(float("nan"))
    eng.launch_x(xrd, y, cfg=cfg)
    return yr, host(y)


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("variant", [1, 2])
def test_fp64_multi_rhs_bitexact(hex32, variant, k):
    """FP64, default and tuned launches, run and run_x: bit-exact at every k."""
    say(f"[gpu] precision fp64 hex_elastic_3d_32 variant={variant} k={k}")
    m = hex32["m"]
    s, L, e64, _ = hex32[variant]
    x = fe.uniform_x(m.ncols, k, seed=8100 + k).reshape(m.ncols, k)
    ref = reference(m, L, variant, m.vals, x)
    xr = np.empty_like(x)
    xr[L.row_map_inv] = x
    xd, xrd = dev(x), dev(xr)
    y = torch.empty((m.nrows, k), dtype=torch.float64, device="cuda")
    dflt = hex32[("default", variant)]
    tuned, _ = e64.tune(xd, y, reps=3)
    tuned_x, _ = e64.tune(xrd, y, op=1, reps=3)
    shapes = [("default", dflt[("f64", 0, k)], dflt[("f64", 1, k)]), ("tuned", tuned, tuned_x)]
    shapes += [(f"shape{i}", c, c) for i, c in enumerate(FP64_SHAPES[k])]
    for name, cfg, cfg_x in shapes:
        yr, _ = run_both(e64, xd, xrd, m.nrows, k, cfg)
        _, yx = run_both(e64, xd, xrd, m.nrows, k, cfg_x)
        assert np.array_equal(yr, ref), (name, cfg)
        assert np.array_equal(yx, ref[L.row_map]), (name, cfg_x, "run_x")


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("variant", [1, 2])
def test_fp32_engines_against_fp64_oracle(hex32, variant, k):
    """FP32 values and x: WK_ACC_F64 bit-exact against the oracle on the
    quantised inputs, WK_ACC_TWOSUM within 1e-10 of the FP64 CSR oracle on
    them; default, tuned and explicit launch shapes, run and run_x."""
    m, vq = hex32["m"], hex32["vq"]
    s, L, _, e32 = hex32[variant]
    x = fe.uniform_x(m.ncols, k, seed=8200 + k).reshape(m.ncols, k)
    xq = x.astype(np.float32)
    xq64 = xq.astype(np.float64)
    ref_q = reference(m, L, variant, vq, xq64)             # F64ACC bar (bit-exact)
    csr_q = O.csr_spmv(m.rowptr, m.colind, vq, xq64)       # TWOSUM bar (1e-10)
    csr_t = O.csr_spmv(m.rowptr, m.colind, m.vals, x)      # quantisation, reported
    xr = np.empty_like(xq)
    xr[L.row_map_inv] = xq
    xd, xrd = dev(xq), dev(xr)
    y = torch.empty((m.nrows, k), dtype=torch.float64, device="cuda")
    dflt = hex32[("default", variant)]
    for acc in (1, 0):
        tuned, _ = e32.tune(xd, y, reps=3, base=dict(accumulate=acc))
        tuned_x, _ = e32.tune(xrd, y, op=1, reps=3, base=dict(accumulate=acc))
        shapes = [("default", dict(dflt[("f32", 0, k)], accumulate=acc),
                   dict(dflt[("f32", 1, k)], accumulate=acc)), ("tuned", tuned, tuned_x)]
        shapes += [(f"shape{i}", dict(c, accumulate=acc), dict(c, accumulate=acc))
                   for i, c in enumerate(FP32_SHAPES[k])]
        for name, cfg, cfg_x in shapes:
            yr, _ = run_both(e32, xd, xrd, m.nrows, k, cfg)
            _, yx = run_both(e32, xd, xrd, m.nrows, k, cfg_x)
            if acc == 1:
                assert np.array_equal(yr, ref_q), (variant, k, name, cfg)
                assert np.array_equal(yx, ref_q[L.row_map]), (variant, k, name, cfg_x, "run_x")
            else:
                e_run = normwise(yr, csr_q)
                e_runx = normwise(yx, csr_q[L.row_map])
                say(f"[gpu] fp32 TwoSum variant={variant} k={k} {name}: error vs FP64 oracle on the "
                    f"quantised inputs {e_run:.3e} (run) {e_runx:.3e} (run_x); quantisation "
                    f"{normwise(yr, csr_t):.3e}")
                assert e_run <= TWOSUM_BOUND, (variant, k, name, cfg, e_run)
                assert e_runx <= TWOSUM_BOUND, (variant, k, name, cfg_x, e_runx)
