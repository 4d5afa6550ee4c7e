#!/usr/bin/env python3
"""WPK SpMV benchmark on one MI355X.

Headline: FP64 WPK1 SpMV (engine::run, group width 64) on the 3D trilinear hex
linear-elasticity matrix, 64^3 elements (823,875 rows, 64,701,513 nnz), one
right-hand side, x ~ U[-1,1].  One step = one SpMV launch with the matrix and x
already resident in HBM.

    GB/s = algorithmic bytes / time,  bytes = nnz*(8+4) + ncols*8 + nrows*8
    (padding, row_map and group metadata are overhead and are not counted)

Prints ONE JSON line.  Extras (skipped with --headline-only) measure the other
parts of the metric: run_x, WPK2, FP32-compensated multi-RHS against plain FP64
(hex elasticity 32^3, k = 1..32) and forward + adjoint time (hex diffusion 48^3).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# unpacked FP32 VALU lane operations per second (MI355X: 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz; the published 157.3 TFLOP/s counts FMA and packed pairs)
VALU_LANE_OPS_PER_S = 256 * 4 * 16 * 2.4e9
METRIC = "SpMV GB/s & % HBM roofline, FP64 vs FP32-comp multi-RHS; fwd+adjoint ms"


def spmv_bytes(nnz, nrows, ncols, k=1, s_val=8, s_x=8, s_y=8):
    return nnz * (s_val + 4) + k * (ncols * s_x + nrows * s_y)


def time_launches(torch, fn, steps, warmup, min_warm_s=0.02):
    """Returns (wall seconds for `steps` launches, mean launch ms from HIP events
    recorded on the launch stream)."""
    # warm-up: `warmup` launches and at least min_warm_s of device time, so
    # that short kernels are not timed while the clocks are still rising (the
    # headline passes 0: exactly --warmup steps)
    t_w = time.perf_counter()
    n = 0
    while n < warmup or time.perf_counter() - t_w < min_warm_s:
        fn()
        n += 1
        if n % 32 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, ev0.elapsed_time(ev1) / steps


def stream_read_ceiling(torch, gib=4, reps=10):
    """The device's streaming-read rate (wk_stream_read over a 4 GiB buffer,
    16-byte nontemporal loads): the practical ceiling of the kernels' HBM
    traffic, reported next to the nominal peak."""
    from warpkernel_amd._capi import check, lib
    nbytes = gib << 30
    buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    _, ms = time_launches(torch, lambda: check(lib().wk_stream_read(buf.data_ptr(), nbytes, None, st)),
                          reps, 3)
    del buf
    return round(nbytes / (ms * 1e-3) / 1e9, 1)


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return float(ln.split(":", 1)[1])
    return None


def cpu_baseline(m, x, budget_s):
    """The oracle's FP64 CSR SpMV (original row order, fma), timed on this host,
    1 thread and all cores; plus the host WPK1 scan and process."""
    from oracle import pyoracle as O
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    b = spmv_bytes(m.nnz, m.nrows, m.ncols)
    O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    t0 = time.perf_counter()
    O.csr_spmv(m.rowptr, m.colind, m.vals, x)
    one = time.perf_counter() - t0
    reps = int(max(5, min(200, 0.4 * budget_s / max(one, 1e-6))))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.csr_spmv(m.rowptr, m.colind, m.vals, x)
        ts.append(time.perf_counter() - t0)
    t1 = float(np.median(ts))
    O.csr_spmv_omp(m.rowptr, m.colind, m.vals, x, cores)
    tsn = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.csr_spmv_omp(m.rowptr, m.colind, m.vals, x, cores)
        tsn.append(time.perf_counter() - t0)
    tn = float(np.median(tsn))
    t0 = time.perf_counter()
    L = O.scan1(m.rowptr, 64)
    t_scan = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.process(L, m.vals, m.colind)
    t_proc = time.perf_counter() - t0
    return {
        "value": round(b / tn / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "port",
        "sample": (f"oracle FP64 CSR SpMV (original row order, fma) on the full {m.name} matrix, "
                   f"median of {reps} reps after a warm-up, OpenMP over rows on {cores} cores; "
                   f"1 thread: {b / t1 / 1e9:.3f} GB/s ({t1 * 1e3:.1f} ms); "
                   f"host WPK1 scan (W=64) {t_scan:.3f} s, host process {t_proc:.3f} s"),
        "ms_per_spmv": round(tn * 1e3, 3), "ms_per_spmv_1thread": round(t1 * 1e3, 3),
    }


def extra_multi_rhs(torch, dev, steps, warmup, tune=True):
    """FP32-compensated multi-RHS vs plain FP64 (hex elasticity 32^3)."""
    from oracle import pyoracle as O
    from warpkernel_amd import engine, fe, structure
    out = {}
    m = fe.hex_elastic_3d(32)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    e64 = engine(s, m.vals, m.colind)
    vq = m.vals.astype(np.float32)
    e32 = engine(s, vq.astype(np.float64), m.colind, dtype="f32")
    rows = []
    base = None
    for k in (1, 4, 8, 16, 32):
        x = fe.uniform_x(m.ncols, k, seed=2100 + k).reshape(m.ncols, k)
        xq = x.astype(np.float32)
        dx64, dx32 = dev(x), dev(xq)
        y64 = torch.empty((m.nrows, k), dtype=torch.float64, device="cuda")
        y32 = torch.empty((m.nrows, k), dtype=torch.float64, device="cuda")
        if tune:
            e64.tune(dx64, y64, reps=5)
            e32.tune(dx32, y32, reps=5)
        _, ms64 = time_launches(torch, lambda: e64.launch(dx64, y64), steps, warmup)
        _, ms32 = time_launches(torch, lambda: e32.launch(dx32, y32), steps, warmup)
        if tune:
            e32.tune(dx32, y32, reps=5, base=dict(accumulate=1))
        _, ms32d = time_launches(torch, lambda: e32.launch(dx32, y32, cfg=dict(accumulate=1)), steps, warmup)
        if tune:
            e32.tune(dx32, y32, reps=5, base=dict(accumulate=0))
        e32.launch(dx32, y32)
        torch.cuda.synchronize()
        ref_q = O.csr_spmv(m.rowptr, m.colind, vq.astype(np.float64), xq.astype(np.float64))
        ref_t = O.csr_spmv(m.rowptr, m.colind, m.vals, x)
        yc = y32.cpu().numpy()
        acc_err = float(np.linalg.norm(yc - ref_q) / np.linalg.norm(ref_q))
        q_err = float(np.linalg.norm(yc - ref_t) / np.linalg.norm(ref_t))
        r64 = m.nnz * k / (ms64 * 1e-3)
        r32 = m.nnz * k / (ms32 * 1e-3)
        if k == 1:
            base = r64
        # bounds: HBM at the nominal peak for the algorithmic bytes, and for the
        # compensated form the FP32 VALU time of its ten unpacked operations
        # per entry and right-hand side (256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz)
        hbm64 = spmv_bytes(m.nnz, m.nrows, m.ncols, k) / (HBM_PEAK_GBS * 1e9) * 1e3
        hbm32 = spmv_bytes(m.nnz, m.nrows, m.ncols, k, 4, 4, 8) / (HBM_PEAK_GBS * 1e9) * 1e3
        valu32 = m.nnz * k * 10 / VALU_LANE_OPS_PER_S * 1e3
        rows.append({
            "k": k, "fp64_ms": round(ms64, 4), "fp32c_ms": round(ms32, 4), "fp32_f64acc_ms": round(ms32d, 4),
            "fp64_hbm_bound_ms": round(hbm64, 4),
            "fp32c_hbm_bound_ms": round(hbm32, 4), "fp32c_valu_bound_ms": round(valu32, 4),
            "fp32c_frac_of_bound": round(max(hbm32, valu32) / ms32, 3),
            "fp64_GBs": round(spmv_bytes(m.nnz, m.nrows, m.ncols, k) / (ms64 * 1e-3) / 1e9, 1),
            "fp32c_GBs": round(spmv_bytes(m.nnz, m.nrows, m.ncols, k, 4, 4, 8) / (ms32 * 1e-3) / 1e9, 1),
            "fp32c_nnz_rhs_per_s_vs_fp64_k1": round(r32 / base, 3),
            "fp32c_vs_fp64_same_k": round(ms64 / ms32, 3),
            "fp32c_err_vs_fp64_oracle_same_inputs": acc_err,
            "fp32c_err_vs_fp64_unquantised": q_err,
        })
    out["multi_rhs_hex_elastic_3d_32"] = rows
    del e64, e32
    return out


def extra_fwd_adjoint(torch, dev, steps, warmup, tune=True):
    """Forward + adjoint (hex diffusion 48^3): y = A x, dx = A^T dy, dvals."""
    from warpkernel_amd import engine, fe, structure
    out = {}
    m = fe.hex_laplace_3d(48)
    s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    eng = engine(s, m.vals, m.colind, adjoint=True)
    fa = {}
    for k in (1, 8):
        x = dev(fe.uniform_x(m.ncols, k, seed=2200 + k).reshape(m.ncols, k))
        dyv = dev(fe.uniform_x(m.nrows, k, seed=3200 + k).reshape(m.nrows, k))
        y = torch.empty((m.nrows, k), dtype=torch.float64, device="cuda")
        dxv = torch.empty((m.ncols, k), dtype=torch.float64, device="cuda")
        dv = torch.empty(m.nnz, dtype=torch.float64, device="cuda")
        if tune:
            eng.tune(x, y, reps=5)
            eng.tune(dyv, dxv, op=2, reps=5)
        _, f = time_launches(torch, lambda: eng.launch(x, y), steps, warmup)
        _, a = time_launches(torch, lambda: eng.run_adjoint(dyv, dxv), steps, warmup)
        _, d = time_launches(torch, lambda: eng.dvals(dyv, x, dv), steps, warmup)
        # the backward as autograd runs it: dx then dvals in one call
        _, bw = time_launches(torch, lambda: eng.backward(dyv, x, dxv, dv), steps, warmup)
        # one training step's SpMV work back to back: forward, then backward
        _, seq = time_launches(torch, lambda: (eng.launch(x, y), eng.backward(dyv, x, dxv, dv)),
                               steps, warmup)
        # the same step with dx and dvals as two launches (for comparison with
        # wk_backward's choice)
        _, seq2 = time_launches(torch, lambda: (eng.launch(x, y), eng.run_adjoint(dyv, dxv),
                                                eng.dvals(dyv, x, dv)), steps, warmup)
        fa[f"k{k}"] = {"fwd_ms": round(f, 4), "dx_ms": round(a, 4), "dvals_ms": round(d, 4),
                       "backward_ms": round(bw, 4), "fwd_adjoint_ms": round(seq, 4),
                       "fwd_adjoint_two_launch_backward_ms": round(seq2, 4),
                       "fwd_GBs": round(spmv_bytes(m.nnz, m.nrows, m.ncols, k) / (f * 1e-3) / 1e9, 1)}
    out["fwd_adjoint_hex_laplace_3d_48"] = fa
    del eng
    return out


def extra_batch(torch, dev, steps, warmup, tune=True):
    """64 mixed FE matrices in one launch (one wave per group, and persistent
    waves), against one launch per matrix and a replayed HIP graph."""
    from warpkernel_amd import engine, fe, structure
    out = {}
    # 64 mixed FE matrices (2D Q1 / 3D hex elasticity, 2k-250k rows): one
    # persistent launch vs one launch per matrix vs a replayed HIP graph
    from warpkernel_amd import batch, graph
    ms = fe.batch_mix(64)
    engs = [engine(structure().scan(mm.nnz, mm.nrows, mm.vals, mm.rowptr, mm.colind), mm.vals, mm.colind)
            for mm in ms]
    xs = [dev(fe.uniform_x(mm.ncols, seed=7000 + i)) for i, mm in enumerate(ms)]
    ys = [torch.empty(mm.nrows, dtype=torch.float64, device="cuda") for mm in ms]
    tot = sum(spmv_bytes(mm.nnz, mm.nrows, mm.ncols) for mm in ms)
    b = batch(engs)
    b.bind(xs, ys)
    _, mb = time_launches(torch, b.run, steps, warmup)
    bp = batch(engs, waves_per_cu=8)
    bp.bind(xs, ys)
    _, mp = time_launches(torch, bp.run, steps, warmup)
    bp.close()

    def each():
        for e_, x_, y_ in zip(engs, xs, ys):
            e_.launch(x_, y_)
    _, me = time_launches(torch, each, steps, warmup)
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    with graph.capture(st) as g:
        b.run(stream=st)
    with torch.cuda.stream(st):
        _, mg = time_launches(torch, lambda: g.launch(st), steps, warmup)
    out["batch_64_fe_matrices"] = {
        "matrices": len(ms), "rows": int(sum(mm.nrows for mm in ms)), "nnz": int(sum(mm.nnz for mm in ms)),
        "one_launch_ms": round(mb, 4), "one_launch_GBs": round(tot / (mb * 1e-3) / 1e9, 1),
        "persistent_8_waves_per_cu_ms": round(mp, 4),
        "persistent_8_waves_per_cu_GBs": round(tot / (mp * 1e-3) / 1e9, 1),
        "graph_replay_ms": round(mg, 4), "graph_replay_GBs": round(tot / (mg * 1e-3) / 1e9, 1),
        "one_launch_per_matrix_ms": round(me, 4), "one_launch_per_matrix_GBs": round(tot / (me * 1e-3) / 1e9, 1),
    }
    g.close()
    # k = 8 right-hand sides in one launch: x windows of 8 values per column
    # (the default) against direct x loads
    xs8 = [dev(fe.uniform_x(mm.ncols, 8, seed=7500 + i).reshape(mm.ncols, 8)) for i, mm in enumerate(ms)]
    ys8 = [torch.empty((mm.nrows, 8), dtype=torch.float64, device="cuda") for mm in ms]
    tot8 = sum(spmv_bytes(mm.nnz, mm.nrows, mm.ncols, 8) for mm in ms)
    rec = out["batch_64_fe_matrices"]
    for key, st in (("k8_one_launch", 5), ("k8_direct_loads", 0)):
        b8 = batch(engs, stage_lds=st)
        b8.bind(xs8, ys8)
        _, m8 = time_launches(torch, b8.run, steps, warmup)
        rec[key + "_ms"] = round(m8, 4)
        rec[key + "_GBs"] = round(tot8 / (m8 * 1e-3) / 1e9, 1)
        b8.close()
    return out


def extra_orderings(torch, dev, steps, warmup, tune=True):
    """P1 Kuhn tets (96^3 cubes, 5% of the elements dropped) in lexicographic
    order and under a random node permutation (x-gather stress), and the 2D Q1
    Laplacian (128^2, cache-resident), WPK1 and WPK2, run and run_x."""
    from warpkernel_amd import engine, fe, structure, structure2
    out = {}
    for name in ("tet_poisson_3d_96", "tet_poisson_3d_96_perm", "q1_laplace_2d_128"):
        m = fe.CONFIGS[name]()
        nb = spmv_bytes(m.nnz, m.nrows, m.ncols)
        x = fe.uniform_x(m.ncols, seed=2300)
        dx = dev(x)
        y = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
        row = {"rows": m.nrows, "nnz": m.nnz, "empty_rows": int((m.row_lengths() == 0).sum())}
        for variant, th in (("wpk1", None), ("wpk2", 5)):
            s = (structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind) if th is None else
                 structure2().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, th))
            eng = engine(s, m.vals, m.colind, kr=True)
            dxr = dev(s.reorder_x(x))
            if tune:
                eng.tune(dx, y, reps=5)
                eng.tune(dxr, y, op=1, reps=5)
            _, ms = time_launches(torch, lambda: eng.launch(dx, y), steps, warmup)
            _, msx = time_launches(torch, lambda: eng.launch_x(dxr, y), steps, warmup)
            key = variant if th is None else f"{variant}_t{th}"
            row[f"{key}_run_GBs"] = round(nb / (ms * 1e-3) / 1e9, 1)
            row[f"{key}_run_x_GBs"] = round(nb / (msx * 1e-3) / 1e9, 1)
            row[f"{key}_run_us"] = round(ms * 1e3, 2)
            del eng
            if th is None:
                # Kr row sort: each row's x gathers in ascending permuted order
                er = engine(s, m.vals, m.colind, kr=True, rowsort=True)
                if tune:
                    er.tune(dxr, y, op=1, reps=5)
                _, msr = time_launches(torch, lambda: er.launch_x(dxr, y), steps, warmup)
                row[f"{key}_rowsort_run_x_GBs"] = round(nb / (msr * 1e-3) / 1e9, 1)
                del er
        out[name] = row
    return out


def extra_device_scan(torch, dev, steps, warmup, tune=True):
    """Scan + process of the 64^3 hex elasticity matrix: on the host (scan,
    then engine() with its process and upload, the reference's path) against
    wk_engine_create_device from a CSR already in HBM; both give the same
    engine, compared here array by array."""
    from warpkernel_amd import _capi as capi, engine, fe, structure, structure2
    del steps, warmup, tune
    m = fe.CONFIGS["hex_elastic_3d_64"]()
    rp, ci, va = dev(m.rowptr), dev(m.colind), dev(m.vals)
    out = {"rows": m.nrows, "nnz": m.nnz}
    for variant in (1, 2):
        th = 27 if variant == 2 else 0
        host_s, dev_s = [], []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if variant == 1:
                s = structure().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
            else:
                s = structure2().scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, th)
            eh = engine(s, m.vals, m.colind, kr=True)
            torch.cuda.synchronize()
            host_s.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            ed = engine.from_device(rp, ci, va, variant=variant, threshold=th, kr=True)
            torch.cuda.synchronize()
            dev_s.append(time.perf_counter() - t0)
        same = all(np.array_equal(eh.download(w), ed.download(w)) for w in
                   (capi.WK_ROW_MAP, capi.WK_REORDER_ROWS, capi.WK_WARP_OFFSETS,
                    capi.WK_ENGINE_VALUES, capi.WK_ENGINE_COLS, capi.WK_ENGINE_COLS_KR,
                    capi.WK_ENGINE_WIN_IDX, capi.WK_ENGINE_WIN_OFF, capi.WK_ENGINE_WIN_COLS,
                    capi.WK_ENGINE_WIN_COLS_KR))
        key = "wpk1" if variant == 1 else f"wpk2_t{th}"
        out[key] = {"host_scan_process_upload_s": round(min(host_s), 4),
                    "device_scan_process_s": round(min(dev_s), 4),
                    "speedup": round(min(host_s) / min(dev_s), 1), "identical_engines": bool(same)}
        eh.close()
        ed.close()
    return out


EXTRAS = {"multi_rhs": extra_multi_rhs, "fwd_adjoint": extra_fwd_adjoint, "batch": extra_batch,
          "orderings": extra_orderings, "device_scan": extra_device_scan}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--variant", choices=["wpk1", "wpk2"], default="wpk1")
    ap.add_argument("--threshold", type=int, default=27)
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (layout default)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--only", choices=sorted(EXTRAS), default=None,
                    help="run one extra section alone (no headline), for profiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the WPK SpMV path here has no multi-GPU form; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from warpkernel_amd import engine, fe, structure, structure2

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a)).cuda()

    if args.only:
        res = EXTRAS[args.only](torch, dev, max(20, args.steps // 5), 10, tune=not args.no_tune)
        print(json.dumps({"section": args.only, "tuned": not args.no_tune, **res}), flush=True)
        return

    print("[bench] building hex_elastic_3d_64", file=sys.stderr, flush=True)
    m = fe.hex_elastic_3d(64)
    x = fe.uniform_x(m.ncols, seed=2002)
    if args.variant == "wpk1":
        s = structure(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind)
    else:
        s = structure2(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, args.threshold)
    eng = engine(s, m.vals, m.colind, kr=not args.headline_only)
    dx = dev(x)
    dy = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
    tuned = None
    if not args.no_tune:
        print("[bench] autotuning launch shape", file=sys.stderr, flush=True)
        tuned, tuned_ms = eng.tune(dx, dy, reps=10)
    print("[bench] timing", file=sys.stderr, flush=True)
    wall, ev_ms = time_launches(torch, lambda: eng.launch(dx, dy), args.steps, args.warmup, 0.0)
    nbytes = spmv_bytes(m.nnz, m.nrows, m.ncols)
    value = nbytes * args.steps / wall / 1e9
    achieved = nbytes / (ev_ms * 1e-3) / 1e9
    # correctness of what was timed
    from oracle import pyoracle as O
    ok = bool(np.array_equal(dy.cpu().numpy(), O.csr_spmv(m.rowptr, m.colind, m.vals, x)))

    workload = "hex_elastic_3d_64"
    traffic = read_traffic(workload)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{workload}: 3D trilinear hex linear elasticity, 64^3 elements, "
                        f"{m.nrows} rows, {m.nnz} nnz, E~lognormal(0,0.5), nu=0.3, x~U[-1,1]",
            "kernel": f"{args.variant.upper()} engine::run, group width 64, k=1, FP64 values, int32 indices",
            "threshold": args.threshold if args.variant == "wpk2" else None,
            "launch": eng.get_launch().as_dict(),
            "launch_source": "autotuner (wk_tune)" if tuned else "layout default",
            "alg_bytes_per_step": nbytes,
            "gflops": round(2 * m.nnz / (wall / args.steps) / 1e9, 1),
            "result_bitexact_vs_oracle": ok,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        },
    }
    # measured context for the nominal peak: the box's streaming-read rate, and
    # the real HBM rate of the kernel (counter traffic over the live duration)
    ceiling = stream_read_ceiling(torch)
    line["hbm_measured"] = {
        "stream_read_GBs": ceiling,
        "kernel_traffic_GBs": round(traffic / (ev_ms * 1e-3) / 1e9, 1) if traffic else None,
    }
    if not args.headline_only:
        dxr = dev(s.reorder_x(x))
        dyr = torch.empty(m.nrows, dtype=torch.float64, device="cuda")
        if not args.no_tune:
            eng.tune(dxr, dyr, op=1, reps=10)
        _, msx = time_launches(torch, lambda: eng.launch_x(dxr, dyr), args.steps, args.warmup)
        ex = {"run_x_GBs": round(nbytes / (msx * 1e-3) / 1e9, 2), "run_x_ms": round(msx, 4)}
        # run_assembly: row sums of the values, bytes = nnz*8 read + nrows*8 written
        _, msa = time_launches(torch, lambda: eng.launch_assembly(dyr), args.steps, args.warmup)
        ex["run_assembly_GBs"] = round((m.nnz * 8 + m.nrows * 8) / (msa * 1e-3) / 1e9, 2)
        ex["run_assembly_ms"] = round(msa, 4)
        del eng
        s2 = structure2(64).scan(m.nnz, m.nrows, m.vals, m.rowptr, m.colind, args.threshold)
        e2 = engine(s2, m.vals, m.colind, kr=True)
        dxr2 = dev(s2.reorder_x(x))
        if not args.no_tune:
            e2.tune(dx, dy, reps=10)
            e2.tune(dxr2, dyr, op=1, reps=10)
        _, ms2 = time_launches(torch, lambda: e2.launch(dx, dy), args.steps, args.warmup)
        _, ms2x = time_launches(torch, lambda: e2.launch_x(dxr2, dyr), args.steps, args.warmup)
        ex[f"wpk2_t{args.threshold}_GBs"] = round(nbytes / (ms2 * 1e-3) / 1e9, 2)
        ex[f"wpk2_t{args.threshold}_run_x_GBs"] = round(nbytes / (ms2x * 1e-3) / 1e9, 2)
        del e2
        print("[bench] extras", file=sys.stderr, flush=True)
        for fn in EXTRAS.values():
            ex.update(fn(torch, dev, max(20, args.steps // 5), 10, tune=not args.no_tune))
        line["extras"] = ex
    else:
        del eng
    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(m, x, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
