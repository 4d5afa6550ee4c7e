"""Layout-level tuning: WPK1 against WPK2 over a threshold sweep, optionally
over group widths and WPK2's lanes-per-row cap too.

The reference picks its layout by sweeping structure2's threshold from min_nz to
max_nz in steps of 10% of the range, and warps/block 2..16, keeping the fastest
verified run (wpk2.cu:140-302).  tune_layout does the same over the C ABI: every
candidate layout gets an engine whose launch shape is tuned on the device
(wk_tune), and the fastest (layout, launch) pair is returned, ready to use.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .api import engine, structure, structure2


@dataclass
class LayoutChoice:
    variant: int                 # 1 = structure, 2 = structure2
    threshold: int | None
    ms: float                    # mean launch time of the tuned launch shape
    launch: dict
    layout: object = field(repr=False)
    engine: object = field(repr=False)
    # every candidate: (variant, threshold, ms, group width, max lanes per row)
    table: list = field(default_factory=list, repr=False)
    group_width: int = 64
    max_tpr: int = 32


def thresholds(min_nz: int, max_nz: int) -> list[int]:
    """The reference's sweep: min_nz..max_nz, step max(1, 10% of the range)."""
    inc = int((max_nz - min_nz) * 0.1) if (max_nz - min_nz) * 0.1 > 0 else 1
    return list(range(int(min_nz), int(max_nz) + 1, max(1, inc)))


def _as_tuple(v):
    return tuple(v) if isinstance(v, (list, tuple)) else (v,)


def tune_layout(rowptr, colind, vals, x, y, ncols: int | None = None, group_width=64,
                max_tpr=32, reps: int = 10, threshold_list=None, dtype: str = "f64",
                kr: bool = False, adjoint: bool = False) -> LayoutChoice:
    """x, y: device tensors of the operator's shape (k right-hand sides allowed).
    group_width and max_tpr may be single values or sequences (e.g. (64, 32) and
    (32, 64)): every combination is a candidate, max_tpr at most the width.
    Returns the fastest layout with its engine (other engines are released)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    nrows = rowptr.size - 1
    nnz = int(rowptr[-1])
    best = None
    table = []

    def consider(variant, th, s, w, tpr):
        nonlocal best
        eng = engine(s, vals, colind, dtype=dtype, kr=kr, adjoint=adjoint)
        cfg, ms = eng.tune(x, y, reps=reps)
        table.append((variant, th, ms, w, tpr))
        if best is None or ms < best.ms:
            best = LayoutChoice(variant, th, ms, cfg, s, eng, group_width=w, max_tpr=tpr)
        else:
            eng.close()

    for w in _as_tuple(group_width):
        consider(1, None, structure(w).scan(nnz, nrows, vals, rowptr, colind, ncols=ncols), w, None)
        for tpr in _as_tuple(max_tpr):
            if tpr > w:
                continue
            probe = structure2(w, tpr).scan(nnz, nrows, vals, rowptr, colind, max(nrows, 1),
                                            ncols=ncols)
            ths = threshold_list if threshold_list is not None else thresholds(probe.min_nz, probe.max_nz)
            for th in ths:
                if th >= probe.max_nz:
                    continue  # no group is split: the WPK1 layout again
                s2 = structure2(w, tpr).scan(nnz, nrows, vals, rowptr, colind, int(th), ncols=ncols)
                consider(2, int(th), s2, w, tpr)
    best.table = table
    return best
