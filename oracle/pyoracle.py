"""ctypes front end of the CPU oracle (oracle/wk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of ``bench.py`` -- always as the checker or the baseline,
never as the thing measured or shipped.  The product package does not import it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libwk_oracle.so")


class _Layout(C.Structure):
    _fields_ = [
        ("variant", C.c_int32), ("W", C.c_int32), ("threshold", C.c_int32),
        ("max_tpr", C.c_int32), ("nrows", C.c_int32), ("nznrows", C.c_int32),
        ("nwarps", C.c_int32), ("nz", C.c_int64), ("allocate_nz", C.c_int64),
        ("min_nz", C.c_uint32), ("max_nz", C.c_uint32),
        ("row_map", C.POINTER(C.c_int32)), ("row_map_inv", C.POINTER(C.c_int32)),
        ("reorder_rows", C.POINTER(C.c_uint32)), ("warp_offsets", C.POINTER(C.c_int64)),
        ("max_nz_rows", C.POINTER(C.c_uint32)), ("threads_per_row", C.POINTER(C.c_uint32)),
        ("row_offset_warp", C.POINTER(C.c_uint32)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.wko_scan1.restype = C.POINTER(_Layout)
        L.wko_scan1.argtypes = [C.c_int32, C.c_int32, C.c_int64, P]
        L.wko_scan2.restype = C.POINTER(_Layout)
        L.wko_scan2.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64, P]
        L.wko_free.argtypes = [C.POINTER(_Layout)]
        L.wko_process.argtypes = [C.POINTER(_Layout), P, P, P, P]
        L.wko_reorder_columns_coalesced.argtypes = [C.POINTER(_Layout), P, P]
        L.wko_reorder_x.argtypes = [C.POINTER(_Layout), P, P]
        L.wko_reorder_columns_rowsort.argtypes = [C.POINTER(_Layout), P, P, P, P]
        L.wko_reorder_values_rowsort.argtypes = [C.POINTER(_Layout), P, P, P]
        L.wko_csr_spmv.argtypes = [C.c_int32, P, P, P, P, P, C.c_int32, C.c_int64, C.c_int64, C.c_int32]
        L.wko_csr_spmv_omp.argtypes = [C.c_int32, P, P, P, P, P, C.c_int32]
        L.wko_wpk_spmv.argtypes = [C.POINTER(_Layout), P, P, P, P, C.c_int32, C.c_int64, C.c_int64, C.c_int32]
        L.wko_transpose.argtypes = [C.c_int32, C.c_int32, P, P, P, P, P]
        L.wko_dvals.argtypes = [C.c_int32, P, P, P, P, P, C.c_int32, C.c_int64, C.c_int64]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


@dataclass
class Layout:
    """Output of the oracle's structure::scan / structure2::scan."""
    variant: int
    W: int
    threshold: int
    max_tpr: int
    nrows: int
    nznrows: int
    nwarps: int
    nz: int
    allocate_nz: int
    min_nz: int
    max_nz: int
    row_map: np.ndarray
    row_map_inv: np.ndarray
    reorder_rows: np.ndarray
    warp_offsets: np.ndarray
    max_nz_rows: np.ndarray
    threads_per_row: np.ndarray | None
    row_offset_warp: np.ndarray | None
    _handle: object = None

    def close(self):
        if self._handle is not None:
            lib().wko_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _wrap(h) -> Layout:
    s = h.contents
    G = s.nwarps
    v2 = s.variant == 2
    return Layout(
        variant=s.variant, W=s.W, threshold=s.threshold, max_tpr=s.max_tpr,
        nrows=s.nrows, nznrows=s.nznrows, nwarps=G, nz=s.nz, allocate_nz=s.allocate_nz,
        min_nz=s.min_nz, max_nz=s.max_nz,
        row_map=_arr(s.row_map, s.nrows, np.int32),
        row_map_inv=_arr(s.row_map_inv, s.nrows, np.int32),
        reorder_rows=_arr(s.reorder_rows, s.nz, np.uint32),
        warp_offsets=_arr(s.warp_offsets, G + 1, np.int64),
        max_nz_rows=_arr(s.max_nz_rows, G, np.uint32),
        threads_per_row=_arr(s.threads_per_row, G, np.uint32) if v2 else None,
        row_offset_warp=_arr(s.row_offset_warp, G + 1, np.uint32) if v2 else None,
        _handle=h,
    )


def scan1(rowptr: np.ndarray, W: int = 64) -> Layout:
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    nrows = rowptr.size - 1
    return _wrap(lib().wko_scan1(W, nrows, int(rowptr[-1]), _p(rowptr)))


def scan2(rowptr: np.ndarray, threshold: int, W: int = 64, max_tpr: int = 32) -> Layout:
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    nrows = rowptr.size - 1
    return _wrap(lib().wko_scan2(W, threshold, max_tpr, nrows, int(rowptr[-1]), _p(rowptr)))


def process(L: Layout, vals: np.ndarray, colind: np.ndarray):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    pv = np.empty(L.allocate_nz, np.float64)
    pc = np.empty(L.allocate_nz, np.int32)
    lib().wko_process(L._handle, _p(vals), _p(colind), _p(pv), _p(pc))
    return pv, pc


def reorder_columns_coalesced(L: Layout, colind: np.ndarray) -> np.ndarray:
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    out = np.empty(L.allocate_nz, np.int32)
    lib().wko_reorder_columns_coalesced(L._handle, _p(colind), _p(out))
    return out


def reorder_x(L: Layout, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().wko_reorder_x(L._handle, _p(x), _p(out))
    return out


def reorder_columns_rowsort(L: Layout, rowptr, colind):
    """Returns (cols_out, reorder_A_rows): CSR-order permuted columns, each row
    sorted ascending, and the source CSR index of every entry."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    out = np.empty(colind.size, np.int32)
    src = np.empty(colind.size, np.int32)
    lib().wko_reorder_columns_rowsort(L._handle, _p(rowptr), _p(colind), _p(out), _p(src))
    return out, src


def reorder_values_rowsort(L: Layout, vals, reorder_A_rows) -> np.ndarray:
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    src = np.ascontiguousarray(reorder_A_rows, dtype=np.int32)
    out = np.empty(L.allocate_nz, np.float64)
    lib().wko_reorder_values_rowsort(L._handle, _p(vals), _p(src), _p(out))
    return out


def csr_spmv(rowptr, colind, vals, x, fused: bool = True) -> np.ndarray:
    """y = A x in original row order; x may be [n] or [n, k] (row-major)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    k = 1 if x.ndim == 1 else x.shape[1]
    nrows = rowptr.size - 1
    y = np.empty((nrows,) if x.ndim == 1 else (nrows, k), np.float64)
    lib().wko_csr_spmv(nrows, _p(rowptr), _p(colind), _p(vals), _p(x), _p(y), k, k, k, int(fused))
    return y


def csr_spmv_omp(rowptr, colind, vals, x, nthreads: int = 0) -> np.ndarray:
    nrows = rowptr.size - 1
    y = np.empty(nrows, np.float64)
    lib().wko_csr_spmv_omp(nrows, _p(rowptr), _p(colind), _p(vals), _p(x), _p(y), nthreads)
    return y


def wpk_spmv(L: Layout, pvals, pcols, x, kr: bool = False) -> np.ndarray:
    """Lane-by-lane emulation of the WPK kernels on the padded arrays."""
    pvals = np.ascontiguousarray(pvals, dtype=np.float64)
    pcols = np.ascontiguousarray(pcols, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    k = 1 if x.ndim == 1 else x.shape[1]
    y = np.empty((L.nrows,) if x.ndim == 1 else (L.nrows, k), np.float64)
    lib().wko_wpk_spmv(L._handle, _p(pvals), _p(pcols), _p(x), _p(y), k, k, k, int(kr))
    return y


def transpose(nrows: int, ncols: int, rowptr, colind):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    nnz = int(rowptr[-1])
    colptr = np.empty(ncols + 1, np.int32)
    rowind = np.empty(nnz, np.int32)
    perm = np.empty(nnz, np.int64)
    lib().wko_transpose(nrows, ncols, _p(rowptr), _p(colind), _p(colptr), _p(rowind), _p(perm))
    return colptr, rowind, perm


def dvals(rowptr, colind, dy, x) -> np.ndarray:
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colind = np.ascontiguousarray(colind, dtype=np.int32)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    k = 1 if x.ndim == 1 else x.shape[1]
    out = np.empty(int(rowptr[-1]), np.float64)
    lib().wko_dvals(rowptr.size - 1, _p(rowptr), _p(colind), _p(dy), _p(x), _p(out), k, k, k)
    return out
