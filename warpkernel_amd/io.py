"""Files: MatrixMarket matrices and WPK layouts on disk, over the C ABI.

read_matrix_market / write_matrix_market stand for the drivers'
cusp::io::read_matrix_market_file + CSR conversion (wpk1.cu:43-48);
save_structure / load_structure for structure::save / load (serialize.cu:135-139).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _capi as capi
from ._capi import check, lib
from .fe import CSR


def read_matrix_market(path) -> CSR:
    """CSR of a coordinate MatrixMarket file, entries ordered by (row, column)."""
    h = C.c_void_p()
    check(lib().wk_mm_read(os.fsencode(path), C.byref(h)))
    try:
        nr, nc, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        rp, ci, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().wk_csr_get(h, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(rp), C.byref(ci),
                               C.byref(va)))

        def grab(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(ptr.value)
            return np.frombuffer(buf, dt).copy()

        return CSR(os.path.basename(os.fspath(path)), nr.value, nc.value,
                   grab(rp, nr.value + 1, np.int32), grab(ci, nnz.value, np.int32),
                   grab(va, nnz.value, np.float64))
    finally:
        lib().wk_csr_destroy(h)


def write_matrix_market(path, m: CSR):
    rp = np.ascontiguousarray(m.rowptr, np.int32)
    ci = np.ascontiguousarray(m.colind, np.int32)
    va = np.ascontiguousarray(m.vals, np.float64)
    check(lib().wk_mm_write(os.fsencode(path), int(m.nrows), int(m.ncols), int(rp[-1]),
                            rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                            va.ctypes.data_as(C.c_void_p)))


def save_structure(s, path, driver_name: str = ""):
    """Writes a scanned structure / structure2 (with reorder_A_rows if set)."""
    if s._h is None:
        raise ValueError("structure has not been scanned")
    check(lib().wk_layout_save(s._h, os.fsencode(path), driver_name.encode()))


def load_structure(path):
    """Reads a layout file into a structure or structure2 (by its variant),
    ready for engine(); returns (structure, driver_name)."""
    from .api import structure, structure2
    h = C.c_void_p()
    name = C.create_string_buffer(256)
    check(lib().wk_layout_load(os.fsencode(path), C.byref(h), name, 256))
    info = capi.LayoutInfo()
    check(lib().wk_layout_get_info(h, C.byref(info)))
    s = structure(info.group_width) if info.variant == 1 else structure2(info.group_width, info.max_tpr)
    if info.variant == 2:
        s.nmax_per_thread = info.threshold
    s._h = h
    s._load_fields()
    s._rowptr = s._array(capi.WK_ROWPTR, np.int32)
    src = s._array(capi.WK_REORDER_A_ROWS, np.int32)
    if src.size:
        s.reorder_A_rows = src
    return s, name.value.decode()
