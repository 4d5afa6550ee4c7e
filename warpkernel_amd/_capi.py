"""ctypes declarations of the C ABI in include/warpkernel_c.h.

The shared library is built in-tree (``python -m warpkernel_amd.build``).  There
is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libwarpkernel.so")

WK_OK = 0
WK_ERR_ARG, WK_ERR_CSR, WK_ERR_LIMIT, WK_ERR_HIP, WK_ERR_STATE, WK_ERR_NOMEM = 1, 2, 3, 4, 5, 6
WK_F64, WK_F32 = 0, 1
WK_ACC_TWOSUM, WK_ACC_F64 = 0, 1
WK_ROW_MAP, WK_ROW_MAP_INV, WK_REORDER_ROWS, WK_WARP_OFFSETS = 0, 1, 2, 3
WK_MAX_NZ_ROWS, WK_THREADS_PER_ROW, WK_ROW_OFFSET_WARP, WK_REORDER_A_ROWS = 4, 5, 6, 7
WK_ROWPTR = 8
WK_ENGINE_KR, WK_ENGINE_ADJOINT, WK_ENGINE_ROWSORT = 1, 2, 4
WK_ENGINE_VALUES, WK_ENGINE_COLS, WK_ENGINE_COLS_KR = 16, 17, 18
WK_ENGINE_WIN_IDX, WK_ENGINE_WIN_OFF, WK_ENGINE_WIN_COLS, WK_ENGINE_WIN_COLS_KR = 19, 20, 21, 22
WK_ENGINE_ND_COLS, WK_ENGINE_ND_CNT, WK_ENGINE_ND_OFF, WK_ENGINE_ND_POS = 23, 24, 25, 26
WK_ENGINE_ND_COLS_KR = 27


class WKError(RuntimeError):
    """A non-zero status returned across the C ABI."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[wk status {code}] {msg}")
        self.code = code


class LayoutInfo(C.Structure):
    _fields_ = [
        ("variant", C.c_int32), ("group_width", C.c_int32), ("threshold", C.c_int32),
        ("max_tpr", C.c_int32), ("nrows", C.c_int32), ("ncols", C.c_int32),
        ("nznrows", C.c_int32), ("nwarps", C.c_int32), ("nz", C.c_int64),
        ("allocate_nz", C.c_int64), ("min_nz", C.c_uint32), ("max_nz", C.c_uint32),
    ]


class Launch(C.Structure):
    _fields_ = [
        ("waves_per_block", C.c_int32), ("unroll", C.c_int32), ("k_tile", C.c_int32),
        ("xcd_remap", C.c_int32), ("nontemporal", C.c_int32), ("accumulate", C.c_int32),
        ("rows_per_lane", C.c_int32), ("blocks_per_cu", C.c_int32), ("stage_lds", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


P = C.c_void_p
I32, I64, U32 = C.c_int32, C.c_int64, C.c_uint32
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "wk_last_error": (C.c_char_p, []),
    "wk_abi_version": (C.c_int, []),
    "wk_scan1": (C.c_int, [I32, I32, I32, I64, P, P, PP]),
    "wk_scan2": (C.c_int, [I32, I32, I32, I32, I32, I64, P, P, PP]),
    "wk_layout_get_info": (C.c_int, [P, C.POINTER(LayoutInfo)]),
    "wk_layout_array": (C.c_int, [P, I32, PP, C.POINTER(I64)]),
    "wk_layout_process_host": (C.c_int, [P, P, P, P, P]),
    "wk_layout_reorder_columns_coalesced": (C.c_int, [P, P, P]),
    "wk_layout_reorder_x": (C.c_int, [P, P, P]),
    "wk_layout_reorder_columns_rowsort": (C.c_int, [P, P, P, P]),
    "wk_layout_reorder_values_rowsort": (C.c_int, [P, P, P]),
    "wk_mm_read": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "wk_csr_get": (C.c_int, [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(I64), C.POINTER(P),
                             C.POINTER(P), C.POINTER(P)]),
    "wk_csr_destroy": (None, [P]),
    "wk_mm_write": (C.c_int, [C.c_char_p, I32, I32, I64, P, P, P]),
    "wk_layout_save": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "wk_layout_load": (C.c_int, [C.c_char_p, C.POINTER(P), C.c_char_p, I64]),
    "wk_layout_from_arrays": (C.c_int, [I32, I32, I32, I32, I32, I32, I32, I32, I64, I64, P, P, P, P, P,
                                        P, P, PP]),
    "wk_layout_destroy": (None, [P]),
    "wk_scatter": (C.c_int, [P, P, P, I64, I64, I32, P]),
    "wk_engine_create": (C.c_int, [P, P, P, P, I32, U32, P, PP]),
    "wk_engine_destroy": (None, [P]),
    "wk_engine_get_info": (C.c_int, [P, C.POINTER(LayoutInfo)]),
    "wk_engine_update_values": (C.c_int, [P, P, P]),
    "wk_engine_set_colinds": (C.c_int, [P, P, I32, P]),
    "wk_engine_x_windows": (C.c_int, [P, C.POINTER(I64)]),
    "wk_engine_set_colinds_noalloc": (C.c_int, [P, P, P]),
    "wk_engine_get_colinds": (C.c_int, [P, P]),
    "wk_engine_create_device": (C.c_int, [I32, I32, I32, I32, I32, I64, P, P, P, I32, C.c_uint32, P,
                                          C.POINTER(P)]),
    "wk_engine_download": (C.c_int, [P, I32, P, C.POINTER(I64)]),
    "wk_launch_default": (None, [C.POINTER(Launch)]),
    "wk_run": (C.c_int, [P, C.POINTER(Launch), P, P, I32, I64, I64, P]),
    "wk_run_x": (C.c_int, [P, C.POINTER(Launch), P, P, I32, I64, I64, P]),
    "wk_run_assembly": (C.c_int, [P, C.POINTER(Launch), P, P]),
    "wk_backward": (C.c_int, [P, C.POINTER(Launch), P, I64, P, I64, P, I64, P, I32, P]),
    "wk_residual": (C.c_int, [P, C.POINTER(Launch), P, I64, P, P, I32, I64, I64, P]),
    "wk_run_adjoint": (C.c_int, [P, C.POINTER(Launch), P, P, I32, I64, I64, P]),
    "wk_dvals": (C.c_int, [P, P, I64, P, I64, I32, P, P]),
    "wk_time_run": (C.c_int, [P, C.POINTER(Launch), I32, P, P, I32, I64, I64, P, C.POINTER(C.c_float)]),
    "wk_stream_read": (C.c_int, [P, I64, P, P]),
    "wk_engine_gather_ceiling": (C.c_int, [P, P, P, P, P]),
    "wk_tune": (C.c_int, [P, I32, C.POINTER(Launch), P, P, I32, I64, I64, I32, P, C.POINTER(Launch),
                          C.POINTER(C.c_float)]),
    "wk_engine_get_launch": (C.c_int, [P, I32, I32, C.POINTER(Launch)]),
    "wk_engine_set_launch": (C.c_int, [P, I32, I32, C.POINTER(Launch)]),
    "wk_batch_create": (C.c_int, [C.POINTER(C.c_void_p), I32, I32, I32, PP]),
    "wk_batch_bind": (C.c_int, [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), I32,
                                C.POINTER(I64), C.POINTER(I64)]),
    "wk_batch_run": (C.c_int, [P, P]),
    "wk_batch_destroy": (None, [P]),
    "wk_graph_begin": (C.c_int, [P]),
    "wk_graph_end": (C.c_int, [P, PP]),
    "wk_graph_launch": (C.c_int, [P, P]),
    "wk_graph_destroy": (None, [P]),
}
WK_OP_RUN, WK_OP_RUN_X, WK_OP_ADJOINT = 0, 1, 2

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP extension not built: {LIB_PATH} is missing "
                "(run `python -m warpkernel_amd.build` or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != WK_OK:
        msg = lib().wk_last_error()
        raise WKError(rc, msg.decode() if msg else "")


def default_launch(**overrides) -> Launch:
    cfg = Launch()
    lib().wk_launch_default(C.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, int(v))
    return cfg
