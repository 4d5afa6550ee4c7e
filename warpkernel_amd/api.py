"""Python mirror of the reference's class API (namespace warpkernel,
warpkernel.hpp:215-959) over the C ABI.

    s = structure();  s.scan(nz, nrows, A, rows, colinds)          # hpp:309-365
    s2 = structure2(); s2.scan(nz, nrows, A, rows, colinds, th)    # hpp:578-736
    eng = engine(s, A, colinds)                                    # hpp:794-809
    t = eng.run(nblocks, blocksize, dx, dy)                        # hpp:811-860
    t = eng.run_x(nblocks, blocksize, dxr, dy)                     # hpp:862-896
    eng.verify(y_ref, y)                                           # hpp:932-943

Differences, all deliberate:
  - group_width (the reference's WARP_SIZE) defaults to 64, one gfx950 wave;
    32 reproduces the reference's layout exactly.
  - dx / dy are torch device tensors (or raw device pointers as ints); run and
    run_x return seconds and block, as the reference's do; launch / launch_x
    are the asynchronous forms with k right-hand sides.
  - Rows without nonzeros are written as 0.
  - verify_x compares orig[row_map[i]] with comp[i] in both of its tests (the
    reference indexes orig[i] in the zero test, hpp:949-950).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import check, lib


def _host_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dev_ptr(t) -> int:
    if t is None:
        return 0
    if isinstance(t, int):
        return t
    return int(t.data_ptr())


def _stream(stream) -> int:
    """The raw hipStream_t: the given stream, or torch's current stream."""
    if stream is None:
        import torch
        raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        if raw is not None:  # 0.1 us instead of 3 us for current_stream().cuda_stream
            return int(raw(torch.cuda.current_device()))
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class structure:
    """WPK1 layout (warpkernel::structure, warpkernel.hpp:215-551)."""

    variant = 1

    def __init__(self, group_width: int = 64):
        self.group_width = int(group_width)
        self._h = None
        self._rowptr = None
        self.nrows = self.ncols = self.nznrows = self.nz = self.allocate_nz = self.nwarps = 0

    # -- scan -----------------------------------------------------------
    def _scan_call(self, nz, nrows, rows, colinds, ncols, extra):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = None if colinds is None else np.ascontiguousarray(colinds, dtype=np.int32)
        if rows.size != int(nrows) + 1:
            raise ValueError("rows must hold nrows + 1 offsets")
        ncols = int(nrows) if ncols is None else int(ncols)
        h = C.c_void_p()
        cptr = None if cols is None else _host_ptr(cols)
        if self.variant == 1:
            check(lib().wk_scan1(self.group_width, int(nrows), ncols, int(nz), _host_ptr(rows), cptr, C.byref(h)))
        else:
            check(lib().wk_scan2(self.group_width, extra[0], extra[1], int(nrows), ncols, int(nz),
                                 _host_ptr(rows), cptr, C.byref(h)))
        self.close()
        self._h = h
        self._rowptr = rows
        self._load_fields()

    def scan(self, nz, nrows, A, rows, colinds, ncols=None):
        """structure::scan(nz, nrows, A, rows, colinds) -- only rows shapes the
        layout; colinds is range-checked.  ncols defaults to nrows."""
        self._scan_call(nz, nrows, rows, colinds, ncols, None)
        return self

    def _array(self, which, dtype):
        data = C.c_void_p()
        n = C.c_int64()
        check(lib().wk_layout_array(self._h, which, C.byref(data), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, dtype=dtype)
        buf = (C.c_char * (n.value * np.dtype(dtype).itemsize)).from_address(data.value)
        return np.frombuffer(buf, dtype=dtype).copy()

    def _load_fields(self):
        info = capi.LayoutInfo()
        check(lib().wk_layout_get_info(self._h, C.byref(info)))
        self.info = info
        self.nrows, self.ncols, self.nznrows = info.nrows, info.ncols, info.nznrows
        self.nz, self.allocate_nz, self.nwarps = info.nz, info.allocate_nz, info.nwarps
        self.row_map = self._array(capi.WK_ROW_MAP, np.int32)
        self.row_map_inv = self._array(capi.WK_ROW_MAP_INV, np.int32)
        self.reorder_rows = self._array(capi.WK_REORDER_ROWS, np.uint32)
        self.warp_offsets = self._array(capi.WK_WARP_OFFSETS, np.int64)
        self.max_nz_rows = self._array(capi.WK_MAX_NZ_ROWS, np.uint32)

    # -- host helpers ---------------------------------------------------
    def process_host(self, A, colinds):
        """structure::reorder (hpp:379-398) on the host: padded (values, colinds)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        colinds = np.ascontiguousarray(colinds, dtype=np.int32)
        pv = np.empty(self.allocate_nz, np.float64)
        pc = np.empty(self.allocate_nz, np.int32)
        check(lib().wk_layout_process_host(self._h, _host_ptr(A), _host_ptr(colinds), _host_ptr(pv), _host_ptr(pc)))
        return pv, pc

    def reorder_columns_coalesced(self, colinds):
        """structure::reorder_columns_coalesced (hpp:446-457): padded Kr columns."""
        colinds = np.ascontiguousarray(colinds, dtype=np.int32)
        out = np.empty(self.allocate_nz, np.int32)
        check(lib().wk_layout_reorder_columns_coalesced(self._h, _host_ptr(colinds), _host_ptr(out)))
        return out

    def reorder_columns_rowsort(self, colinds, rows=None):
        """structure::reorder_columns_rowsort(colinds, rows) (hpp:414-444): CSR-order
        permuted columns, each row ascending; sets self.reorder_A_rows (the CSR
        index of the entry now at each position).  rows defaults to the scanned
        rowptr."""
        colinds = np.ascontiguousarray(colinds, dtype=np.int32)
        rows = self._rowptr if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
        out = np.empty(colinds.size, np.int32)
        check(lib().wk_layout_reorder_columns_rowsort(self._h, _host_ptr(rows), _host_ptr(colinds),
                                                      _host_ptr(out)))
        self.reorder_A_rows = self._array(capi.WK_REORDER_A_ROWS, np.int32)
        return out

    def reorder_values_coalesced_rowsort(self, values):
        """reorder_values_coalesced_rowsort (hpp:505-515): padded values,
        new[reorder_rows[i]] = values[reorder_A_rows[i]]."""
        values = np.ascontiguousarray(values, dtype=np.float64)
        out = np.empty(self.allocate_nz, np.float64)
        check(lib().wk_layout_reorder_values_rowsort(self._h, _host_ptr(values), _host_ptr(out)))
        return out

    def reorder_x(self, x):
        """structure::reorder_x (hpp:517-530): x'[row_map_inv[r]] = x[r]."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        check(lib().wk_layout_reorder_x(self._h, _host_ptr(x), _host_ptr(out)))
        return out

    def save(self, path, driver_name: str = ""):
        """structure::save (serialize.cu:135): the layout to a file."""
        from .io import save_structure
        save_structure(self, path, driver_name)

    def close(self):
        if self._h is not None:
            lib().wk_layout_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class structure2(structure):
    """WPK2 layout (warpkernel::structure2, warpkernel.hpp:553-771)."""

    variant = 2

    def __init__(self, group_width: int = 64, max_tpr: int = 32):
        super().__init__(group_width)
        self.max_tpr = int(max_tpr)
        self.nmax_per_thread = 0

    def scan(self, nz, nrows, A, rows, colinds, threshold, ncols=None):
        """structure2::scan(nz, nrows, A, rows, colinds, threshold)."""
        self.nmax_per_thread = int(threshold)
        self._scan_call(nz, nrows, rows, colinds, ncols, (int(threshold), self.max_tpr))
        return self

    def _load_fields(self):
        super()._load_fields()
        self.min_nz, self.max_nz = self.info.min_nz, self.info.max_nz
        self.threads_per_row = self._array(capi.WK_THREADS_PER_ROW, np.uint32)
        self.row_offset_warp = self._array(capi.WK_ROW_OFFSET_WARP, np.uint32)


class engine:
    """Device arrays + kernel launches (warpkernel::engine, warpkernel.hpp:773-959).

    engine(k, A, colinds) runs process (hpp:794-809).  dtype 'f64' or 'f32'
    (FP32 values and x, FP64 y); kr=True also builds the Kr columns for run_x;
    adjoint=True builds the transposed layout for run_adjoint and dvals;
    rowsort=True orders each row's entries by permuted column
    (reorder_columns_rowsort, hpp:414-444) for both run and run_x.
    """

    def __init__(self, k: structure, A, colinds, dtype: str = "f64", kr: bool = False,
                 adjoint: bool = False, rowsort: bool = False, stream=None):
        if k._h is None:
            raise ValueError("structure has not been scanned")
        self.kernel = k
        self.dtype = {"f64": capi.WK_F64, "f32": capi.WK_F32}[dtype]
        A = np.ascontiguousarray(A, dtype=np.float64)
        colinds = np.ascontiguousarray(colinds, dtype=np.int32)
        flags = ((capi.WK_ENGINE_KR if kr else 0) | (capi.WK_ENGINE_ADJOINT if adjoint else 0)
                 | (capi.WK_ENGINE_ROWSORT if rowsort else 0))
        h = C.c_void_p()
        check(lib().wk_engine_create(k._h, _host_ptr(k._rowptr), _host_ptr(colinds), _host_ptr(A),
                                     self.dtype, flags, _stream(stream), C.byref(h)))
        self._h = h
        self.nrows, self.ncols, self.nz = k.nrows, k.ncols, k.nz
        self.has_kr, self.has_adjoint = kr, adjoint

    @classmethod
    def from_device(cls, rowptr, colind, vals, ncols: int | None = None, variant: int = 1,
                    threshold: int = 0, max_tpr: int = 32, dtype: str = "f64", kr: bool = False,
                    stream=None) -> "engine":
        """Scan and process on the GPU: rowptr / colind (int32) and vals (float64)
        are device tensors of one CSR.  The layout is structure's (variant 1) or
        structure2's with `threshold` (variant 2) at group width 64, the same bit
        for bit as a host scan followed by engine(); it stays on the device
        (engine.kernel exposes its scalars and downloads its arrays)."""
        import torch
        for t, dt in ((rowptr, torch.int32), (colind, torch.int32), (vals, torch.float64)):
            if not (t.is_cuda and t.dtype == dt and t.is_contiguous()):
                raise ValueError("rowptr/colind must be contiguous int32 and vals float64 device tensors")
        self = cls.__new__(cls)
        self._h = None
        self.dtype = {"f64": capi.WK_F64, "f32": capi.WK_F32}[dtype]
        nrows = rowptr.numel() - 1
        nnz = colind.numel()
        ncols = nrows if ncols is None else int(ncols)
        if vals.numel() != nnz:
            raise ValueError("vals and colind differ in length")
        h = C.c_void_p()
        check(lib().wk_engine_create_device(int(variant), int(threshold), int(max_tpr), nrows, ncols,
                                            nnz, _dev_ptr(rowptr), _dev_ptr(colind), _dev_ptr(vals),
                                            self.dtype, capi.WK_ENGINE_KR if kr else 0,
                                            _stream(stream), C.byref(h)))
        self._h = h
        self.kernel = _EngineLayout(self)
        self.nrows, self.ncols, self.nz = nrows, ncols, nnz
        self.has_kr, self.has_adjoint = kr, False
        return self

    def download(self, which: int) -> np.ndarray:
        """One device array of the forward layout (WK_ROW_MAP .. WK_ROW_OFFSET_WARP,
        WK_ENGINE_VALUES / COLS / COLS_KR, the x windows WK_ENGINE_WIN_* and the
        node windows WK_ENGINE_ND_*); WK_REORDER_ROWS is the CSR-to-slot map."""
        dt = {capi.WK_ROW_MAP: np.int32, capi.WK_ROW_MAP_INV: np.int32,
              capi.WK_REORDER_ROWS: np.uint32, capi.WK_WARP_OFFSETS: np.int64,
              capi.WK_MAX_NZ_ROWS: np.uint32, capi.WK_THREADS_PER_ROW: np.uint32,
              capi.WK_ROW_OFFSET_WARP: np.uint32,
              capi.WK_ENGINE_VALUES: np.float64 if self.dtype == capi.WK_F64 else np.float32,
              capi.WK_ENGINE_COLS: np.int32, capi.WK_ENGINE_COLS_KR: np.int32,
              capi.WK_ENGINE_WIN_IDX: np.uint16, capi.WK_ENGINE_WIN_OFF: np.uint32,
              capi.WK_ENGINE_WIN_COLS: np.int32, capi.WK_ENGINE_WIN_COLS_KR: np.int32,
              capi.WK_ENGINE_ND_COLS: np.int32, capi.WK_ENGINE_ND_CNT: np.uint32,
              capi.WK_ENGINE_ND_OFF: np.uint32, capi.WK_ENGINE_ND_POS: np.uint32,
              capi.WK_ENGINE_ND_COLS_KR: np.int32}[which]
        n = C.c_int64()
        check(lib().wk_engine_download(self._h, which, None, C.byref(n)))
        out = np.empty(n.value, dt)
        if n.value:
            check(lib().wk_engine_download(self._h, which, _host_ptr(out), C.byref(n)))
        return out

    # -- launch configuration ---------------------------------------------
    def get_launch(self, op: int = capi.WK_OP_RUN, k: int = 1) -> capi.Launch:
        """The configuration used for (op, k) when no cfg is passed."""
        c = capi.Launch()
        check(lib().wk_engine_get_launch(self._h, op, k, C.byref(c)))
        return c

    def set_launch(self, cfg, op: int = capi.WK_OP_RUN, k: int = 1):
        c = cfg if isinstance(cfg, capi.Launch) else self._cfg(op, k, cfg)
        check(lib().wk_engine_set_launch(self._h, op, k, C.byref(c)))

    def _cfg(self, op, k, cfg):
        """None -> NULL (engine's own choice); dict -> overrides on top of it."""
        if cfg is None:
            return None
        if isinstance(cfg, capi.Launch):
            return cfg
        c = self.get_launch(op, k)
        for key, v in cfg.items():
            if not hasattr(c, key):
                raise KeyError(key)
            setattr(c, key, int(v))
        return c

    @staticmethod
    def _ref(c):
        return None if c is None else C.byref(c)

    def tune(self, x, y, op: int = capi.WK_OP_RUN, reps: int = 10, base=None, stream=None):
        """Time every candidate launch shape for (op, k = columns of x) and keep
        the fastest in the engine.  Returns (config dict, mean ms)."""
        k, ldx = self._k_ld(x)
        _, ldy = self._k_ld(y)
        best = capi.Launch()
        ms = C.c_float()
        b = self._cfg(op, k, base)
        check(lib().wk_tune(self._h, op, self._ref(b), _dev_ptr(x), _dev_ptr(y), k, ldx, ldy, reps,
                            _stream(stream), C.byref(best), C.byref(ms)))
        return best.as_dict(), ms.value

    # -- reference-shaped, blocking, returns seconds --------------------
    def _timed(self, which, blocksize, dx, dy, cfg, stream):
        op = capi.WK_OP_RUN_X if which == 1 else capi.WK_OP_RUN
        c = self._cfg(op, 1, cfg or {})
        if blocksize:
            c.waves_per_block = max(1, min(8, int(blocksize) // 64))
        sec = C.c_float()
        check(lib().wk_time_run(self._h, C.byref(c), which, _dev_ptr(dx), _dev_ptr(dy), 1, 1, 1,
                                _stream(stream), C.byref(sec)))
        return sec.value

    def run(self, nblocks, blocksize, dx, dy, cfg=None, stream=None) -> float:
        """engine::run<cache>(nblocks, blocksize, dx, dy): seconds of one launch.
        The grid is derived from the layout; nblocks is accepted for the
        reference's signature.  blocksize / 64 = waves per block (1..8)."""
        del nblocks
        return self._timed(0, blocksize, dx, dy, cfg, stream)

    def run_x(self, nblocks, blocksize, dx, dy, cfg=None, stream=None) -> float:
        """engine::run_x<cache>: the Kr form, y'[i] for x' = reorder_x(x)."""
        del nblocks
        return self._timed(1, blocksize, dx, dy, cfg, stream)

    def run_assembly(self, nblocks, blocksize, dy, cfg=None, stream=None) -> float:
        """engine::run_assembly<cache> (hpp:898-929): row sums of the padded
        values into dy (original row order); seconds of one launch."""
        del nblocks
        return self._timed(2, blocksize, None, dy, cfg, stream)

    # -- asynchronous forms ---------------------------------------------
    @staticmethod
    def _k_ld(t):
        if t is None or isinstance(t, int):
            return 1, 1
        if t.dim() == 1:
            return 1, 1
        return int(t.shape[1]), int(t.stride(0))

    def launch(self, x, y, cfg=None, stream=None):
        """y = A x (k right-hand sides if x, y are [n, k] row-major)."""
        k, ldx = self._k_ld(x)
        _, ldy = self._k_ld(y)
        c = self._cfg(capi.WK_OP_RUN, k, cfg)
        check(lib().wk_run(self._h, self._ref(c), _dev_ptr(x), _dev_ptr(y), k, ldx, ldy, _stream(stream)))

    def launch_x(self, x, y, cfg=None, stream=None):
        k, ldx = self._k_ld(x)
        _, ldy = self._k_ld(y)
        c = self._cfg(capi.WK_OP_RUN_X, k, cfg)
        check(lib().wk_run_x(self._h, self._ref(c), _dev_ptr(x), _dev_ptr(y), k, ldx, ldy, _stream(stream)))

    def launch_assembly(self, y, cfg=None, stream=None):
        """Row sums of the padded values into y (FP64, original row order)."""
        c = self._cfg(capi.WK_OP_RUN, 1, cfg)
        check(lib().wk_run_assembly(self._h, self._ref(c), _dev_ptr(y), _stream(stream)))

    def residual(self, b, x, r, cfg=None, stream=None):
        """r = b - A x in FP64."""
        k, ldx = self._k_ld(x)
        _, ldb = self._k_ld(b)
        _, ldr = self._k_ld(r)
        c = self._cfg(capi.WK_OP_RUN, k, cfg)
        check(lib().wk_residual(self._h, self._ref(c), _dev_ptr(b), ldb, _dev_ptr(x), _dev_ptr(r), k,
                                ldx, ldr, _stream(stream)))

    def run_adjoint(self, dy, dx, cfg=None, stream=None):
        """dx = A^T dy through the transposed layout."""
        k, lddy = self._k_ld(dy)
        _, lddx = self._k_ld(dx)
        c = self._cfg(capi.WK_OP_ADJOINT, k, cfg)
        check(lib().wk_run_adjoint(self._h, self._ref(c), _dev_ptr(dy), _dev_ptr(dx), k, lddy, lddx,
                                   _stream(stream)))

    def dvals(self, dy, x, out, stream=None):
        """out[j] = sum_r dy[row_j, r] * x[col_j, r] (CSR order)."""
        k, lddy = self._k_ld(dy)
        _, ldx = self._k_ld(x)
        check(lib().wk_dvals(self._h, _dev_ptr(dy), lddy, _dev_ptr(x), ldx, k, _dev_ptr(out), _stream(stream)))

    def backward(self, dy, x, dx, dvals, cfg=None, stream=None):
        """dx = A^T dy and dvals in one call (FP64 engines; one launch for a single
        right-hand side on short groups)."""
        k, lddy = self._k_ld(dy)
        _, ldx = self._k_ld(x)
        _, lddx = self._k_ld(dx)
        c = self._cfg(capi.WK_OP_ADJOINT, k, cfg)
        check(lib().wk_backward(self._h, self._ref(c), _dev_ptr(dy), lddy, _dev_ptr(x), ldx, _dev_ptr(dx),
                                lddx, _dev_ptr(dvals), k, _stream(stream)))

    def update_values(self, vals_dev, stream=None):
        """GPU-side process of new CSR values (same pattern)."""
        check(lib().wk_engine_update_values(self._h, _dev_ptr(vals_dev), _stream(stream)))

    @property
    def device_colinds(self) -> np.ndarray:
        out = np.empty(self.kernel.allocate_nz, np.int32)
        check(lib().wk_engine_get_colinds(self._h, _host_ptr(out)))
        return out

    @device_colinds.setter
    def device_colinds(self, pcols):
        """eng.device_colinds = reorder_cols (wpk1.cu:175): host array or device tensor."""
        if isinstance(pcols, np.ndarray):
            pcols = np.ascontiguousarray(pcols, dtype=np.int32)
            if pcols.size != self.kernel.allocate_nz:
                raise ValueError("padded column array has the wrong length")
            check(lib().wk_engine_set_colinds(self._h, _host_ptr(pcols), 0, _stream(None)))
        else:
            check(lib().wk_engine_set_colinds(self._h, _dev_ptr(pcols), 1, _stream(None)))

    def x_window_entries(self) -> int:
        """Entries over all x windows the forward launches use (0: none, or the
        columns were replaced by device_colinds)."""
        n = C.c_int64()
        check(lib().wk_engine_x_windows(self._h, C.byref(n)))
        return n.value

    def gather_ceiling(self, x, y, cfg=None, stream=None):
        """The forward direct kernel without its values (every slot 1.0): the
        column stream and the x gathers alone, timed as the ceiling of layouts
        whose gathers miss the caches.  y = per-row sums of x over the slots."""
        c = self._cfg(capi.WK_OP_RUN, 1, cfg)
        check(lib().wk_engine_gather_ceiling(self._h, self._ref(c), _dev_ptr(x), _dev_ptr(y), _stream(stream)))

    # -- verification (host) ------------------------------------------------
    def verify(self, orig, comp) -> bool:
        """engine::verify (hpp:932-943): relative 1e-5 where |orig| > 1e-14,
        absolute 1e-14 elsewhere; stops at the first failure."""
        return _verify(np.asarray(orig), np.asarray(comp), None)

    def verify_x(self, orig, comp) -> bool:
        """engine::verify_x (hpp:946-957): compares orig[row_map[i]] with comp[i]."""
        return _verify(np.asarray(orig), np.asarray(comp), self.kernel.row_map)

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().wk_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _EngineLayout:
    """The layout of an engine scanned on the device: structure's scalars, and
    its arrays downloaded on first use."""

    _ARRAYS = {"row_map": capi.WK_ROW_MAP, "row_map_inv": capi.WK_ROW_MAP_INV,
               "reorder_rows": capi.WK_REORDER_ROWS, "warp_offsets": capi.WK_WARP_OFFSETS,
               "max_nz_rows": capi.WK_MAX_NZ_ROWS, "threads_per_row": capi.WK_THREADS_PER_ROW,
               "row_offset_warp": capi.WK_ROW_OFFSET_WARP}

    def __init__(self, eng: "engine"):
        self._eng = eng
        info = capi.LayoutInfo()
        check(lib().wk_engine_get_info(eng._h, C.byref(info)))
        self.info = info
        self.variant, self.group_width = info.variant, info.group_width
        self.nrows, self.ncols, self.nznrows = info.nrows, info.ncols, info.nznrows
        self.nz, self.allocate_nz, self.nwarps = info.nz, info.allocate_nz, info.nwarps
        self.min_nz, self.max_nz = info.min_nz, info.max_nz

    def __getattr__(self, name):
        if name in _EngineLayout._ARRAYS:
            a = self._eng.download(_EngineLayout._ARRAYS[name])
            setattr(self, name, a)
            return a
        raise AttributeError(name)


def _verify(orig, comp, row_map) -> bool:
    o = orig if row_map is None else orig[row_map]
    c = comp
    small = np.abs(o) <= 1e-14
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.abs((o - c) / o)
    bad = (small & (np.abs(o - c) > 1e-14)) | (~small & (rel > 1e-5))
    if bad.any():
        i = int(np.argmax(bad))
        print(f"verify: {o[i]}\t{c[i]}\t{i}")
        return False
    return True


class batch:
    """Many engines in one launch (wk_batch_*).

        b = batch([eng0, eng1, ...]);  b.bind(xs, ys);  b.run()

    xs / ys: lists of device tensors, [ncols_i] / [nrows_i] or [n, k].
    waves_per_cu = 0: one wave per group (default); > 0: that many persistent
    waves per compute unit over equal-cost ranges.  stage_lds 5 (default):
    x windows with packed positions where a matrix has them."""

    def __init__(self, engines, waves_per_cu: int = 0, stage_lds: int = 5):
        self.engines = list(engines)  # keep them alive
        arr = (C.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        h = C.c_void_p()
        check(lib().wk_batch_create(arr, len(self.engines), int(waves_per_cu), int(stage_lds), C.byref(h)))
        self._h = h

    def bind(self, xs, ys):
        n = len(self.engines)
        if len(xs) != n or len(ys) != n:
            raise ValueError("one x and one y per engine")
        k, _ = engine._k_ld(xs[0])
        xp = (C.c_void_p * n)(*[_dev_ptr(t) for t in xs])
        yp = (C.c_void_p * n)(*[_dev_ptr(t) for t in ys])
        ldx = (C.c_int64 * n)(*[engine._k_ld(t)[1] for t in xs])
        ldy = (C.c_int64 * n)(*[engine._k_ld(t)[1] for t in ys])
        check(lib().wk_batch_bind(self._h, xp, yp, k, ldx, ldy))
        self._bound = (list(xs), list(ys))

    def run(self, stream=None):
        check(lib().wk_batch_run(self._h, _stream(stream)))

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().wk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class graph:
    """HIP graph of library launches captured on a (non-null) stream:

        with graph.capture(stream) as g:
            eng.launch(x, y, stream=stream); b.run(stream)
        g.launch(stream)
    """

    def __init__(self):
        self._h = None

    @classmethod
    def capture(cls, stream):
        g = cls()
        g._stream = stream
        return g

    def __enter__(self):
        check(lib().wk_graph_begin(_stream(self._stream)))
        return self

    def __exit__(self, et, ev, tb):
        h = C.c_void_p()
        rc = lib().wk_graph_end(_stream(self._stream), C.byref(h))
        if et is None:
            check(rc)
            self._h = h
        return False

    def launch(self, stream=None):
        check(lib().wk_graph_launch(self._h, _stream(stream if stream is not None else self._stream)))

    def close(self):
        if self._h is not None:
            lib().wk_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
