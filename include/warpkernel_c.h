/*
 * warpkernel_c.h -- C ABI of the MI355X (gfx950) WarpKernel SpMV library.
 *
 * The reference exposes one C++ class API, namespace warpkernel in
 * warpkernel.hpp: structure (WPK1 layout), structure2 (WPK2 layout) and
 * engine<V,I> (device arrays + kernel launches).  Every entry point below names
 * the reference member it replaces.  include/warpkernel.hpp rebuilds that class
 * API on top of these functions; Python binds them with ctypes.
 *
 * Conventions
 *   - Every function returns WK_OK (0) or a WK_ERR_* code; wk_last_error()
 *     returns the message of the last failure on the calling thread.  Nothing
 *     aborts the process (the reference asserts, warpkernel.hpp:346, 359, 614,
 *     625, 669, 699).
 *   - Plain pointers and sizes only.  "host" pointers are CPU memory, "dev"
 *     pointers are device memory of the current HIP device.  `stream` is a
 *     hipStream_t passed as void* (NULL = the null stream).
 *   - Indices are int32 (the reference's IndexType int); nnz < 2^31,
 *     allocate_nz < 2^32.  warp_offsets are int64.
 *   - Multi-RHS vectors are row-major: x is [ncols][ldx], y is [nrows][ldy],
 *     right-hand side j at column j.  k = 1 with ldx = ldy = 1 is a plain vector.
 *   - No global mutable state: different engines (or the same engine with
 *     different output buffers) may be used from different streams at once.
 */
#ifndef WARPKERNEL_C_H
#define WARPKERNEL_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WK_ABI_VERSION 1

/* status codes */
#define WK_OK 0
#define WK_ERR_ARG 1      /* bad argument (NULL handle, unsupported option) */
#define WK_ERR_CSR 2      /* malformed CSR input */
#define WK_ERR_LIMIT 3    /* size exceeds a documented limit */
#define WK_ERR_HIP 4      /* a HIP runtime call failed */
#define WK_ERR_STATE 5    /* call not valid in the handle's current state */
#define WK_ERR_NOMEM 6    /* host allocation failed */

/* value / vector element types */
#define WK_F64 0          /* FP64 values, x and y */
#define WK_F32 1          /* FP32 values and x, FP64 y (see accumulate) */

/* accumulation of an FP32 engine */
#define WK_ACC_TWOSUM 0   /* FP32 TwoProd + TwoSum (compensated), FP64 result */
#define WK_ACC_F64 1      /* exact FP32 products summed by FP64 fma */

/* layout arrays (wk_layout_array) */
#define WK_ROW_MAP 0         /* int32 [nrows]      y[row_map[i]] = y'[i]          hpp:276 */
#define WK_ROW_MAP_INV 1     /* int32 [nrows]                                     hpp:277 */
#define WK_REORDER_ROWS 2    /* uint32 [nz]        new_values[reorder_rows[i]]=A[i] hpp:278 */
#define WK_WARP_OFFSETS 3    /* int64 [nwarps+1]                                  hpp:280 */
#define WK_MAX_NZ_ROWS 4     /* uint32 [nwarps]                                   hpp:281 */
#define WK_THREADS_PER_ROW 5 /* uint32 [nwarps]    structure2 only                hpp:562 */
#define WK_ROW_OFFSET_WARP 6 /* uint32 [nwarps+1]  structure2 only                hpp:561 */
#define WK_REORDER_A_ROWS 7  /* int32 [nz] after wk_layout_reorder_columns_rowsort  hpp:279 */
#define WK_ROWPTR 8          /* int32 [nrows+1]    the scanned CSR's row offsets */

/* wk_engine_create flags */
#define WK_ENGINE_KR 1u       /* also build the Kr (run_x) column indices */
#define WK_ENGINE_ADJOINT 2u  /* also build the transposed layout (run_adjoint, dvals) */
#define WK_ENGINE_ROWSORT 4u  /* each row's entries in permuted-column order (square
                                 matrices; see wk_layout_reorder_columns_rowsort) */

const char *wk_last_error(void);
int wk_abi_version(void);

/* ------------------------------------------------------------------ layout */

typedef struct wk_layout wk_layout;

/* Scalars of a scan: structure's public fields (warpkernel.hpp:282-286) and
 * structure2's (warpkernel.hpp:561-565). */
typedef struct wk_layout_info {
  int32_t variant;     /* 1 = structure (WPK1), 2 = structure2 (WPK2) */
  int32_t group_width; /* W = WARP_SIZE: 64 (one gfx950 wave) or 32 */
  int32_t threshold;   /* nmax_per_thread (variant 2) */
  int32_t max_tpr;     /* thmax_per_row, the literal 32 at hpp:643 (variant 2) */
  int32_t nrows;
  int32_t ncols;
  int32_t nznrows;
  int32_t nwarps;      /* number of groups */
  int64_t nz;
  int64_t allocate_nz;
  uint32_t min_nz;     /* variant 2 */
  uint32_t max_nz;     /* variant 2 */
} wk_layout_info;

/* structure::scan (warpkernel.hpp:309-365).  Only rowptr shapes the layout;
 * colind (may be NULL) is range-checked against ncols when given. */
int wk_scan1(int32_t group_width, int32_t nrows, int32_t ncols, int64_t nnz,
             const int32_t *rowptr_host, const int32_t *colind_host, wk_layout **out);

/* structure2::scan(..., threshold) (warpkernel.hpp:578-736).  max_tpr is a
 * power of two in [2, group_width]; 32 reproduces the reference. */
int wk_scan2(int32_t group_width, int32_t threshold, int32_t max_tpr, int32_t nrows,
             int32_t ncols, int64_t nnz, const int32_t *rowptr_host,
             const int32_t *colind_host, wk_layout **out);

int wk_layout_get_info(const wk_layout *layout, wk_layout_info *info);

/* Borrowed pointer to one of the arrays above (valid until destroy). */
int wk_layout_array(const wk_layout *layout, int32_t which, const void **data, int64_t *count);

/* structure::reorder (warpkernel.hpp:379-398), host side: zero-filled padded
 * arrays of allocate_nz entries, pvals[reorder_rows[i]] = vals[i]. */
int wk_layout_process_host(const wk_layout *layout, const double *vals_host,
                           const int32_t *colind_host, double *pvals_host, int32_t *pcols_host);

/* structure::reorder_columns_coalesced (warpkernel.hpp:446-457). */
int wk_layout_reorder_columns_coalesced(const wk_layout *layout, const int32_t *colind_host,
                                        int32_t *pcols_kr_host);

/* structure::reorder_x (warpkernel.hpp:517-530): xout[row_map_inv[r]] = x[r]. */
int wk_layout_reorder_x(const wk_layout *layout, const double *x_host, double *xout_host);

/* structure::reorder_columns_rowsort (warpkernel.hpp:414-444), for the Kr form
 * of a square matrix: within each row the entries are ordered by permuted
 * column row_map_inv[col] ascending (equal columns keep CSR order, where the
 * reference's std::sort leaves them unspecified).  cols_out_host[n] (nz
 * entries, CSR order) receives the permuted columns; the layout keeps
 * reorder_A_rows[n] = the CSR index of the entry now at n (WK_REORDER_A_ROWS).
 * An engine made from vals[reorder_A_rows] and these columns runs run_x in
 * that order (optimize.cu:645-704); WK_ENGINE_ROWSORT builds the same engine
 * from the original CSR. */
int wk_layout_reorder_columns_rowsort(wk_layout *layout, const int32_t *rowptr_host,
                                      const int32_t *colind_host, int32_t *cols_out_host);

/* structure::reorder_values_coalesced_rowsort (warpkernel.hpp:505-515): zero-
 * filled padded values, pvals[reorder_rows[i]] = vals[reorder_A_rows[i]]. */
int wk_layout_reorder_values_rowsort(const wk_layout *layout, const double *vals_host,
                                     double *pvals_host);

/* A layout from its arrays: the state structure::store / storemap /
 * storereorder and structure2::store_param leave in a structure's public
 * fields (warpkernel.hpp:288-307, 567-576), together with the fields the
 * caller sets (warp_offsets, max_nz_rows, nwarps, nznrows, allocate_nz) and the
 * CSR row offsets (nrows + 1).  row_map_inv is derived; min_nz / max_nz
 * (variant 2) are computed from rowptr.  Every invariant the kernels index by
 * is checked, as for a loaded layout file (WK_ERR_CSR).  threads_per_row and
 * row_offset_warp are read for variant 2 only. */
int wk_layout_from_arrays(int32_t variant, int32_t group_width, int32_t threshold, int32_t max_tpr,
                          int32_t nrows, int32_t ncols, int32_t nznrows, int32_t nwarps, int64_t nz,
                          int64_t allocate_nz, const int32_t *rowptr_host, const int32_t *row_map_host,
                          const uint32_t *reorder_rows_host, const int64_t *warp_offsets_host,
                          const uint32_t *max_nz_rows_host, const uint32_t *threads_per_row_host,
                          const uint32_t *row_offset_warp_host, wk_layout **out);

void wk_layout_destroy(wk_layout *layout);

/* thrust::scatter as the reference's reorder_columns_coalesced_GPU and
 * reorder_values_coalesced_GPU use it (warpkernel.hpp:459-502): dst[index[i]]
 * = src[i] for i < n, device arrays of 4- or 8-byte elements, asynchronous on
 * `stream`.  Indices at or beyond dst_count are skipped (the reference does
 * not check them). */
int wk_scatter(const void *src_dev, const uint32_t *index_dev, void *dst_dev, int64_t n,
               int64_t dst_count, int32_t elem_bytes, void *stream);

/* ------------------------------------------------------------------- files */

/* A host CSR matrix read from a file (int32 indices, FP64 values). */
typedef struct wk_csr wk_csr;

/* cusp::io::read_matrix_market_file + the COO -> CSR conversion of the drivers
 * (wpk1.cu:43-48, wpk2.cu:55-59).  Coordinate format; field real, double,
 * integer or pattern (entries 1); symmetry general, symmetric, skew-symmetric
 * (mirrored entries negated) or hermitian (real: as symmetric).  Entries are
 * ordered by (row, column), stably; duplicates are kept. */
int wk_mm_read(const char *path, wk_csr **out);
/* Borrowed views (valid until wk_csr_destroy); any output may be NULL. */
int wk_csr_get(const wk_csr *csr, int32_t *nrows, int32_t *ncols, int64_t *nnz,
               const int32_t **rowptr, const int32_t **colind, const double **vals);
void wk_csr_destroy(wk_csr *csr);
/* Writes "coordinate real general", entries in CSR order, each value as the
 * shortest decimal that reads back to the same double. */
int wk_mm_write(const char *path, int32_t nrows, int32_t ncols, int64_t nnz,
                const int32_t *rowptr_host, const int32_t *colind_host, const double *vals_host);

/* structure::save / load (serialize.cu:135-139) over the fields of
 * structure::serialize (warpkernel.hpp:252-270) plus structure2's: a versioned
 * little-endian binary file with a checksum (the reference's boost archive is
 * not reproduced).  load checks the layout's invariants (WK_ERR_CSR on a
 * damaged file); driver_name (may be NULL) receives the stored name, cut to
 * name_capacity - 1 bytes. */
int wk_layout_save(const wk_layout *layout, const char *path, const char *driver_name);
int wk_layout_load(const char *path, wk_layout **out, char *driver_name, int64_t name_capacity);

/* ------------------------------------------------------------------ engine */

typedef struct wk_engine wk_engine;

/* engine(structure&, A, colinds) / engine(structure2&, A, colinds)
 * (warpkernel.hpp:794-809) -> structure::process (hpp:532-548) and
 * structure2::process (hpp:749-769): the padded arrays and the layout metadata
 * are built and copied to the current device.  vals_host is FP64 CSR (cast to
 * FP32 when values_dtype is WK_F32); rowptr_host may be NULL (the layout's
 * own offsets, WK_ROWPTR).  The engine copies what it needs; the
 * layout may be destroyed afterwards. */
int wk_engine_create(const wk_layout *layout, const int32_t *rowptr_host,
                     const int32_t *colind_host, const double *vals_host, int32_t values_dtype,
                     uint32_t flags, void *stream, wk_engine **out);

void wk_engine_destroy(wk_engine *engine);

int wk_engine_get_info(const wk_engine *engine, wk_layout_info *info);

/* GPU-side process: replace the values, keeping the pattern.  vals_dev is CSR
 * order on the device (FP64, or FP32 for a WK_F32 engine), scattered into the
 * padded layout(s) on `stream` (the thrust::scatter of warpkernel.hpp:459-502,
 * without atomics). */
int wk_engine_update_values(wk_engine *engine, const void *vals_dev, void *stream);

/* The reference's drivers overwrite the public engine::device_colinds with the
 * Kr columns and restore them (wpk1.cu:160-197).  pcols has allocate_nz
 * entries in padded order; where = 0 host, 1 device.  Every entry must lie in
 * [0, ncols) (checked before the copy; WK_ERR_ARG otherwise, nothing changed).
 * The new columns feed wk_run, wk_residual and run_assembly's layout, and
 * wk_run_x on engines without their own Kr columns (WK_ENGINE_KR engines
 * keep running run_x on those).  The engine's x windows are built from its
 * scanned columns: they are used while the columns equal those (an exact check
 * on the device after each call), so restoring the original columns restores
 * the windowed kernels.  Engines built with WK_ENGINE_ADJOINT refuse the call
 * (WK_ERR_STATE): their transposed layout and dvals follow the scanned
 * pattern.  Allocates a small check buffer and synchronises `stream`. */
int wk_engine_set_colinds(wk_engine *engine, const int32_t *pcols, int32_t where, void *stream);
int wk_engine_get_colinds(const wk_engine *engine, int32_t *pcols_host);

/* wk_engine_set_colinds with device columns (where = 1), allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate (loops
 * that swap an engine's columns while serving, harnesses that forbid
 * hipMalloc).  The column checks write their flags into a fixed device array
 * of the library, 1024 int32 (4 KB), one per device, held until the process
 * ends; the checks never need more flags, whatever allocate_nz, so no input
 * exceeds it.  Every call on one device shares that array: calls that overlap
 * on one device race.  Same results, errors and window handling as
 * wk_engine_set_colinds; synchronises `stream`. */
int wk_engine_set_colinds_noalloc(wk_engine *engine, const int32_t *pcols_dev, void *stream);

/* Entries over all x windows the engine's forward launches use: 0 when it has
 * none, or while its columns differ from the scanned ones (see
 * wk_engine_set_colinds).  No reference counterpart. */
int wk_engine_x_windows(const wk_engine *engine, int64_t *entries);

/* structure::scan + process (warpkernel.hpp:309-365, 532-548) or
 * structure2::scan(.., threshold) + process (hpp:578-736, 749-769) with the
 * CSR already on the device: the sort by row length, the grouping and the
 * padded arrays are computed on the GPU (reorder.cu:138-166 moves only the
 * scatter there).  The result equals wk_scan1 / wk_scan2 at group width 64
 * followed by wk_engine_create, bit for bit.  variant 1 or 2 (threshold and
 * max_tpr as wk_scan2; ignored for 1); flags: WK_ENGINE_KR only.  The CSR is
 * checked on the device (WK_ERR_CSR); rows longer than 8191 entries give
 * WK_ERR_LIMIT.  Synchronises `stream`. */
int wk_engine_create_device(int32_t variant, int32_t threshold, int32_t max_tpr, int32_t nrows,
                            int32_t ncols, int64_t nnz, const int32_t *rowptr_dev,
                            const int32_t *colind_dev, const double *vals_dev,
                            int32_t values_dtype, uint32_t flags, void *stream, wk_engine **out);

/* engine arrays (wk_engine_download), besides WK_ROW_MAP .. WK_ROW_OFFSET_WARP
 * (WK_REORDER_ROWS gives the CSR-to-slot map; WK_ROW_MAP_INV exists for
 * engines built by wk_engine_create_device) */
#define WK_ENGINE_VALUES 16   /* double or float [allocate_nz]  device_values   hpp:782 */
#define WK_ENGINE_COLS 17     /* int32 [allocate_nz]            device_colinds  hpp:783 */
#define WK_ENGINE_COLS_KR 18  /* int32 [allocate_nz]            Kr columns (WK_ENGINE_KR) */
#define WK_ENGINE_WIN_IDX 19  /* uint16 [allocate_nz]  x-window position of each slot */
#define WK_ENGINE_WIN_OFF 20  /* uint32 [nwarps+1]     window offsets per group */
#define WK_ENGINE_WIN_COLS 21 /* int32 [WIN_OFF[nwarps]] window columns, ascending per group */
#define WK_ENGINE_WIN_COLS_KR 22 /* int32, the window columns in Kr numbering */
/* node windows (stage_lds 7), present where every triple of a lane's steps
 * reads columns 3n, 3n+1, 3n+2 or one window position three times: */
#define WK_ENGINE_ND_COLS 23  /* int32 [WIN_OFF[nwarps]] from WIN_OFF[g]: group g's ND_CNT[g]
                                 entries, e >= 0 for x[e..e+2], e < 0 for x[-e-1] thrice */
#define WK_ENGINE_ND_CNT 24   /* uint32 [nwarps] entries per group */
#define WK_ENGINE_ND_OFF 25   /* uint32 [nwarps+1] first position word of each group */
#define WK_ENGINE_ND_POS 26   /* uint32 [ND_OFF[nwarps] * 64] word w of lane l of group g at
                                 (ND_OFF[g] + w) * 64 + l: the 10-bit entries of steps
                                 9w+3f .. 9w+3f+2 in bits 10f .. 10f+9 */
#define WK_ENGINE_ND_COLS_KR 27 /* int32 [3 * WIN_OFF[nwarps]] the entries in Kr numbering
                                   (WK_ENGINE_KR engines): three columns each, entry j of
                                   group g at 3 * (WIN_OFF[g] + j) */

/* Copies one of the engine's device arrays of the forward layout to the host.
 * host_dst NULL: *count receives the array's length (0 for an array the engine
 * lacks); else *count must equal it. */
int wk_engine_download(const wk_engine *engine, int32_t which, void *host_dst, int64_t *count);

/* Launch options; wk_launch_default fills the built-in choice. */
typedef struct wk_launch {
  int32_t waves_per_block; /* 1, 2, 4 or 8 (the drivers' warps/block sweep).
                              With stage_lds 3 and a k-tile above 1, the
                              block's waves share one group's x window, each
                              summing an equal share of the k-tile's
                              right-hand sides (at most one wave per 16
                              bytes of a window entry; FP32 engines with
                              stage_lds 4: at most one per right-hand side) */
  int32_t unroll;          /* slots in flight per lane: 4, 8 or 16 (k_tile 1;
                              and 16 with stage_lds 4 at k_tiles 4 and 8) */
  int32_t k_tile;          /* right-hand sides per pass: 1, 2, 4, 8, 16 or 32.
                              k is run as whole k-tiles, then the rest in
                              halving tiles; with register-fed x windows
                              (stage_lds 4 or 7, unroll 8, k-tiles 4 and 8,
                              waves_per_block 1 or 2; run and the residual)
                              a rest of 3 or more columns runs as one masked
                              tile of the next power of two (its missing
                              columns of x read as 0, of y not written), so
                              k_tile may exceed k */
  int32_t xcd_remap;       /* 1: contiguous group ranges per XCD */
  int32_t nontemporal;     /* 1: streaming (nontemporal) loads for the values;
                              2: for the column indices too (direct loads at
                              k_tile 1, unroll 8, rows_per_lane 1 or 2;
                              elsewhere as 1), which leaves more of L2 to
                              scattered x gathers; 0: plain loads (direct
                              loads at k_tile 1, unroll 8, rows_per_lane 1
                              only: other shapes are refused) */
  int32_t accumulate;      /* WK_ACC_* for FP32 engines */
  int32_t rows_per_lane;   /* 1, 2 or 4 groups per wave (k_tile 1, group width 64) */
  int32_t blocks_per_cu;   /* 0: no cap; else resident blocks per CU are capped
                              by reserving LDS (occupancy control) */
  int32_t stage_lds;       /* 1: copy value/index batches through LDS with
                              16-byte loads; 2: the same with the next batch's
                              loads issued before the current batch's FMAs;
                              3: as 2, and x comes from an LDS window of the
                              group's distinct columns with 16-bit positions
                              per slot (engines whose groups have at most 2048
                              distinct columns; otherwise as 2).  Applies to
                              group width 64, rows_per_lane 1, nontemporal 1,
                              and k_tile 1 with unroll 8/16; stage 3 also to
                              k_tiles 2, 4 and 8 (unroll 8), all k-tiles in
                              one launch; 4: as 3, with values and window
                              positions loaded straight into registers
                              instead of through LDS; 5 (k_tile 1): as 3,
                              with window positions packed 10 bits each
                              where every window has at most 1024 entries
                              (unroll 8: 12-step batches, 16: 24-step),
                              otherwise as 3; 6 (several right-hand
                              sides, engines whose windows have packed
                              positions and whose groups have at most 84
                              steps): each wave loads its group's values
                              and positions into registers once and makes
                              one pass per tile of 4 (FP64) or 8 (FP32)
                              right-hand sides through a window of that
                              many values per column; the rest of k (fewer
                              than one tile), and engines it does not fit,
                              run as 4; 7 (k_tile 1, engines with node
                              windows; run, run_x and the residual): as 5, with a
                              window of node entries (three x values at
                              columns 3n .. 3n+2, or one column thrice for
                              pads; run_x: three Kr columns each) and one
                              10-bit entry index per three steps (unroll
                              16: 27-step batches, 8: 18-step); otherwise
                              as 5 */
} wk_launch;

void wk_launch_default(wk_launch *cfg);

/* engine::run (warpkernel.hpp:811-860): y[row_map[i]] = sum over group slots.
 * Asynchronous on `stream`.  Rows without nonzeros are written as 0 (the
 * reference leaves them untouched and its drivers pre-zero y, wpk1.cu:146).
 * x: [ncols][ldx] (FP64, or FP32 for a WK_F32 engine); y: FP64 [nrows][ldy].
 * cfg may be NULL. */
int wk_run(wk_engine *engine, const wk_launch *cfg, const void *x_dev, double *y_dev,
           int32_t k, int64_t ldx, int64_t ldy, void *stream);

/* engine::run_x (warpkernel.hpp:862-896), the Kr form: y'[i] = sum, with x and
 * the column indices in the permuted order (reorder_x, reorder_columns_coalesced).
 * Uses the Kr columns built by WK_ENGINE_KR (or set by the caller through
 * wk_engine_set_colinds when the engine has none). */
int wk_run_x(wk_engine *engine, const wk_launch *cfg, const void *x_dev, double *y_dev,
             int32_t k, int64_t ldx, int64_t ldy, void *stream);

/* r = b - A x in FP64 (FP64 engines).  No reference counterpart. */
int wk_residual(wk_engine *engine, const wk_launch *cfg, const double *b_dev, int64_t ldb,
                const double *x_dev, double *r_dev, int32_t k, int64_t ldx, int64_t ldr,
                void *stream);

/* The backward of y = A(vals) x in one call: dx = A^T dy (as wk_run_adjoint)
 * and dvals (as wk_dvals) on `stream`, FP64 engines, bit-identical to those two
 * calls.  For one contiguous right-hand side and short groups (mean length
 * below 40) both run in one launch whose blocks interleave the two kinds of
 * work; otherwise two launches, the adjoint with cfg.  No reference
 * counterpart (the reference has no adjoint). */
int wk_backward(wk_engine *engine, const wk_launch *cfg, const void *dy_dev, int64_t lddy,
                const void *x_dev, int64_t ldx, double *dx_dev, int64_t lddx, double *dvals_dev,
                int32_t k, void *stream);

/* engine::run_assembly<cache> (warpkernel.hpp:898-929, kernel hpp:78-100):
 * y[row_map[p]] = the padded values of sorted row p summed in slot order
 * (first value, then one add per slot, pads included) in the engine's value
 * type; y is FP64 [nrows], rows without entries get 0.  WPK2 layouts (no
 * reference kernel) sum each lane's chunk and combine the lanes of a row in
 * warpKernel2's pairwise order.  cfg (NULL = the engine's run configuration)
 * supplies waves_per_block and xcd_remap. */
int wk_run_assembly(wk_engine *engine, const wk_launch *cfg, double *y_dev, void *stream);

/* dx = A^T dy through the transposed WPK layout (WK_ENGINE_ADJOINT); no
 * atomics.  dy: [nrows][lddy], dx: [ncols][lddx].  No reference counterpart. */
int wk_run_adjoint(wk_engine *engine, const wk_launch *cfg, const void *dy_dev, double *dx_dev,
                   int32_t k, int64_t lddy, int64_t lddx, void *stream);

/* dvals[j] = sum_r dy[row_j][r] * x[col_j][r] in CSR order (FP64): the
 * gradient of y = A(vals) x with respect to the values.  No reference
 * counterpart. */
int wk_dvals(wk_engine *engine, const double *dy_dev, int64_t lddy, const double *x_dev,
             int64_t ldx, int32_t k, double *dvals_dev, void *stream);

/* ---------------------------------------------------------------- tuning */

/* Operations an engine launches; the engine keeps one launch configuration
 * per (operation, k), used whenever cfg is NULL. */
#define WK_OP_RUN 0
#define WK_OP_RUN_X 1
#define WK_OP_ADJOINT 2

/* Autotuner (no reference counterpart; the reference sweeps warps/block and
 * the WPK2 threshold in its drivers, wpk1.cu:138-257, wpk2.cu:156-302).  Times
 * every candidate launch configuration of operation `op` on this engine with
 * the given device vectors (k right-hand sides, reps launches each, HIP
 * events on `stream`), times the three fastest twice more (each keeps its best
 * mean) beside the incumbent (base, or the engine's launch for (op, k)), and
 * stores the fastest in the engine for (op, k), or the incumbent unless the
 * fastest beats it by more than 1%; returns it with its mean time.  Only
 * performance knobs vary; base (may be
 * NULL) fixes the numerics (accumulate).  Results are identical for every
 * candidate.  y_dev is overwritten. */
int wk_tune(wk_engine *engine, int32_t op, const wk_launch *base, const void *x_dev,
            double *y_dev, int32_t k, int64_t ldx, int64_t ldy, int32_t reps, void *stream,
            wk_launch *best, float *best_ms);

/* Read / replace the configuration the engine uses for (op, k) when cfg is
 * NULL (before tuning: a choice from the layout's mean row length). */
int wk_engine_get_launch(const wk_engine *engine, int32_t op, int32_t k, wk_launch *cfg);
int wk_engine_set_launch(wk_engine *engine, int32_t op, int32_t k, const wk_launch *cfg);

/* ------------------------------------------------------ batches and graphs */

/* Many matrices in one launch (no reference counterpart; the reference
 * launches one kernel per matrix and per call).  The groups of all engines
 * (and the chunks of their empty rows) form one work list, longest groups
 * first; no atomics.  Engines must share the value type, use group width 64
 * and outlive the batch. */
typedef struct wk_batch wk_batch;

/* waves_per_cu: 0 (the default) runs one wave per work item and lets the
 * dispatcher balance them; > 0 (at most 1024) runs that many persistent waves
 * per compute unit, each over a contiguous range of equal cost.  stage_lds: as
 * in wk_launch, for batches with one right-hand side (0 direct loads, 1
 * staged, 2 staged with the next batch in flight, 3 as 2 plus x windows, 5 as
 * 3 with packed window positions where a matrix has them, or node windows
 * (wk_launch stage_lds 7, 27-step batches) where it has those and groups of
 * at least 40 steps on average, one wave per block); staging applies to the matrices whose mean group length is at
 * least 24 and windows to those with windows whose entries serve at least 4.5
 * slots each on average (3 with packed positions at stage_lds 5); the others
 * use direct loads.  With several
 * right-hand sides, stage_lds 5 gives the matrices whose window entries serve
 * at least 3 slots windows of k-tile values per column (values and positions
 * straight into registers; at k-tile 8 with waves_per_cu 0, two waves per
 * item share its window, each summing four right-hand sides), the others
 * direct loads; every other stage_lds means direct loads. */
int wk_batch_create(wk_engine *const *engines, int32_t count, int32_t waves_per_cu,
                    int32_t stage_lds, wk_batch **out);

/* Binds the vectors of every matrix: x_devs[i] ([ncols_i][ldx[i]]) and y_devs[i]
 * ([nrows_i][ldy[i]]) for k right-hand sides (1..32).  Host arrays of device
 * pointers; copied to the device here (synchronously), so bind outside graph
 * capture and re-bind when the vectors move.  The batch's device arrays are
 * allocated once, by wk_batch_create, and rewritten in place by every bind: a
 * graph that captured wk_batch_run keeps working after a re-bind with the same
 * k (it then reads the new vectors), but must be captured again after a bind
 * that changes k (the launch shape and its arguments depend on k). */
int wk_batch_bind(wk_batch *batch, const void *const *x_devs, double *const *y_devs, int32_t k,
                  const int64_t *ldx, const int64_t *ldy);

/* y_i = A_i x_i for every matrix of the batch, one kernel launch on `stream`. */
int wk_batch_run(wk_batch *batch, void *stream);

void wk_batch_destroy(wk_batch *batch);

/* HIP graph of any sequence of launches of this library (wk_run, wk_run_x,
 * wk_residual, wk_run_adjoint, wk_dvals, wk_engine_update_values,
 * wk_batch_run) on one stream: begin capture, issue the calls, end.  None of
 * those calls allocates, copies from the host or sets memory. */
typedef struct wk_graph wk_graph;
int wk_graph_begin(void *stream);
int wk_graph_end(void *stream, wk_graph **out);
int wk_graph_launch(wk_graph *graph, void *stream);
void wk_graph_destroy(wk_graph *graph);

/* Blocking timing helper with the reference's contract (engine::run returns
 * the seconds of one launch, warpkernel.hpp:829-843): launches once, waits,
 * returns the elapsed seconds measured with HIP events in *seconds.
 * which: 0 run, 1 run_x, 2 run_assembly (x unused, k = 1). */
int wk_time_run(wk_engine *engine, const wk_launch *cfg, int32_t which, const void *x_dev,
                double *y_dev, int32_t k, int64_t ldx, int64_t ldy, void *stream,
                float *seconds);

/* Measurement helper, no counterpart in the reference: reads `bytes` bytes
 * of device memory at `buf` (16-byte aligned, size a multiple of 16) with
 * 16-byte nontemporal loads, eight per thread in flight, on `stream`.  Its
 * duration gives the device's streaming-read rate, the practical ceiling of
 * the SpMV kernels' HBM traffic.  `sink` (4 bytes of device memory, or NULL)
 * is written only if the XOR of the data hits a fixed value. */
int wk_stream_read(const void *buf, int64_t bytes, uint32_t *sink, void *stream);

/* Measurement helper, no counterpart in the reference: the forward SpMV's
 * direct kernel (group width 64, one right-hand side) with the values not
 * loaded, every slot counting 1.0: it reads the column stream and gathers
 * x[col] for every slot exactly as wk_run's direct kernel does with the same
 * unroll (4 or 8), rows_per_lane, waves_per_block and xcd_remap (cfg, or the
 * engine's launch for WK_OP_RUN at k = 1 when NULL), and writes y = the sum
 * of x over each row's slots, pads included.  Its duration is the ceiling of
 * a layout whose x gathers miss the caches (randomly numbered meshes).  FP64
 * engines; x [ncols], y [nrows] device arrays. */
int wk_engine_gather_ceiling(wk_engine *engine, const wk_launch *cfg, const double *x_dev,
                             double *y_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
