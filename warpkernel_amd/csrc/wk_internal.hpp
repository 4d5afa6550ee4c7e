// wk_internal.hpp -- types shared by the host scan, the C ABI and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/warpkernel_c.h"

// Host-side WPK layout: the result of structure::scan / structure2::scan
// (warpkernel.hpp:276-286, 561-565), generic in the group width W.
struct wk_layout {
  int32_t variant = 1;
  int32_t W = 64;
  int32_t threshold = 0;
  int32_t max_tpr = 32;
  int32_t nrows = 0, ncols = 0, nznrows = 0, nwarps = 0;
  int64_t nz = 0, allocate_nz = 0;
  uint32_t min_nz = 0, max_nz = 0;
  std::vector<int32_t> row_map, row_map_inv;
  std::vector<uint32_t> reorder_rows;
  std::vector<int64_t> warp_offsets;
  std::vector<uint32_t> max_nz_rows;
  std::vector<uint32_t> threads_per_row;  // variant 2
  std::vector<uint32_t> row_offset_warp;  // variant 2
  std::vector<int32_t> reorder_A_rows;    // after reorder_columns_rowsort
  std::vector<int32_t> rowptr;            // the scanned CSR's row offsets [nrows+1]
};

namespace wk {

// Error plumbing: every C entry point catches wk::Error and returns its code.
struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string &msg);
void hip_check(hipError_t e, const char *what);
void set_last_error(const std::string &msg);

void validate_csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rowptr,
                  const int32_t *colind);
void scan1(wk_layout &L, int32_t W, int32_t nrows, int32_t ncols, int64_t nnz,
           const int32_t *rowptr);
void scan2(wk_layout &L, int32_t W, int32_t threshold, int32_t max_tpr, int32_t nrows,
           int32_t ncols, int64_t nnz, const int32_t *rowptr);
// Host scatter through reorder_rows into zero-filled padded arrays (T = double or float).
template <class T>
void process_host(const wk_layout &L, const double *vals, const int32_t *colind, T *pvals,
                  int32_t *pcols);
void reorder_columns_coalesced(const wk_layout &L, const int32_t *colind, int32_t *out);
// reorder_columns_rowsort (warpkernel.hpp:414-444): per row, entries ordered by
// permuted column (ties in CSR order); cols_out[n] = permuted column,
// src[n] = CSR index of the entry placed at n.  Needs a square matrix.
void reorder_columns_rowsort(const wk_layout &L, const int32_t *rowptr, const int32_t *colind,
                             int32_t *cols_out, int32_t *src);
// CSR of A^T with the entries of each column in increasing row order;
// perm[t] = CSR index of transposed entry t.
void transpose_csr(int32_t nrows, int32_t ncols, const int32_t *rowptr, const int32_t *colind,
                   std::vector<int32_t> &colptr, std::vector<int32_t> &rowind,
                   std::vector<int64_t> &perm);
// Fails with WK_ERR_CSR ("<what>: ...") unless the layout satisfies every
// invariant the kernels and the host helpers index by.
void check_layout(const wk_layout &L, const char *what);
// Run fn(begin, end) over [0, n) on up to 16 host threads.
void parallel_for(int64_t n, int64_t min_chunk, const std::function<void(int64_t, int64_t)> &fn);

// One WPK layout resident on the device.  All pointers are device memory owned
// by the engine.
struct DevLayout {
  int32_t variant = 1, W = 64;
  int32_t nrows = 0;      // rows of the product (ncols of A for the transpose)
  int32_t nznrows = 0, ngroups = 0;
  int64_t allocate_nz = 0;
  void *vals = nullptr;           // [allocate_nz] double or float
  int32_t *cols = nullptr;        // [allocate_nz]
  int32_t *cols_kr = nullptr;     // [allocate_nz] Kr columns (optional)
  int32_t *row_map = nullptr;     // [nrows]
  int32_t *row_map_inv = nullptr; // [nrows] (engines scanned on the device)
  int64_t *warp_offsets = nullptr;   // [ngroups+1]
  uint32_t *max_nz_rows = nullptr;   // [ngroups]
  uint32_t *tpr = nullptr;           // [ngroups] (variant 2)
  uint32_t *row_off = nullptr;       // [ngroups+1] (variant 2)
  uint32_t *slot_of_nz = nullptr;    // [nnz] CSR index -> padded slot (GPU process)
  std::vector<uint32_t> h_max_nz_rows;  // host copy (batch work partition)
  // x windows (stage_lds 3): each group's distinct columns, ascending, and per
  // slot the 16-bit position of its column in that list.  win_max = 0: none.
  uint16_t *win_idx = nullptr;     // [allocate_nz]
  uint32_t *win_off = nullptr;     // [ngroups+1] into win_cols
  int32_t *win_cols = nullptr;     // [win_off[ngroups]]
  int32_t *win_cols_kr = nullptr;  // the same columns in Kr numbering (row_map_inv)
  int32_t win_max = 0;             // largest window; 0 when the layout has none
  int64_t win_total = 0;           // win_off[ngroups]: entries over all windows
  // the values above as built, kept while a column swap turns the windows off
  // (they describe the engine's own columns only)
  int32_t win_max_built = 0;
  int64_t win_total_built = 0;
  // the same positions packed 10 bits each, three per 32-bit word, when every
  // window has at most 1024 entries: per group ceil(len/12) blocks of 64 lanes
  // x 16 bytes, lane l's block holding its 12 positions of those 12 steps
  uint32_t *win_pk = nullptr;      // [win_pk_off[ngroups] * 64 * 4]
  uint32_t *win_pk_off = nullptr;  // [ngroups+1] first 12-step block of each group
  uint32_t pk_longest = 0;         // the longest group (steps), with win_pk
  // node windows (stage_lds 7), built from the packed windows when every
  // group's length is a multiple of three and each lane reads its window
  // three steps at a time either at three consecutive positions holding
  // columns 3n, 3n+1, 3n+2 (an elasticity node) or three times at one
  // position (pads).  A group's window is then a list of node entries, each
  // three x values in LDS: entry e >= 0 holds x[e], x[e+1], x[e+2]; e < 0
  // holds x[-e-1] three times.  Per lane, one 10-bit entry index per three
  // steps, three to a 32-bit word, three words per 27-step block.
  int32_t *nd_cols = nullptr;      // [win_total] group g's entries from win_off[g]
  // the same entries in Kr numbering (engines with Kr windows), as three
  // explicit columns each: entry j of group g at 3 * (win_off[g] + j)
  int32_t *nd_cols_kr = nullptr;   // [3 * win_total]
  uint32_t *nd_cnt = nullptr;      // [ngroups] entries per group
  uint32_t *nd_pos = nullptr;      // [nd_off[ngroups] * 64] words, word-major, lane fastest
  uint32_t *nd_off = nullptr;      // [ngroups+1] first word of each group
  int32_t nd_max = 0;              // most entries in one group; 0: no node windows
};

// Window positions fit the packed form up to this many entries per window.
constexpr int32_t kPackedWindow = 1024;

// Largest x window (distinct columns of one group) that the window kernel
// stages; layouts with a larger group get no windows.
constexpr int32_t kMaxWindow = 2048;

// Builds the x windows of a layout from one key per padded slot: the column of
// an entry, -1 for a pad (a pad reads column 0 of whichever numbering the
// product uses, which in the Kr form differs from an entry of column 0).
// keys_out lists each group's distinct keys ascending.  Returns false, leaving
// the outputs empty, when a group has more than kMaxWindow distinct keys.
bool build_windows(const std::vector<int64_t> &warp_offsets, const std::vector<uint32_t> &max_nz_rows,
                   int32_t W, const int32_t *keys, std::vector<uint16_t> &idx,
                   std::vector<uint32_t> &off, std::vector<int32_t> &keys_out, int32_t &win_max);

}  // namespace wk

struct wk_engine {
  int32_t dtype = WK_F64;
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  wk_layout_info info{};
  wk::DevLayout fwd;
  bool has_adj = false;
  wk::DevLayout adj;                // layout of A^T
  int32_t *d_rowptr = nullptr;      // CSR pattern for dvals (adjoint engines)
  int32_t *d_colind = nullptr;
  int32_t *d_tile_row = nullptr;    // dvals tile -> first row
  int32_t *d_rowind = nullptr;      // row of every CSR entry (dvals)
  std::vector<void *> owned;        // every hipMalloc of this engine
  int device = 0;
  std::mutex cfg_mu;                // guards launch_cfg
  std::map<int64_t, wk_launch> launch_cfg;  // (op << 32 | k) -> configuration
};

namespace wk {

// Kernel-side argument block of one SpMV launch.
struct SpmvArgs {
  const void *vals;
  const int32_t *cols;
  const int32_t *row_map;   // nullptr: store in sorted order (run_x)
  const int64_t *warp_offsets;
  const uint32_t *max_nz_rows;
  const uint32_t *tpr;      // variant 2
  const uint32_t *row_off;  // variant 2
  const void *x;
  double *y;
  const double *b;          // residual: y = b - A x
  int64_t ldx, ldy, ldb;
  int32_t ngroups, nznrows, nrows, k;
};

enum Out { OUT_RUN = 0, OUT_RUNX = 1, OUT_RESID = 2 };

// Launches one SpMV over a device layout.  prec: WK_F64, or WK_F32 with
// cfg.accumulate.  Throws wk::Error on unsupported options.
void launch_spmv(const DevLayout &D, int32_t dtype, const wk_launch &cfg, int out,
                 const void *x, int64_t ldx, double *y, int64_t ldy, const double *b,
                 int64_t ldb, int32_t k, hipStream_t stream);

// Zero-fill then scatter CSR values into the padded slots (GPU process).
void launch_scatter_values(const DevLayout &D, int32_t dtype, int64_t nnz, const void *vals_csr,
                           hipStream_t stream);
void launch_stream_read(const void *p, int64_t bytes, uint32_t *sink, hipStream_t stream);
void launch_gather_ceiling(const DevLayout &D, const wk_launch &cfg, const double *x, double *y,
                           hipStream_t stream);
// dst[index[i]] = src[i] for i < n with index[i] < dst_count (others skipped);
// elements of elem_bytes (4 or 8) bytes.
void launch_scatter(const void *src, const uint32_t *index, void *dst, int64_t n, int64_t dst_count,
                    int32_t elem_bytes, hipStream_t stream);
// Builds D.win_pk / D.win_pk_off from D.win_idx when the windows fit the
// packed form (width 64, win_max <= kPackedWindow); otherwise leaves them null.
// Then builds the node windows (D.nd_*) where the layout has them.
void build_packed_positions(wk_engine *E, DevLayout &D, hipStream_t stream);
void launch_fill_zero(void *p, int64_t bytes, hipStream_t stream);
// Flags (int32) the column swap's check kernels write, one per block.
constexpr int32_t kColFlags = 1024;
// Replaces the forward layout's padded columns with pcols (allocate_nz
// entries; where = 0 host, 1 device) after checking every entry lies in
// [0, ncols) (WK_ERR_ARG otherwise, nothing changed); the x windows stay in use
// exactly while the columns equal the ones they were built from.  flags:
// kColFlags int32 of device memory for the checks.  Engines with the adjoint
// layout are refused (WK_ERR_STATE).  Synchronises the stream.
void swap_colinds(wk_engine *E, const int32_t *pcols, int32_t where, hipStream_t stream,
                  int32_t *flags);
// Row sums of the padded values (run_assembly); cfg supplies waves_per_block
// and xcd_remap.
void launch_rowsum(const DevLayout &D, int32_t dtype, const wk_launch &cfg, double *y,
                   hipStream_t stream);
// dvals in CSR order: k = 1 on tiles of consecutive entries (tile_row[t] =
// row of entry t*tile_size, tile_row[tiles] = nrows - 1), k > 1 from the row
// of every entry (rowind[e]).
void launch_dvals(const int32_t *rowptr, const int32_t *colind, const int32_t *tile_row,
                  const int32_t *rowind, int64_t nnz, const double *dy, int64_t lddy,
                  const double *x, int64_t ldx, int32_t k, double *dvals, hipStream_t stream);
// dx = A^T dy and dvals in one launch (FP64, one contiguous right-hand side,
// group width 64), for layouts whose groups are short.
void launch_backward(const DevLayout &T, const double *dy, double *dx, const int32_t *rowptr,
                     const int32_t *colind, const int32_t *tile_row, int64_t nnz, const double *x,
                     double *dvals, hipStream_t stream);
int64_t dvals_tiles(int64_t nnz);
int64_t dvals_tile_size();

// The configuration an engine uses for (op, k) when the caller passes none:
// the tuned one if any, else a choice from the layout.
wk_launch engine_launch(wk_engine *E, int32_t op, int32_t k);
wk_launch heuristic_launch(const DevLayout &D, int32_t dtype, int32_t k);

// WPK2's group split (warpkernel.hpp:633-691) over the longest row of each
// W-row group of sorted rows (aw0); nzr = non-empty rows.  Outputs per
// (sub-)group: A_w, rows per group for threads_per_row, rows it advances
// row_offset_warp by.
void split_groups(const std::vector<uint32_t> &aw0, int32_t nzr, int32_t W, int32_t threshold,
                  int32_t max_tpr, std::vector<uint32_t> &max_nz_rows,
                  std::vector<uint32_t> &rows_nominal, std::vector<uint32_t> &rows_actual);

// Runs f and turns what it throws into a status code and wk_last_error().
template <class F>
int guarded(F &&f) {
  try {
    f();
    return WK_OK;
  } catch (const Error &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return WK_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return WK_ERR_ARG;
  }
}

// Device allocations owned (and freed) by an engine.
template <class T>
T *dev_alloc(wk_engine *E, int64_t count) {
  if (count <= 0) return nullptr;
  void *p = nullptr;
  hip_check(hipMalloc(&p, sizeof(T) * (size_t)count), "hipMalloc");
  E->owned.push_back(p);
  return static_cast<T *>(p);
}

template <class T>
T *dev_upload(wk_engine *E, const T *src, int64_t count, hipStream_t s) {
  T *d = dev_alloc<T>(E, count);
  if (d) hip_check(hipMemcpyAsync(d, src, sizeof(T) * (size_t)count, hipMemcpyHostToDevice, s), "upload");
  return d;
}

}  // namespace wk
