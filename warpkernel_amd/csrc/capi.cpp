// capi.cpp -- the extern "C" boundary (include/warpkernel_c.h): error codes,
// the host layout object (structure / structure2) and the device engine
// (engine<V,I>, warpkernel.hpp:773-959).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>

#include "wk_internal.hpp"

namespace wk {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

[[noreturn]] void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(WK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

void free_engine(wk_engine *E) {
  if (!E) return;
  // teardown is best effort: errors here have nowhere to go
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(E->device);
  for (void *p : E->owned) (void)hipFree(p);
  (void)hipSetDevice(cur);
  delete E;
}

// Copies one host layout (and its padded values/columns) to the device.
// vals/colind are the layout's own CSR; slot_of_nz (optional) replaces
// reorder_rows as the device scatter map (the transpose is scattered from A's
// CSR order).
void upload_layout(wk_engine *E, DevLayout &D, const wk_layout &L, const double *vals,
                   const int32_t *colind, int32_t dtype, bool kr, const uint32_t *slot_of_nz,
                   hipStream_t s) {
  D.variant = L.variant;
  D.W = L.W;
  D.nrows = L.nrows;
  D.nznrows = L.nznrows;
  D.ngroups = L.nwarps;
  D.allocate_nz = L.allocate_nz;
  std::vector<int32_t> pcols((size_t)L.allocate_nz);
  if (dtype == WK_F64) {
    std::vector<double> pv((size_t)L.allocate_nz);
    process_host<double>(L, vals, colind, pv.data(), pcols.data());
    D.vals = dev_upload(E, pv.data(), L.allocate_nz, s);
    hip_check(hipStreamSynchronize(s), "upload");
  } else {
    std::vector<float> pv((size_t)L.allocate_nz);
    process_host<float>(L, vals, colind, pv.data(), pcols.data());
    D.vals = dev_upload(E, pv.data(), L.allocate_nz, s);
    hip_check(hipStreamSynchronize(s), "upload");
  }
  D.cols = dev_upload(E, pcols.data(), L.allocate_nz, s);
  hip_check(hipStreamSynchronize(s), "upload");
  if (kr) {
    reorder_columns_coalesced(L, colind, pcols.data());
    D.cols_kr = dev_upload(E, pcols.data(), L.allocate_nz, s);
    hip_check(hipStreamSynchronize(s), "upload");
  }
  D.row_map = dev_upload(E, L.row_map.data(), L.nrows, s);
  D.warp_offsets = dev_upload(E, L.warp_offsets.data(), (int64_t)L.warp_offsets.size(), s);
  D.max_nz_rows = dev_upload(E, L.max_nz_rows.data(), (int64_t)L.max_nz_rows.size(), s);
  D.h_max_nz_rows = L.max_nz_rows;
  if (L.variant == 2) {
    D.tpr = dev_upload(E, L.threads_per_row.data(), (int64_t)L.threads_per_row.size(), s);
    D.row_off = dev_upload(E, L.row_offset_warp.data(), (int64_t)L.row_offset_warp.size(), s);
  }
  D.slot_of_nz = dev_upload(E, slot_of_nz ? slot_of_nz : L.reorder_rows.data(), L.nz, s);
  hip_check(hipStreamSynchronize(s), "upload");
  if (L.W == 64) {
    // x windows: one key per slot, the entry's column or -1 for a pad
    std::vector<int32_t> keys((size_t)L.allocate_nz, -1);
    parallel_for(L.nz, 1 << 20, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; i++) keys[L.reorder_rows[i]] = colind[i];
    });
    std::vector<uint16_t> idx;
    std::vector<uint32_t> off;
    std::vector<int32_t> wkeys;
    if (build_windows(L.warp_offsets, L.max_nz_rows, L.W, keys.data(), idx, off, wkeys, D.win_max)) {
      std::vector<int32_t>().swap(keys);
      std::vector<int32_t> wc(wkeys.size());
      for (size_t j = 0; j < wkeys.size(); j++) wc[j] = wkeys[j] < 0 ? 0 : wkeys[j];
      D.win_idx = dev_upload(E, idx.data(), (int64_t)idx.size(), s);
      D.win_off = dev_upload(E, off.data(), (int64_t)off.size(), s);
      D.win_cols = dev_upload(E, wc.data(), (int64_t)wc.size(), s);
      D.win_total = (int64_t)wc.size();
      hip_check(hipStreamSynchronize(s), "upload");
      if (kr) {
        for (size_t j = 0; j < wkeys.size(); j++) wc[j] = wkeys[j] < 0 ? 0 : L.row_map_inv[wkeys[j]];
        D.win_cols_kr = dev_upload(E, wc.data(), (int64_t)wc.size(), s);
        hip_check(hipStreamSynchronize(s), "upload");
      }
      build_packed_positions(E, D, s);
    } else {
      D.win_max = 0;
    }
  }
}

wk_launch resolve(wk_engine *E, int32_t op, int32_t k, const wk_launch *cfg) {
  return cfg ? *cfg : engine_launch(E, op, k);
}

void check_engine(const wk_engine *E) {
  if (!E) fail(WK_ERR_ARG, "engine is NULL");
}

}  // namespace
}  // namespace wk

using namespace wk;

extern "C" {

const char *wk_last_error(void) { return g_last_error.c_str(); }

int wk_abi_version(void) { return WK_ABI_VERSION; }

int wk_scan1(int32_t group_width, int32_t nrows, int32_t ncols, int64_t nnz,
             const int32_t *rowptr_host, const int32_t *colind_host, wk_layout **out) {
  return guarded([&] {
    if (!out) fail(WK_ERR_ARG, "out is NULL");
    *out = nullptr;
    validate_csr(nrows, ncols, nnz, rowptr_host, colind_host);
    auto *L = new wk_layout();
    try {
      scan1(*L, group_width, nrows, ncols, nnz, rowptr_host);
    } catch (...) {
      delete L;
      throw;
    }
    *out = L;
  });
}

int wk_scan2(int32_t group_width, int32_t threshold, int32_t max_tpr, int32_t nrows,
             int32_t ncols, int64_t nnz, const int32_t *rowptr_host, const int32_t *colind_host,
             wk_layout **out) {
  return guarded([&] {
    if (!out) fail(WK_ERR_ARG, "out is NULL");
    *out = nullptr;
    validate_csr(nrows, ncols, nnz, rowptr_host, colind_host);
    auto *L = new wk_layout();
    try {
      scan2(*L, group_width, threshold, max_tpr, nrows, ncols, nnz, rowptr_host);
    } catch (...) {
      delete L;
      throw;
    }
    *out = L;
  });
}

int wk_layout_get_info(const wk_layout *L, wk_layout_info *info) {
  return guarded([&] {
    if (!L || !info) fail(WK_ERR_ARG, "NULL argument");
    info->variant = L->variant;
    info->group_width = L->W;
    info->threshold = L->threshold;
    info->max_tpr = L->max_tpr;
    info->nrows = L->nrows;
    info->ncols = L->ncols;
    info->nznrows = L->nznrows;
    info->nwarps = L->nwarps;
    info->nz = L->nz;
    info->allocate_nz = L->allocate_nz;
    info->min_nz = L->min_nz;
    info->max_nz = L->max_nz;
  });
}

int wk_layout_array(const wk_layout *L, int32_t which, const void **data, int64_t *count) {
  return guarded([&] {
    if (!L || !data || !count) fail(WK_ERR_ARG, "NULL argument");
    switch (which) {
      case WK_ROW_MAP: *data = L->row_map.data(); *count = (int64_t)L->row_map.size(); break;
      case WK_ROW_MAP_INV: *data = L->row_map_inv.data(); *count = (int64_t)L->row_map_inv.size(); break;
      case WK_REORDER_ROWS: *data = L->reorder_rows.data(); *count = (int64_t)L->reorder_rows.size(); break;
      case WK_WARP_OFFSETS: *data = L->warp_offsets.data(); *count = (int64_t)L->warp_offsets.size(); break;
      case WK_MAX_NZ_ROWS: *data = L->max_nz_rows.data(); *count = (int64_t)L->max_nz_rows.size(); break;
      case WK_THREADS_PER_ROW:
        if (L->variant != 2) fail(WK_ERR_STATE, "threads_per_row exists for structure2 only");
        *data = L->threads_per_row.data(); *count = (int64_t)L->threads_per_row.size(); break;
      case WK_ROW_OFFSET_WARP:
        if (L->variant != 2) fail(WK_ERR_STATE, "row_offset_warp exists for structure2 only");
        *data = L->row_offset_warp.data(); *count = (int64_t)L->row_offset_warp.size(); break;
      case WK_REORDER_A_ROWS:
        *data = L->reorder_A_rows.data(); *count = (int64_t)L->reorder_A_rows.size(); break;
      case WK_ROWPTR: *data = L->rowptr.data(); *count = (int64_t)L->rowptr.size(); break;
      default: fail(WK_ERR_ARG, "unknown layout array");
    }
  });
}

int wk_layout_process_host(const wk_layout *L, const double *vals_host, const int32_t *colind_host,
                           double *pvals_host, int32_t *pcols_host) {
  return guarded([&] {
    if (!L || (!pvals_host && !pcols_host)) fail(WK_ERR_ARG, "NULL argument");
    if ((pvals_host && !vals_host) || (pcols_host && !colind_host)) fail(WK_ERR_ARG, "NULL input");
    process_host<double>(*L, vals_host, colind_host, pvals_host, pcols_host);
  });
}

int wk_layout_reorder_columns_coalesced(const wk_layout *L, const int32_t *colind_host,
                                        int32_t *pcols_kr_host) {
  return guarded([&] {
    if (!L || !colind_host || !pcols_kr_host) fail(WK_ERR_ARG, "NULL argument");
    if (L->nrows != L->ncols) fail(WK_ERR_ARG, "the Kr form needs a square matrix");
    reorder_columns_coalesced(*L, colind_host, pcols_kr_host);
  });
}

int wk_layout_reorder_x(const wk_layout *L, const double *x_host, double *xout_host) {
  return guarded([&] {
    if (!L || !x_host || !xout_host) fail(WK_ERR_ARG, "NULL argument");
    for (int32_t r = 0; r < L->nrows; r++) xout_host[L->row_map_inv[r]] = x_host[r];
  });
}

int wk_layout_reorder_columns_rowsort(wk_layout *L, const int32_t *rowptr_host,
                                      const int32_t *colind_host, int32_t *cols_out_host) {
  return guarded([&] {
    if (!L || !rowptr_host || !cols_out_host || (L->nz && !colind_host)) fail(WK_ERR_ARG, "NULL argument");
    if (L->nrows != L->ncols) fail(WK_ERR_ARG, "the row sort needs a square matrix");
    validate_csr(L->nrows, L->ncols, L->nz, rowptr_host, colind_host);
    std::vector<int32_t> src((size_t)L->nz);
    reorder_columns_rowsort(*L, rowptr_host, colind_host, cols_out_host, src.data());
    L->reorder_A_rows.swap(src);
  });
}

int wk_layout_reorder_values_rowsort(const wk_layout *L, const double *vals_host,
                                     double *pvals_host) {
  return guarded([&] {
    if (!L || !pvals_host || (L->nz && !vals_host)) fail(WK_ERR_ARG, "NULL argument");
    if ((int64_t)L->reorder_A_rows.size() != L->nz)
      fail(WK_ERR_STATE, "call wk_layout_reorder_columns_rowsort first");
    std::memset(pvals_host, 0, sizeof(double) * (size_t)L->allocate_nz);
    for (int64_t i = 0; i < L->nz; i++) pvals_host[L->reorder_rows[i]] = vals_host[L->reorder_A_rows[i]];
  });
}

int wk_layout_from_arrays(int32_t variant, int32_t group_width, int32_t threshold, int32_t max_tpr,
                          int32_t nrows, int32_t ncols, int32_t nznrows, int32_t nwarps, int64_t nz,
                          int64_t allocate_nz, const int32_t *rowptr, const int32_t *row_map,
                          const uint32_t *reorder_rows, const int64_t *warp_offsets,
                          const uint32_t *max_nz_rows, const uint32_t *threads_per_row,
                          const uint32_t *row_offset_warp, wk_layout **out) {
  return guarded([&] {
    if (!out) fail(WK_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (nrows < 0 || nwarps < 0 || nz < 0) fail(WK_ERR_ARG, "negative size");
    if (!rowptr || (nrows && !row_map) || (nz && !reorder_rows) || !warp_offsets ||
        (nwarps && !max_nz_rows) || (variant == 2 && nwarps && (!threads_per_row || !row_offset_warp)))
      fail(WK_ERR_ARG, "NULL array");
    auto L = std::make_unique<wk_layout>();
    L->variant = variant;
    L->W = group_width;
    L->threshold = threshold;
    L->max_tpr = max_tpr;
    L->nrows = nrows;
    L->ncols = ncols;
    L->nznrows = nznrows;
    L->nwarps = nwarps;
    L->nz = nz;
    L->allocate_nz = allocate_nz;
    L->rowptr.assign(rowptr, rowptr + nrows + 1);
    L->row_map.assign(row_map, row_map + nrows);
    L->row_map_inv.assign((size_t)nrows, -1);
    for (int32_t i = 0; i < nrows; i++) {
      const int32_t r = row_map[i];
      if (r < 0 || r >= nrows || L->row_map_inv[r] != -1) fail(WK_ERR_CSR, "layout arrays: row_map is not a permutation");
      L->row_map_inv[r] = i;
    }
    L->reorder_rows.assign(reorder_rows, reorder_rows + nz);
    L->warp_offsets.assign(warp_offsets, warp_offsets + nwarps + 1);
    L->max_nz_rows.assign(max_nz_rows, max_nz_rows + nwarps);
    if (variant == 2) {
      L->threads_per_row.assign(threads_per_row, threads_per_row + nwarps);
      L->row_offset_warp.assign(row_offset_warp, row_offset_warp + nwarps + 1);
    }
    check_layout(*L, "layout arrays");
    if (ncols < 0) fail(WK_ERR_CSR, "layout arrays: ncols");
    if (variant == 2) {
      // over the rows with entries, the shortest starting from nrows (hpp:607)
      L->min_nz = (uint32_t)nrows;
      L->max_nz = 0;
      for (int32_t r = 0; r < nrows; r++) {
        const uint32_t len = (uint32_t)(rowptr[r + 1] - rowptr[r]);
        if (len) L->min_nz = std::min(L->min_nz, len);
        L->max_nz = std::max(L->max_nz, len);
      }
    }
    *out = L.release();
  });
}

void wk_layout_destroy(wk_layout *L) { delete L; }

int wk_engine_create(const wk_layout *L, const int32_t *rowptr_host, const int32_t *colind_host,
                     const double *vals_host, int32_t values_dtype, uint32_t flags, void *stream,
                     wk_engine **out) {
  return guarded([&] {
    if (!out) fail(WK_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!L || (L->nz && (!colind_host || !vals_host))) fail(WK_ERR_ARG, "NULL argument");
    // rowptr_host NULL: the offsets the layout was scanned (or loaded) with
    if (!rowptr_host) rowptr_host = L->rowptr.data();
    if ((int64_t)L->rowptr.size() != (int64_t)L->nrows + 1 && rowptr_host == L->rowptr.data())
      fail(WK_ERR_ARG, "rowptr is NULL and the layout holds none");
    if (values_dtype != WK_F64 && values_dtype != WK_F32) fail(WK_ERR_ARG, "unknown values dtype");
    if (values_dtype == WK_F32 && L->W != 64) fail(WK_ERR_ARG, "FP32 engines need group_width 64");
    if (flags & ~(uint32_t)(WK_ENGINE_KR | WK_ENGINE_ADJOINT | WK_ENGINE_ROWSORT))
      fail(WK_ERR_ARG, "unknown engine flag");
    // the kernels gather x[col]: every column must be in range
    validate_csr(L->nrows, L->ncols, L->nz, rowptr_host, colind_host);
    const bool kr = flags & WK_ENGINE_KR, adj = flags & WK_ENGINE_ADJOINT;
    const bool rowsort = flags & WK_ENGINE_ROWSORT;
    if ((kr || rowsort) && L->nrows != L->ncols)
      fail(WK_ERR_ARG, "the Kr form and the row sort need a square matrix");
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto *E = new wk_engine();
    try {
      hip_check(hipGetDevice(&E->device), "hipGetDevice");
      E->dtype = values_dtype;
      E->nrows = L->nrows;
      E->ncols = L->ncols;
      E->nnz = L->nz;
      wk_layout_get_info(L, &E->info);
      if (rowsort) {
        // reorder_columns_rowsort + reorder_values_coalesced_rowsort (hpp:414-444,
        // 505-515): each row's entries in permuted-column order; the CSR ->
        // slot map follows every entry to its new slot
        const int64_t nz = L->nz;
        std::vector<int32_t> kcols((size_t)nz), src((size_t)nz), cols2((size_t)nz);
        std::vector<double> vals2((size_t)nz);
        std::vector<uint32_t> slot((size_t)nz);
        reorder_columns_rowsort(*L, rowptr_host, colind_host, kcols.data(), src.data());
        for (int64_t n = 0; n < nz; n++) {
          vals2[n] = vals_host[src[n]];
          cols2[n] = colind_host[src[n]];
          slot[src[n]] = L->reorder_rows[n];
        }
        upload_layout(E, E->fwd, *L, vals2.data(), cols2.data(), values_dtype, kr, slot.data(), s);
      } else {
        upload_layout(E, E->fwd, *L, vals_host, colind_host, values_dtype, kr, nullptr, s);
      }
      if (adj) {
        // A^T in CSR with the same scan options; its values are A's values
        // permuted, and slot_of_nz maps A's CSR index to the transposed slot.
        std::vector<int32_t> colptr, rowind;
        std::vector<int64_t> perm;
        transpose_csr(L->nrows, L->ncols, rowptr_host, colind_host, colptr, rowind, perm);
        std::vector<double> tv(perm.size());
        for (size_t t = 0; t < perm.size(); t++) tv[t] = vals_host[perm[t]];
        wk_layout LT;
        if (L->variant == 2)
          scan2(LT, L->W, L->threshold, L->max_tpr, L->ncols, L->nrows, L->nz, colptr.data());
        else
          scan1(LT, L->W, L->ncols, L->nrows, L->nz, colptr.data());
        std::vector<uint32_t> slot_t((size_t)L->nz);
        for (size_t t = 0; t < perm.size(); t++) slot_t[perm[t]] = LT.reorder_rows[t];
        upload_layout(E, E->adj, LT, tv.data(), rowind.data(), values_dtype, false, slot_t.data(), s);
        E->has_adj = true;
        E->d_rowptr = dev_upload(E, rowptr_host, (int64_t)L->nrows + 1, s);
        E->d_colind = dev_upload(E, colind_host, L->nz, s);
        // dvals tiles: row of each tile's first entry (rows with entries only)
        const int64_t nt = dvals_tiles(L->nz), ts = dvals_tile_size();
        std::vector<int32_t> tile_row((size_t)nt + 1);
        int32_t r = 0;
        for (int64_t t = 0; t < nt; t++) {
          while (r + 1 < L->nrows && (int64_t)rowptr_host[r + 1] <= t * ts) r++;
          tile_row[t] = r;
        }
        tile_row[nt] = L->nrows > 0 ? L->nrows - 1 : 0;
        E->d_tile_row = dev_upload(E, tile_row.data(), nt + 1, s);
        // the row of every CSR entry (dvals)
        std::vector<int32_t> coo_row((size_t)L->nz);
        for (int32_t rr = 0; rr < L->nrows; rr++)
          for (int64_t e = rowptr_host[rr]; e < rowptr_host[rr + 1]; e++) coo_row[e] = rr;
        E->d_rowind = dev_upload(E, coo_row.data(), L->nz, s);
        hip_check(hipStreamSynchronize(s), "upload");
      }
    } catch (...) {
      free_engine(E);
      throw;
    }
    *out = E;
  });
}

void wk_engine_destroy(wk_engine *E) { free_engine(E); }

int wk_engine_get_info(const wk_engine *E, wk_layout_info *info) {
  return guarded([&] {
    check_engine(E);
    if (!info) fail(WK_ERR_ARG, "info is NULL");
    *info = E->info;
  });
}

int wk_engine_update_values(wk_engine *E, const void *vals_dev, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (!vals_dev && E->nnz) fail(WK_ERR_ARG, "vals_dev is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);
    launch_scatter_values(E->fwd, E->dtype, E->nnz, vals_dev, s);
    if (E->has_adj) launch_scatter_values(E->adj, E->dtype, E->nnz, vals_dev, s);
  });
}

int wk_engine_set_colinds(wk_engine *E, const int32_t *pcols, int32_t where, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (!pcols && E->fwd.allocate_nz) fail(WK_ERR_ARG, "pcols is NULL");
    if (!E->fwd.allocate_nz) return;
    // device memory for the check kernels' flags, freed on every path
    struct Flags {
      int32_t *p = nullptr;
      ~Flags() {
        if (p) (void)hipFree(p);
      }
    } flags;
    hip_check(hipMalloc(reinterpret_cast<void **>(&flags.p), sizeof(int32_t) * kColFlags), "hipMalloc");
    swap_colinds(E, pcols, where, static_cast<hipStream_t>(stream), flags.p);
  });
}

int wk_engine_x_windows(const wk_engine *E, int64_t *entries) {
  return guarded([&] {
    check_engine(E);
    if (!entries) fail(WK_ERR_ARG, "entries is NULL");
    *entries = E->fwd.win_max > 0 ? E->fwd.win_total : 0;
  });
}

int wk_engine_get_colinds(const wk_engine *E, int32_t *pcols_host) {
  return guarded([&] {
    check_engine(E);
    const size_t bytes = sizeof(int32_t) * (size_t)E->fwd.allocate_nz;
    if (!bytes) return;
    if (!pcols_host) fail(WK_ERR_ARG, "pcols_host is NULL");
    hip_check(hipMemcpy(pcols_host, E->fwd.cols, bytes, hipMemcpyDeviceToHost), "get_colinds");
  });
}

void wk_launch_default(wk_launch *cfg) {
  if (!cfg) return;
  cfg->waves_per_block = 4;
  cfg->unroll = 8;
  cfg->k_tile = 8;
  cfg->xcd_remap = 1;
  cfg->nontemporal = 1;
  cfg->accumulate = WK_ACC_TWOSUM;
  cfg->rows_per_lane = 1;
  cfg->blocks_per_cu = 0;
  cfg->stage_lds = 0;
}

int wk_run(wk_engine *E, const wk_launch *cfg, const void *x_dev, double *y_dev, int32_t k,
           int64_t ldx, int64_t ldy, void *stream) {
  return guarded([&] {
    check_engine(E);
    if ((!x_dev && E->ncols) || (!y_dev && E->nrows)) fail(WK_ERR_ARG, "NULL vector");
    if (ldx < k || ldy < k) fail(WK_ERR_ARG, "leading dimension smaller than k");
    launch_spmv(E->fwd, E->dtype, resolve(E, WK_OP_RUN, k, cfg), OUT_RUN, x_dev, ldx, y_dev, ldy, nullptr, 0, k,
                static_cast<hipStream_t>(stream));
  });
}

int wk_run_x(wk_engine *E, const wk_launch *cfg, const void *x_dev, double *y_dev, int32_t k,
             int64_t ldx, int64_t ldy, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (E->nrows != E->ncols) fail(WK_ERR_ARG, "the Kr form needs a square matrix");
    if ((!x_dev && E->ncols) || (!y_dev && E->nrows)) fail(WK_ERR_ARG, "NULL vector");
    if (ldx < k || ldy < k) fail(WK_ERR_ARG, "leading dimension smaller than k");
    launch_spmv(E->fwd, E->dtype, resolve(E, WK_OP_RUN_X, k, cfg), OUT_RUNX, x_dev, ldx, y_dev, ldy, nullptr, 0, k,
                static_cast<hipStream_t>(stream));
  });
}

int wk_residual(wk_engine *E, const wk_launch *cfg, const double *b_dev, int64_t ldb,
                const double *x_dev, double *r_dev, int32_t k, int64_t ldx, int64_t ldr,
                void *stream) {
  return guarded([&] {
    check_engine(E);
    if (E->dtype != WK_F64) fail(WK_ERR_ARG, "the residual is computed by FP64 engines");
    if ((!x_dev && E->ncols) || ((!r_dev || !b_dev) && E->nrows)) fail(WK_ERR_ARG, "NULL vector");
    if (ldx < k || ldr < k || ldb < k) fail(WK_ERR_ARG, "leading dimension smaller than k");
    launch_spmv(E->fwd, E->dtype, resolve(E, WK_OP_RUN, k, cfg), OUT_RESID, x_dev, ldx, r_dev, ldr, b_dev, ldb, k,
                static_cast<hipStream_t>(stream));
  });
}

int wk_run_adjoint(wk_engine *E, const wk_launch *cfg, const void *dy_dev, double *dx_dev,
                   int32_t k, int64_t lddy, int64_t lddx, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (!E->has_adj) fail(WK_ERR_STATE, "engine was created without WK_ENGINE_ADJOINT");
    if ((!dy_dev && E->nrows) || (!dx_dev && E->ncols)) fail(WK_ERR_ARG, "NULL vector");
    if (lddy < k || lddx < k) fail(WK_ERR_ARG, "leading dimension smaller than k");
    launch_spmv(E->adj, E->dtype, resolve(E, WK_OP_ADJOINT, k, cfg), OUT_RUN, dy_dev, lddy, dx_dev, lddx, nullptr, 0,
                k, static_cast<hipStream_t>(stream));
  });
}

int wk_dvals(wk_engine *E, const double *dy_dev, int64_t lddy, const double *x_dev, int64_t ldx,
             int32_t k, double *dvals_dev, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (!E->has_adj) fail(WK_ERR_STATE, "engine was created without WK_ENGINE_ADJOINT");
    if (k < 1 || lddy < k || ldx < k) fail(WK_ERR_ARG, "bad k or leading dimension");
    if (E->nnz && (!dy_dev || !x_dev || !dvals_dev)) fail(WK_ERR_ARG, "NULL vector");
    launch_dvals(E->d_rowptr, E->d_colind, E->d_tile_row, E->d_rowind, E->nnz, dy_dev, lddy, x_dev,
                 ldx, k, dvals_dev, static_cast<hipStream_t>(stream));
  });
}

int wk_backward(wk_engine *E, const wk_launch *cfg, const void *dy_dev, int64_t lddy,
                const void *x_dev, int64_t ldx, double *dx_dev, int64_t lddx, double *dvals_dev,
                int32_t k, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (!E->has_adj) fail(WK_ERR_STATE, "engine was created without WK_ENGINE_ADJOINT");
    if (k < 1 || lddy < k || ldx < k || lddx < k) fail(WK_ERR_ARG, "bad k or leading dimension");
    if ((!dy_dev && E->nrows) || (!dx_dev && E->ncols) || (E->nnz && (!x_dev || !dvals_dev)))
      fail(WK_ERR_ARG, "NULL vector");
    if (E->dtype != WK_F64) fail(WK_ERR_ARG, "the combined backward is computed by FP64 engines");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // short groups: both halves in one launch (measured 9% faster on 27-entry
    // rows); long groups run the adjoint's own (windowed) configuration
    const DevLayout &T = E->adj;
    const double mean_len = T.ngroups ? (double)T.allocate_nz / ((double)T.ngroups * T.W) : 0.0;
    if (k == 1 && lddy == 1 && ldx == 1 && lddx == 1 && T.W == 64 && mean_len < 40.0) {
      launch_backward(T, static_cast<const double *>(dy_dev), dx_dev, E->d_rowptr, E->d_colind,
                      E->d_tile_row, E->nnz, static_cast<const double *>(x_dev), dvals_dev, s);
      return;
    }
    const wk_launch c = resolve(E, WK_OP_ADJOINT, k, cfg);
    launch_spmv(E->adj, E->dtype, c, OUT_RUN, dy_dev, lddy, dx_dev, lddx, nullptr, 0, k, s);
    launch_dvals(E->d_rowptr, E->d_colind, E->d_tile_row, E->d_rowind, E->nnz,
                 static_cast<const double *>(dy_dev), lddy, static_cast<const double *>(x_dev), ldx,
                 k, dvals_dev, s);
  });
}

int wk_run_assembly(wk_engine *E, const wk_launch *cfg, double *y_dev, void *stream) {
  return guarded([&] {
    check_engine(E);
    if (!y_dev && E->nrows) fail(WK_ERR_ARG, "NULL vector");
    launch_rowsum(E->fwd, E->dtype, resolve(E, WK_OP_RUN, 1, cfg), y_dev,
                  static_cast<hipStream_t>(stream));
  });
}

int wk_scatter(const void *src_dev, const uint32_t *index_dev, void *dst_dev, int64_t n,
               int64_t dst_count, int32_t elem_bytes, void *stream) {
  return guarded([&] {
    if (n < 0 || dst_count < 0) fail(WK_ERR_ARG, "negative size");
    if (elem_bytes != 4 && elem_bytes != 8) fail(WK_ERR_ARG, "elem_bytes must be 4 or 8");
    if (n && (!src_dev || !index_dev || (dst_count && !dst_dev))) fail(WK_ERR_ARG, "NULL array");
    launch_scatter(src_dev, index_dev, dst_dev, n, dst_count, elem_bytes,
                   static_cast<hipStream_t>(stream));
  });
}

int wk_stream_read(const void *buf, int64_t bytes, uint32_t *sink, void *stream) {
  return guarded([&] {
    if (!buf && bytes > 0) fail(WK_ERR_ARG, "NULL buffer");
    launch_stream_read(buf, bytes, sink, static_cast<hipStream_t>(stream));
  });
}

int wk_engine_gather_ceiling(wk_engine *E, const wk_launch *cfg, const double *x_dev, double *y_dev,
                             void *stream) {
  return guarded([&] {
    check_engine(E);
    if (E->dtype != WK_F64) fail(WK_ERR_ARG, "the gather ceiling runs on FP64 engines");
    if ((!x_dev && E->ncols) || (!y_dev && E->nrows)) fail(WK_ERR_ARG, "NULL vector");
    launch_gather_ceiling(E->fwd, resolve(E, WK_OP_RUN, 1, cfg), x_dev, y_dev,
                          static_cast<hipStream_t>(stream));
  });
}

int wk_time_run(wk_engine *E, const wk_launch *cfg, int32_t which, const void *x_dev,
                double *y_dev, int32_t k, int64_t ldx, int64_t ldy, void *stream, float *seconds) {
  return guarded([&] {
    check_engine(E);
    if (!seconds) fail(WK_ERR_ARG, "seconds is NULL");
    if (which < 0 || which > 2) fail(WK_ERR_ARG, "which must be 0 (run), 1 (run_x) or 2 (run_assembly)");
    if (which == 2 && k != 1) fail(WK_ERR_ARG, "run_assembly has one output column");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the two events are destroyed on every path out of this scope
    struct Events {
      hipEvent_t e[2] = {nullptr, nullptr};
      ~Events() {
        for (hipEvent_t ev : e)
          if (ev) (void)hipEventDestroy(ev);
      }
    } ev;
    hip_check(hipEventCreate(&ev.e[0]), "event");
    hip_check(hipEventCreate(&ev.e[1]), "event");
    hip_check(hipEventRecord(ev.e[0], s), "event");
    const int rc = which == 0   ? wk_run(E, cfg, x_dev, y_dev, k, ldx, ldy, stream)
                   : which == 1 ? wk_run_x(E, cfg, x_dev, y_dev, k, ldx, ldy, stream)
                                : wk_run_assembly(E, cfg, y_dev, stream);
    if (rc != WK_OK) fail(rc, g_last_error);
    hip_check(hipEventRecord(ev.e[1], s), "event");
    hip_check(hipEventSynchronize(ev.e[1]), "event");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "timing events");
    *seconds = ms * 1e-3f;
  });
}

}  // extern "C"
