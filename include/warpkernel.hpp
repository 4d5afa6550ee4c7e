// warpkernel.hpp -- drop-in C++ API for the MI355X WarpKernel SpMV library.
//
// Rebuilds the reference's class API (namespace warpkernel, warpkernel.hpp:215-959
// of the reference) on top of the C ABI in warpkernel_c.h.  Same class names,
// member names and argument order:
//
//   warpkernel::structure    scan (pointers or a CSR container) / reorder /
//                            process / reorder_columns / reorder_columns_rowsort /
//                            reorder_columns_coalesced[_GPU] /
//                            reorder_values_coalesced[_GPU|_rowsort] / reorder_x /
//                            storemap / storereorder / store; public fields
//                            row_map, row_map_inv, reorder_rows, reorder_A_rows,
//                            warp_offsets, max_nz_rows, nrows, nznrows, nz,
//                            allocate_nz, nwarps                (hpp:215-551)
//   warpkernel::structure2   scan(..., threshold) (pointers or a CSR container),
//                            process with 7 arrays, store_param, store; fields
//                            row_offset_warp, threads_per_row, nmax_per_thread,
//                            min_nz, max_nz                     (hpp:553-771)
//   warpkernel::engine<V,I>  engine(k), engine(k, A, colinds); run<cache>;
//                            run_x<cache>; run_assembly<cache>; verify; verify_x;
//                            public device_values, device_colinds,
//                            device_row_map, device_max_nz_per_row,
//                            device_warp_offsets, device_threads_per_row,
//                            device_row_offset_warp              (hpp:773-959)
//
// What changes for a caller of the reference:
//   - cusp containers become std::vector (host) and warpkernel::device_vector
//     (device, RAII over hipMalloc); any CSR type with the members num_rows,
//     num_cols, num_entries, row_offsets, column_indices and values (cusp's
//     names; warpkernel::csr_matrix is one) can be scanned;
//   - the group width is a constructor argument: 64 (one gfx950 wavefront, the
//     default) or 32 (the reference's WARP_SIZE, identical layouts);
//   - warp_offsets hold int64_t and convert to the reference's std::vector<int>
//     (range-checked); process() fills an int or an int64_t device array;
//   - run / run_x take a blocksize that is a multiple of 64 (waves of 64);
//     nblocks is accepted and ignored (the grid follows from the layout);
//   - run<cache>: cache = true streams the matrix with nontemporal loads and
//     keeps x in the caches, cache = false uses plain loads for everything
//     (the reference's texture cache for x has no counterpart on CDNA);
//   - the engine's device arrays are views of the library's engine: reading
//     one downloads it, assigning values or columns replaces them on the
//     device, and the layout arrays (row map, group arrays) must equal the
//     structure's;
//   - rows without nonzeros are written as 0;
//   - errors throw warpkernel::error instead of asserting.
// engine<float, int> stores FP32 values and takes FP32 x; y is always double.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "warpkernel_c.h"

namespace warpkernel {

typedef unsigned int uint;

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != WK_OK) throw error(rc, wk_last_error());
}

inline void hip_check(hipError_t e) {
  if (e != hipSuccess) throw error(WK_ERR_HIP, hipGetErrorString(e));
}

// Element-wise copy into another integer type, refusing values it cannot hold.
template <class D, class S>
std::vector<D> narrow_copy(const std::vector<S> &s, const char *what) {
  if constexpr (std::is_same<D, S>::value) {
    return s;
  } else {
    std::vector<D> d(s.size());
    for (size_t i = 0; i < s.size(); i++) {
      if ((long double)s[i] > (long double)std::numeric_limits<D>::max() ||
          (long double)s[i] < (long double)std::numeric_limits<D>::lowest())
        throw error(WK_ERR_LIMIT, std::string(what) + " does not fit the array's element type");
      d[i] = (D)s[i];
    }
    return d;
  }
}

// Minimal device array (the reference's cusp::array1d<T, device_memory>).
template <class T>
class device_vector {
 public:
  typedef T value_type;
  device_vector() = default;
  explicit device_vector(size_t n, T fill = T()) { assign(std::vector<T>(n, fill)); }
  device_vector(const std::vector<T> &h) { assign(h); }
  device_vector(const device_vector &o) { *this = o; }
  device_vector(device_vector &&o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  ~device_vector() { release(); }
  device_vector &operator=(const std::vector<T> &h) {
    assign(h);
    return *this;
  }
  device_vector &operator=(const device_vector &o) {
    if (this != &o) {
      resize(o.n_);
      if (n_) hip_check(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice));
    }
    return *this;
  }
  operator std::vector<T>() const {
    std::vector<T> h(n_);
    if (n_) hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
  // resize keeps no contents (cusp's resize(n, value) fills: see resize_fill)
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) hip_check(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)));
    n_ = n;
  }
  void resize_fill(size_t n, T fill) { assign(std::vector<T>(n, fill)); }
  T *data() { return p_; }
  const T *data() const { return p_; }
  size_t size() const { return n_; }

 private:
  void assign(const std::vector<T> &h) {
    resize(h.size());
    if (n_) hip_check(hipMemcpy(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T *p_ = nullptr;
  size_t n_ = 0;
};

// Assigns a host array to any destination the reference's process() fills:
// a device_vector, a std::vector or an engine's device array, converting the
// element type where they differ (warp offsets: int64_t here, int there).
template <class Dst, class S>
void assign_array(Dst &dst, const std::vector<S> &src, const char *what) {
  dst = narrow_copy<typename Dst::value_type>(src, what);
}

// warp_offsets (hpp:280): int64_t entries here, so layouts past 2^31 slots work;
// converts to the reference's std::vector<int> when every offset fits.
class offsets_vector : public std::vector<int64_t> {
 public:
  offsets_vector() = default;
  offsets_vector(const std::vector<int64_t> &v) : std::vector<int64_t>(v) {}
  offsets_vector &operator=(const std::vector<int64_t> &v) {
    std::vector<int64_t>::operator=(v);
    return *this;
  }
  offsets_vector &operator=(const std::vector<int> &v) {
    assign(v.begin(), v.end());
    return *this;
  }
  operator std::vector<int>() const { return narrow_copy<int>(*this, "warp_offsets"); }
};

// The host CSR the drivers convert their MatrixMarket input to
// (cusp::csr_matrix<int, double, host_memory>, wpk1.cu:48), same field names.
struct csr_matrix {
  uint num_rows = 0, num_cols = 0, num_entries = 0;
  std::vector<int> row_offsets, column_indices;
  std::vector<double> values;
};

// cusp::io::read_matrix_market_file + the COO -> CSR conversion
// (wpk1.cu:43-48): coordinate files, entries ordered by (row, column).
inline void read_matrix_market_file(csr_matrix &A, const char *path) {
  wk_csr *h = nullptr;
  check(wk_mm_read(path, &h));
  int32_t nr = 0, nc = 0;
  int64_t nnz = 0;
  const int32_t *rp = nullptr, *ci = nullptr;
  const double *va = nullptr;
  const int rc = wk_csr_get(h, &nr, &nc, &nnz, &rp, &ci, &va);
  if (rc == WK_OK) {
    A.num_rows = (uint)nr;
    A.num_cols = (uint)nc;
    A.num_entries = (uint)nnz;
    A.row_offsets.assign(rp, rp + nr + 1);
    A.column_indices.assign(ci, ci + nnz);
    A.values.assign(va, va + nnz);
  }
  wk_csr_destroy(h);
  check(rc);
}

inline void write_matrix_market_file(const csr_matrix &A, const char *path) {
  check(wk_mm_write(path, (int32_t)A.num_rows, (int32_t)A.num_cols, (int64_t)A.num_entries,
                    A.row_offsets.data(), A.column_indices.data(), A.values.data()));
}

// WPK1 layout (hpp:215-551).  The public fields are the layout; after a scan or
// a load they mirror the library's layout object, and after storemap /
// storereorder / store (or a caller's own edits followed by one of them) the
// object is rebuilt from them, with every invariant checked, when next used.
class structure {
 public:
  explicit structure(int group_width = 64) : group_width_(group_width) {}
  // structure(nrows, nz, alloc) (hpp:273-274): fields only, no layout yet
  structure(const uint &nrows_, const uint &nz_, const uint &alloc_, int group_width = 64)
      : nrows(nrows_), nz(nz_), allocate_nz(alloc_), group_width_(group_width) {}
  structure(const structure &) = delete;
  structure &operator=(const structure &) = delete;
  virtual ~structure() { wk_layout_destroy(h_); }

  std::vector<int> row_map;      // y[row_map[i]] = y'[i]
  std::vector<int> row_map_inv;  // new_colinds[n] = row_map_inv[colinds[n]]
  std::vector<uint> reorder_rows;  // new_values[reorder_rows[i]] = A[i]
  std::vector<int> reorder_A_rows;  // only used for reordering A
  offsets_vector warp_offsets;
  std::vector<uint> max_nz_rows;
  uint nrows = 0, nznrows = 0, nz = 0, allocate_nz = 0, nwarps = 0;
  uint ncols = 0;  // 0: square (ncols = nrows)

  // storemap (hpp:288-296): row_map, nrows and row_map_inv from a row map.
  virtual void storemap(std::vector<int> &row_map_) {
    row_map = row_map_;
    nrows = (uint)row_map.size();
    row_map_inv.assign(nrows, 0);
    for (uint i = 0; i < nrows; i++) {
      if ((uint)row_map[i] >= nrows) throw error(WK_ERR_ARG, "row_map entry out of range");
      row_map_inv[row_map[i]] = (int)i;
    }
    stale_ = true;
  }
  // storereorder (hpp:298-302)
  void storereorder(std::vector<uint> &reorder_rows_) {
    reorder_rows = reorder_rows_;
    nz = (uint)reorder_rows.size();
    if (allocate_nz < nz) allocate_nz = nz;
    stale_ = true;
  }
  // store (hpp:304-307)
  void store(std::vector<int> &row_map_, std::vector<uint> &reorder_rows_) {
    storemap(row_map_);
    storereorder(reorder_rows_);
  }

  // structure::scan (hpp:309-365); only rows shapes the layout.
  template <typename ValueType, typename IndexType>
  void scan(uint &nz_, uint &nrows_, ValueType *A, IndexType *rows, IndexType *colinds) {
    static_assert(std::is_same<IndexType, int>::value, "IndexType must be int");
    (void)A;
    rows_.assign(rows, rows + nrows_ + 1);
    wk_layout *h = nullptr;
    check(wk_scan1(group_width_, (int32_t)nrows_, (int32_t)(ncols ? ncols : nrows_), nz_,
                   rows_.data(), colinds, &h));
    adopt(h);
  }

  // scan(nz, nrows, A) on a CSR container (hpp:369-377).
  template <typename CSR>
  auto scan(uint &nz_, uint &nrows_, CSR &A)
      -> decltype((void)A.row_offsets, (void)A.column_indices, (void)A.values, void()) {
    ncols = (uint)A.num_cols;
    scan(nz_, nrows_, A.values.data(), A.row_offsets.data(), A.column_indices.data());
  }

  // structure::reorder (hpp:379-398): zero-filled padded values and columns,
  // new[reorder_rows[i]] = A[i], colinds[i].
  template <typename IndexType, typename ValueType, class DV, class DC>
  void reorder(const ValueType *A, const IndexType *colinds, DV &device_values, DC &device_colinds) {
    std::vector<ValueType> pv;
    std::vector<IndexType> pc;
    process(A, colinds, pv, pc);
    assign_array(device_values, pv, "values");
    assign_array(device_colinds, pc, "colinds");
  }

  // The padded arrays on the host (structure::reorder without the upload).
  template <typename IndexType, typename ValueType>
  void process(const ValueType *A, const IndexType *colinds, std::vector<ValueType> &new_values,
               std::vector<IndexType> &new_colinds) const {
    std::vector<double> vals(A, A + nz), pv(allocate_nz);
    new_colinds.assign(allocate_nz, 0);
    check(wk_layout_process_host(handle(), vals.data(), colinds, pv.data(), new_colinds.data()));
    new_values.assign(pv.begin(), pv.end());
  }

  // structure::process (hpp:532-548): the padded arrays, the row map, the
  // group lengths and the group offsets, into the five device arrays.
  template <typename IndexType, typename ValueType, class DV, class DC, class DR, class DM, class DW>
  void process(const ValueType *A, const IndexType *colinds, DV &device_values, DC &device_colinds,
               DR &device_row_map, DM &device_max_nz_per_row, DW &device_warp_offsets) {
    reorder(A, colinds, device_values, device_colinds);
    assign_array(device_row_map, row_map, "row_map");
    assign_array(device_max_nz_per_row, max_nz_rows, "max_nz_rows");
    assign_array(device_warp_offsets, static_cast<const std::vector<int64_t> &>(warp_offsets),
                 "warp_offsets");
  }

  // reorder_columns (hpp:401-411): colinds[n] = row_map_inv[colinds[n]], CSR order.
  template <typename IndexType>
  void reorder_columns(std::vector<IndexType> &colinds) const {
    std::vector<IndexType> out(nz);
    for (uint n = 0; n < nz; n++) {
      if ((uint)colinds[n] >= row_map_inv.size()) throw error(WK_ERR_ARG, "column out of range");
      out[n] = row_map_inv[colinds[n]];
    }
    colinds.swap(out);
  }

  // reorder_columns_coalesced (hpp:446-457): colinds becomes the padded Kr columns.
  template <typename IndexType>
  void reorder_columns_coalesced(std::vector<IndexType> &colinds) const {
    std::vector<IndexType> out(allocate_nz);
    check(wk_layout_reorder_columns_coalesced(handle(), colinds.data(), out.data()));
    colinds.swap(out);
  }

  // reorder_columns_coalesced_GPU (hpp:459-466): thrust::scatter of device
  // arrays, dvalues[reorder_rows[i]] = values[i] (entries of dvalues that no
  // index names keep their contents).
  template <typename IndexType>
  void reorder_columns_coalesced_GPU(device_vector<IndexType> &values, device_vector<IndexType> &dvalues,
                                     device_vector<uint> &reorder_rows_) const {
    static_assert(sizeof(IndexType) == 4, "4-byte indices");
    scatter(values.data(), values.size(), reorder_rows_, dvalues.data(), dvalues.size(), 4);
  }

  // reorder_values_coalesced (hpp:469-492): zero-filled padded values on the device.
  template <typename ValueType, typename Array>
  void reorder_values_coalesced(const Array &values, device_vector<ValueType> &dvalues) const {
    std::vector<ValueType> pv(allocate_nz, ValueType(0));
    for (uint i = 0; i < nz; i++) pv[reorder_rows[i]] = (ValueType)values[i];
    dvalues = pv;
  }

  // reorder_values_coalesced_GPU (hpp:494-502): thrust::scatter of the values.
  template <typename ValueType>
  void reorder_values_coalesced_GPU(device_vector<ValueType> &values, device_vector<ValueType> &dvalues,
                                    device_vector<uint> &reorder_rows_) const {
    static_assert(sizeof(ValueType) == 4 || sizeof(ValueType) == 8, "4- or 8-byte values");
    scatter(values.data(), values.size(), reorder_rows_, dvalues.data(), dvalues.size(),
            (int)sizeof(ValueType));
  }

  // reorder_columns_rowsort (hpp:414-444): colinds (CSR order, nz entries)
  // becomes the permuted columns with each row ascending; reorder_A_rows[n]
  // is the CSR index of the entry now at n.  Equal columns keep CSR order.
  template <typename IndexType>
  void reorder_columns_rowsort(std::vector<IndexType> &colinds, const std::vector<IndexType> &rows) {
    static_assert(std::is_same<IndexType, int>::value, "IndexType must be int");
    std::vector<IndexType> out(colinds.size());
    check(wk_layout_reorder_columns_rowsort(handle_mut(), rows.data(), colinds.data(), out.data()));
    colinds.swap(out);
    reorder_A_rows = array<int>(WK_REORDER_A_ROWS);
  }

  // reorder_values_coalesced_rowsort (hpp:505-515): padded values,
  // dvalues[reorder_rows[i]] = values[reorder_A_rows[i]].
  template <typename ValueType>
  void reorder_values_coalesced_rowsort(const std::vector<ValueType> &values,
                                        device_vector<ValueType> &dvalues) const {
    std::vector<double> v(values.begin(), values.end()), pv(allocate_nz);
    check(wk_layout_reorder_values_rowsort(handle(), v.data(), pv.data()));
    dvalues = std::vector<ValueType>(pv.begin(), pv.end());
  }

  // reorder_x (hpp:517-530): device_x[row_map_inv[r]] = x[r].
  template <typename ValueType>
  void reorder_x(const std::vector<ValueType> &x, device_vector<ValueType> &device_x) const {
    std::vector<ValueType> out(x.size());
    for (uint r = 0; r < nrows; r++) out[row_map_inv[r]] = x[r];
    device_x = out;
  }

  // structure::save / load (serialize.cu:135-139): the layout in a file (a
  // versioned binary format of this library; load checks its invariants).
  void save(const char *path) const { check(wk_layout_save(handle(), path, driver_name.c_str())); }
  void load(const char *path) {
    wk_layout *h = nullptr;
    char name[256];
    check(wk_layout_load(path, &h, name, sizeof name));
    wk_layout_info info;
    const int rc = wk_layout_get_info(h, &info);
    if (rc != WK_OK || info.variant != variant()) {
      wk_layout_destroy(h);
      throw error(rc != WK_OK ? rc : WK_ERR_ARG, "layout file holds the other structure variant");
    }
    group_width_ = info.group_width;
    driver_name = name;
    adopt(h);
    rows_ = array<int>(WK_ROWPTR);
    reorder_A_rows = array<int>(WK_REORDER_A_ROWS);
  }
  std::string driver_name;  // hpp:218, stored with the layout

  // The library's layout object of the public fields (rebuilt after store*).
  const wk_layout *handle() const {
    if (stale_) rebuild();
    return h_;
  }
  const std::vector<int> &rowptr() const {
    if (stale_) rebuild();
    return rows_;
  }
  int group_width() const { return group_width_; }

 protected:
  wk_layout *handle_mut() {
    if (stale_) rebuild();
    return h_;
  }
  template <class T>
  std::vector<T> array(int which) const {
    const void *d = nullptr;
    int64_t n = 0;
    check(wk_layout_array(h_, which, &d, &n));
    const T *p = static_cast<const T *>(d);
    return std::vector<T>(p, p + n);
  }
  virtual int variant() const { return 1; }
  virtual void adopt(wk_layout *h) {
    wk_layout_destroy(h_);
    h_ = h;
    stale_ = false;
    wk_layout_info info;
    check(wk_layout_get_info(h_, &info));
    nrows = info.nrows;
    ncols = info.ncols;
    nznrows = info.nznrows;
    nz = (uint)info.nz;
    allocate_nz = (uint)info.allocate_nz;
    nwarps = info.nwarps;
    row_map = array<int>(WK_ROW_MAP);
    row_map_inv = array<int>(WK_ROW_MAP_INV);
    reorder_rows = array<uint>(WK_REORDER_ROWS);
    warp_offsets = array<int64_t>(WK_WARP_OFFSETS);
    max_nz_rows = array<uint>(WK_MAX_NZ_ROWS);
  }
  // Group g and sorted row of padded slot s (variant 1: lane = row).
  virtual uint sorted_row(uint g, uint lane) const { return g * (uint)group_width_ + lane; }
  virtual std::vector<uint> tpr_array() const { return {}; }
  virtual std::vector<uint> row_off_array() const { return {}; }
  virtual int32_t threshold_value() const { return 0; }
  virtual int32_t max_tpr_value() const { return 32; }
  // The layout object from the public fields.  Without scanned row offsets,
  // each row's length is counted from the slots its entries occupy.
  void rebuild() const {
    std::vector<int> rows = rows_;
    if (rows.size() != (size_t)nrows + 1 || (nrows && (uint)rows[nrows] != nz)) {
      std::vector<int64_t> len((size_t)nrows, 0);
      for (uint i = 0; i < nz; i++) {
        const int64_t s = reorder_rows[i];
        const auto it = std::upper_bound(warp_offsets.begin(), warp_offsets.end(), s);
        if (it == warp_offsets.begin() || it == warp_offsets.end())
          throw error(WK_ERR_CSR, "layout arrays: reorder_rows outside the groups");
        const uint g = (uint)(it - warp_offsets.begin() - 1);
        const uint p = sorted_row(g, (uint)((s - warp_offsets[g]) % group_width_));
        if (p >= row_map.size()) throw error(WK_ERR_CSR, "layout arrays: slot of no row");
        len[row_map[p]]++;
      }
      rows.assign((size_t)nrows + 1, 0);
      for (uint r = 0; r < nrows; r++) rows[r + 1] = (int)(rows[r] + len[r]);
    }
    const std::vector<uint> tpr = tpr_array(), roff = row_off_array();
    wk_layout *h = nullptr;
    check(wk_layout_from_arrays(variant(), group_width_, threshold_value(), max_tpr_value(),
                                (int32_t)nrows, (int32_t)(ncols ? ncols : nrows), (int32_t)nznrows,
                                (int32_t)nwarps, nz, allocate_nz, rows.data(), row_map.data(),
                                reorder_rows.data(), warp_offsets.data(), max_nz_rows.data(),
                                tpr.empty() ? nullptr : tpr.data(),
                                roff.empty() ? nullptr : roff.data(), &h));
    wk_layout_destroy(h_);
    h_ = h;
    rows_ = rows;
    stale_ = false;
  }
  static void scatter(const void *src, size_t n, device_vector<uint> &index, void *dst, size_t dst_n,
                      int elem_bytes) {
    if (index.size() < n) throw error(WK_ERR_ARG, "scatter: fewer indices than values");
    check(wk_scatter(src, index.data(), dst, (int64_t)n, (int64_t)dst_n, elem_bytes, nullptr));
    hip_check(hipDeviceSynchronize());
  }

  int group_width_;
  mutable wk_layout *h_ = nullptr;
  mutable std::vector<int> rows_;
  mutable bool stale_ = false;
};

// WPK2 layout (hpp:553-771).
class structure2 : public structure {
 public:
  explicit structure2(int group_width = 64, int max_tpr = 32)
      : structure(group_width), max_tpr_(max_tpr) {}
  // structure2(nrows, nz, alloc, threshold) (hpp:558-560)
  structure2(const uint &nrows_, const uint &nz_, const uint &alloc_, const uint &threshold_,
             int group_width = 64, int max_tpr = 32)
      : structure(nrows_, nz_, alloc_, group_width), nmax_per_thread(threshold_), max_tpr_(max_tpr) {}

  std::vector<uint> row_offset_warp;
  std::vector<uint> threads_per_row;
  uint nmax_per_thread = 0;  // nz threshold per thread
  uint min_nz = 0, max_nz = 0;

  // store_param (hpp:567-570)
  void store_param(std::vector<uint> &row_offset_warp_, std::vector<uint> &tpr_) {
    row_offset_warp = row_offset_warp_;
    threads_per_row = tpr_;
    stale_ = true;
  }
  // store (hpp:572-576)
  void store(std::vector<int> &row_map_, std::vector<uint> &reorder_rows_,
             std::vector<uint> &row_offsets_, std::vector<uint> &tpr_) {
    structure::store(row_map_, reorder_rows_);
    store_param(row_offsets_, tpr_);
  }

  // structure2::scan(nz, nrows, A, rows, colinds, threshold) (hpp:578-736).
  template <typename ValueType, typename IndexType>
  void scan(uint &nz_, uint &nrows_, ValueType *A, IndexType *rows, IndexType *colinds,
            IndexType threshold) {
    static_assert(std::is_same<IndexType, int>::value, "IndexType must be int");
    (void)A;
    nmax_per_thread = (uint)threshold;
    rows_.assign(rows, rows + nrows_ + 1);
    wk_layout *h = nullptr;
    check(wk_scan2(group_width_, threshold, max_tpr_, (int32_t)nrows_,
                   (int32_t)(ncols ? ncols : nrows_), nz_, rows_.data(), colinds, &h));
    adopt(h);
  }

  // scan(nz, nrows, A, threshold) on a CSR container (hpp:738-747).
  template <typename CSR, typename IndexType>
  auto scan(uint &nz_, uint &nrows_, CSR &A, IndexType threshold)
      -> decltype((void)A.row_offsets, (void)A.column_indices, (void)A.values, void()) {
    ncols = (uint)A.num_cols;
    scan(nz_, nrows_, A.values.data(), A.row_offsets.data(), A.column_indices.data(), (int)threshold);
  }

  // structure2::process (hpp:749-769): the five arrays of structure::process,
  // then the lanes per row and the first row of every group.
  template <typename IndexType, typename ValueType, class DV, class DC, class DR, class DM, class DW,
            class DT, class DO>
  void process(const ValueType *A, const IndexType *colinds, DV &device_values, DC &device_colinds,
               DR &device_row_map, DM &device_max_nz_per_row, DW &device_warp_offsets,
               DT &device_threads_per_row, DO &device_row_offset_warp) {
    structure::process(A, colinds, device_values, device_colinds, device_row_map,
                       device_max_nz_per_row, device_warp_offsets);
    assign_array(device_threads_per_row, threads_per_row, "threads_per_row");
    assign_array(device_row_offset_warp, row_offset_warp, "row_offset_warp");
  }
  using structure::process;

 protected:
  int variant() const override { return 2; }
  void adopt(wk_layout *h) override {
    structure::adopt(h);
    wk_layout_info info;
    check(wk_layout_get_info(h_, &info));
    min_nz = info.min_nz;
    max_nz = info.max_nz;
    row_offset_warp = array<uint>(WK_ROW_OFFSET_WARP);
    threads_per_row = array<uint>(WK_THREADS_PER_ROW);
  }
  uint sorted_row(uint g, uint lane) const override {
    if (g >= threads_per_row.size() || g >= row_offset_warp.size() || !threads_per_row[g])
      throw error(WK_ERR_CSR, "layout arrays: group arrays");
    return row_offset_warp[g] + lane / threads_per_row[g];
  }
  std::vector<uint> tpr_array() const override { return threads_per_row; }
  std::vector<uint> row_off_array() const override { return row_offset_warp; }
  int32_t threshold_value() const override { return (int32_t)nmax_per_thread; }
  int32_t max_tpr_value() const override { return max_tpr_; }

 private:
  int max_tpr_;
};

// Device arrays + kernel launches (hpp:773-959).
template <typename ValueType, typename IndexType>
class engine {
  static_assert(std::is_same<IndexType, int>::value, "IndexType must be int");
  static_assert(std::is_same<ValueType, double>::value || std::is_same<ValueType, float>::value,
                "ValueType must be double or float");

 public:
  // The reference's public device arrays (hpp:782-788), as views of the
  // engine.  Reading one copies it to the host (or to a device_vector).
  // Assigning: values (padded order, pads 0) and columns replace the
  // engine's, or, for an engine made by engine(structure&), become the
  // arrays it is built from at its first launch; the layout arrays must equal
  // the structure's.
  template <class T, int Which>
  class device_array {
   public:
    typedef T value_type;
    explicit device_array(engine *e) : e_(e) {}
    device_array(const device_array &) = delete;
    device_array &operator=(const device_array &o) { return *this = std::vector<T>(o); }
    device_array &operator=(const std::vector<T> &h) {
      e_->assign(Which, h);
      return *this;
    }
    device_array &operator=(const device_vector<T> &d) { return *this = std::vector<T>(d); }
    operator std::vector<T>() const { return e_->template read<T>(Which); }
    operator device_vector<T>() const { return device_vector<T>(std::vector<T>(*this)); }
    size_t size() const { return std::vector<T>(*this).size(); }

   private:
    engine *e_;
  };

  device_array<ValueType, WK_ENGINE_VALUES> device_values{this};                 // nz
  device_array<IndexType, WK_ENGINE_COLS> device_colinds{this};                  // nz
  device_array<IndexType, WK_ROW_MAP> device_row_map{this};                      // nrows
  device_array<uint, WK_MAX_NZ_ROWS> device_max_nz_per_row{this};                // nwarps
  device_array<IndexType, WK_WARP_OFFSETS> device_warp_offsets{this};            // nwarps + 1
  device_array<uint, WK_THREADS_PER_ROW> device_threads_per_row{this};           // nwarps
  device_array<uint, WK_ROW_OFFSET_WARP> device_row_offset_warp{this};           // nwarps + 1

  // engine(structure&) (hpp:790-792): no device arrays yet; they are assigned
  // (for instance by k.process(A, colinds, eng.device_values, ...)) before the
  // first launch.
  explicit engine(structure &k) : kernel(&k) {}

  // engine(structure&, A, colinds) and engine(structure2&, A, colinds)
  // (hpp:794-809): process and upload.  flags: WK_ENGINE_KR, WK_ENGINE_ADJOINT.
  engine(structure &k, const ValueType *A, const IndexType *colinds, uint32_t flags = 0)
      : kernel(&k) {
    create(A, colinds, flags);
  }
  engine(const engine &) = delete;
  engine &operator=(const engine &) = delete;
  ~engine() { wk_engine_destroy(h_); }

  // engine::run<cache> (hpp:811-860): one launch, blocking, returns seconds.
  template <bool cache>
  float run(int nblocks, int blocksize, const ValueType *dx, double *dy) {
    return timed<cache>(0, nblocks, blocksize, dx, dy);
  }
  // engine::run_x<cache> (hpp:862-896): the Kr form.
  template <bool cache>
  float run_x(int nblocks, int blocksize, const ValueType *dx, double *dy) {
    return timed<cache>(1, nblocks, blocksize, dx, dy);
  }
  // engine::run_assembly<cache> (hpp:898-929): row sums of the padded values.
  template <bool cache>
  float run_assembly(int nblocks, int blocksize, double *dy) {
    return timed<cache>(2, nblocks, blocksize, nullptr, dy);
  }

  // Asynchronous forms with k right-hand sides (x [ncols][ldx], y [nrows][ldy]).
  void launch(const ValueType *dx, double *dy, int k = 1, int64_t ldx = 1, int64_t ldy = 1,
              hipStream_t s = nullptr, const wk_launch *cfg = nullptr) {
    check(wk_run(handle(), cfg, dx, dy, k, ldx, ldy, s));
  }
  void launch_x(const ValueType *dx, double *dy, int k = 1, int64_t ldx = 1, int64_t ldy = 1,
                hipStream_t s = nullptr, const wk_launch *cfg = nullptr) {
    check(wk_run_x(handle(), cfg, dx, dy, k, ldx, ldy, s));
  }
  wk_launch tune(int op, const ValueType *dx, double *dy, int k = 1, int64_t ldx = 1,
                 int64_t ldy = 1, int reps = 10, hipStream_t s = nullptr) {
    wk_launch best;
    float ms = 0;
    check(wk_tune(handle(), op, nullptr, dx, dy, k, ldx, ldy, reps, s, &best, &ms));
    return best;
  }

  // engine::verify (hpp:932-943): relative 1e-5 where |orig| > 1e-14, absolute
  // 1e-14 elsewhere; prints and stops at the first failure.
  template <typename R>
  bool verify(R orig, R comp) {
    for (uint i = 0; i < kernel->nrows; i++)
      if (bad(orig[i], comp[i])) {
        std::cout << "verify: " << orig[i] << "\t" << comp[i] << "\t" << i << std::endl;
        return false;
      }
    return true;
  }
  // engine::verify_x (hpp:946-957): orig[row_map[i]] against comp[i] (in both
  // tests; the reference's zero test reads orig[i]).
  template <typename R>
  bool verify_x(R orig, R comp) {
    for (uint i = 0; i < kernel->nrows; i++)
      if (bad(orig[kernel->row_map[i]], comp[i])) {
        std::cout << "verify_x: " << orig[kernel->row_map[i]] << "\t" << comp[i] << "\t" << i << std::endl;
        return false;
      }
    return true;
  }

  // The library's engine (built now from the assigned arrays for an engine
  // made by engine(structure&)).
  wk_engine *handle() {
    if (!h_) build_from_arrays();
    return h_;
  }
  structure *kernel;

 private:
  static bool bad(double o, double c) {
    return (std::fabs(o) <= 1e-14 && std::fabs(o - c) > 1e-14) ||
           (std::fabs(o) > 1e-14 && std::fabs((o - c) / o) > 1e-5);
  }
  void create(const ValueType *A, const IndexType *colinds, uint32_t flags) {
    std::vector<double> vals(A, A + kernel->nz);
    check(wk_engine_create(kernel->handle(), kernel->rowptr().data(), colinds, vals.data(),
                           std::is_same<ValueType, double>::value ? WK_F64 : WK_F32, flags,
                           nullptr, &h_));
  }
  // The CSR of padded arrays (pads must hold 0): entry i sits at reorder_rows[i].
  void build_from_arrays() {
    const size_t an = kernel->allocate_nz;
    if (staged_vals_.size() != an || staged_cols_.size() != an)
      throw error(WK_ERR_STATE, "engine(structure&) needs device_values and device_colinds "
                                "(allocate_nz entries each) before its first launch");
    std::vector<ValueType> vals(kernel->nz);
    std::vector<IndexType> cols(kernel->nz);
    for (uint i = 0; i < kernel->nz; i++) {
      vals[i] = staged_vals_[kernel->reorder_rows[i]];
      cols[i] = staged_cols_[kernel->reorder_rows[i]];
    }
    check_pads(staged_vals_);
    create(vals.data(), cols.data(), 0);
    staged_vals_.clear();
    staged_cols_.clear();
  }
  // the kernels treat pads as (value 0, column 0); other pads are refused
  void check_pads(const std::vector<ValueType> &pv) const {
    std::vector<char> entry(pv.size(), 0);
    for (uint i = 0; i < kernel->nz; i++) entry[kernel->reorder_rows[i]] = 1;
    for (size_t s = 0; s < pv.size(); s++)
      if (!entry[s] && pv[s] != ValueType(0)) throw error(WK_ERR_ARG, "padded slots must hold 0");
  }
  template <class T>
  void assign(int which, const std::vector<T> &h) {
    if (which == WK_ENGINE_VALUES) {
      if (h.size() != kernel->allocate_nz) throw error(WK_ERR_ARG, "padded value array has the wrong length");
      std::vector<ValueType> pv(h.begin(), h.end());
      if (!h_) {
        staged_vals_ = pv;
        return;
      }
      check_pads(pv);
      // back to CSR order, then the library's device scatter
      std::vector<ValueType> vals(kernel->nz);
      for (uint i = 0; i < kernel->nz; i++) vals[i] = pv[kernel->reorder_rows[i]];
      device_vector<ValueType> d(vals);
      check(wk_engine_update_values(h_, d.data(), nullptr));
      hip_check(hipDeviceSynchronize());
    } else if (which == WK_ENGINE_COLS) {
      if (h.size() != kernel->allocate_nz) throw error(WK_ERR_ARG, "padded column array has the wrong length");
      std::vector<IndexType> pc(h.begin(), h.end());
      if (!h_) {
        staged_cols_ = pc;
        return;
      }
      check(wk_engine_set_colinds(h_, pc.data(), 0, nullptr));
    } else {
      // the layout arrays are the structure's
      if (narrow_copy<int64_t>(h, "array") != narrow_copy<int64_t>(read<T>(which), "array"))
        throw error(WK_ERR_ARG, "an engine's layout arrays are the structure's and cannot be replaced");
    }
  }
  template <class T>
  std::vector<T> read(int which) const {
    if (!h_) {
      // no device engine yet: what was assigned, or the structure's arrays
      if (which == WK_ENGINE_VALUES) return std::vector<T>(staged_vals_.begin(), staged_vals_.end());
      if (which == WK_ENGINE_COLS) return std::vector<T>(staged_cols_.begin(), staged_cols_.end());
      structure2 *k2 = dynamic_cast<structure2 *>(kernel);
      switch (which) {
        case WK_ROW_MAP: return narrow_copy<T>(kernel->row_map, "row_map");
        case WK_MAX_NZ_ROWS: return narrow_copy<T>(kernel->max_nz_rows, "max_nz_rows");
        case WK_WARP_OFFSETS:
          return narrow_copy<T>(static_cast<const std::vector<int64_t> &>(kernel->warp_offsets),
                                "warp_offsets");
        case WK_THREADS_PER_ROW: return k2 ? narrow_copy<T>(k2->threads_per_row, "tpr") : std::vector<T>();
        case WK_ROW_OFFSET_WARP: return k2 ? narrow_copy<T>(k2->row_offset_warp, "row_offset_warp") : std::vector<T>();
      }
      return {};
    }
    int64_t n = 0;
    check(wk_engine_download(h_, which, nullptr, &n));
    if (which == WK_WARP_OFFSETS) {
      std::vector<int64_t> w((size_t)n);
      if (n) check(wk_engine_download(h_, which, w.data(), &n));
      return narrow_copy<T>(w, "warp_offsets");
    }
    std::vector<T> out((size_t)n);
    if (n) check(wk_engine_download(h_, which, out.data(), &n));
    return out;
  }
  template <bool cache>
  float timed(int which, int nblocks, int blocksize, const ValueType *dx, double *dy) {
    (void)nblocks;
    if (blocksize <= 0 || blocksize % 64) throw error(WK_ERR_ARG, "blocksize must be a multiple of 64");
    wk_launch c;
    check(wk_engine_get_launch(handle(), which == 1 ? WK_OP_RUN_X : WK_OP_RUN, 1, &c));
    c.waves_per_block = blocksize / 64 > 8 ? 8 : blocksize / 64;
    if (!cache) {
      // plain loads for everything: the direct kernel, one row per lane
      c.nontemporal = 0;
      c.stage_lds = 0;
      c.unroll = 8;
      c.rows_per_lane = 1;
    }
    float sec = 0;
    check(wk_time_run(h_, &c, which, dx, dy, 1, 1, 1, nullptr, &sec));
    return sec;
  }
  wk_engine *h_ = nullptr;
  std::vector<ValueType> staged_vals_;
  std::vector<IndexType> staged_cols_;
};

}  // namespace warpkernel
