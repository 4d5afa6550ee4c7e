// warpkernel.hpp -- drop-in C++ API for the MI355X WarpKernel SpMV library.
//
// Rebuilds the reference's class API (namespace warpkernel, warpkernel.hpp:35-1024
// of the reference) on top of the C ABI in warpkernel_c.h.  Same class names,
// member names and argument order:
//
//   warpkernel::structure    scan / process / reorder_columns_coalesced / reorder_x,
//                            public fields row_map, row_map_inv, reorder_rows,
//                            warp_offsets, max_nz_rows, nrows, nznrows, nz,
//                            allocate_nz, nwarps                  (hpp:215-551)
//   warpkernel::structure2   scan(..., threshold); row_offset_warp, threads_per_row,
//                            nmax_per_thread, min_nz, max_nz      (hpp:553-771)
//   warpkernel::engine<V,I>  engine(k, A, colinds); run<cache>; run_x<cache>;
//                            verify; verify_x; device_colinds    (hpp:773-959)
//
// What changes for a caller of the reference:
//   - cusp containers become std::vector (host) and warpkernel::device_vector
//     (device, RAII over hipMalloc);
//   - the group width is a constructor argument: 64 (one gfx950 wavefront, the
//     default) or 32 (the reference's WARP_SIZE, identical layouts);
//   - warp_offsets are int64_t;
//   - run / run_x take a blocksize that is a multiple of 64 (waves of 64);
//     nblocks is accepted and ignored (the grid follows from the layout);
//   - run<cache>: cache = true streams the matrix with nontemporal loads and
//     keeps x in the caches, cache = false uses plain loads for everything
//     (the reference's texture cache for x has no counterpart on CDNA);
//   - rows without nonzeros are written as 0;
//   - errors throw warpkernel::error instead of asserting.
// engine<float, int> stores FP32 values and takes FP32 x; y is always double.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
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

// Minimal device array (the reference's cusp::array1d<T, device_memory>).
template <class T>
class device_vector {
 public:
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
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) hip_check(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)));
    n_ = n;
  }
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

// WPK1 layout (hpp:215-551).
class structure {
 public:
  explicit structure(int group_width = 64) : group_width_(group_width) {}
  structure(const structure &) = delete;
  structure &operator=(const structure &) = delete;
  virtual ~structure() { wk_layout_destroy(h_); }

  std::vector<int> row_map;      // y[row_map[i]] = y'[i]
  std::vector<int> row_map_inv;  // new_colinds[n] = row_map_inv[colinds[n]]
  std::vector<uint> reorder_rows;  // new_values[reorder_rows[i]] = A[i]
  std::vector<int64_t> warp_offsets;
  std::vector<uint> max_nz_rows;
  std::vector<int> reorder_A_rows;  // after reorder_columns_rowsort (hpp:279)
  uint nrows = 0, nznrows = 0, nz = 0, allocate_nz = 0, nwarps = 0;
  uint ncols = 0;

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

  // structure::reorder / process on the host (hpp:379-398): padded arrays.
  template <typename IndexType, typename ValueType>
  void process(const ValueType *A, const IndexType *colinds, std::vector<ValueType> &new_values,
               std::vector<IndexType> &new_colinds) const {
    std::vector<double> vals(A, A + nz), pv(allocate_nz);
    new_colinds.assign(allocate_nz, 0);
    check(wk_layout_process_host(h_, vals.data(), colinds, pv.data(), new_colinds.data()));
    new_values.assign(pv.begin(), pv.end());
  }

  // reorder_columns_coalesced (hpp:446-457): colinds becomes the padded Kr columns.
  template <typename IndexType>
  void reorder_columns_coalesced(std::vector<IndexType> &colinds) const {
    std::vector<IndexType> out(allocate_nz);
    check(wk_layout_reorder_columns_coalesced(h_, colinds.data(), out.data()));
    colinds.swap(out);
  }

  // reorder_columns_rowsort (hpp:414-444): colinds (CSR order, nz entries)
  // becomes the permuted columns with each row ascending; reorder_A_rows[n]
  // is the CSR index of the entry now at n.  Equal columns keep CSR order.
  template <typename IndexType>
  void reorder_columns_rowsort(std::vector<IndexType> &colinds, const std::vector<IndexType> &rows) {
    static_assert(std::is_same<IndexType, int>::value, "IndexType must be int");
    std::vector<IndexType> out(colinds.size());
    check(wk_layout_reorder_columns_rowsort(h_, rows.data(), colinds.data(), out.data()));
    colinds.swap(out);
    reorder_A_rows = array<int>(WK_REORDER_A_ROWS);
  }

  // reorder_values_coalesced_rowsort (hpp:505-515): padded values,
  // dvalues[reorder_rows[i]] = values[reorder_A_rows[i]].
  template <typename ValueType>
  void reorder_values_coalesced_rowsort(const std::vector<ValueType> &values,
                                        device_vector<ValueType> &dvalues) const {
    std::vector<double> v(values.begin(), values.end()), pv(allocate_nz);
    check(wk_layout_reorder_values_rowsort(h_, v.data(), pv.data()));
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
  void save(const char *path) const { check(wk_layout_save(h_, path, driver_name.c_str())); }
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

  const wk_layout *handle() const { return h_; }
  const std::vector<int> &rowptr() const { return rows_; }
  int group_width() const { return group_width_; }

 protected:
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

  int group_width_;
  wk_layout *h_ = nullptr;
  std::vector<int> rows_;
};

// WPK2 layout (hpp:553-771).
class structure2 : public structure {
 public:
  explicit structure2(int group_width = 64, int max_tpr = 32)
      : structure(group_width), max_tpr_(max_tpr) {}

  std::vector<uint> row_offset_warp;
  std::vector<uint> threads_per_row;
  uint nmax_per_thread = 0;
  uint min_nz = 0, max_nz = 0;

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
  // The reference's public device_colinds, which its drivers swap for the Kr
  // columns and back (wpk1.cu:160-197).
  class colinds_proxy {
   public:
    explicit colinds_proxy(engine *e) : e_(e) {}
    colinds_proxy &operator=(const std::vector<IndexType> &h) {
      if (h.size() != e_->kernel->allocate_nz) throw error(WK_ERR_ARG, "padded column array has the wrong length");
      check(wk_engine_set_colinds(e_->h_, h.data(), 0, nullptr));
      return *this;
    }
    colinds_proxy &operator=(const device_vector<IndexType> &d) {
      check(wk_engine_set_colinds(e_->h_, d.data(), 1, nullptr));
      return *this;
    }
    operator std::vector<IndexType>() const {
      std::vector<IndexType> h(e_->kernel->allocate_nz);
      check(wk_engine_get_colinds(e_->h_, h.data()));
      return h;
    }

   private:
    engine *e_;
  };

  colinds_proxy device_colinds{this};

  // engine(structure&, A, colinds) and engine(structure2&, A, colinds)
  // (hpp:794-809): process and upload.  flags: WK_ENGINE_KR, WK_ENGINE_ADJOINT.
  engine(structure &k, const ValueType *A, const IndexType *colinds, uint32_t flags = 0)
      : kernel(&k) {
    std::vector<double> vals(A, A + k.nz);
    check(wk_engine_create(k.handle(), k.rowptr().data(), colinds, vals.data(),
                           std::is_same<ValueType, double>::value ? WK_F64 : WK_F32, flags,
                           nullptr, &h_));
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
    check(wk_run(h_, cfg, dx, dy, k, ldx, ldy, s));
  }
  void launch_x(const ValueType *dx, double *dy, int k = 1, int64_t ldx = 1, int64_t ldy = 1,
                hipStream_t s = nullptr, const wk_launch *cfg = nullptr) {
    check(wk_run_x(h_, cfg, dx, dy, k, ldx, ldy, s));
  }
  wk_launch tune(int op, const ValueType *dx, double *dy, int k = 1, int64_t ldx = 1,
                 int64_t ldy = 1, int reps = 10, hipStream_t s = nullptr) {
    wk_launch best;
    float ms = 0;
    check(wk_tune(h_, op, nullptr, dx, dy, k, ldx, ldy, reps, s, &best, &ms));
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

  wk_engine *handle() { return h_; }
  structure *kernel;

 private:
  static bool bad(double o, double c) {
    return (std::fabs(o) <= 1e-14 && std::fabs(o - c) > 1e-14) ||
           (std::fabs(o) > 1e-14 && std::fabs((o - c) / o) > 1e-5);
  }
  template <bool cache>
  float timed(int which, int nblocks, int blocksize, const ValueType *dx, double *dy) {
    (void)nblocks;
    if (blocksize <= 0 || blocksize % 64) throw error(WK_ERR_ARG, "blocksize must be a multiple of 64");
    wk_launch c;
    check(wk_engine_get_launch(h_, which == 1 ? WK_OP_RUN_X : WK_OP_RUN, 1, &c));
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
};

}  // namespace warpkernel
