// Drop-in check: the reference drivers' flow (wpk1.cu:114-242, wpk2.cu:136-290)
// written against include/warpkernel.hpp, checked against the oracle.
//   dropin_wpk [n]   3D hex elasticity on an n^3 element grid (default 16)
// Exit status 0 when every result matches; 1 otherwise.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/warpkernel.hpp"
#include "../../oracle/wk_oracle.h"

extern "C" {
int64_t fe_q1_rowptr(int dim, int ex, int ey, int ez, int ndof, int32_t *rowptr);
void fe_q1_fill(int dim, int ex, int ey, int ez, int ndof, double nu, const double *coef,
                const int32_t *rowptr, int32_t *colind, double *vals);
}

using warpkernel::uint;

static int failures = 0;
static void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    failures++;
  }
}

// The drivers' environment: their cusp / thrust names over this library's
// containers (the drivers define ValueType and friends after their includes,
// wpk1.cu:17-21).
namespace thrust {
template <class T>
T *raw_pointer_cast(T *p) {
  return p;
}
}  // namespace thrust
namespace cusp {
struct host_memory {};
struct device_memory {};
template <class T, class Space>
using array1d = typename std::conditional<std::is_same<Space, device_memory>::value,
                                          warpkernel::device_vector<T>, std::vector<T>>::type;
template <class I, class V, class Space>
using csr_matrix = warpkernel::csr_matrix;
}  // namespace cusp
#define ValueType double
#define IndexType int
#define DeviceSpace cusp::device_memory
#define CPUSpace cusp::host_memory

template <class T>
static std::vector<T> host_of(const warpkernel::device_vector<T> &d) {
  return std::vector<T>(d);
}

// wpk1.cu:118-135, wpk2.cu:136-165 and serialize.cu:76-98, their lines kept as
// the drivers have them, then every array checked against the oracle.
static void drivers_verbatim(const std::vector<int> &rowptr, const std::vector<int> &colind,
                             const std::vector<double> &vals, const std::vector<double> &xin,
                             const std::vector<double> &y, uint M, uint N, uint nz) {
  cusp::csr_matrix<IndexType, ValueType, CPUSpace> A;
  A.num_rows = M;
  A.num_cols = N;
  A.num_entries = nz;
  A.row_offsets = rowptr;
  A.column_indices = colind;
  A.values = vals;
  cusp::array1d<ValueType, CPUSpace> x = xin;
  wko_layout *L1 = wko_scan1(64, (int32_t)M, nz, rowptr.data());
  std::vector<double> opv(L1->allocate_nz);
  std::vector<int32_t> opc(L1->allocate_nz);
  wko_process(L1, vals.data(), colind.data(), opv.data(), opc.data());

  // wpk1.cu:120-135
  {
    cusp::array1d<ValueType, DeviceSpace> dx = x;
    warpkernel::structure kernel1;
    kernel1.scan(nz, M, A);

    cusp::array1d<IndexType, CPUSpace> reorder_cols = A.column_indices;
    kernel1.reorder_columns_coalesced(reorder_cols);

    cusp::array1d<ValueType, DeviceSpace> dreordered_x;
    kernel1.reorder_x(x, dreordered_x);

    warpkernel::engine<ValueType, IndexType> eng(kernel1,
						 &(A.values[0]),
						 &(A.column_indices[0]));
    uint warps_per_block = 4;
    uint nblocks = (kernel1.nwarps + warps_per_block - 1) / warps_per_block;
    uint blocksize = warps_per_block * 64;
    cusp::array1d<ValueType, DeviceSpace> dy(N, 0);
    eng.run<true>(nblocks, blocksize, dx.data(), dy.data());
    expect(host_of(dy) == y, "wpk1.cu flow: scan(nz, M, A), engine, run bit-exact");
    expect(std::vector<uint>(L1->reorder_rows, L1->reorder_rows + nz) == kernel1.reorder_rows,
           "wpk1.cu flow: layout matches the oracle");
    // warp_offsets into the reference's std::vector<int>
    std::vector<int> wo = kernel1.warp_offsets;
    expect(std::vector<int64_t>(wo.begin(), wo.end()) ==
               std::vector<int64_t>(L1->warp_offsets, L1->warp_offsets + L1->nwarps + 1),
           "warp_offsets as std::vector<int>");
    // the engine's public device arrays (hpp:782-788)
    cusp::array1d<IndexType, DeviceSpace> restore_col = eng.device_colinds;
    expect(host_of(restore_col) == opc, "engine.device_colinds read");
    std::vector<double> dv = eng.device_values;
    expect(dv == opv, "engine.device_values read");
    std::vector<int> rm = eng.device_row_map, dwo = eng.device_warp_offsets;
    std::vector<uint> mx = eng.device_max_nz_per_row;
    expect(rm == kernel1.row_map && dwo == wo && mx == kernel1.max_nz_rows, "engine layout arrays read");
    // new values through the public array (padded order) then run
    for (auto &v : dv) v *= 2.0;
    eng.device_values = dv;
    cusp::array1d<ValueType, DeviceSpace> dy2(N, 0);
    eng.run<true>(nblocks, blocksize, dx.data(), dy2.data());
    std::vector<double> y2 = host_of(dy2);
    bool twice = true;
    for (uint i = 0; i < M; i++) twice = twice && y2[i] == 2.0 * y[i];
    expect(twice, "engine.device_values = 2 A: y doubles exactly");
    try {
      std::vector<int> bad_map = kernel1.row_map;
      std::swap(bad_map[0], bad_map[1]);
      eng.device_row_map = bad_map;
      expect(false, "a layout array other than the structure's is refused");
    } catch (const warpkernel::error &) {
    }
  }

  // serialize.cu:79-98: process into five device arrays
  {
    warpkernel::structure kernel1;
    kernel1.scan(nz, N, A);

    // allocate arrays
    cusp::array1d<ValueType, DeviceSpace> device_values; // nz
    cusp::array1d<IndexType, DeviceSpace> device_colinds; // nz
    cusp::array1d<IndexType, DeviceSpace> device_row_map; // nrows
    cusp::array1d<uint, DeviceSpace> device_max_nz_per_row; // nwarps
    cusp::array1d<IndexType, DeviceSpace> device_warp_offsets; // nwarps
    cusp::array1d<ValueType, DeviceSpace> dx = x;
    cusp::array1d<ValueType, DeviceSpace> dy(N);

    kernel1.process(thrust::raw_pointer_cast(&A.values[0]),
    		    thrust::raw_pointer_cast(&A.column_indices[0]),
    		    device_values, device_colinds,
    		    device_row_map, device_max_nz_per_row,
    		    device_warp_offsets);

    expect(host_of(device_values) == opv && host_of(device_colinds) == opc,
           "serialize.cu flow: padded arrays match the oracle's process");
    expect(host_of(device_row_map) == std::vector<int>(L1->row_map, L1->row_map + M),
           "serialize.cu flow: row map");
    expect(host_of(device_max_nz_per_row) == std::vector<uint>(L1->max_nz_rows, L1->max_nz_rows + L1->nwarps),
           "serialize.cu flow: group lengths");
    std::vector<int> wo = host_of(device_warp_offsets);
    expect(std::vector<int64_t>(wo.begin(), wo.end()) ==
               std::vector<int64_t>(L1->warp_offsets, L1->warp_offsets + L1->nwarps + 1),
           "serialize.cu flow: group offsets as int");

    // engine(structure&) (hpp:790-792), filled by process, then run
    warpkernel::engine<ValueType, IndexType> eng(kernel1);
    kernel1.process(thrust::raw_pointer_cast(&A.values[0]),
    		    thrust::raw_pointer_cast(&A.column_indices[0]),
    		    eng.device_values, eng.device_colinds,
    		    eng.device_row_map, eng.device_max_nz_per_row,
    		    eng.device_warp_offsets);
    eng.run<true>((kernel1.nwarps + 3) / 4, 256, dx.data(), dy.data());
    expect(host_of(dy) == y, "engine(structure&) + process into its arrays: run bit-exact");

    // reorder_columns (hpp:401-411) and the _GPU scatters (hpp:459-502)
    cusp::array1d<IndexType, CPUSpace> kr = A.column_indices;
    kernel1.reorder_columns(kr);
    bool perm = true;
    for (uint i = 0; i < nz; i++) perm = perm && kr[i] == kernel1.row_map_inv[A.column_indices[i]];
    expect(perm, "reorder_columns");
    cusp::array1d<IndexType, CPUSpace> krc = A.column_indices;
    kernel1.reorder_columns_coalesced(krc);
    cusp::array1d<IndexType, DeviceSpace> dkr = kr, dkrc(kernel1.allocate_nz, 0);
    cusp::array1d<uint, DeviceSpace> drr = kernel1.reorder_rows;
    kernel1.reorder_columns_coalesced_GPU(dkr, dkrc, drr);
    expect(host_of(dkrc) == krc, "reorder_columns_coalesced_GPU equals the host form");
    cusp::array1d<ValueType, DeviceSpace> dval1, dval2(kernel1.allocate_nz, 0.0);
    kernel1.reorder_values_coalesced(A.values, dval1);
    cusp::array1d<ValueType, DeviceSpace> dvals = A.values;
    kernel1.reorder_values_coalesced_GPU(dvals, dval2, drr);
    expect(host_of(dval1) == opv && host_of(dval2) == opv, "reorder_values_coalesced (host and _GPU)");

    // store / storemap / storereorder (hpp:288-307): a structure rebuilt from
    // its saved maps plus the group fields, then an engine on it
    warpkernel::structure restored(kernel1.nrows, kernel1.nz, kernel1.allocate_nz);
    std::vector<int> map = kernel1.row_map;
    std::vector<uint> rr = kernel1.reorder_rows;
    restored.store(map, rr);
    restored.warp_offsets = kernel1.warp_offsets;
    restored.max_nz_rows = kernel1.max_nz_rows;
    restored.nwarps = kernel1.nwarps;
    restored.nznrows = kernel1.nznrows;
    warpkernel::engine<ValueType, IndexType> eng2(restored, &(A.values[0]), &(A.column_indices[0]));
    cusp::array1d<ValueType, DeviceSpace> dy3(N, 0);
    eng2.run<true>(1, 256, dx.data(), dy3.data());
    expect(host_of(dy3) == y, "structure from store(): run bit-exact");
    // a damaged stored map is refused when used
    warpkernel::structure broken(kernel1.nrows, kernel1.nz, kernel1.allocate_nz);
    std::vector<uint> rr_bad = rr;
    rr_bad[0] = kernel1.allocate_nz + 5;
    broken.store(map, rr_bad);
    broken.warp_offsets = kernel1.warp_offsets;
    broken.max_nz_rows = kernel1.max_nz_rows;
    broken.nwarps = kernel1.nwarps;
    broken.nznrows = kernel1.nznrows;
    try {
      warpkernel::engine<ValueType, IndexType> e3(broken, &(A.values[0]), &(A.column_indices[0]));
      expect(false, "store() with a slot outside the padded arrays is refused");
    } catch (const warpkernel::error &) {
    }
  }

  // wpk2.cu:139-165 with structure2::process into seven arrays (hpp:749-769)
  {
    cusp::array1d<ValueType, DeviceSpace> dx = x;

    warpkernel::structure2 kernel2;
    kernel2.scan(nz, N, A, (int)N); // use maximum threshold

    int max_nz = kernel2.max_nz;
    int min_nz = kernel2.min_nz;
    int inc = (max_nz-min_nz)*0.1 > 0 ? (max_nz-min_nz)*0.1 : 1;
    for (int threshold = min_nz; threshold <= max_nz; threshold+=inc) {
	kernel2.scan(nz, N, A, threshold); // use maximum threshold

	cusp::array1d<IndexType, CPUSpace> reorder_cols = A.column_indices;
	kernel2.reorder_columns_coalesced(reorder_cols);

	cusp::array1d<ValueType, DeviceSpace> dreordered_x;
	kernel2.reorder_x(x, dreordered_x);

	warpkernel::engine<ValueType, IndexType> eng(kernel2,
						     &(A.values[0]),
						     &(A.column_indices[0]));

      wko_layout *L2 = wko_scan2(64, threshold, 32, (int32_t)N, nz, rowptr.data());
      std::vector<double> pv(L2->allocate_nz), yw(M);
      std::vector<int32_t> pc(L2->allocate_nz);
      wko_process(L2, vals.data(), colind.data(), pv.data(), pc.data());
      wko_wpk_spmv(L2, pv.data(), pc.data(), xin.data(), yw.data(), 1, 1, 1, 0);
      cusp::array1d<ValueType, DeviceSpace> dy(N, 0);
      eng.run<true>((kernel2.nwarps + 3) / 4, 256, dx.data(), dy.data());
      expect(host_of(dy) == yw, "wpk2.cu flow: scan(nz, N, A, threshold), run bit-exact");

      cusp::array1d<ValueType, DeviceSpace> device_values;
      cusp::array1d<IndexType, DeviceSpace> device_colinds, device_row_map, device_warp_offsets;
      cusp::array1d<uint, DeviceSpace> device_max_nz_per_row, device_threads_per_row,
          device_row_offset_warp;
      kernel2.process(&(A.values[0]), &(A.column_indices[0]), device_values, device_colinds,
                      device_row_map, device_max_nz_per_row, device_warp_offsets,
                      device_threads_per_row, device_row_offset_warp);
      expect(host_of(device_values) == pv && host_of(device_colinds) == pc,
             "structure2::process: padded arrays match the oracle");
      expect(host_of(device_threads_per_row) ==
                     std::vector<uint>(L2->threads_per_row, L2->threads_per_row + L2->nwarps) &&
                 host_of(device_row_offset_warp) ==
                     std::vector<uint>(L2->row_offset_warp, L2->row_offset_warp + L2->nwarps + 1),
             "structure2::process: threads_per_row and row_offset_warp");
      std::vector<uint> tpr = eng.device_threads_per_row, row_off = eng.device_row_offset_warp;
      expect(tpr == kernel2.threads_per_row && row_off == kernel2.row_offset_warp,
             "structure2 engine arrays read");

      // store / store_param (hpp:567-576)
      warpkernel::structure2 restored(kernel2.nrows, kernel2.nz, kernel2.allocate_nz, threshold);
      std::vector<int> map = kernel2.row_map;
      std::vector<uint> rr = kernel2.reorder_rows, roff = kernel2.row_offset_warp,
                        tp = kernel2.threads_per_row;
      restored.store(map, rr, roff, tp);
      restored.warp_offsets = kernel2.warp_offsets;
      restored.max_nz_rows = kernel2.max_nz_rows;
      restored.nwarps = kernel2.nwarps;
      restored.nznrows = kernel2.nznrows;
      warpkernel::engine<ValueType, IndexType> e2(restored, &(A.values[0]), &(A.column_indices[0]));
      cusp::array1d<ValueType, DeviceSpace> dy2(N, 0);
      e2.run<true>(1, 256, dx.data(), dy2.data());
      expect(host_of(dy2) == yw, "structure2 from store(): run bit-exact");
      wko_free(L2);
    }
  }
  wko_free(L1);
}

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 16;
  const int nodes = (n + 1) * (n + 1) * (n + 1);
  uint M = 3 * nodes, N = M;
  std::vector<int> rowptr(M + 1);
  uint nz = (uint)fe_q1_rowptr(3, n, n, n, 3, rowptr.data());
  std::vector<int> colind(nz);
  std::vector<double> vals(nz), E((size_t)n * n * n);
  for (size_t e = 0; e < E.size(); e++) E[e] = 1.0 + 0.5 * ((e * 2654435761u) % 1000) / 1000.0;
  fe_q1_fill(3, n, n, n, 3, 0.3, E.data(), rowptr.data(), colind.data(), vals.data());

  // x as in the reference drivers: rand() % 100 / 100 - 0.3 (wpk1.cu:24-29)
  std::vector<double> x(N);
  for (auto &v : x) v = (double)(std::rand() % 100) / 100. - 0.3;
  std::vector<double> y(M);
  wko_csr_spmv((int32_t)M, rowptr.data(), colind.data(), vals.data(), x.data(), y.data(), 1, 1, 1, 1);
  std::printf("hex elasticity n=%d: %u rows, %u nnz\n", n, M, nz);

  warpkernel::device_vector<double> dx = x;
  for (int W : {64, 32}) {
    warpkernel::structure kernel1(W);
    kernel1.scan(nz, M, vals.data(), rowptr.data(), colind.data());
    std::vector<int> reorder_cols = colind;
    kernel1.reorder_columns_coalesced(reorder_cols);
    warpkernel::device_vector<double> dreordered_x;
    kernel1.reorder_x(x, dreordered_x);
    warpkernel::engine<double, int> eng(kernel1, vals.data(), colind.data());
    for (int warps_per_block = 1; warps_per_block <= 8; warps_per_block *= 2) {
      uint nblocks = (kernel1.nwarps + warps_per_block - 1) / warps_per_block;
      uint blocksize = warps_per_block * 64;
      for (bool cache : {true, false}) {
        warpkernel::device_vector<double> dy(N, 0.);
        float t = cache ? eng.run<true>(nblocks, blocksize, dx.data(), dy.data())
                        : eng.run<false>(nblocks, blocksize, dx.data(), dy.data());
        std::vector<double> ycheck = dy;
        expect(eng.verify(y, ycheck), "WPK1 verify");
        expect(ycheck == y, "WPK1 bit-exact");
        if (cache && warps_per_block == 4) std::printf("W=%d WPK1 %e s\n", W, t);
      }
      std::vector<int> restore_col = eng.device_colinds;
      {
        warpkernel::device_vector<double> dy(N, 0);
        eng.device_colinds = reorder_cols;
        eng.run_x<true>(nblocks, blocksize, dreordered_x.data(), dy.data());
        std::vector<double> ycheck = dy;
        expect(eng.verify_x(y, ycheck), "WPK1_r verify_x");
        for (uint i = 0; i < M; i++)
          if (ycheck[i] != y[kernel1.row_map[i]]) {
            expect(false, "WPK1_r bit-exact");
            break;
          }
        eng.device_colinds = restore_col;
      }
      {
        warpkernel::device_vector<double> dy(N, 0);
        eng.device_colinds = reorder_cols;
        eng.run<true>(nblocks, blocksize, dreordered_x.data(), dy.data());
        std::vector<double> ycheck = dy;
        expect(ycheck == y, "WPK1_rx bit-exact");
        eng.device_colinds = restore_col;
      }
      expect(std::vector<int>(eng.device_colinds) == restore_col, "colinds restored");
    }
  }

  // MatrixMarket in, layout saved and loaded again (wpk1.cu:43-48, serialize.cu:130-160)
  {
    const char *tmp = std::getenv("TMPDIR");
    const std::string dir = tmp && *tmp ? tmp : "/tmp";
    const std::string mtx = dir + "/wk_dropin_" + std::to_string(n) + ".mtx";
    const std::string lay = dir + "/wk_dropin_" + std::to_string(n) + ".k1";
    warpkernel::csr_matrix A;
    A.num_rows = M;
    A.num_cols = N;
    A.num_entries = nz;
    A.row_offsets = rowptr;
    A.column_indices = colind;
    A.values = vals;
    warpkernel::write_matrix_market_file(A, mtx.c_str());
    warpkernel::csr_matrix B;
    warpkernel::read_matrix_market_file(B, mtx.c_str());
    expect(B.num_rows == M && B.row_offsets == rowptr && B.column_indices == colind && B.values == vals,
           "MatrixMarket round trip");
    warpkernel::structure kernel1;
    kernel1.driver_name = "dropin";
    kernel1.scan(B.num_entries, B.num_rows, B.values.data(), B.row_offsets.data(), B.column_indices.data());
    kernel1.save(lay.c_str());
    warpkernel::structure kernel2;
    kernel2.load(lay.c_str());
    expect(kernel2.reorder_rows == kernel1.reorder_rows && kernel2.row_map == kernel1.row_map &&
               kernel2.warp_offsets == kernel1.warp_offsets && kernel2.driver_name == "dropin",
           "layout file round trip");
    warpkernel::engine<double, int> eng(kernel2, B.values.data(), B.column_indices.data());
    warpkernel::device_vector<double> dy(N, 0.);
    eng.run<true>((kernel2.nwarps + 3) / 4, 256, dx.data(), dy.data());
    expect(std::vector<double>(dy) == y, "engine from a loaded layout bit-exact");
    // run_assembly (hpp:898-929): row sums, the CSR sum at x = 1
    std::vector<double> ones(N, 1.0), rs(M);
    wko_csr_spmv((int32_t)M, rowptr.data(), colind.data(), vals.data(), ones.data(), rs.data(), 1, 1, 1, 1);
    warpkernel::device_vector<double> ds(N, 0.);
    eng.run_assembly<true>((kernel2.nwarps + 3) / 4, 256, ds.data());
    expect(std::vector<double>(ds) == rs, "run_assembly bit-exact");
    std::remove(mtx.c_str());
    std::remove(lay.c_str());
  }

  // Kr row sort (optimize.cu:645-704): each row in permuted-column order
  {
    warpkernel::structure kernel1;
    kernel1.scan(nz, N, vals.data(), rowptr.data(), colind.data());
    std::vector<int> reorder_cols = colind;
    kernel1.reorder_columns_rowsort(reorder_cols, rowptr);
    warpkernel::device_vector<double> dreordered_x;
    kernel1.reorder_x(x, dreordered_x);
    std::vector<double> A_new_values(nz);
    std::vector<int> cols_sorted(nz);
    for (uint i = 0; i < nz; i++) {
      A_new_values[i] = vals[kernel1.reorder_A_rows[i]];
      cols_sorted[i] = colind[kernel1.reorder_A_rows[i]];
    }
    warpkernel::engine<double, int> eng(kernel1, A_new_values.data(), reorder_cols.data());
    warpkernel::device_vector<double> dy(N, 0.);
    eng.run_x<true>((kernel1.nwarps + 3) / 4, 256, dreordered_x.data(), dy.data());
    std::vector<double> ycheck = dy, ys(M);
    wko_csr_spmv((int32_t)M, rowptr.data(), cols_sorted.data(), A_new_values.data(), x.data(), ys.data(), 1, 1, 1, 1);
    bool same = true;
    for (uint i = 0; i < M; i++) same = same && ycheck[i] == ys[kernel1.row_map[i]];
    expect(same, "rowsort run_x bit-exact vs the row-sorted CSR order");
    expect(eng.verify_x(y, ycheck), "rowsort verify_x");
  }

  // WPK2 threshold sweep (wpk2.cu:140-157)
  warpkernel::structure2 kernel2;
  kernel2.scan(nz, N, vals.data(), rowptr.data(), colind.data(), (int)N);
  int max_nz = kernel2.max_nz, min_nz = kernel2.min_nz;
  int inc = (max_nz - min_nz) * 0.1 > 0 ? (int)((max_nz - min_nz) * 0.1) : 1;
  std::printf("min_nz = %d, max_nz %d\n", min_nz, max_nz);
  for (int threshold = min_nz; threshold <= max_nz; threshold += inc) {
    kernel2.scan(nz, N, vals.data(), rowptr.data(), colind.data(), threshold);
    wko_layout *L = wko_scan2(64, threshold, 32, (int32_t)N, nz, rowptr.data());
    std::vector<double> pv(L->allocate_nz), yw(M);
    std::vector<int32_t> pc(L->allocate_nz);
    wko_process(L, vals.data(), colind.data(), pv.data(), pc.data());
    wko_wpk_spmv(L, pv.data(), pc.data(), x.data(), yw.data(), 1, 1, 1, 0);
    expect(std::vector<uint>(L->reorder_rows, L->reorder_rows + nz) == kernel2.reorder_rows,
           "WPK2 layout matches the oracle");
    wko_free(L);
    warpkernel::engine<double, int> eng(kernel2, vals.data(), colind.data());
    warpkernel::device_vector<double> dy(N, 0.);
    eng.run<true>((kernel2.nwarps + 3) / 4, 256, dx.data(), dy.data());
    std::vector<double> ycheck = dy;
    expect(eng.verify(y, ycheck), "WPK2 verify");
    expect(ycheck == yw, "WPK2 bit-exact vs the lane-order oracle");
  }

  // FP32 values, FP64 output
  {
    warpkernel::structure kernel1;
    kernel1.scan(nz, M, vals.data(), rowptr.data(), colind.data());
    std::vector<float> vf(vals.begin(), vals.end()), xf(x.begin(), x.end());
    warpkernel::engine<float, int> eng(kernel1, vf.data(), colind.data());
    warpkernel::device_vector<float> dxf = xf;
    warpkernel::device_vector<double> dy(N, 0.);
    eng.run<true>(1, 256, dxf.data(), dy.data());
    std::vector<double> ycheck = dy, vq(vf.begin(), vf.end()), xq(xf.begin(), xf.end()), yq(M);
    wko_csr_spmv((int32_t)M, rowptr.data(), colind.data(), vq.data(), xq.data(), yq.data(), 1, 1, 1, 1);
    double num = 0, den = 0;
    for (uint i = 0; i < M; i++) {
      num += (ycheck[i] - yq[i]) * (ycheck[i] - yq[i]);
      den += yq[i] * yq[i];
    }
    std::printf("FP32 compensated: normwise error vs FP64 on the same inputs %.3e\n", std::sqrt(num / den));
    expect(std::sqrt(num / den) <= 1e-10, "FP32 compensated error");
  }

  drivers_verbatim(rowptr, colind, vals, x, y, M, N, nz);

  // errors surface as exceptions
  try {
    warpkernel::structure bad(48);
    bad.scan(nz, M, vals.data(), rowptr.data(), colind.data());
    expect(false, "group width 48 rejected");
  } catch (const warpkernel::error &e) {
    std::printf("expected error: %s\n", e.what());
  }
  std::printf(failures ? "DROPIN FAILED (%d)\n" : "DROPIN OK\n", failures);
  return failures ? 1 : 0;
}
