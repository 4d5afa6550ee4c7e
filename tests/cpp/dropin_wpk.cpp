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
