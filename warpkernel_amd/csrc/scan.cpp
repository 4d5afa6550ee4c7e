// scan.cpp -- host-side WPK1 / WPK2 layout builder (the reference's
// structure::scan, warpkernel.hpp:309-365, and structure2::scan,
// warpkernel.hpp:578-736), generic in the group width W.
//
// Design differences from the reference, same output bit for bit:
//   - rows are ordered by a counting sort on row length (descending), stable in
//     row id, instead of std::sort with checkPairOrder (hpp:222-225, 249): the
//     comparator is a strict total order, so both give the one permutation;
//   - WPK2's split (hpp:633-691, vector inserts driven by three iterators) is a
//     single forward pass that emits the sub-groups directly;
//   - the reorder map and the padded scatter run on several host threads.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <thread>

#include "wk_internal.hpp"

namespace wk {

void parallel_for(int64_t n, int64_t min_chunk, const std::function<void(int64_t, int64_t)> &fn) {
  if (n <= 0) return;
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / std::max<int64_t>(1, min_chunk)));
  if (nt <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve((size_t)nt);
  for (int64_t t = 0; t < nt; t++) th.emplace_back([&, t] { fn(n * t / nt, n * (t + 1) / nt); });
  for (auto &x : th) x.join();
}

void validate_csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rowptr,
                  const int32_t *colind) {
  if (nrows < 0 || ncols < 0) fail(WK_ERR_ARG, "negative matrix dimension");
  if (nnz < 0 || nnz > INT32_MAX) fail(WK_ERR_LIMIT, "nnz must be in [0, 2^31)");
  if (!rowptr) fail(WK_ERR_ARG, "rowptr is NULL");
  if (rowptr[0] != 0) fail(WK_ERR_CSR, "rowptr[0] != 0");
  for (int32_t r = 0; r < nrows; r++)
    if (rowptr[r + 1] < rowptr[r])
      fail(WK_ERR_CSR, "rowptr decreases at row " + std::to_string(r));
  if ((int64_t)rowptr[nrows] != nnz) fail(WK_ERR_CSR, "rowptr[nrows] != nnz");
  if (colind) {
    std::atomic<int64_t> first_bad{INT64_MAX};
    parallel_for(nnz, 1 << 20, [&](int64_t a, int64_t b) {
      for (int64_t i = a; i < b; i++)
        if ((uint32_t)colind[i] >= (uint32_t)ncols) {
          int64_t cur = first_bad.load();
          while (i < cur && !first_bad.compare_exchange_weak(cur, i)) {
          }
          return;
        }
    });
    if (first_bad.load() != INT64_MAX)
      fail(WK_ERR_CSR, "column index out of range at entry " + std::to_string(first_bad.load()));
  }
}

namespace {

// row_map: sorted position -> row, by length descending then row ascending.
// Returns the row lengths in sorted order through `slen`.
void sort_rows(wk_layout &L, const int32_t *rowptr, std::vector<uint32_t> &slen) {
  const int32_t n = L.nrows;
  uint32_t maxlen = 0;
  for (int32_t r = 0; r < n; r++) maxlen = std::max<uint32_t>(maxlen, (uint32_t)(rowptr[r + 1] - rowptr[r]));
  std::vector<int64_t> start((size_t)maxlen + 2, 0);
  for (int32_t r = 0; r < n; r++) start[rowptr[r + 1] - rowptr[r]]++;
  // start[len] = rows longer than len
  int64_t acc = 0;
  for (int64_t len = maxlen; len >= 0; len--) {
    int64_t c = start[len];
    start[len] = acc;
    acc += c;
  }
  L.row_map.assign(n, 0);
  L.row_map_inv.assign(n, 0);
  slen.assign(n, 0);
  for (int32_t r = 0; r < n; r++) {
    uint32_t len = (uint32_t)(rowptr[r + 1] - rowptr[r]);
    int64_t p = start[len]++;
    L.row_map[p] = r;
    L.row_map_inv[r] = (int32_t)p;
    slen[p] = len;
  }
  L.nznrows = 0;
  while (L.nznrows < n && slen[L.nznrows] > 0) L.nznrows++;
}

void finish_offsets(wk_layout &L) {
  const int32_t G = L.nwarps;
  L.warp_offsets.assign((size_t)G + 1, 0);
  for (int32_t w = 0; w < G; w++)
    L.warp_offsets[w + 1] = L.warp_offsets[w] + (int64_t)L.max_nz_rows[w] * L.W;
  L.allocate_nz = L.warp_offsets[G];
  if (L.allocate_nz > (int64_t)UINT32_MAX)
    fail(WK_ERR_LIMIT, "padded layout exceeds 2^32 slots");
}

void check_width(int32_t W) {
  if (W != 32 && W != 64) fail(WK_ERR_ARG, "group_width must be 32 or 64");
}

}  // namespace

void scan1(wk_layout &L, int32_t W, int32_t nrows, int32_t ncols, int64_t nnz,
           const int32_t *rowptr) {
  check_width(W);
  L = wk_layout();
  L.variant = 1;
  L.W = W;
  L.nrows = nrows;
  L.ncols = ncols;
  L.nz = nnz;
  L.rowptr.assign(rowptr, rowptr + (size_t)nrows + 1);
  std::vector<uint32_t> slen;
  sort_rows(L, rowptr, slen);
  L.nwarps = (L.nznrows + W - 1) / W;
  L.max_nz_rows.resize(L.nwarps);
  for (int32_t w = 0; w < L.nwarps; w++) L.max_nz_rows[w] = slen[(size_t)w * W];  // longest first
  finish_offsets(L);
  L.reorder_rows.assign((size_t)nnz, 0);
  const int32_t nzr = L.nznrows;
  parallel_for(nzr, 4096, [&](int64_t p0, int64_t p1) {
    for (int64_t p = p0; p < p1; p++) {
      const int32_t r = L.row_map[p];
      const int64_t base = L.warp_offsets[p / W] + p % W;
      const int32_t b = rowptr[r], len = (int32_t)slen[p];
      for (int32_t i = 0; i < len; i++) L.reorder_rows[(size_t)b + i] = (uint32_t)(base + (int64_t)i * W);
    }
  });
}

void split_groups(const std::vector<uint32_t> &aw0, int32_t nzr, int32_t W, int32_t threshold,
                  int32_t max_tpr, std::vector<uint32_t> &max_nz_rows,
                  std::vector<uint32_t> &rows_nominal, std::vector<uint32_t> &rows_actual) {
  // Forward pass over the W-row groups of sorted rows.  A group whose longest
  // row exceeds T gets i lanes per row (smallest power of two with
  // A_w <= cap*i, cap = T or ceil(A_w/max_tpr)), each lane a contiguous chunk
  // of nz_per_thread = ceil(A_w/i) entries.  A full group becomes i sub-groups
  // of W/i rows; the last, partial group is split the same way unless its rows
  // times i fall short of W, in which case it stays one group with W/i nominal
  // rows (hpp:660-662).  Sub-groups of a split count W/i rows each for
  // row_offset_warp (hpp:666), even in a partial last group.
  const uint32_t T = (uint32_t)threshold;
  max_nz_rows.clear();
  rows_nominal.clear();
  rows_actual.clear();
  const int32_t G0 = (int32_t)aw0.size();
  for (int32_t g = 0; g < G0; g++) {
    const uint32_t aw = aw0[g];
    const uint32_t rows = (uint32_t)std::min<int64_t>(W, nzr - (int64_t)g * W);
    // smallest i in {2, 4, .., max_tpr} with aw <= cap*i (uint32 arithmetic as
    // in the reference); none qualifies only if cap*i wraps, and then the
    // group is left whole, as the reference's loop leaves it
    uint32_t i = 0;
    if (aw > T) {
      const uint32_t cap = aw > T * (uint32_t)max_tpr ? (aw + max_tpr - 1) / max_tpr : T;
      for (uint32_t c = 2; c <= (uint32_t)max_tpr; c *= 2)
        if (aw <= cap * c) {
          i = c;
          break;
        }
    }
    if (i) {
      const uint32_t nzpt = (aw + i - 1) / i;
      if (rows * i < (uint32_t)W) {
        max_nz_rows.push_back(nzpt);
        rows_nominal.push_back(W / i);
        rows_actual.push_back(rows);
      } else {
        for (uint32_t j = 0; j < i; j++) {
          max_nz_rows.push_back(nzpt);
          rows_nominal.push_back(W / i);
          rows_actual.push_back(W / i);
        }
      }
    } else {
      max_nz_rows.push_back(aw);
      rows_nominal.push_back(W);
      rows_actual.push_back(rows);
    }
  }
}

void scan2(wk_layout &L, int32_t W, int32_t threshold, int32_t max_tpr, int32_t nrows,
           int32_t ncols, int64_t nnz, const int32_t *rowptr) {
  check_width(W);
  if (threshold < 0) fail(WK_ERR_ARG, "threshold must be >= 0");
  if (max_tpr < 2 || max_tpr > W || (max_tpr & (max_tpr - 1)))
    fail(WK_ERR_ARG, "max_tpr must be a power of two in [2, group_width]");
  L = wk_layout();
  L.variant = 2;
  L.W = W;
  L.threshold = threshold;
  L.max_tpr = max_tpr;
  L.nrows = nrows;
  L.ncols = ncols;
  L.nz = nnz;
  L.rowptr.assign(rowptr, rowptr + (size_t)nrows + 1);
  std::vector<uint32_t> slen;
  sort_rows(L, rowptr, slen);
  const int32_t nzr = L.nznrows;
  // min_nz starts at nrows (hpp:607), so it never exceeds nrows even when a
  // row is longer than that (a case the reference asserts against, hpp:614)
  L.min_nz = nzr > 0 ? std::min<uint32_t>(slen[nzr - 1], (uint32_t)nrows) : (uint32_t)nrows;
  L.max_nz = nzr > 0 ? slen[0] : 0;

  const int32_t G0 = (nzr + W - 1) / W;
  std::vector<uint32_t> aw0((size_t)G0);
  for (int32_t g = 0; g < G0; g++) aw0[g] = slen[(size_t)g * W];
  std::vector<uint32_t> rows_nominal, rows_actual;
  split_groups(aw0, nzr, W, threshold, max_tpr, L.max_nz_rows, rows_nominal, rows_actual);
  L.nwarps = (int32_t)L.max_nz_rows.size();
  const int32_t G = L.nwarps;
  finish_offsets(L);
  L.threads_per_row.resize(G);
  L.row_offset_warp.assign((size_t)G + 1, 0);
  for (int32_t w = 0; w < G; w++) {
    L.threads_per_row[w] = (uint32_t)W / rows_nominal[w];
    L.row_offset_warp[w + 1] = L.row_offset_warp[w] + rows_actual[w];
  }
  L.row_offset_warp[G] = (uint32_t)nrows;

  // CSR entry m of sorted row rs (slot r of sub-group w) goes to lane
  // tid + r*tpr, step m % A, with tid = m / A.
  L.reorder_rows.assign((size_t)nnz, 0);
  parallel_for(G, 256, [&](int64_t w0, int64_t w1) {
    for (int64_t w = w0; w < w1; w++) {
      const uint32_t tpr = L.threads_per_row[w], A = L.max_nz_rows[w];
      const uint32_t first = L.row_offset_warp[w];
      for (uint32_t r = 0; r < rows_nominal[w] && (int64_t)first + r < nzr; r++) {
        const int64_t rs = (int64_t)first + r;
        const int32_t row = L.row_map[rs];
        const uint32_t len = slen[rs];
        const int64_t base = L.warp_offsets[w] + (int64_t)r * tpr;
        const int32_t b = rowptr[row];
        for (uint32_t m = 0; m < len && m < tpr * A; m++)
          L.reorder_rows[(size_t)b + m] = (uint32_t)(base + m / A + (int64_t)(m % A) * W);
      }
    }
  });
}

template <class T>
void process_host(const wk_layout &L, const double *vals, const int32_t *colind, T *pvals,
                  int32_t *pcols) {
  parallel_for(L.allocate_nz, 1 << 20, [&](int64_t a, int64_t b) {
    if (pvals) std::memset(pvals + a, 0, sizeof(T) * (size_t)(b - a));
    if (pcols) std::memset(pcols + a, 0, sizeof(int32_t) * (size_t)(b - a));
  });
  parallel_for(L.nz, 1 << 20, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; i++) {
      const uint32_t s = L.reorder_rows[i];
      if (pvals) pvals[s] = (T)vals[i];
      if (pcols) pcols[s] = colind[i];
    }
  });
}

template void process_host<double>(const wk_layout &, const double *, const int32_t *, double *,
                                   int32_t *);
template void process_host<float>(const wk_layout &, const double *, const int32_t *, float *,
                                  int32_t *);

void reorder_columns_rowsort(const wk_layout &L, const int32_t *rowptr, const int32_t *colind,
                             int32_t *cols_out, int32_t *src) {
  // sort keys (permuted column << 32 | position in row) make the order total,
  // which is std::sort by the column with CSR order kept among equal columns
  parallel_for(L.nrows, 2048, [&](int64_t r0, int64_t r1) {
    std::vector<uint64_t> key;
    for (int64_t r = r0; r < r1; r++) {
      const int32_t b = rowptr[r], n = rowptr[r + 1] - rowptr[r];
      key.resize((size_t)n);
      for (int32_t i = 0; i < n; i++)
        key[i] = ((uint64_t)(uint32_t)L.row_map_inv[colind[b + i]] << 32) | (uint32_t)i;
      std::sort(key.begin(), key.end());
      for (int32_t i = 0; i < n; i++) {
        cols_out[b + i] = (int32_t)(key[i] >> 32);
        src[b + i] = b + (int32_t)(key[i] & 0xffffffffu);
      }
    }
  });
}

bool build_windows(const std::vector<int64_t> &warp_offsets, const std::vector<uint32_t> &max_nz_rows,
                   int32_t W, const int32_t *keys, std::vector<uint16_t> &idx,
                   std::vector<uint32_t> &off, std::vector<int32_t> &cols, int32_t &win_max) {
  const int32_t G = (int32_t)max_nz_rows.size();
  idx.clear();
  off.clear();
  cols.clear();
  win_max = 0;
  if (!G) return false;
  // pass 1: distinct columns per group (sorted), in per-thread chunks
  std::vector<std::vector<int32_t>> uniq((size_t)G);
  std::atomic<bool> too_big{false};
  parallel_for(G, 64, [&](int64_t g0, int64_t g1) {
    for (int64_t g = g0; g < g1 && !too_big.load(std::memory_order_relaxed); g++) {
      const int64_t a = warp_offsets[g], b = a + (int64_t)max_nz_rows[g] * W;
      std::vector<int32_t> &u = uniq[g];
      u.assign(keys + a, keys + b);
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      if ((int64_t)u.size() > kMaxWindow) too_big.store(true);
    }
  });
  if (too_big.load()) return false;
  off.assign((size_t)G + 1, 0);
  for (int32_t g = 0; g < G; g++) {
    off[g + 1] = off[g] + (uint32_t)uniq[g].size();
    win_max = std::max<int32_t>(win_max, (int32_t)uniq[g].size());
  }
  cols.resize(off[G]);
  idx.resize((size_t)warp_offsets[G]);
  // pass 2: each slot's position in its group's list
  parallel_for(G, 64, [&](int64_t g0, int64_t g1) {
    for (int64_t g = g0; g < g1; g++) {
      const std::vector<int32_t> &u = uniq[g];
      std::copy(u.begin(), u.end(), cols.begin() + off[g]);
      const int64_t a = warp_offsets[g], b = a + (int64_t)max_nz_rows[g] * W;
      for (int64_t s = a; s < b; s++)
        idx[s] = (uint16_t)(std::lower_bound(u.begin(), u.end(), keys[s]) - u.begin());
      std::vector<int32_t>().swap(uniq[g]);
    }
  });
  return true;
}

void reorder_columns_coalesced(const wk_layout &L, const int32_t *colind, int32_t *out) {
  parallel_for(L.allocate_nz, 1 << 20, [&](int64_t a, int64_t b) {
    std::memset(out + a, 0, sizeof(int32_t) * (size_t)(b - a));
  });
  parallel_for(L.nz, 1 << 20, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; i++) out[L.reorder_rows[i]] = L.row_map_inv[colind[i]];
  });
}

void transpose_csr(int32_t nrows, int32_t ncols, const int32_t *rowptr, const int32_t *colind,
                   std::vector<int32_t> &colptr, std::vector<int32_t> &rowind,
                   std::vector<int64_t> &perm) {
  const int64_t nnz = rowptr[nrows];
  colptr.assign((size_t)ncols + 1, 0);
  for (int64_t i = 0; i < nnz; i++) colptr[colind[i] + 1]++;
  for (int32_t c = 0; c < ncols; c++) colptr[c + 1] += colptr[c];
  std::vector<int32_t> fill(colptr.begin(), colptr.end() - 1);
  rowind.resize((size_t)nnz);
  perm.resize((size_t)nnz);
  for (int32_t r = 0; r < nrows; r++)
    for (int32_t t = rowptr[r]; t < rowptr[r + 1]; t++) {
      const int32_t d = fill[colind[t]]++;
      rowind[d] = r;
      perm[d] = t;
    }
}

}  // namespace wk
