// dscan.hip -- the scan/process step on the GPU: a device CSR becomes an
// engine without the matrix visiting the host.
//
// The result is the layout of structure::scan / structure2::scan
// (warpkernel.hpp:309-365, 578-736) at group width 64 and the padded arrays of
// process (hpp:379-398, 749-769), bit for bit the host scan's; only the way
// they are computed differs:
//   1. one pass checks the CSR (rowptr monotone from 0 to nnz, columns in
//      range) and finds the longest row; nothing else runs on a malformed CSR;
//   2. per block of rows (one wave), a histogram of row lengths in LDS; the
//      histograms in (length descending, block ascending) order are
//      prefix-summed on the host: the start of every (length, block) run of
//      the sorted order;
//   3. each block walks its rows in order, 64 at a time, and ranks the rows of
//      one length with ballots, so placement is stable: length descending, row
//      ascending (checkPairOrder, hpp:222-225);
//   4. the longest row of each group comes back to the host (G values), where
//      WPK2's split rule (shared with the host scan) and the group offsets are
//      computed; one kernel then writes the padded values, columns, Kr columns
//      and the CSR-to-slot map with coalesced stores.
// Limits: group width 64, rows of at most 8191 entries (the LDS histogram),
// no transposed layout (the adjoint engine is built by wk_engine_create).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "wk_internal.hpp"

namespace wk {
namespace {

constexpr int32_t kMaxLen = 8191;
constexpr int64_t kMaxCounts = int64_t(1) << 24;  // histogram entries, all blocks

// part[b] = longest row seen by block b, part[nb + b] = error bits
// (1: rowptr ends, 2: rowptr decreases, 4: column out of range)
__global__ __launch_bounds__(256) void csr_check_kernel(const int32_t *__restrict__ rowptr,
                                                        int32_t nrows,
                                                        const int32_t *__restrict__ colind,
                                                        int64_t nnz, int32_t ncols,
                                                        int32_t *__restrict__ part) {
  __shared__ int32_t red_max[256], red_bad[256];
  int32_t m = 0, bad = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = t0; r < nrows; r += stride) {
    const int32_t len = rowptr[r + 1] - rowptr[r];
    if (len < 0) bad |= 2;
    m = max(m, len);
  }
  for (int64_t e = t0; e < nnz; e += stride)
    if ((uint32_t)colind[e] >= (uint32_t)ncols) bad |= 4;
  if (t0 == 0 && (rowptr[0] != 0 || (int64_t)rowptr[nrows] != nnz)) bad |= 1;
  red_max[threadIdx.x] = m;
  red_bad[threadIdx.x] = bad;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red_max[threadIdx.x] = max(red_max[threadIdx.x], red_max[threadIdx.x + o]);
      red_bad[threadIdx.x] |= red_bad[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = red_max[0];
    part[gridDim.x + blockIdx.x] = red_bad[0];
  }
}

// counts[(maxlen - len) * nblk + b] = rows of length len among block b's rows
__global__ __launch_bounds__(64) void len_hist_kernel(const int32_t *__restrict__ rowptr,
                                                      int32_t nrows, int32_t rpb, int32_t maxlen,
                                                      int32_t nblk, int32_t *__restrict__ counts) {
  extern __shared__ int32_t hist[];
  const int32_t b = blockIdx.x;
  for (int32_t l = threadIdx.x; l <= maxlen; l += 64) hist[l] = 0;
  __syncthreads();
  const int64_t r0 = (int64_t)b * rpb;
  for (int32_t t = threadIdx.x; t < rpb; t += 64) {
    const int64_t r = r0 + t;
    if (r < nrows) atomicAdd(&hist[rowptr[r + 1] - rowptr[r]], 1);
  }
  __syncthreads();
  for (int32_t l = threadIdx.x; l <= maxlen; l += 64)
    counts[(int64_t)(maxlen - l) * nblk + b] = hist[l];
}

// Stable placement: starts[(maxlen - len) * nblk + b] is the sorted position
// of block b's first row of length len.  One wave per block; per chunk of 64
// rows, one round per distinct length among them.
__global__ __launch_bounds__(64) void place_kernel(const int32_t *__restrict__ rowptr,
                                                   int32_t nrows, int32_t rpb, int32_t maxlen,
                                                   int32_t nblk, const int32_t *__restrict__ starts,
                                                   int32_t *__restrict__ row_map,
                                                   int32_t *__restrict__ row_map_inv,
                                                   uint32_t *__restrict__ slen) {
  extern __shared__ int32_t next[];  // next free sorted position per length
  const int32_t b = blockIdx.x;
  const int32_t lane = threadIdx.x;
  for (int32_t l = lane; l <= maxlen; l += 64) next[l] = starts[(int64_t)(maxlen - l) * nblk + b];
  __syncthreads();
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes under this one
  for (int32_t c = 0; c < rpb; c += 64) {
    const int64_t r = (int64_t)b * rpb + c + lane;
    const int32_t len = r < nrows ? rowptr[r + 1] - rowptr[r] : -1;
    bool pending = r < nrows;
    uint64_t act;
    while ((act = __ballot(pending)) != 0) {
      const int32_t src = __ffsll((unsigned long long)act) - 1;
      const int32_t v = __shfl(len, src);  // this round's length, wave-uniform
      const bool mine = pending && len == v;
      const uint64_t same = __ballot(mine);
      const int32_t base = next[v];
      if (mine) {
        const int32_t pos = base + (int32_t)__popcll(same & below);
        row_map[pos] = (int32_t)r;
        row_map_inv[r] = pos;
        slen[pos] = (uint32_t)len;
        pending = false;
      }
      __builtin_amdgcn_wave_barrier();
      if (lane == src) next[v] = base + (int32_t)__popcll(same);
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// aw[g] = length of sorted row 64 g, the longest of group g
__global__ __launch_bounds__(256) void group_len_kernel(const uint32_t *__restrict__ slen,
                                                        int32_t ngroups, uint32_t *__restrict__ aw) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ngroups) aw[g] = slen[g * 64];
}

// One wave per (sub-)group w: lane l holds chunk l % tpr of sorted row
// row_off[w] + l / tpr (WPK1: tpr 1, rows 64 w + l); its slots are
// warp_offsets[w] + l + 64 i, i < A_w; slots past the chunk are pads (0, col 0).
template <class T>
__global__ __launch_bounds__(256) void fill_kernel(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
    const double *__restrict__ vals, const int32_t *__restrict__ row_map,
    const int32_t *__restrict__ row_map_inv, const int64_t *__restrict__ wo,
    const uint32_t *__restrict__ aw, const uint32_t *__restrict__ tpr_a,
    const uint32_t *__restrict__ row_off, int32_t ngroups, int32_t nznrows, T *__restrict__ pv,
    int32_t *__restrict__ pc, int32_t *__restrict__ pkr, uint32_t *__restrict__ slot_of_nz) {
  const int32_t lane = threadIdx.x & 63;
  const int32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= ngroups) return;
  const uint32_t A = aw[w];
  const uint32_t tpr = tpr_a ? tpr_a[w] : 1u;
  const uint32_t p = (tpr_a ? row_off[w] : (uint32_t)w * 64u) + (uint32_t)lane / tpr;
  const uint32_t chunk = (uint32_t)lane & (tpr - 1);
  int32_t rb = 0, n = 0;
  if (p < (uint32_t)nznrows) {
    const int32_t r = row_map[p];
    const int32_t b1 = rowptr[r + 1];
    rb = rowptr[r] + (int32_t)(chunk * A);
    n = max(0, min(b1 - rb, (int32_t)A));
  }
  const int64_t base = wo[w] + lane;
  for (uint32_t i = 0; i < A; i++) {
    const int64_t slot = base + (int64_t)i * 64;
    if ((int32_t)i < n) {
      const int32_t e = rb + (int32_t)i;
      const int32_t c = colind[e];
      pv[slot] = (T)vals[e];
      pc[slot] = c;
      if (pkr) pkr[slot] = row_map_inv[c];
      slot_of_nz[e] = (uint32_t)slot;
    } else {
      pv[slot] = T(0);
      pc[slot] = 0;
      if (pkr) pkr[slot] = 0;
    }
  }
}

// ---------------------------------------------------------------- x windows
// The windows of build_windows (scan.cpp), built on the device: one block per
// group sorts the group's slot keys (the entry's column, -1 for a pad) in LDS,
// counts the distinct ones (pass 1), then writes them and each slot's
// position among them (pass 2).  Groups of more than kWinSlots slots are not
// handled here: the engine then has no windows.
constexpr int32_t kWinSlots = 8192;
constexpr int kWinThreads = 256;

struct WinBuild {
  const int32_t *rowptr, *row_map, *row_map_inv, *cols;
  const int64_t *wo;
  const uint32_t *aw, *tpr, *row_off;
  int32_t nznrows;
};

// Sorted keys of group g in keys[0, N) (INT_MAX past the S slots); returns S,
// or -1 when the group has too many slots.
__device__ int32_t sort_group_keys(const WinBuild &b, int32_t g, int32_t *keys, int32_t *lane_n,
                                   int32_t &N) {
  const uint32_t A = b.aw[g];
  const int64_t S64 = (int64_t)A * 64;
  if (S64 > kWinSlots) return -1;
  const int32_t S = (int32_t)S64;
  const int32_t tid = threadIdx.x;
  if (tid < 64) {
    // entries of lane tid (the fill kernel's chunk rule)
    const uint32_t tpr = b.tpr ? b.tpr[g] : 1u;
    const uint32_t p = (b.tpr ? b.row_off[g] : (uint32_t)g * 64u) + (uint32_t)tid / tpr;
    const uint32_t chunk = (uint32_t)tid & (tpr - 1);
    int32_t n = 0;
    if (p < (uint32_t)b.nznrows) {
      const int32_t r = b.row_map[p];
      const int32_t rb = b.rowptr[r] + (int32_t)(chunk * A);
      n = max(0, min(b.rowptr[r + 1] - rb, (int32_t)A));
    }
    lane_n[tid] = n;
  }
  N = 64;
  while (N < S) N <<= 1;
  __syncthreads();
  const int64_t base = b.wo[g];
  for (int32_t t = tid; t < N; t += kWinThreads)
    keys[t] = t < S ? ((t >> 6) < lane_n[t & 63] ? b.cols[base + t] : -1) : INT32_MAX;
  __syncthreads();
  for (int32_t k = 2; k <= N; k <<= 1)
    for (int32_t j = k >> 1; j > 0; j >>= 1) {
      for (int32_t t = tid; t < N; t += kWinThreads) {
        const int32_t o = t ^ j;
        if (o > t) {
          const int32_t a = keys[t], c = keys[o];
          if ((a > c) == ((t & k) == 0)) {
            keys[t] = c;
            keys[o] = a;
          }
        }
      }
      __syncthreads();
    }
  return S;
}

__global__ __launch_bounds__(kWinThreads) void window_count_kernel(WinBuild b, int32_t *count) {
  __shared__ int32_t keys[kWinSlots];
  __shared__ int32_t lane_n[64];
  __shared__ int32_t part[kWinThreads];
  const int32_t g = blockIdx.x;
  int32_t N = 0;
  const int32_t S = sort_group_keys(b, g, keys, lane_n, N);
  if (S < 0) {
    if (threadIdx.x == 0) count[g] = -1;
    return;
  }
  int32_t c = 0;
  for (int32_t t = threadIdx.x; t < N; t += kWinThreads)
    c += keys[t] != INT32_MAX && (t == 0 || keys[t] != keys[t - 1]);
  part[threadIdx.x] = c;
  __syncthreads();
  for (int o = kWinThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) count[g] = part[0] > kMaxWindow ? -1 : part[0];
}

__global__ __launch_bounds__(kWinThreads) void window_write_kernel(WinBuild b, const uint32_t *off,
                                                                   int32_t *win_cols,
                                                                   int32_t *win_cols_kr,
                                                                   uint16_t *win_idx) {
  __shared__ int32_t keys[kWinSlots];
  __shared__ int32_t uniq[kMaxWindow];
  __shared__ int32_t lane_n[64];
  __shared__ int32_t part[kWinThreads];
  const int32_t g = blockIdx.x;
  const int32_t tid = threadIdx.x;
  int32_t N = 0;
  const int32_t S = sort_group_keys(b, g, keys, lane_n, N);
  if (S < 0) return;  // not reached: the host builds windows only when every group fits
  // ranks of the distinct keys: each thread counts a contiguous chunk of the
  // sorted keys, then an inclusive scan of the chunk counts
  const int32_t per = N / kWinThreads > 0 ? N / kWinThreads : 1;
  const int32_t t0 = min(N, tid * per), t1 = min(N, t0 + per);
  int32_t c = 0;
  for (int32_t t = t0; t < t1; t++) c += keys[t] != INT32_MAX && (t == 0 || keys[t] != keys[t - 1]);
  part[tid] = c;
  __syncthreads();
  for (int o = 1; o < kWinThreads; o <<= 1) {
    const int32_t v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int32_t rank = part[tid] - c;
  const uint32_t w0 = off[g];
  for (int32_t t = t0; t < t1; t++) {
    const int32_t k = keys[t];
    if (k != INT32_MAX && (t == 0 || k != keys[t - 1])) {
      uniq[rank] = k;
      win_cols[w0 + rank] = k < 0 ? 0 : k;
      if (win_cols_kr) win_cols_kr[w0 + rank] = k < 0 ? 0 : b.row_map_inv[k];
      rank++;
    }
  }
  __syncthreads();
  // each slot's position among the distinct keys (ascending; a pad's key -1,
  // when present, is the first)
  const int32_t n = part[kWinThreads - 1];
  const int64_t base = b.wo[g];
  for (int32_t t = tid; t < S; t += kWinThreads) {
    const int32_t key = (t >> 6) < lane_n[t & 63] ? b.cols[base + t] : -1;
    int32_t lo = 0, hi = n;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (uniq[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
    win_idx[base + t] = (uint16_t)lo;
  }
}

template <class T>
std::vector<T> download(const T *d, int64_t n, hipStream_t s) {
  std::vector<T> h((size_t)n);
  if (n) {
    hip_check(hipMemcpyAsync(h.data(), d, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, s),
              "download");
    hip_check(hipStreamSynchronize(s), "download");
  }
  return h;
}

template <class T>
void launch_fill(const DevLayout &D, const int32_t *rowptr, const int32_t *colind,
                 const double *vals, hipStream_t s) {
  if (!D.ngroups) return;
  hipLaunchKernelGGL(fill_kernel<T>, dim3((uint32_t)((D.ngroups + 3) / 4)), dim3(256), 0, s,
                     rowptr, colind, vals, D.row_map, D.row_map_inv, D.warp_offsets,
                     D.max_nz_rows, D.tpr, D.row_off, D.ngroups, D.nznrows,
                     static_cast<T *>(D.vals), D.cols, D.cols_kr, D.slot_of_nz);
  hip_check(hipGetLastError(), "fill");
}

// Scan + process of one device CSR into E->fwd.
void device_scan(wk_engine *E, int32_t variant, int32_t threshold, int32_t max_tpr, bool kr,
                 const int32_t *rowptr, const int32_t *colind, const double *vals,
                 hipStream_t s) {
  const int32_t nrows = E->nrows, ncols = E->ncols;
  const int64_t nnz = E->nnz;
  DevLayout &D = E->fwd;
  D.variant = variant;
  D.W = 64;
  D.nrows = nrows;

  // 1. check the CSR, find the longest row
  const int64_t work = std::max<int64_t>(nrows, nnz);
  const int32_t nb = (int32_t)std::min<int64_t>(std::max<int64_t>((work + 1023) / 1024, 1), 1024);
  int32_t *part = dev_alloc<int32_t>(E, 2 * (int64_t)nb);
  hipLaunchKernelGGL(csr_check_kernel, dim3(nb), dim3(256), 0, s, rowptr, nrows, colind, nnz,
                     ncols, part);
  hip_check(hipGetLastError(), "csr_check");
  const std::vector<int32_t> hp = download(part, 2 * (int64_t)nb, s);
  int32_t maxlen = 0, bad = 0;
  for (int32_t b = 0; b < nb; b++) {
    maxlen = std::max(maxlen, hp[b]);
    bad |= hp[nb + b];
  }
  if (bad & 1) fail(WK_ERR_CSR, "rowptr must start at 0 and end at nnz");
  if (bad & 2) fail(WK_ERR_CSR, "rowptr decreases");
  if (bad & 4) fail(WK_ERR_CSR, "column index out of range");
  if (maxlen > kMaxLen)
    fail(WK_ERR_LIMIT, "rows longer than " + std::to_string(kMaxLen) +
                           " entries: scan on the host (wk_scan1 / wk_scan2)");
  if (nrows == 0) return;

  // 2. histograms of row lengths per block; run starts
  int32_t rpb = 1024;
  while ((int64_t)(maxlen + 1) * ((nrows + (int64_t)rpb - 1) / rpb) > kMaxCounts) rpb *= 2;
  const int32_t nblk = (int32_t)((nrows + (int64_t)rpb - 1) / rpb);
  const int64_t ncount = (int64_t)(maxlen + 1) * nblk;
  const size_t lds = sizeof(int32_t) * (size_t)(maxlen + 1);
  D.row_map = dev_alloc<int32_t>(E, nrows);
  D.row_map_inv = dev_alloc<int32_t>(E, nrows);
  uint32_t *slen = dev_alloc<uint32_t>(E, nrows);
  int32_t *counts = dev_alloc<int32_t>(E, ncount);
  hipLaunchKernelGGL(len_hist_kernel, dim3(nblk), dim3(64), lds, s, rowptr, nrows, rpb, maxlen,
                     nblk, counts);
  hip_check(hipGetLastError(), "len_hist");
  std::vector<int32_t> starts = download(counts, ncount, s);
  int64_t acc = 0;
  for (int64_t t = 0; t < ncount; t++) {
    const int32_t c = starts[t];
    starts[t] = (int32_t)acc;
    acc += c;
  }
  if (acc != nrows) fail(WK_ERR_STATE, "row histogram does not add up to nrows");
  // rows of length 0 form the last run
  const int32_t nznrows = starts[(int64_t)maxlen * nblk];
  hip_check(hipMemcpyAsync(counts, starts.data(), sizeof(int32_t) * (size_t)ncount,
                           hipMemcpyHostToDevice, s),
            "upload");
  hipLaunchKernelGGL(place_kernel, dim3(nblk), dim3(64), lds, s, rowptr, nrows, rpb, maxlen, nblk,
                     counts, D.row_map, D.row_map_inv, slen);
  hip_check(hipGetLastError(), "place");
  D.nznrows = nznrows;

  // 3. groups and their split
  const int32_t G0 = (nznrows + 63) / 64;
  std::vector<uint32_t> aw0;
  uint32_t shortest = 0;
  if (G0) {
    uint32_t *daw = dev_alloc<uint32_t>(E, G0);
    hipLaunchKernelGGL(group_len_kernel, dim3((uint32_t)((G0 + 255) / 256)), dim3(256), 0, s, slen,
                       G0, daw);
    hip_check(hipGetLastError(), "group_len");
    aw0 = download(daw, G0, s);
    shortest = download(slen + nznrows - 1, 1, s)[0];
  }
  std::vector<uint32_t> A, rows_nominal, rows_actual;
  if (variant == 2)
    split_groups(aw0, nznrows, 64, threshold, max_tpr, A, rows_nominal, rows_actual);
  else
    A = aw0;
  const int32_t G = (int32_t)A.size();
  std::vector<int64_t> wo((size_t)G + 1, 0);
  for (int32_t g = 0; g < G; g++) wo[g + 1] = wo[g] + (int64_t)A[g] * 64;
  if (wo[G] > (int64_t)UINT32_MAX) fail(WK_ERR_LIMIT, "padded layout exceeds 2^32 slots");
  D.ngroups = G;
  D.allocate_nz = wo[G];
  D.h_max_nz_rows = A;
  D.warp_offsets = dev_upload(E, wo.data(), (int64_t)wo.size(), s);
  D.max_nz_rows = dev_upload(E, A.data(), G, s);
  if (variant == 2) {
    std::vector<uint32_t> tpr((size_t)G), roff((size_t)G + 1, 0);
    for (int32_t w = 0; w < G; w++) {
      tpr[w] = 64u / rows_nominal[w];
      roff[w + 1] = roff[w] + rows_actual[w];
    }
    roff[G] = (uint32_t)nrows;  // sentinel (hpp:716)
    D.tpr = dev_upload(E, tpr.data(), G, s);
    D.row_off = dev_upload(E, roff.data(), G + 1, s);
    hip_check(hipStreamSynchronize(s), "upload");
  }

  // 4. padded arrays and the CSR -> slot map
  D.cols = dev_alloc<int32_t>(E, D.allocate_nz);
  D.cols_kr = kr ? dev_alloc<int32_t>(E, D.allocate_nz) : nullptr;
  D.slot_of_nz = dev_alloc<uint32_t>(E, nnz);
  if (E->dtype == WK_F64) {
    D.vals = dev_alloc<double>(E, D.allocate_nz);
    launch_fill<double>(D, rowptr, colind, vals, s);
  } else {
    D.vals = dev_alloc<float>(E, D.allocate_nz);
    launch_fill<float>(D, rowptr, colind, vals, s);
  }
  hip_check(hipStreamSynchronize(s), "device scan");

  // 5. x windows, when every group fits
  if (G) {
    WinBuild wb{rowptr, D.row_map, D.row_map_inv, D.cols, D.warp_offsets, D.max_nz_rows,
                D.tpr, D.row_off, nznrows};
    int32_t *dcount = dev_alloc<int32_t>(E, G);
    hipLaunchKernelGGL(window_count_kernel, dim3((uint32_t)G), dim3(kWinThreads), 0, s, wb, dcount);
    hip_check(hipGetLastError(), "window count");
    const std::vector<int32_t> cnt = download(dcount, G, s);
    bool fits = true;
    std::vector<uint32_t> off((size_t)G + 1, 0);
    int32_t most = 0;
    for (int32_t g = 0; g < G && fits; g++) {
      if (cnt[g] < 0) fits = false;
      off[g + 1] = off[g] + (uint32_t)std::max(cnt[g], 0);
      most = std::max(most, cnt[g]);
    }
    if (fits) {
      D.win_off = dev_upload(E, off.data(), (int64_t)off.size(), s);
      D.win_cols = dev_alloc<int32_t>(E, off[G]);
      D.win_cols_kr = kr ? dev_alloc<int32_t>(E, off[G]) : nullptr;
      D.win_idx = dev_alloc<uint16_t>(E, D.allocate_nz);
      hipLaunchKernelGGL(window_write_kernel, dim3((uint32_t)G), dim3(kWinThreads), 0, s, wb,
                         D.win_off, D.win_cols, D.win_cols_kr, D.win_idx);
      hip_check(hipGetLastError(), "window write");
      hip_check(hipStreamSynchronize(s), "window write");
      D.win_max = most;
      D.win_total = off[G];
      build_packed_positions(E, D, s);
    }
  }

  wk_layout_info &I = E->info;
  I.nznrows = nznrows;
  I.nwarps = G;
  I.allocate_nz = D.allocate_nz;
  if (variant == 2) {
    // min_nz starts at nrows (hpp:607)
    I.min_nz = nznrows ? std::min<uint32_t>(shortest, (uint32_t)nrows) : (uint32_t)nrows;
    I.max_nz = (uint32_t)maxlen;
  }
}

}  // namespace

}  // namespace wk

using namespace wk;

extern "C" int wk_engine_create_device(int32_t variant, int32_t threshold, int32_t max_tpr,
                                       int32_t nrows, int32_t ncols, int64_t nnz,
                                       const int32_t *rowptr_dev, const int32_t *colind_dev,
                                       const double *vals_dev, int32_t values_dtype,
                                       uint32_t flags, void *stream, wk_engine **out) {
  return guarded([&] {
    if (!out) fail(WK_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (variant != 1 && variant != 2) fail(WK_ERR_ARG, "variant must be 1 or 2");
    if (nrows < 0 || ncols < 0) fail(WK_ERR_ARG, "negative matrix dimension");
    if (nnz < 0 || nnz > INT32_MAX) fail(WK_ERR_LIMIT, "nnz must be in [0, 2^31)");
    if (!rowptr_dev || (nnz && (!colind_dev || !vals_dev))) fail(WK_ERR_ARG, "NULL argument");
    if (values_dtype != WK_F64 && values_dtype != WK_F32) fail(WK_ERR_ARG, "unknown values dtype");
    if (flags & ~(uint32_t)WK_ENGINE_KR)
      fail(WK_ERR_ARG, "only WK_ENGINE_KR is supported here (the adjoint layout is built by "
                       "wk_engine_create)");
    const bool kr = flags & WK_ENGINE_KR;
    if (kr && nrows != ncols) fail(WK_ERR_ARG, "the Kr form needs a square matrix");
    if (variant == 2) {
      if (threshold < 0) fail(WK_ERR_ARG, "threshold must be >= 0");
      if (max_tpr < 2 || max_tpr > 64 || (max_tpr & (max_tpr - 1)))
        fail(WK_ERR_ARG, "max_tpr must be a power of two in [2, 64]");
    }
    auto *E = new wk_engine();
    try {
      hip_check(hipGetDevice(&E->device), "hipGetDevice");
      E->dtype = values_dtype;
      E->nrows = nrows;
      E->ncols = ncols;
      E->nnz = nnz;
      wk_layout_info &I = E->info;
      I.variant = variant;
      I.group_width = 64;
      I.threshold = variant == 2 ? threshold : 0;
      I.max_tpr = variant == 2 ? max_tpr : 32;
      I.nrows = nrows;
      I.ncols = ncols;
      I.nz = nnz;
      if (variant == 2) I.min_nz = (uint32_t)nrows;
      device_scan(E, variant, threshold, max_tpr, kr, rowptr_dev, colind_dev, vals_dev,
                  static_cast<hipStream_t>(stream));
    } catch (...) {
      wk_engine_destroy(E);
      throw;
    }
    *out = E;
  });
}

extern "C" int wk_engine_download(const wk_engine *E, int32_t which, void *host_dst,
                                  int64_t *count) {
  return guarded([&] {
    if (!E) fail(WK_ERR_ARG, "engine is NULL");
    const DevLayout &D = E->fwd;
    const void *src = nullptr;
    size_t esz = 4;
    int64_t n = 0;
    switch (which) {
      case WK_ROW_MAP: src = D.row_map; n = D.row_map ? D.nrows : 0; break;
      case WK_ROW_MAP_INV: src = D.row_map_inv; n = D.row_map_inv ? D.nrows : 0; break;
      case WK_REORDER_ROWS: src = D.slot_of_nz; n = D.slot_of_nz ? E->nnz : 0; break;
      case WK_WARP_OFFSETS: src = D.warp_offsets; n = D.warp_offsets ? D.ngroups + 1 : 0; esz = 8; break;
      case WK_MAX_NZ_ROWS: src = D.max_nz_rows; n = D.max_nz_rows ? D.ngroups : 0; break;
      case WK_THREADS_PER_ROW: src = D.tpr; n = D.tpr ? D.ngroups : 0; break;
      case WK_ROW_OFFSET_WARP: src = D.row_off; n = D.row_off ? D.ngroups + 1 : 0; break;
      case WK_ENGINE_VALUES:
        src = D.vals;
        n = D.vals ? D.allocate_nz : 0;
        esz = E->dtype == WK_F64 ? 8 : 4;
        break;
      case WK_ENGINE_COLS: src = D.cols; n = D.cols ? D.allocate_nz : 0; break;
      case WK_ENGINE_COLS_KR: src = D.cols_kr; n = D.cols_kr ? D.allocate_nz : 0; break;
      case WK_ENGINE_WIN_IDX: src = D.win_idx; n = D.win_idx ? D.allocate_nz : 0; esz = 2; break;
      case WK_ENGINE_WIN_OFF: src = D.win_off; n = D.win_off ? D.ngroups + 1 : 0; break;
      case WK_ENGINE_WIN_COLS:
      case WK_ENGINE_WIN_COLS_KR: {
        src = which == WK_ENGINE_WIN_COLS ? D.win_cols : D.win_cols_kr;
        uint32_t total = 0;
        if (src && D.win_off)
          hip_check(hipMemcpy(&total, D.win_off + D.ngroups, sizeof total, hipMemcpyDeviceToHost), "download");
        n = src ? total : 0;
        break;
      }
      case WK_ENGINE_ND_COLS: src = D.nd_cols; n = D.nd_cols ? D.win_total : 0; break;
      case WK_ENGINE_ND_CNT: src = D.nd_cnt; n = D.nd_cnt ? D.ngroups : 0; break;
      case WK_ENGINE_ND_COLS_KR: src = D.nd_cols_kr; n = D.nd_cols_kr ? 3 * D.win_total : 0; break;
      case WK_ENGINE_ND_OFF: src = D.nd_off; n = D.nd_off ? D.ngroups + 1 : 0; break;
      case WK_ENGINE_ND_POS: {
        src = D.nd_pos;
        uint32_t words = 0;
        if (src && D.nd_off)
          hip_check(hipMemcpy(&words, D.nd_off + D.ngroups, sizeof words, hipMemcpyDeviceToHost), "download");
        n = src ? (int64_t)words * 64 : 0;
        break;
      }
      default: fail(WK_ERR_ARG, "unknown engine array");
    }
    if (!count) fail(WK_ERR_ARG, "count is NULL");
    if (!host_dst) {
      *count = n;
      return;
    }
    if (*count != n) fail(WK_ERR_ARG, "count must be " + std::to_string(n) + " for this array");
    if (!n) return;
    int cur = 0;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    hip_check(hipSetDevice(E->device), "hipSetDevice");
    const hipError_t e = hipMemcpy(host_dst, src, esz * (size_t)n, hipMemcpyDeviceToHost);
    (void)hipSetDevice(cur);
    hip_check(e, "download");
  });
}
