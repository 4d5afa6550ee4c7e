// colinds.hip -- replacing an engine's padded column indices: the reference's
// public engine::device_colinds, which its drivers overwrite with the Kr
// columns and restore (wpk1.cu:160-197), and which engine::run then gathers x
// through (warpkernel.hpp:53-75).
//
// Every new column is checked against [0, ncols) before it lands, because the
// kernels gather x[col].  The x windows were built from the engine's own
// columns; they stay in use only while the columns equal those: after each
// swap one kernel compares every slot's column with the column its window
// entry records (pads: column 0 in both), which is exact, so restoring the
// original columns brings the windows back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "wk_internal.hpp"

namespace wk {
namespace {

constexpr int kFlagThreads = 256;
// flags[b] = 1 if block b saw an index outside [0, bound)
__global__ __launch_bounds__(kFlagThreads) void range_check_kernel(const int32_t *__restrict__ a,
                                                                   int64_t n, int32_t bound,
                                                                   int32_t *__restrict__ flags) {
  __shared__ int32_t red[kFlagThreads];
  int32_t bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    bad |= (uint32_t)a[i] >= (uint32_t)bound;
  red[threadIdx.x] = bad;
  __syncthreads();
  for (int o = kFlagThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] |= red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) flags[blockIdx.x] = red[0];
}

// flags[b] = 1 if a slot of the groups block b visits holds a column other
// than its window entry's: one wave per group (grid-stride over groups), lane l
// walks slots warp_offsets[g] + l + 64 i, i < max_nz_rows[g] (group width 64).
__global__ __launch_bounds__(kFlagThreads) void window_match_kernel(
    const int32_t *__restrict__ cols, const int64_t *__restrict__ warp_offsets,
    const uint32_t *__restrict__ max_nz_rows, const uint16_t *__restrict__ win_idx,
    const uint32_t *__restrict__ win_off, const int32_t *__restrict__ win_cols, int32_t ngroups,
    int32_t *__restrict__ flags) {
  __shared__ int32_t red[kFlagThreads];
  constexpr int kWaves = kFlagThreads / 64;
  const int32_t lane = threadIdx.x & 63;
  int32_t bad = 0;
  for (int64_t g = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6); g < ngroups;
       g += (int64_t)gridDim.x * kWaves) {
    const int64_t base = warp_offsets[g] + lane;
    const uint32_t len = max_nz_rows[g];
    const int32_t *wc = win_cols + win_off[g];
    for (uint32_t i = 0; i < len; i++) {
      const int64_t s = base + 64 * (int64_t)i;
      bad |= cols[s] != wc[win_idx[s]];
    }
  }
  red[threadIdx.x] = bad;
  __syncthreads();
  for (int o = kFlagThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] |= red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) flags[blockIdx.x] = red[0];
}

// Runs `launch` (which writes flags[0..nb)) and returns whether any flag is set.
// The flags are read back into host memory on the stack and the stream is
// synchronised.
template <class F>
bool any_flag(int32_t *flags, int32_t nb, hipStream_t s, const char *what, F &&launch) {
  if (nb > kColFlags) fail(WK_ERR_LIMIT, "more check blocks than flags");
  launch();
  hipError_t e = hipGetLastError();
  int32_t h[kColFlags];
  if (e == hipSuccess)
    e = hipMemcpyAsync(h, flags, sizeof(int32_t) * (size_t)nb, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hip_check(e, what);
  return std::any_of(h, h + nb, [](int32_t v) { return v != 0; });
}

}  // namespace

void swap_colinds(wk_engine *E, const int32_t *pcols, int32_t where, hipStream_t s,
                  int32_t *flags) {
  DevLayout &D = E->fwd;
  const int64_t n = D.allocate_nz;
  if (n <= 0) return;
  if (E->has_adj)
    fail(WK_ERR_STATE, "the columns of an engine with the adjoint layout cannot be replaced "
                       "(its transposed layout and dvals follow the scanned pattern)");
  // the kernels gather x[col]: reject out-of-range columns before they land
  if (where) {
    const int32_t nb = (int32_t)std::min<int64_t>((n + 4095) / 4096, kColFlags);
    if (any_flag(flags, nb, s, "index range check", [&] {
          hipLaunchKernelGGL(range_check_kernel, dim3(nb), dim3(kFlagThreads), 0, s, pcols, n,
                             E->ncols, flags);
        }))
      fail(WK_ERR_ARG, "padded column out of range [0, " + std::to_string(E->ncols) + ")");
  } else {
    for (int64_t i = 0; i < n; i++)
      if ((uint32_t)pcols[i] >= (uint32_t)E->ncols)
        fail(WK_ERR_ARG, "padded column out of range at slot " + std::to_string(i));
  }
  hip_check(hipMemcpyAsync(D.cols, pcols, sizeof(int32_t) * (size_t)n,
                           where ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s),
            "set_colinds");
  // windows: in use while the columns are the ones they were built from
  if (D.win_max > 0) {
    D.win_max_built = D.win_max;
    D.win_total_built = D.win_total;
  }
  if (D.win_max_built > 0 && D.win_idx && D.win_off && D.win_cols) {
    const int32_t nb = (int32_t)std::min<int64_t>(((int64_t)D.ngroups + 3) / 4, kColFlags);
    const bool differ = any_flag(flags, nb, s, "window match", [&] {
      hipLaunchKernelGGL(window_match_kernel, dim3(nb), dim3(kFlagThreads), 0, s, D.cols,
                         D.warp_offsets, D.max_nz_rows, D.win_idx, D.win_off, D.win_cols,
                         D.ngroups, flags);
    });
    D.win_max = differ ? 0 : D.win_max_built;
    D.win_total = differ ? 0 : D.win_total_built;
  }
  hip_check(hipStreamSynchronize(s), "set_colinds");
}

}  // namespace wk
