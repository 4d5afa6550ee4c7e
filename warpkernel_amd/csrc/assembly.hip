// assembly.hip -- row sums of the padded values: warpKernel_assembly /
// engine::run_assembly (warpkernel.hpp:78-100, 898-929).
//
// y[row_map[p]] = the values of sorted row p summed in slot order: the first
// value, then one add per further slot, trailing pads (value 0) included, in
// the engine's value type (double, or float for an FP32 engine, stored as
// double).  The reference has this kernel for WPK1 only (its WPK2 branch is
// commented out, hpp:901-915); on a WPK2 layout each lane sums its chunk and
// the lanes of a row combine in the pairwise order of warpKernel2
// (hpp:154-162).  Rows without entries get 0, as in run.
//
// The kernel streams values only (no columns, no x): one wave per group
// (W = 64) or per two groups (W = 32), U slots per lane in flight.
#include <hip/hip_runtime.h>

#include "wk_internal.hpp"

namespace wk {
namespace {

struct RowsumArgs {
  const void *vals;
  const int32_t *row_map;
  const int64_t *warp_offsets;
  const uint32_t *max_nz_rows;
  const uint32_t *tpr;      // variant 2
  const uint32_t *row_off;  // variant 2
  double *y;
  int32_t ngroups, nznrows, nrows;
};

__device__ __forceinline__ int32_t rowsum_work_id(int32_t b, int32_t nwork, int32_t remap) {
  if (!remap) return b;
  const int32_t q = nwork >> 3, r = nwork & 7;
  const int32_t xcd = b & 7, idx = b >> 3;
  return xcd * q + (xcd < r ? xcd : r) + idx;
}

template <class V, int W, bool SPLIT, int U>
__global__ __launch_bounds__(512) void rowsum_kernel(RowsumArgs a, int32_t nwork, int32_t remap) {
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  if ((int32_t)blockIdx.x >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(blockIdx.x - nwork) * blockDim.x + threadIdx.x;
    if (pos < a.nrows) a.y[a.row_map[pos]] = 0.0;
    return;
  }
  const int32_t wave = __builtin_amdgcn_readfirstlane(
      rowsum_work_id(blockIdx.x, nwork, remap) * wpb + (int32_t)(threadIdx.x >> 6));
  const int32_t g = W == 64 ? wave : 2 * wave + (lane >> 5);
  const int32_t l = lane & (W - 1);
  const bool valid = g < a.ngroups;
  if (W == 64 ? !valid : 2 * wave >= a.ngroups) return;
  const uint32_t len = valid ? a.max_nz_rows[g] : 0u;
  const V *vp = static_cast<const V *>(a.vals) + (valid ? a.warp_offsets[g] : 0) + l;

  // -0 + v is v bit for bit, so starting at -0 and adding every slot is "first
  // value, then add".  Two U-step batches alternate: one loads while the
  // other is summed (81.0 us against 82.5 with one batch at a time on the 64^3
  // elasticity values; 16-byte loads through an LDS slice measured 82-84);
  // the last, partial batch issues all its loads first.
  V s = V(-0.0);
  auto load = [&](uint32_t at, V (&v)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(vp + (int64_t)(at + u) * W);
  };
  auto sum = [&](const V (&v)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) s = s + v[u];
  };
  const uint32_t nb = len / U;
  uint32_t b = 0;
  V va[U], vb[U];
  if (nb > 0) load(0, va);
  for (; b + 3 <= nb; b += 2) {
    load((b + 1) * U, vb);
    sum(va);
    load((b + 2) * U, va);
    sum(vb);
  }
  if (b + 2 == nb) {
    load((b + 1) * U, vb);
    sum(va);
    sum(vb);
  } else if (b + 1 == nb) {
    sum(va);
  }
  const uint32_t i = nb * U;
  if (i < len) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) v[u] = __builtin_nontemporal_load(vp + (int64_t)(i + u) * W);
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) s = s + v[u];
  }

  uint32_t tpr = 1, row0 = (uint32_t)g * (uint32_t)W;
  if constexpr (SPLIT) {
    tpr = valid ? a.tpr[g] : 1u;
    row0 = valid ? a.row_off[g] : 0u;
    // tpr is uniform per group; with W = 32 the two halves may differ, and a
    // lane only ever adds partials of its own row (width W keeps halves apart)
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const V other = __shfl_down(s, (int)o, W);
      if (o < tpr) s = s + other;
    }
  }
  if (!valid) return;
  const uint32_t rowid = row0 + (uint32_t)l / tpr;
  if (((uint32_t)l & (tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  a.y[a.row_map[rowid]] = (double)s;
}

template <class V, int W, bool SPLIT>
void launch_rowsum_t(const RowsumArgs &a, int32_t wpb, int32_t remap, hipStream_t s) {
  const int64_t waves = W == 64 ? (int64_t)a.ngroups : ((int64_t)a.ngroups + 1) / 2;
  const int64_t nwork = (waves + wpb - 1) / wpb;
  const int64_t nfill = ((int64_t)a.nrows - a.nznrows + wpb * 64 - 1) / (wpb * 64);
  if (nwork + nfill > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
  if (nwork + nfill == 0) return;
  hipLaunchKernelGGL((rowsum_kernel<V, W, SPLIT, 8>), dim3((uint32_t)(nwork + nfill)),
                     dim3(wpb * 64), 0, s, a, (int32_t)nwork, remap);
  hip_check(hipGetLastError(), "run_assembly launch");
}

}  // namespace

void launch_rowsum(const DevLayout &D, int32_t dtype, const wk_launch &cfg, double *y,
                   hipStream_t s) {
  const int wpb = cfg.waves_per_block;
  if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8)
    fail(WK_ERR_ARG, "waves_per_block must be 1, 2, 4 or 8");
  RowsumArgs a{};
  a.vals = D.vals;
  a.row_map = D.row_map;
  a.warp_offsets = D.warp_offsets;
  a.max_nz_rows = D.max_nz_rows;
  a.tpr = D.tpr;
  a.row_off = D.row_off;
  a.y = y;
  a.ngroups = D.ngroups;
  a.nznrows = D.nznrows;
  a.nrows = D.nrows;
  const int32_t remap = cfg.xcd_remap ? 1 : 0;
  const bool split = D.variant == 2;
#define WK_ROWSUM(V)                                                       \
  do {                                                                     \
    if (D.W == 64)                                                         \
      split ? launch_rowsum_t<V, 64, true>(a, wpb, remap, s)               \
            : launch_rowsum_t<V, 64, false>(a, wpb, remap, s);             \
    else                                                                   \
      split ? launch_rowsum_t<V, 32, true>(a, wpb, remap, s)               \
            : launch_rowsum_t<V, 32, false>(a, wpb, remap, s);             \
  } while (0)
  if (dtype == WK_F64)
    WK_ROWSUM(double);
  else
    WK_ROWSUM(float);
#undef WK_ROWSUM
}

}  // namespace wk
