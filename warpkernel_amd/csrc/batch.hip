// batch.hip -- many WPK matrices in one launch, and HIP graph capture of any
// sequence of this library's launches.
//
// Work list.  Every group of every matrix is one item (cost: its steps times
// the k-tiles), every 64 rows without nonzeros one fill item.
//
// Two launch shapes, no atomics and no inter-wave communication in either:
//   - one wave per item (waves_per_cu = 0, the default): items ordered by
//     cost, longest first (stable: matrix, then group order among equal
//     costs), so the cheap groups of short-row matrices fill the end of the
//     launch; the hardware dispatcher balances the load;
//   - persistent: waves_per_cu waves per compute unit; items in matrix and
//     group order, cut into contiguous ranges of equal cost, one per wave.
// On 64 mixed FE matrices one wave per item measured 1.9x faster than 8
// persistent waves per CU: a static partition leaves waves with less memory
// parallelism than the dispatcher keeps in flight.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "spmv_device.hpp"
#include "wk_internal.hpp"

namespace wk {

struct BatchDesc {
  const void *vals;
  const int32_t *cols;
  const int32_t *row_map;
  const int64_t *warp_offsets;
  const uint32_t *max_nz_rows;
  const uint32_t *tpr;
  const uint32_t *row_off;
  const void *x;
  double *y;
  int64_t ldx, ldy;
  int32_t nznrows, nrows, ngroups, split;
  int32_t staged;    // stream this matrix's groups through LDS (long rows)
  int32_t windowed;  // and read x from the group's LDS window (stage 3)
  int32_t packed;    // windows with packed positions (stage 5)
  int32_t nodes;     // node windows, 27-step batches (stage 5, k-tile 1)
  const uint16_t *win_idx;
  const uint32_t *win_off;
  const int32_t *win_cols;
  const uint32_t *win_pk;
  const uint32_t *win_pk_off;
  const int32_t *nd_cols;
  const uint32_t *nd_cnt;
  const uint32_t *nd_pos;
  const uint32_t *nd_off;
};

struct BatchItem {
  int32_t m;  // matrix
  int32_t g;  // group (>= 0) or fill chunk -1-g
};

}  // namespace wk

struct wk_batch {
  std::vector<wk_engine *> engines;
  int32_t dtype = WK_F64;
  int device = 0;
  int32_t nwaves = 0;       // persistent waves; 0: one wave per item
  int32_t k = 0;
  int32_t kt = 1;
  int32_t stage = 2;
  std::vector<wk::BatchDesc> desc;
  wk::BatchDesc *d_desc = nullptr;
  // device arrays of a fixed size, allocated at creation and freed only at
  // destruction, so a graph captured after a bind stays valid across re-binds
  // with the same k (they rewrite the arrays in place)
  wk::BatchItem *d_items = nullptr;
  int32_t *d_wstart = nullptr;
  int64_t nitems = 0;       // work items: every group and fill chunk of every engine
  int32_t win_bytes = 0;  // LDS window per wave (largest windowed matrix), 16-byte multiple
};

struct wk_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace wk {
namespace {

// Packed-position batches per 12-step block: two for FP64 values, one for FP32
// (the single-matrix defaults for long groups).
template <class P>
constexpr int batch_nb() {
  return sizeof(typename P::V) == 8 ? 2 : 1;
}

// LDS slice per wave behind the window: the largest of the staged batch, the
// window batch (values and 16-bit positions) and, at STAGE 5, the packed and
// the node-window (27-step) batches.
template <class P, int KT, int U, int STAGE>
constexpr int batch_slice() {
  if constexpr (STAGE == 0 || KT != 1) return 16;
  constexpr int st = staged_bytes<P, U>();
  constexpr int win = U * 64 * (int)(sizeof(typename P::V) + sizeof(uint16_t));
  constexpr int pk = STAGE == 5 ? 12 * batch_nb<P>() * 64 * (int)sizeof(typename P::V) : 0;
  constexpr int nd = STAGE == 5 ? WinBatchNd<P, 3>::U * 64 * (int)sizeof(typename P::V) : 0;
  constexpr int a = st > win ? st : win;
  constexpr int b = pk > nd ? pk : nd;
  return a > b ? a : b;
}

// Waves per block: one at STAGE 5 (a window and a packed slice per block, as
// the single-matrix default), else four.
template <int STAGE>
constexpr int batch_wpb() {
  return STAGE == 5 ? 1 : 4;
}

// Waves per item: at k-tile 8 with windows (stage 5) and one item per block,
// the block's two waves fill the item's window together and each sums four
// of the eight right-hand sides (the single-matrix shape for k = 8); one
// otherwise.
template <int KT, int STAGE>
constexpr int batch_wpi() {
  return KT == 8 && STAGE == 5 ? 2 : 1;
}

// One group of one matrix for right-hand sides [j0, j0+KT): the summation
// order of warpKernel / warpKernel2 (every slot by fma onto -0, i.e. first
// product then fma; the pairwise lane tree for split rows).  With STAGE > 0
// and KT = 1 the group streams through the wave's LDS slice (16-byte loads);
// with STAGE 3 or 5 a windowed matrix reads x from the wave's LDS window (the
// win_bytes in front of the slice), at STAGE 5 with packed positions where the
// matrix has them.
//
// WPI waves per item (k-tiles above 1): wave wib of the block sums right-hand
// sides [j0 + wib*KW, j0 + (wib+1)*KW) of the k-tile, KW = KT / WPI; with a
// window the block's waves fill it together (one barrier; the block runs this
// one item only).
template <class P, int KT, int U, int STAGE, int WPI = 1>
__device__ __forceinline__ void batch_group(const BatchDesc &d, int32_t g, int32_t j0, int32_t lane,
                                            int32_t wib, char *region, int32_t win_bytes) {
  constexpr int KW = KT / WPI;
  static_assert(KW * WPI == KT && (WPI == 1 || KT > 1), "waves split the k-tile evenly");
  char *slice = region + win_bytes;
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  const int64_t base = d.warp_offsets[g];
  const uint32_t len = d.max_nz_rows[g];
  const uint32_t tpr = d.split ? d.tpr[g] : 1u;
  const uint32_t row0 = d.split ? d.row_off[g] : (uint32_t)g * 64u;
  const V *vbase = static_cast<const V *>(d.vals) + base;
  const int32_t *cbase = d.cols + base;
  const X *xb = static_cast<const X *>(d.x) + j0;
  const uint32_t t0 = WPI > 1 ? (uint32_t)wib * KW : 0u;
  Acc acc[KW];
  bool done = false;
  if constexpr (STAGE == 5 && KT > 1) {
    // several right-hand sides: the window holds KT values per column, values
    // and positions go straight into registers (two 8-step batches
    // alternating), as the single-matrix default
    if (d.windowed) {
      const uint32_t w0 = d.win_off[g], w1 = d.win_off[g + 1];
      WinFill<X, KT, 64 * WPI> f;
      const int32_t first = WPI > 1 ? wib * 64 + lane : lane;
      f.cols(d.win_cols + w0, w1 - w0, first);
      f.gather(region, d.win_cols + w0, w1 - w0, xb, d.ldx, first);
      if constexpr (WPI > 1) __syncthreads();  // the window is the block's
      window_group_regs<P, U, KT, KW>(acc, vbase, d.win_idx + base, region, len, t0, lane);
      done = true;
    }
  }
  if constexpr (STAGE == 5 && KT == 1) {
    if (d.windowed && d.nodes) {
      // node windows (three x values per entry, one position per three steps)
      const uint32_t *pw = d.nd_pos + (int64_t)d.nd_off[g] * 64 + lane;
      const int32_t *nc = d.nd_cols + d.win_off[g];
      const uint32_t nn = d.nd_cnt[g];
      NdFill<X, false> f;
      f.cols(nc, nn, lane);
      WinBatchNd<P, 3> bt;
      const bool early = len >= (uint32_t)WinBatchNd<P, 3>::U;
      X xv[4][3];
      f.issue_x(xv, xb, d.ldx);  // the first x gathers ahead of the first batch
      if (early) bt.issue(vbase, pw, 0, lane);  // in flight while the window fills
      f.write_x(region, xv, nn, lane);
      f.rest(region, nc, nn, xb, d.ldx, lane);
      window_group_nd<P, 3>(acc[0], vbase, pw, region, len, bt, early, slice, lane);
      done = true;
    }
    if (!done && d.windowed && d.packed) {
      constexpr int NB = batch_nb<P>();
      char *xs = region;
      const u32x4 *pk = reinterpret_cast<const u32x4 *>(d.win_pk) + (int64_t)d.win_pk_off[g] * 64;
      const uint32_t w0 = d.win_off[g], w1 = d.win_off[g + 1];
      WinFill<X, 1> f;
      f.cols(d.win_cols + w0, w1 - w0, lane);
      WinBatchPk<P, NB> bt;
      const bool early = len >= (uint32_t)WinBatchPk<P, NB>::U;
      if (early) bt.issue(vbase, pk, 0, lane);  // in flight while the window fills
      f.gather(xs, d.win_cols + w0, w1 - w0, xb, d.ldx, lane);
      window_group_pk<P, NB>(acc[0], vbase, pk, xs, len, bt, early, slice, lane);
      done = true;
    }
  }
  if constexpr ((STAGE == 3 || STAGE == 5) && KT == 1) {
    if (!done && d.windowed) {
      char *xs = region;
      const uint16_t *ibase = d.win_idx + base;
      WinBatch<P, U> bt;
      const bool early = len >= (uint32_t)U;
      if (early) bt.issue(vbase, ibase, 0, lane);  // in flight while the window fills
      const uint32_t w0 = d.win_off[g], w1 = d.win_off[g + 1];
      fill_window<X, 1>(xs, d.win_cols + w0, w1 - w0, xb, d.ldx, lane);
      window_group<P, U, 1>(acc, vbase, ibase, xs, len, bt, early, slice, lane);
      done = true;
    }
  }
  if constexpr (STAGE > 0 && KT == 1) {
    if (!done && d.staged) {
      acc[0] = staged_group<P, U, STAGE >= 2>(vbase, cbase, xb, d.ldx, len, slice, lane);
      done = true;
    }
  }
  if (!done) {
    (void)slice;
#pragma unroll
    for (int j = 0; j < KW; j++) P::init(acc[j]);
    const V *const vp[1] = {vbase + lane};
    const int32_t *const cp[1] = {cbase + lane};
    Acc ar[1][KW];
#pragma unroll
    for (int j = 0; j < KW; j++) ar[0][j] = acc[j];
    const uint32_t i = stream_steps<P, 64, KW, U, true, 1>(ar, vp, cp, xb + t0, d.ldx, 0u, len);
    stream_tail<P, 64, KW, U, true>(ar[0], vp[0], cp[0], xb + t0, d.ldx, i, len);
#pragma unroll
    for (int j = 0; j < KW; j++) acc[j] = ar[0][j];
  }
  for (uint32_t o = 1; o < tpr; o <<= 1) {
#pragma unroll
    for (int j = 0; j < KW; j++) {
      Acc other = P::shfl_down(acc[j], (int)o, 64);
      P::add(acc[j], other);
    }
  }
  const uint32_t rowid = row0 + (uint32_t)lane / tpr;
  if (((uint32_t)lane & (tpr - 1)) != 0 || rowid >= (uint32_t)d.nznrows) return;
  double *yr = d.y + (int64_t)d.row_map[rowid] * d.ldy + j0 + t0;
#pragma unroll
  for (int j = 0; j < KW; j++) yr[j] = P::result(acc[j]);
}

template <class P, int KT, int U, int STAGE, int WPI>
__global__ __launch_bounds__(256) void batch_kernel(const BatchDesc *__restrict__ desc,
                                                    const BatchItem *__restrict__ items,
                                                    const int32_t *__restrict__ wstart,
                                                    int32_t nwaves, int32_t ktiles,
                                                    int32_t win_bytes) {
  // wstart == nullptr: wave w runs item w alone.  Per wave, LDS holds the
  // window (win_bytes, stage 3 and 5) and then the staging slice.  WPI > 1:
  // block w runs item w (no wstart), its waves sharing one region.
  constexpr int kSlice = batch_slice<P, KT, U, STAGE>();
  extern __shared__ __attribute__((aligned(16))) char lds[];
  const int32_t region = win_bytes + kSlice;
  const int32_t lane = threadIdx.x & 63;
  const int32_t wib = threadIdx.x >> 6;
  const int32_t w = __builtin_amdgcn_readfirstlane(
      WPI > 1 ? (int32_t)blockIdx.x : (int32_t)(blockIdx.x * (blockDim.x >> 6) + wib));
  if (w >= nwaves) return;  // WPI > 1: the same for the whole block
  const int32_t it0 = wstart ? wstart[w] : w;
  const int32_t it1 = wstart ? wstart[w + 1] : w + 1;
  for (int32_t it = it0; it < it1; it++) {
    const BatchItem item = items[it];
    const BatchDesc d = desc[item.m];
    if (item.g < 0) {
      const int64_t pos = (int64_t)d.nznrows + (int64_t)(-1 - item.g) * 64 + lane;
      if (pos < d.nrows && (WPI == 1 || wib == 0)) {  // WPI > 1: the block's first wave
        double *yr = d.y + (int64_t)d.row_map[pos] * d.ldy;
        for (int32_t j = 0; j < ktiles * KT; j++) yr[j] = 0.0;
      }
      continue;
    }
    for (int32_t kt = 0; kt < ktiles; kt++) {
      if (WPI > 1 && kt > 0) __syncthreads();  // the previous k-tile's window is read
      batch_group<P, KT, U, STAGE, WPI>(d, item.g, kt * KT, lane, wib,
                                        lds + (WPI > 1 ? 0 : wib) * region, win_bytes);
    }
  }
}

using BatchFn = void (*)(const wk_batch *, int32_t, hipStream_t);

template <class P, int KT, int STAGE>
void launch_batch(const wk_batch *B, int32_t ktiles, hipStream_t s) {
  const bool flat = B->nwaves == 0;
  const int64_t waves = flat ? B->nitems : B->nwaves;
  constexpr int kSlice = batch_slice<P, KT, 8, STAGE>();
  const int32_t win = STAGE == 3 || STAGE == 5 ? B->win_bytes : 0;
  constexpr int wpi = batch_wpi<KT, STAGE>();
  if (wpi > 1 && flat) {
    // one item per block, its waves sharing the window
    if (waves > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
    hipLaunchKernelGGL((batch_kernel<P, KT, 8, STAGE, wpi>), dim3((uint32_t)waves), dim3(wpi * 64),
                       (size_t)(win + kSlice), s, B->d_desc, B->d_items, nullptr, (int32_t)waves,
                       ktiles, win);
    return;
  }
  constexpr int wpb = batch_wpb<STAGE>();
  const int64_t blocks = (waves + wpb - 1) / wpb;
  if (blocks > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
  const size_t lds = wpb * (size_t)(win + kSlice);
  hipLaunchKernelGGL((batch_kernel<P, KT, 8, STAGE, 1>), dim3((uint32_t)blocks), dim3(wpb * 64), lds, s,
                     B->d_desc, B->d_items, flat ? nullptr : B->d_wstart, (int32_t)waves, ktiles,
                     win);
}

template <class P>
BatchFn pick_batch_kt(int kt, int stage) {
  switch (kt) {
    case 1:
      if (stage == 5) return &launch_batch<P, 1, 5>;
      if (stage == 3) return &launch_batch<P, 1, 3>;
      if (stage == 2) return &launch_batch<P, 1, 2>;
      if (stage == 1) return &launch_batch<P, 1, 1>;
      return &launch_batch<P, 1, 0>;
    case 2: return stage == 5 ? &launch_batch<P, 2, 5> : &launch_batch<P, 2, 0>;
    case 4: return stage == 5 ? &launch_batch<P, 4, 5> : &launch_batch<P, 4, 0>;
    case 8: return stage == 5 ? &launch_batch<P, 8, 5> : &launch_batch<P, 8, 0>;
  }
  return nullptr;
}

void free_batch(wk_batch *B) {
  if (!B) return;
  (void)hipFree(B->d_desc);
  (void)hipFree(B->d_items);
  (void)hipFree(B->d_wstart);
  delete B;
}

}  // namespace
}  // namespace wk

using namespace wk;

extern "C" {

int wk_batch_create(wk_engine *const *engines, int32_t count, int32_t waves_per_cu,
                    int32_t stage_lds, wk_batch **out) {
  return guarded([&] {
    if (!out || !engines || count < 1) fail(WK_ERR_ARG, "bad argument");
    *out = nullptr;
    if (waves_per_cu < 0 || waves_per_cu > 1024) fail(WK_ERR_ARG, "waves_per_cu must be in [0, 1024]");
    if (stage_lds < 0 || stage_lds > 5 || stage_lds == 4)
      fail(WK_ERR_ARG, "stage_lds must be 0, 1, 2, 3 or 5");
    auto *B = new wk_batch();
    B->stage = stage_lds;
    try {
      hip_check(hipGetDevice(&B->device), "hipGetDevice");
      for (int32_t i = 0; i < count; i++) {
        wk_engine *E = engines[i];
        if (!E) fail(WK_ERR_ARG, "NULL engine in batch");
        if (E->fwd.W != 64) fail(WK_ERR_ARG, "batched engines need group width 64");
        if (i && E->dtype != engines[0]->dtype) fail(WK_ERR_ARG, "batched engines must share the value type");
        if (E->device != B->device) fail(WK_ERR_ARG, "engine lives on another device");
        B->engines.push_back(E);
      }
      B->dtype = engines[0]->dtype;
      int cus = 0;
      hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->device), "attr");
      B->nwaves = waves_per_cu ? cus * waves_per_cu : 0;
      B->desc.resize(count);
      hip_check(hipMalloc(&B->d_desc, sizeof(BatchDesc) * count), "hipMalloc");
      // the work list has one item per group and per 64 rows without nonzeros
      // of every engine, whatever the bound vectors and k
      int64_t nitems = 0;
      for (const wk_engine *E : B->engines)
        nitems += (int64_t)E->fwd.ngroups + (E->fwd.nrows - E->fwd.nznrows + 63) / 64;
      if (nitems > (int64_t)INT32_MAX) fail(WK_ERR_LIMIT, "too many work items");
      B->nitems = nitems;
      if (nitems) hip_check(hipMalloc(&B->d_items, sizeof(BatchItem) * (size_t)nitems), "hipMalloc");
      if (B->nwaves) {
        hip_check(hipMalloc(&B->d_wstart, sizeof(int32_t) * (B->nwaves + 1)), "hipMalloc");
      }
    } catch (...) {
      free_batch(B);
      throw;
    }
    *out = B;
  });
}

int wk_batch_bind(wk_batch *B, const void *const *x_devs, double *const *y_devs, int32_t k,
                  const int64_t *ldx, const int64_t *ldy) {
  return guarded([&] {
    if (!B || !x_devs || !y_devs || !ldx || !ldy) fail(WK_ERR_ARG, "NULL argument");
    if (k < 1 || k > 32) fail(WK_ERR_ARG, "k must be in [1, 32]");
    // (k-tile 4 windows for k = 8 measured slower than k-tile 8: 0.737 against
    // 0.659 ms on 64 FE matrices)
    const int32_t kt = k % 8 == 0 ? 8 : k % 4 == 0 ? 4 : k % 2 == 0 ? 2 : 1;
    const int32_t ktiles = k / kt;
    std::vector<BatchItem> items;
    int32_t win_bytes = 0;
    std::vector<int64_t> cost;
    for (size_t i = 0; i < B->engines.size(); i++) {
      const wk_engine *E = B->engines[i];
      const DevLayout &D = E->fwd;
      if (ldx[i] < k || ldy[i] < k) fail(WK_ERR_ARG, "leading dimension smaller than k");
      BatchDesc &d = B->desc[i];
      d.vals = D.vals;
      d.cols = D.cols;
      d.row_map = D.row_map;
      d.warp_offsets = D.warp_offsets;
      d.max_nz_rows = D.max_nz_rows;
      d.tpr = D.tpr;
      d.row_off = D.row_off;
      d.x = x_devs[i];
      d.y = y_devs[i];
      d.ldx = ldx[i];
      d.ldy = ldy[i];
      d.nznrows = D.nznrows;
      d.nrows = D.nrows;
      d.ngroups = D.ngroups;
      d.split = D.variant == 2;
      // staging pays off for long rows only (it adds LDS traffic per batch)
      const double mean_len = D.ngroups ? (double)D.allocate_nz / ((double)D.ngroups * 64.0) : 0.0;
      d.staged = D.ngroups && mean_len >= 24.0;
      // x windows where each entry serves enough slots, as in the
      // single-matrix default: one right-hand side from 3 slots per entry with
      // packed positions (stage 5), 4.5 without; k-tiles above 1 (stage 5
      // only) from 3
      const bool pk5 = B->stage == 5 && D.win_pk != nullptr;
      d.windowed = D.win_max > 0 && D.win_total > 0 &&
                   (kt == 1 ? (B->stage == 3 || B->stage == 5) &&
                                  (double)D.allocate_nz >= (pk5 ? 3.0 : 4.5) * (double)D.win_total
                            : B->stage == 5 && (double)D.allocate_nz >= 3.0 * (double)D.win_total);
      d.packed = B->stage == 5 && D.win_pk != nullptr;
      // node windows at k-tile 1 where the matrix has them (long groups: the
      // batch runs 27-step batches only)
      d.nodes = B->stage == 5 && kt == 1 && D.nd_max > 0 && mean_len >= 40.0;
      d.win_idx = D.win_idx;
      d.win_off = D.win_off;
      d.win_cols = D.win_cols;
      d.win_pk = D.win_pk;
      d.win_pk_off = D.win_pk_off;
      d.nd_cols = D.nd_cols;
      d.nd_cnt = D.nd_cnt;
      d.nd_pos = D.nd_pos;
      d.nd_off = D.nd_off;
      if (d.windowed) {
        // a block's waves, each with this window and its slice, in 64 KB of
        // LDS; a matrix whose window does not fit reads x directly
        const bool f64 = E->dtype == WK_F64;
        const int64_t entries = d.nodes ? 3 * (int64_t)D.nd_max : (int64_t)D.win_max;
        const int64_t win = (entries * kt * (f64 ? 8 : 4) + 15) / 16 * 16;
        const int64_t slice =
            kt > 1 ? 16
                   : B->stage == 5 ? (f64 ? batch_slice<PF64, 1, 8, 5>() : batch_slice<PF32C, 1, 8, 5>())
                                   : (f64 ? batch_slice<PF64, 1, 8, 3>() : batch_slice<PF32C, 1, 8, 3>());
        const int64_t wpb = B->stage == 5 ? batch_wpb<5>() : batch_wpb<3>();
        if (wpb * (win + slice) <= 65536)
          win_bytes = std::max<int32_t>(win_bytes, (int32_t)win);
        else
          d.windowed = 0;
      }
      if ((D.nrows && !y_devs[i]) || (E->ncols && !x_devs[i])) fail(WK_ERR_ARG, "NULL vector");
      for (int32_t g = 0; g < D.ngroups; g++) {
        items.push_back({(int32_t)i, g});
        cost.push_back((int64_t)D.h_max_nz_rows[g] * ktiles);
      }
      const int32_t nfill = (D.nrows - D.nznrows + 63) / 64;
      for (int32_t c = 0; c < nfill; c++) {
        items.push_back({(int32_t)i, -1 - c});
        cost.push_back(1);
      }
    }
    if (items.size() > (size_t)INT32_MAX) fail(WK_ERR_LIMIT, "too many work items");
    // one wave per item: longest items first, so short ones fill the end.
    // Persistent waves keep matrix order, so a wave's range stays on
    // neighbouring groups of one matrix (x locality).
    if (!B->nwaves) {
      std::vector<int32_t> ord(items.size());
      for (size_t t = 0; t < ord.size(); t++) ord[t] = (int32_t)t;
      std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
      std::vector<BatchItem> sorted(items.size());
      std::vector<int64_t> scost(items.size());
      for (size_t t = 0; t < ord.size(); t++) {
        sorted[t] = items[ord[t]];
        scost[t] = cost[ord[t]];
      }
      items.swap(sorted);
      cost.swap(scost);
    }
    // persistent waves: contiguous ranges of equal cost
    std::vector<int32_t> wstart;
    if (B->nwaves) {
      int64_t total = 0;
      for (int64_t c : cost) total += c;
      wstart.assign((size_t)B->nwaves + 1, 0);
      int64_t acc = 0;
      int32_t w = 0;
      for (size_t it = 0; it < items.size(); it++) {
        while (w < B->nwaves && acc >= total * (w + 1) / B->nwaves) wstart[++w] = (int32_t)it;
        acc += cost[it];
      }
      while (w < B->nwaves) wstart[++w] = (int32_t)items.size();
    }
    if ((int64_t)items.size() != B->nitems) fail(WK_ERR_STATE, "work list changed size since creation");
    if (!items.empty())
      hip_check(hipMemcpy(B->d_items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice), "bind");
    hip_check(hipMemcpy(B->d_desc, B->desc.data(), sizeof(BatchDesc) * B->desc.size(), hipMemcpyHostToDevice), "bind");
    if (!wstart.empty())
      hip_check(hipMemcpy(B->d_wstart, wstart.data(), sizeof(int32_t) * wstart.size(), hipMemcpyHostToDevice), "bind");
    B->win_bytes = win_bytes;
    B->k = k;
    B->kt = kt;
  });
}

int wk_batch_run(wk_batch *B, void *stream) {
  return guarded([&] {
    if (!B) fail(WK_ERR_ARG, "batch is NULL");
    if (!B->k) fail(WK_ERR_STATE, "batch has no bound vectors (wk_batch_bind)");
    if (!B->nitems) return;
    BatchFn fn = B->dtype == WK_F64 ? pick_batch_kt<PF64>(B->kt, B->stage)
                                    : pick_batch_kt<PF32C>(B->kt, B->stage);
    if (!fn) fail(WK_ERR_ARG, "unsupported batch configuration");
    fn(B, B->k / B->kt, static_cast<hipStream_t>(stream));
    hip_check(hipGetLastError(), "batch launch");
  });
}

void wk_batch_destroy(wk_batch *B) { free_batch(B); }

int wk_graph_begin(void *stream) {
  return guarded([&] {
    if (!stream) fail(WK_ERR_ARG, "graph capture needs a non-null stream");
    hip_check(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
  });
}

int wk_graph_end(void *stream, wk_graph **out) {
  return guarded([&] {
    if (!stream || !out) fail(WK_ERR_ARG, "NULL argument");
    *out = nullptr;
    auto *G = new wk_graph();
    hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &G->graph);
    if (e != hipSuccess) {
      delete G;
      hip_check(e, "hipStreamEndCapture");
    }
    e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(G->graph);
      delete G;
      hip_check(e, "hipGraphInstantiate");
    }
    *out = G;
  });
}

int wk_graph_launch(wk_graph *G, void *stream) {
  return guarded([&] {
    if (!G) fail(WK_ERR_ARG, "graph is NULL");
    hip_check(hipGraphLaunch(G->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  });
}

void wk_graph_destroy(wk_graph *G) {
  if (!G) return;
  (void)hipGraphExecDestroy(G->exec);
  (void)hipGraphDestroy(G->graph);
  delete G;
}

}  // extern "C"
