// kernels.hip -- WPK SpMV kernels for gfx950 (CDNA4, wave64).
//
// One template covers the reference's four SpMV kernels:
//   warpKernel            (warpkernel.hpp:53-75)    SPLIT=false, OUT_RUN
//   warpKernel_reorderx   (warpkernel.hpp:103-125)  SPLIT=false, OUT_RUNX
//   warpKernel2           (warpkernel.hpp:127-168)  SPLIT=true,  OUT_RUN
//   warpKernel2_reorderx  (warpkernel.hpp:170-211)  SPLIT=true,  OUT_RUNX
// plus the residual r = b - A x, k right-hand sides per pass and FP32 values
// with compensated or FP64 accumulation.
//
// Mapping.  With W = 64 one wavefront owns one group: lane l streams sorted
// row (group*64 + l) through slots warp_offsets[g] + l + 64*i, so each step of
// the wave is one contiguous 512 B (FP64) + 256 B (index) read.  Group
// metadata is wave-uniform and lives in SGPRs.  W = 32 layouts (the
// reference's own) run two groups per wave, one per half.  WPK2's shared-memory
// tree (hpp:154-162) becomes cross-lane shuffles in the same pairwise order.
//
// Summation order per lane is the reference's: first product rounded, then one
// fma per further slot (nvcc contracts `sum += A*x`), trailing pads included
// (value 0, column 0).  This file is compiled with -ffp-contract=off, so every
// fusion below is an explicit fma.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <type_traits>

#include "spmv_device.hpp"
#include "wk_internal.hpp"

namespace wk {
namespace {

constexpr int kMaxThreads = 512;

// ----------------------------------------------------------- block mapping
// Workgroups are handed to the 8 XCDs round-robin by id.  With remap, the
// blocks an XCD receives cover one contiguous range of work ids, so the x
// window of neighbouring groups stays in that XCD's L2.
__device__ __forceinline__ int32_t work_id(int32_t b, int32_t nwork, int32_t remap) {
  if (!remap) return b;
  const int32_t q = nwork >> 3, r = nwork & 7;
  const int32_t xcd = b & 7, idx = b >> 3;
  return xcd * q + (xcd < r ? xcd : r) + idx;
}

// ----------------------------------------------------------------- kernel
// One group's slice of a wave: base offset, steps, lanes per row, first row.
struct GroupMeta {
  int64_t base;
  uint32_t len, tpr, row0;
  bool valid;
};

template <int W, bool SPLIT>
__device__ __forceinline__ GroupMeta load_meta(const SpmvArgs &a, int32_t g) {
  GroupMeta m;
  m.valid = g < a.ngroups;
  m.base = m.valid ? a.warp_offsets[g] : 0;
  m.len = m.valid ? a.max_nz_rows[g] : 0u;
  m.tpr = SPLIT && m.valid ? a.tpr[g] : 1u;
  m.row0 = SPLIT ? (m.valid ? a.row_off[g] : 0u) : (uint32_t)g * (uint32_t)W;
  return m;
}

// grid.x = nwork + fill blocks.  Work block -> (group block, k-tile), k-tile
// fastest, so blocks running together stream the same groups.  With W = 64 a
// wave owns R consecutive groups ("rows per lane" = R); with W = 32 it owns
// two groups, one per half-wave.  Fill blocks write the rows without nonzeros
// (sorted positions nznrows..nrows-1).
template <class P, int W, bool SPLIT, int OUT, int KT, int U, int NT, int R>
__global__ __launch_bounds__(kMaxThreads) void spmv_kernel(SpmvArgs a, int32_t nwork,
                                                           int32_t ktiles, int32_t remap) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  static_assert(W == 64 || R == 1, "half-wave groups take one group per half");
  extern __shared__ char occupancy_cap[];  // sized by the launcher only to limit blocks per CU
  (void)occupancy_cap;
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  if ((int32_t)blockIdx.x >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(blockIdx.x - nwork) * blockDim.x + threadIdx.x;
    if (pos >= a.nrows) return;
    const int64_t dst = OUT == OUT_RUNX ? pos : (int64_t)a.row_map[pos];
    for (int32_t j = 0; j < ktiles * KT; j++)
      a.y[dst * a.ldy + j] = OUT == OUT_RESID ? a.b[dst * a.ldb + j] : 0.0;
    return;
  }
  const int32_t w = work_id(blockIdx.x, nwork, remap);
  const int32_t gb = w / ktiles, kt = w - gb * ktiles;
  const int32_t j0 = kt * KT;
  const int32_t wave = __builtin_amdgcn_readfirstlane(gb * wpb + (int32_t)(threadIdx.x >> 6));

  int32_t l;
  GroupMeta m[R];
  if constexpr (W == 64) {
    l = lane;
    if (wave * R >= a.ngroups) return;
#pragma unroll
    for (int r = 0; r < R; r++) m[r] = load_meta<W, SPLIT>(a, wave * R + r);
  } else {
    l = lane & 31;
    if (2 * wave >= a.ngroups) return;
    m[0] = load_meta<W, SPLIT>(a, 2 * wave + (lane >> 5));
  }

  const V *vp[R];
  const int32_t *cp[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    vp[r] = static_cast<const V *>(a.vals) + m[r].base + l;
    cp[r] = a.cols + m[r].base + l;
  }
  const V *const(&vpc)[R] = vp;
  const int32_t *const(&cpc)[R] = cp;
  const X *__restrict__ xb = static_cast<const X *>(a.x) + j0;
  const int64_t ldx = a.ldx;

  // every slot, the first included, by fma onto -0 (see PF64::init)
  Acc acc[R][KT];
  uint32_t common = 0xffffffffu;
#pragma unroll
  for (int r = 0; r < R; r++) {
#pragma unroll
    for (int j = 0; j < KT; j++) P::init(acc[r][j]);
    common = m[r].len < common ? m[r].len : common;
  }
  // all R groups together up to the shortest, then each group alone
  uint32_t i = stream_steps<P, W, KT, U, NT, R>(acc, vpc, cpc, xb, ldx, 0u, common);
#pragma unroll
  for (int r = 0; r < R; r++) {
    const V *const vr[1] = {vp[r]};
    const int32_t *const cr[1] = {cp[r]};
    Acc ar[1][KT];
#pragma unroll
    for (int j = 0; j < KT; j++) ar[0][j] = acc[r][j];
    const uint32_t ii = stream_steps<P, W, KT, U, NT, 1>(ar, vr, cr, xb, ldx, i, m[r].len);
    stream_tail<P, W, KT, U, NT>(ar[0], vp[r], cp[r], xb, ldx, ii, m[r].len);
#pragma unroll
    for (int j = 0; j < KT; j++) acc[r][j] = ar[0][j];
  }

#pragma unroll
  for (int r = 0; r < R; r++) {
    if constexpr (SPLIT) {
      // lockstep pairwise tree of hpp:157-162: lane t adds lane t+o's partial
      for (uint32_t o = 1; o < m[r].tpr; o <<= 1) {
#pragma unroll
        for (int j = 0; j < KT; j++) {
          Acc other = P::shfl_down(acc[r][j], (int)o, W);
          P::add(acc[r][j], other);
        }
      }
    }
    if (!m[r].valid) continue;
    const uint32_t tpr = m[r].tpr;
    const uint32_t rowid = m[r].row0 + (uint32_t)l / tpr;
    if (((uint32_t)l & (tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) continue;
    const int64_t dst = OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
    double *yr = a.y + dst * a.ldy + j0;
#pragma unroll
    for (int j = 0; j < KT; j++) {
      const double res = P::result(acc[r][j]);
      if constexpr (OUT == OUT_RESID)
        yr[j] = a.b[dst * a.ldb + j0 + j] - res;
      else
        yr[j] = res;
    }
  }
}

// ------------------------------------------------------- LDS-staged kernel
// Same result as spmv_kernel<P, 64, SPLIT, OUT, 1, U, true, 1>, different load
// shape.  A batch of U steps of one group is a contiguous block of U*64 values
// and U*64 indices; the wave copies it with 16-byte-per-lane loads (1 KB per
// wave instruction instead of 512 B / 256 B) into its own LDS slice, and each
// lane reads its slots back (ds_read, bank-conflict free).  Narrow per-lane
// loads are limited by the texture-address rate, wide ones reach the HBM
// ceiling.  Only the wave touches its slice, so no barrier is needed: LDS
// operations of one wave complete in order.
template <class P, bool SPLIT, int OUT, int U, bool PIPE>
__device__ __forceinline__ void staged_block(const SpmvArgs &a, int32_t bid, int32_t nwork,
                                             int32_t remap, char *lds) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  constexpr int W = 64;
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  if (bid >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(bid - nwork) * blockDim.x + threadIdx.x;
    if (pos >= a.nrows) return;
    const int64_t dst = OUT == OUT_RUNX ? pos : (int64_t)a.row_map[pos];
    a.y[dst * a.ldy] = OUT == OUT_RESID ? a.b[dst * a.ldb] : 0.0;
    return;
  }
  const int32_t wib = threadIdx.x >> 6;
  const int32_t wave = __builtin_amdgcn_readfirstlane(work_id(bid, nwork, remap) * wpb + wib);
  if (wave >= a.ngroups) return;
  const GroupMeta m = load_meta<W, SPLIT>(a, wave);
  const V *vbase = static_cast<const V *>(a.vals) + m.base;
  const int32_t *cbase = a.cols + m.base;
  const X *__restrict__ xb = static_cast<const X *>(a.x);
  const int64_t ldx = a.ldx;

  Acc acc = staged_group<P, U, PIPE>(vbase, cbase, xb, ldx, m.len,
                                      lds + wib * staged_bytes<P, U>(), lane);
  if constexpr (SPLIT) {
    for (uint32_t o = 1; o < m.tpr; o <<= 1) {
      Acc other = P::shfl_down(acc, (int)o, W);
      P::add(acc, other);
    }
  }
  const uint32_t rowid = m.row0 + (uint32_t)lane / m.tpr;
  if (((uint32_t)lane & (m.tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  const int64_t dst = OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
  const double res = P::result(acc);
  if constexpr (OUT == OUT_RESID)
    a.y[dst * a.ldy] = a.b[dst * a.ldb] - res;
  else
    a.y[dst * a.ldy] = res;
}

template <class P, bool SPLIT, int OUT, int U, bool PIPE>
__global__ __launch_bounds__(kMaxThreads) void spmv_staged_kernel(SpmvArgs a, int32_t nwork,
                                                                  int32_t remap) {
  extern __shared__ __attribute__((aligned(16))) char lds[];
  staged_block<P, SPLIT, OUT, U, PIPE>(a, (int32_t)blockIdx.x, nwork, remap, lds);
}

// ------------------------------------------------------- x-window kernel
// Same result as spmv_staged_kernel; each wave first gathers x at its group's
// distinct columns into its own LDS window (the group's columns come from a
// few L2 gathers instead of one per slot), then streams values and 16-bit
// window positions (10 bytes per slot instead of 12) and reads x from LDS.
struct WinArgs {
  const uint16_t *idx;   // [allocate_nz] window position of each slot
  const uint32_t *off;   // [ngroups+1]
  const int32_t *cols;   // window columns in the product's numbering
  int32_t wave_bytes;    // LDS per wave: window, then slice
  int32_t win_bytes;     // window part (16-byte multiple)
  const uint32_t *pk;      // packed positions (spmv_winpk_kernel)
  const uint32_t *pk_off;  // [ngroups+1] first 12-step block of each group
  const int32_t *nd_cols;  // node entries from off[g] (spmv_winnd_kernel)
  const int32_t *nd_cols_kr;  // the same, three Kr columns each, from 3 * off[g]
  const uint32_t *nd_cnt;  // [ngroups]
  const uint32_t *nd_pos;  // node positions, word-major
  const uint32_t *nd_off;  // [ngroups+1] first word of each group
};

// Work block -> (group block, k-tile), k-tile fastest: the blocks of one
// group's k-tiles run together and share its values and positions in L2.
// WPW = 1: each wave of the block takes its own group and window.  WPW > 1
// (k-tiles above 1): the block's WPW waves take one group, fill its window
// together (one barrier), and wave w sums values [w*KT/WPW, (w+1)*KT/WPW) of
// the k-tile, so a window serves WPW waves.  LDS: per wave [window | slice]
// (WPW = 1), or [window | slice of wave 0 | ... ] (WPW > 1).  D batches of
// values and positions in flight through the slices, or D = 0: loaded
// straight into registers, no slices.
template <class P, bool SPLIT, int OUT, int U, int KT, int D, int WPW, bool MASK = false>
__global__ __launch_bounds__(kMaxThreads) void spmv_win_kernel(SpmvArgs a, WinArgs w, int32_t nwork,
                                                               int32_t ktiles, int32_t remap) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  constexpr int W = 64;
  constexpr int KW = KT / WPW;  // right-hand sides per wave
  static_assert(KW * WPW == KT, "waves split the k-tile evenly");
  extern __shared__ __attribute__((aligned(16))) char lds[];
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  const int32_t bid = blockIdx.x;
  if (bid >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(bid - nwork) * blockDim.x + threadIdx.x;
    if (pos >= a.nrows) return;
    const int64_t dst = OUT == OUT_RUNX ? pos : (int64_t)a.row_map[pos];
    for (int32_t t = 0; t < a.k; t++)
      a.y[dst * a.ldy + t] = OUT == OUT_RESID ? a.b[dst * a.ldb + t] : 0.0;
    return;
  }
  const int32_t wid = work_id(bid, nwork, remap);
  const int32_t gb = wid / ktiles, kt = wid - gb * ktiles;
  const int32_t wib = threadIdx.x >> 6;
  // a.k columns in all; MASK: the last k-tile may hold fewer than KT
  // the group, and the first of this wave's right-hand sides in the tile
  const int32_t wave = __builtin_amdgcn_readfirstlane(WPW > 1 ? gb : gb * wpb + wib);
  const uint32_t t0 = WPW > 1 ? (uint32_t)wib * KW : 0u;
  const int32_t j0 = kt * KT;
  if (wave >= a.ngroups) return;  // WPW > 1: the same for the whole block
  const GroupMeta m = load_meta<W, SPLIT>(a, wave);
  const V *vbase = static_cast<const V *>(a.vals) + m.base;
  const uint16_t *ibase = w.idx + m.base;
  char *xs = WPW > 1 ? lds : lds + wib * w.wave_bytes;
  char *slice = WPW > 1 ? lds + w.win_bytes + wib * win_slice_bytes<P, U>() : xs + w.win_bytes;
  (void)slice;  // D = 0: no slice
  const uint32_t w0 = w.off[wave], w1 = w.off[wave + 1];
  const X *xk = static_cast<const X *>(a.x) + j0;
  const int32_t kv = MASK && a.k - j0 < KT ? a.k - j0 : KT;  // columns of this tile in x
  Acc acc[KW];
  // window columns first, then the first value batches, then the x gathers
  WinFill<X, KT, 64 * WPW> f;
  const int32_t first = WPW > 1 ? wib * 64 + lane : lane;
  f.cols(w.cols + w0, w1 - w0, first);
  if constexpr (D == 0) {
    // values and positions straight into registers (no slice)
    f.template gather<MASK>(xs, w.cols + w0, w1 - w0, xk, a.ldx, first, kv);
    if constexpr (WPW > 1) __syncthreads();  // the window is the block's
    window_group_regs<P, U, KT, KW>(acc, vbase, ibase, xs, m.len, t0, lane);
  } else if constexpr (D == 1) {
    static_assert(WPW == 1, "one batch in flight only with a window per wave");
    WinBatch<P, U> bt;
    const bool early = m.len >= (uint32_t)U;
    if (early) bt.issue(vbase, ibase, 0, lane);  // in flight while the window fills
    f.template gather<MASK>(xs, w.cols + w0, w1 - w0, xk, a.ldx, first, kv);
    window_group<P, U, KT>(acc, vbase, ibase, xs, m.len, bt, early, slice, lane);
  } else {
    WinPipe<P, U, D> pp;
    pp.prologue(vbase, ibase, m.len, lane);  // in flight while the window fills
    f.template gather<MASK>(xs, w.cols + w0, w1 - w0, xk, a.ldx, first, kv);
    if constexpr (WPW > 1) __syncthreads();  // the window is the block's
    window_group_deep<P, U, KT, KW, D>(acc, vbase, ibase, xs, m.len, pp, slice, t0, lane);
  }
  if constexpr (SPLIT) {
    for (uint32_t o = 1; o < m.tpr; o <<= 1) {
#pragma unroll
      for (int t = 0; t < KW; t++) {
        Acc other = P::shfl_down(acc[t], (int)o, W);
        P::add(acc[t], other);
      }
    }
  }
  const uint32_t rowid = m.row0 + (uint32_t)lane / m.tpr;
  if (((uint32_t)lane & (m.tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  const int64_t dst = OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
  const int64_t c0 = (int64_t)j0 + t0;
#pragma unroll
  for (int t = 0; t < KW; t++) {
    if constexpr (MASK)
      if (c0 + t >= a.k) break;  // masked columns of a last k-tile
    const double res = P::result(acc[t]);
    if constexpr (OUT == OUT_RESID)
      a.y[dst * a.ldy + c0 + t] = a.b[dst * a.ldb + c0 + t] - res;
    else
      a.y[dst * a.ldy + c0 + t] = res;
  }
}

// One right-hand side through the x window with packed 10-bit positions:
// spmv_win_kernel<P, SPLIT, OUT, U, 1, 1, 1> with batches of NB 12-step
// blocks whose positions (16 bytes per lane per block) go straight into
// registers; the index stream shrinks from 2 to 1.33 bytes per slot.
template <class P, bool SPLIT, int OUT, int NB>
__global__ __launch_bounds__(kMaxThreads) void spmv_winpk_kernel(SpmvArgs a, WinArgs w, int32_t nwork,
                                                                 int32_t remap) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  constexpr int W = 64;
  constexpr int U = WinBatchPk<P, NB>::U;
  extern __shared__ __attribute__((aligned(16))) char lds[];
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  const int32_t bid = blockIdx.x;
  if (bid >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(bid - nwork) * blockDim.x + threadIdx.x;
    if (pos >= a.nrows) return;
    const int64_t dst = OUT == OUT_RUNX ? pos : (int64_t)a.row_map[pos];
    a.y[dst * a.ldy] = OUT == OUT_RESID ? a.b[dst * a.ldb] : 0.0;
    return;
  }
  const int32_t wib = threadIdx.x >> 6;
  const int32_t wave = __builtin_amdgcn_readfirstlane(work_id(bid, nwork, remap) * wpb + wib);
  if (wave >= a.ngroups) return;
  const GroupMeta m = load_meta<W, SPLIT>(a, wave);
  const V *vbase = static_cast<const V *>(a.vals) + m.base;
  const u32x4 *pk = reinterpret_cast<const u32x4 *>(w.pk) + (int64_t)w.pk_off[wave] * 64;
  char *xs = lds + wib * w.wave_bytes;
  const uint32_t w0 = w.off[wave], w1 = w.off[wave + 1];
  WinFill<X, 1> f;
  f.cols(w.cols + w0, w1 - w0, lane);
  WinBatchPk<P, NB> bt;
  const bool early = m.len >= (uint32_t)U;
  if (early) bt.issue(vbase, pk, 0, lane);  // in flight while the window fills
  f.gather(xs, w.cols + w0, w1 - w0, static_cast<const X *>(a.x), a.ldx, lane);
  Acc acc;
  window_group_pk<P, NB>(acc, vbase, pk, xs, m.len, bt, early, xs + w.win_bytes, lane);
  if constexpr (SPLIT) {
    for (uint32_t o = 1; o < m.tpr; o <<= 1) {
      Acc other = P::shfl_down(acc, (int)o, W);
      P::add(acc, other);
    }
  }
  const uint32_t rowid = m.row0 + (uint32_t)lane / m.tpr;
  if (((uint32_t)lane & (m.tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  const int64_t dst = OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
  const double res = P::result(acc);
  if constexpr (OUT == OUT_RESID)
    a.y[dst * a.ldy] = a.b[dst * a.ldb] - res;
  else
    a.y[dst * a.ldy] = res;
}

// One right-hand side through a node window (layouts whose slots read x in
// aligned column triples, as elasticity's three unknowns per node): the
// same results as spmv_winpk_kernel, with one 4-byte window entry per node
// instead of three and 10 bits of position per three slots instead of per
// slot.  Batches of 9 * NW steps of values through the slice, positions in
// registers; NW = 0 (layouts whose groups have at most 27 steps): the group's
// values and positions straight into registers, no slice.
template <class P, bool SPLIT, int OUT, int NW>
__global__ __launch_bounds__(kMaxThreads) void spmv_winnd_kernel(SpmvArgs a, WinArgs w, int32_t nwork,
                                                                 int32_t remap) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  constexpr int W = 64;
  extern __shared__ __attribute__((aligned(16))) char lds[];
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  const int32_t bid = blockIdx.x;
  if (bid >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(bid - nwork) * blockDim.x + threadIdx.x;
    if (pos >= a.nrows) return;
    const int64_t dst = OUT == OUT_RUNX ? pos : (int64_t)a.row_map[pos];
    a.y[dst * a.ldy] = OUT == OUT_RESID ? a.b[dst * a.ldb] : 0.0;
    return;
  }
  const int32_t wib = threadIdx.x >> 6;
  const int32_t wave = __builtin_amdgcn_readfirstlane(work_id(bid, nwork, remap) * wpb + wib);
  if (wave >= a.ngroups) return;
  const GroupMeta m = load_meta<W, SPLIT>(a, wave);
  const V *vbase = static_cast<const V *>(a.vals) + m.base;
  const uint32_t *pw = w.nd_pos + (int64_t)w.nd_off[wave] * 64 + lane;
  char *xs = lds + wib * w.wave_bytes;
  constexpr bool KR = OUT == OUT_RUNX;  // entries as three Kr columns each
  const int32_t *nc = KR ? w.nd_cols_kr + 3 * (int64_t)w.off[wave] : w.nd_cols + w.off[wave];
  const uint32_t nn = w.nd_cnt[wave];
  NdFill<X, KR> f;
  f.cols(nc, nn, lane);
  Acc acc;
  if constexpr (NW == 0) {
    // one batch of at most 27 steps (m.len <= 27, checked by the launcher),
    // in flight while the window fills
    // (issuing the window's x gathers ahead of the values measured slower
    // here: WPK2 at threshold 27, 64^3, 102.5 us against 101.8)
    V v[27];
#pragma unroll
    for (int u = 0; u < 27; u++)
      if ((uint32_t)u < m.len) v[u] = __builtin_nontemporal_load(vbase + (int64_t)u * W + lane);
    uint32_t p[3] = {0u, 0u, 0u};
#pragma unroll
    for (int q = 0; q < 3; q++)
      if ((uint32_t)(9 * q) < m.len) p[q] = pw[q * 64];
    f.gather(xs, nc, nn, static_cast<const X *>(a.x), a.ldx, lane);
    P::init(acc);  // every slot by fma onto -0
#pragma unroll
    for (int t = 0; t < 9; t++)
      if ((uint32_t)(3 * t) < m.len) {
        const X *e = reinterpret_cast<const X *>(xs + WinBatchNd<P, 3>::entry(p, t));
#pragma unroll
        for (int d = 0; d < 3; d++) P::step(acc, v[3 * t + d], e[d]);
      }
  } else {
    constexpr int U = WinBatchNd<P, NW>::U;
    WinBatchNd<P, NW> bt;
    const bool early = m.len >= (uint32_t)U;
    // the window's first x gathers ahead of the first value batch (hex
    // elasticity 64^3: 88.2-88.6 us against 89.1-89.6 with the batch first),
    // the batch in flight while the window fills
    X xv[4][3];
    f.issue_x(xv, static_cast<const X *>(a.x), a.ldx);
    if (early) bt.issue(vbase, pw, 0, lane);
    f.write_x(xs, xv, nn, lane);
    f.rest(xs, nc, nn, static_cast<const X *>(a.x), a.ldx, lane);
    window_group_nd<P, NW>(acc, vbase, pw, xs, m.len, bt, early, xs + w.win_bytes, lane);
  }
  if constexpr (SPLIT) {
    for (uint32_t o = 1; o < m.tpr; o <<= 1) {
      Acc other = P::shfl_down(acc, (int)o, W);
      P::add(acc, other);
    }
  }
  const uint32_t rowid = m.row0 + (uint32_t)lane / m.tpr;
  if (((uint32_t)lane & (m.tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  const int64_t dst = OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
  const double res = P::result(acc);
  if constexpr (OUT == OUT_RESID)
    a.y[dst * a.ldy] = a.b[dst * a.ldb] - res;
  else
    a.y[dst * a.ldy] = res;
}

// Several right-hand sides with the group's values held in registers: each
// wave loads its lane's values and packed window positions once (at most
// 12 * NB steps), then makes one pass per k-tile of KT right-hand sides: fill
// the x window with the tile's KT values per column (KT * sizeof(X) = 32
// bytes per entry), sum every slot from registers and LDS, store the tile.
// The values stream from HBM once whatever k, the window is a quarter (FP64)
// of the 8-wide one, and the only LDS traffic is the x reads and the fill.
// Per lane and right-hand side the order is the other kernels' (every slot by
// fma onto -0, pads included), so results are bit-identical.  A pass reuses
// the window of the previous one: the wave's LDS operations complete in
// order, so the fill's writes follow the previous pass's reads.  All of a
// group's passes stay in one wave: dealing them to two to eight waves (each
// loading the group's values) measured slower at every size (hex elasticity
// 32^3 k=32: 99.6 us in one wave, 111.4 to 214 split; 64^3 k=32: 803
// against 860 to 1746).
template <class P, bool SPLIT, int OUT, int NB>
__global__ __launch_bounds__(kMaxThreads) void spmv_winres_kernel(SpmvArgs a, WinArgs w, int32_t nwork,
                                                                  int32_t ntiles, int32_t remap) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  constexpr int W = 64;
  constexpr int KT = 32 / (int)sizeof(X);  // right-hand sides per pass
  constexpr int NS = 12 * NB;              // steps held in registers
  constexpr int kAhead = 4;
  extern __shared__ __attribute__((aligned(16))) char lds[];
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  const int32_t bid = blockIdx.x;
  if (bid >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(bid - nwork) * blockDim.x + threadIdx.x;
    if (pos >= a.nrows) return;
    const int64_t dst = OUT == OUT_RUNX ? pos : (int64_t)a.row_map[pos];
    for (int32_t t = 0; t < ntiles * KT; t++)
      a.y[dst * a.ldy + t] = OUT == OUT_RESID ? a.b[dst * a.ldb + t] : 0.0;
    return;
  }
  const int32_t wib = threadIdx.x >> 6;
  const int32_t wave = __builtin_amdgcn_readfirstlane(work_id(bid, nwork, remap) * wpb + wib);
  if (wave >= a.ngroups) return;
  const GroupMeta m = load_meta<W, SPLIT>(a, wave);  // m.len <= NS (checked by the launcher)
  const V *vp = static_cast<const V *>(a.vals) + m.base + lane;
  const u32x4 *pk = reinterpret_cast<const u32x4 *>(w.pk) + (int64_t)w.pk_off[wave] * 64 + lane;
  // the lane's positions and values, all loads issued before the first use
  u32x4 pw[NB];
  V v[NS];
#pragma unroll
  for (int q = 0; q < NB; q++)
    if ((uint32_t)(12 * q) < m.len) pw[q] = pk[q * 64];
#pragma unroll
  for (int i = 0; i < NS; i++)
    if ((uint32_t)i < m.len) v[i] = __builtin_nontemporal_load(vp + (int64_t)i * W);
  char *xs = lds + wib * w.wave_bytes;
  const uint32_t w0 = w.off[wave], nwin = w.off[wave + 1] - w0;
  const uint32_t rowid = m.row0 + (uint32_t)lane / m.tpr;
  const bool writer = ((uint32_t)lane & (m.tpr - 1)) == 0 && rowid < (uint32_t)a.nznrows;
  const int64_t dst = !writer ? 0 : OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
  for (int32_t t = 0; t < ntiles; t++) {
    // the positions as if new each pass: otherwise every step's window
    // address is hoisted out of this loop and held in registers
#pragma unroll
    for (int q = 0; q < NB; q++) {
      uint32_t c0 = pw[q].x, c1 = pw[q].y, c2 = pw[q].z, c3 = pw[q].w;
      asm volatile("" : "+v"(c0), "+v"(c1), "+v"(c2), "+v"(c3));
      pw[q] = u32x4{c0, c1, c2, c3};
    }
    WinFill<X, KT, 64, 32> f;
    f.cols(w.cols + w0, nwin, lane);
    f.gather(xs, w.cols + w0, nwin, static_cast<const X *>(a.x) + t * KT, a.ldx, lane);
    Acc acc[KT];
#pragma unroll
    for (int j = 0; j < KT; j++) P::init(acc[j]);  // every slot by fma onto -0
#pragma unroll
    for (int i = 0; i < NS; i++) {
      if ((uint32_t)i < m.len) {  // wave-uniform
        X xv[KT];
        WinEntry<X, KT>::read(xs, WinBatchPk<P, NB>::pos(pw[i / 12], i % 12), xv);
#pragma unroll
        for (int j = 0; j < KT; j++) P::step(acc[j], v[i], xv[j]);
      }
      // at most kAhead steps of window reads scheduled ahead of their FMAs:
      // the values already take 2 * NS registers (FP64)
      if (i % kAhead == kAhead - 1) __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (SPLIT) {
      for (uint32_t o = 1; o < m.tpr; o <<= 1) {
#pragma unroll
        for (int j = 0; j < KT; j++) {
          Acc other = P::shfl_down(acc[j], (int)o, W);
          P::add(acc[j], other);
        }
      }
    }
    if (writer) {
      const int64_t c0 = (int64_t)t * KT;
#pragma unroll
      for (int j = 0; j < KT; j++) {
        const double res = P::result(acc[j]);
        if constexpr (OUT == OUT_RESID)
          a.y[dst * a.ldy + c0 + j] = a.b[dst * a.ldb + c0 + j] - res;
        else
          a.y[dst * a.ldy + c0 + j] = res;
      }
    }
  }
}

// ------------------------------------------------------------ small kernels
// Streaming read of a buffer: 16-byte nontemporal loads, 8 per thread issued
// before use.  The XOR of the data only decides whether one word is stored,
// which keeps the loads; its duration gives the device's streaming-read rate.
__global__ __launch_bounds__(256) void stream_read_kernel(const u32x4 *__restrict__ p, int64_t n,
                                                          uint32_t *sink) {
  constexpr int U = 8;
  uint32_t s = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * U;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x * U + threadIdx.x; base < n; base += stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + (int64_t)u * blockDim.x;
      v[u] = i < n ? __builtin_nontemporal_load(p + i) : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < U; u++) s ^= v[u].x ^ v[u].w;
  }
  if (s == 0x9e3779b9u && sink) *sink = s;
}

__global__ void fill_zero_kernel(uint32_t *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

template <class T>
__global__ void scatter_kernel(const T *__restrict__ src, const uint32_t *__restrict__ slot,
                               T *__restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[slot[i]] = src[i];
}

// thrust::scatter of the reference's _GPU helpers (warpkernel.hpp:459-502),
// bounded: indices at or beyond dst_count are skipped.
template <class T>
__global__ void scatter_bounded_kernel(const T *__restrict__ src, const uint32_t *__restrict__ index,
                                       T *__restrict__ dst, int64_t n, int64_t dst_count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t j = index[i];
    if ((int64_t)j < dst_count) dst[j] = src[i];
  }
}

// dvals[t] = sum_j dy[row_t][j] * x[col_t][j], CSR order.  With one
// right-hand side each block takes a tile of kDvTile consecutive entries;
// tile_row[b] (built with the engine) is the row holding the tile's first
// entry.  Each thread loads its entries' columns and issues their x gathers
// first; then the block's threads walk the rows the tile spans and stamp
// each entry slot of an LDS array with its row's dy value, so after one
// barrier every entry finds its dy with one conflict-free LDS read.  The x
// gathers are in flight during the tile_row -> rowptr -> dy chain, and no
// global load follows the barrier.  Column reads and dvals writes are
// coalesced, one entry per lane (consecutive lanes, consecutive entries: a
// wave's x gathers fall on the fewest lines; four entries per lane with
// 16-byte loads and stores measured 13% slower on hex diffusion 48^3).
// Stamped row numbers with a dy gather after the barrier measured 5% slower
// (11.0 against 10.45 us at 48^3), and a stored row per entry slower still
// (it adds 4 bytes per entry); with several right-hand sides the stored rows
// pay, and dvals_coo_kernel uses them.
constexpr int kDvTile = 2048;
constexpr int kDvThreads = 256;

struct DvArgs {
  const int32_t *rowptr, *colind, *tile_row;
  int64_t nnz;
  const double *dy;
  int64_t lddy;
  const double *x;
  int64_t ldx;
  int32_t k;
  double *dv;
};

// One right-hand side (several go through dvals_coo_kernel), a tile of M
// table tiles (M * kDvTile entries, sdy as many): its rows lie between the
// table rows of its first tile and of the tile after its last.
template <int M>
__device__ __forceinline__ void dvals_tile(const DvArgs &d, int32_t tile, double *sdy) {
  constexpr int TILE = M * kDvTile;
  constexpr int PER = TILE / kDvThreads;
  const int64_t e0 = (int64_t)tile * TILE;
  const int64_t e1 = e0 + TILE < d.nnz ? e0 + TILE : d.nnz;
  const int64_t ntab = (d.nnz + kDvTile - 1) / kDvTile;  // table tiles
  const int64_t tnext = (int64_t)(tile + 1) * M < ntab ? (int64_t)(tile + 1) * M : ntab;
  const int32_t r0 = d.tile_row[(int64_t)tile * M];
  const int32_t r1 = d.tile_row[tnext];
  int32_t c[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int64_t e = e0 + threadIdx.x + (int64_t)q * kDvThreads;
    c[q] = e < e1 ? d.colind[e] : 0;
  }
  double xv[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) xv[q] = d.x[(int64_t)c[q] * d.ldx];
  // rows of the tile lie in [r0, r1]: the next tile starts in row r1 or later.
  // Each row is stamped by P = kDvThreads / rows threads (at least 1), each
  // writing every P-th entry, so long rows do not serialise on one thread
  // (hex elasticity 32^3, 81-entry rows: 22.2 -> 20.1 us)
  const int32_t nr = r1 - r0 + 1;
  const int32_t P = nr >= kDvThreads ? 1 : kDvThreads / nr;
  for (int32_t idx = (int32_t)threadIdx.x; idx < nr * P; idx += kDvThreads) {
    const int32_t r = r0 + idx / P, part = idx % P;
    const int64_t a = d.rowptr[r] > e0 ? d.rowptr[r] : e0;
    const int64_t b = d.rowptr[r + 1] < e1 ? d.rowptr[r + 1] : e1;
    if (a + part >= b) continue;
    const double dyr = d.dy[(int64_t)r * d.lddy];
    for (int64_t e = a + part; e < b; e += P) sdy[e - e0] = dyr;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int64_t e = e0 + threadIdx.x + (int64_t)q * kDvThreads;
    if (e < e1) d.dv[e] = sdy[e - e0] * xv[q];
  }
}

// dvals with k right-hand sides, from the row of every entry (COO rows): no
// row lookup, no LDS; for each entry the first product, then one fma per
// further pair (by fma onto -0, as in the SpMV), 8 pairs per chunk with all
// 16 loads of a chunk issued before its fmas (hex diffusion 48^3, k=8: 45 us
// with a per-pair loop, 28.6 through the tiles, 25.2 here).
constexpr int kCooPer = 8;
__global__ __launch_bounds__(kDvThreads) void dvals_coo_kernel(DvArgs d, const int32_t *rowind) {
  constexpr int KC = 8;
  const int64_t e0 = (int64_t)blockIdx.x * (kDvThreads * kCooPer) + threadIdx.x;
  int32_t r[kCooPer], c[kCooPer];
#pragma unroll
  for (int q = 0; q < kCooPer; q++) {
    const int64_t e = e0 + (int64_t)q * kDvThreads;
    const bool ok = e < d.nnz;
    r[q] = ok ? rowind[e] : 0;
    c[q] = ok ? d.colind[e] : 0;
  }
#pragma unroll 2
  for (int q = 0; q < kCooPer; q++) {
    const int64_t e = e0 + (int64_t)q * kDvThreads;
    if (e >= d.nnz) continue;
    const double *xr = d.x + (int64_t)c[q] * d.ldx;
    const double *dr = d.dy + (int64_t)r[q] * d.lddy;
    double s = -0.0;
    int32_t j = 0;
    for (; j + KC <= d.k; j += KC) {
      double xv[KC], dyv[KC];
#pragma unroll
      for (int u = 0; u < KC; u++) {
        xv[u] = xr[j + u];
        dyv[u] = dr[j + u];
      }
#pragma unroll
      for (int u = 0; u < KC; u++) s = __builtin_fma(dyv[u], xv[u], s);
    }
    if (j < d.k) {
      double xv[KC - 1], dyv[KC - 1];
#pragma unroll
      for (int u = 0; u < KC - 1; u++)
        if (j + u < d.k) {
          xv[u] = xr[j + u];
          dyv[u] = dr[j + u];
        }
#pragma unroll
      for (int u = 0; u < KC - 1; u++)
        if (j + u < d.k) s = __builtin_fma(dyv[u], xv[u], s);
    }
    d.dv[e] = s;
  }
}

__global__ __launch_bounds__(kDvThreads) void dvals_kernel(DvArgs d) {
  __shared__ double sdy[kDvTile];
  dvals_tile<1>(d, (int32_t)blockIdx.x, sdy);
}

// Tiles of the combined backward: one table tile (2048 entries, 16 KB of dy
// stamps) each.  Two tiles per block (32 KB) leave fewer blocks resident:
// hex diffusion 48^3 dx and dvals 21.2 us against 18.6 with one.
constexpr int kBwdTiles = 1;

// dx = A^T dy and dvals in one launch, for layouts of short groups: the
// adjoint blocks (the direct body, one right-hand side, no LDS) and the dvals
// tiles (kBwdTiles x 16 KB of LDS) are spread evenly over the grid, so one kind's
// latency-bound phases overlap the other's streaming and the launch and tail
// costs are paid once.  Block b is dvals tile floor(b*ndv/T) when that count
// steps at b, else adjoint block b - floor(b*ndv/T), T = nspmv + ndv.
template <bool SPLIT>
__device__ __forceinline__ void direct_block(const SpmvArgs &a, int32_t bid, int32_t nwork) {
  using P = PF64;
  using Acc = P::Acc;
  constexpr int W = 64, U = 8;
  const int32_t lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  if (bid >= nwork) {
    const int64_t pos = (int64_t)a.nznrows + (int64_t)(bid - nwork) * blockDim.x + threadIdx.x;
    if (pos < a.nrows) a.y[(int64_t)a.row_map[pos] * a.ldy] = 0.0;
    return;
  }
  const int32_t wave = __builtin_amdgcn_readfirstlane(bid * wpb + (int32_t)(threadIdx.x >> 6));
  if (wave >= a.ngroups) return;
  const GroupMeta m = load_meta<W, SPLIT>(a, wave);
  const double *vp[1] = {static_cast<const double *>(a.vals) + m.base + lane};
  const int32_t *cp[1] = {a.cols + m.base + lane};
  const double *const(&vpc)[1] = vp;
  const int32_t *const(&cpc)[1] = cp;
  const double *xb = static_cast<const double *>(a.x);
  Acc acc[1][1];
  P::init(acc[0][0]);
  const uint32_t i = stream_steps<P, W, 1, U, true, 1>(acc, vpc, cpc, xb, a.ldx, 0u, m.len);
  stream_tail<P, W, 1, U, true>(acc[0], vp[0], cp[0], xb, a.ldx, i, m.len);
  Acc r = acc[0][0];
  if constexpr (SPLIT) {
    for (uint32_t o = 1; o < m.tpr; o <<= 1) {
      Acc other = P::shfl_down(r, (int)o, W);
      P::add(r, other);
    }
  }
  const uint32_t rowid = m.row0 + (uint32_t)lane / m.tpr;
  if (((uint32_t)lane & (m.tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  a.y[(int64_t)a.row_map[rowid] * a.ldy] = P::result(r);
}

template <bool SPLIT>
__global__ __launch_bounds__(kDvThreads) void backward_kernel(SpmvArgs a, int32_t nwork,
                                                              int32_t nspmv, DvArgs d,
                                                              int32_t ndv) {
  __shared__ double sdy[kBwdTiles * kDvTile];
  const int64_t T = (int64_t)nspmv + ndv;
  const int64_t b = blockIdx.x;
  const int64_t before = b * ndv / T;
  const int64_t upto = (b + 1) * ndv / T;
  if (upto > before)
    dvals_tile<kBwdTiles>(d, (int32_t)before, sdy);
  else
    direct_block<SPLIT>(a, (int32_t)(b - before), nwork);
}

// ------------------------------------------------------------- dispatch
using LaunchFn = void (*)(const SpmvArgs &, dim3, dim3, size_t, int32_t, int32_t, int32_t,
                          hipStream_t);

template <class P, int W, bool SPLIT, int OUT, int KT, int U, int NT, int R>
void launch_one(const SpmvArgs &a, dim3 grid, dim3 block, size_t lds, int32_t nwork,
                int32_t ktiles, int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_kernel<P, W, SPLIT, OUT, KT, U, NT, R>), grid, block, lds, s, a,
                     nwork, ktiles, remap);
}

// Instantiated shapes.  k_tile 1 (the bandwidth-bound case) has the unroll,
// rows-per-lane and load-policy choices; wider k-tiles carry more registers
// per slot and keep one shape each.
template <class P, int W, bool SPLIT, int OUT>
LaunchFn pick_kt(int kt, int u, int nt, int r) {
  if constexpr (W == 32) {
    // the reference's own layout width: parity path, one right-hand side per pass
    return kt == 1 && r == 1 ? &launch_one<P, 32, SPLIT, OUT, 1, 8, true, 1> : nullptr;
  } else {
    if (kt == 1) {
      if (nt == 0) return r == 1 && u == 8 ? &launch_one<P, W, SPLIT, OUT, 1, 8, 0, 1> : nullptr;
      // nontemporal index loads at unroll 8, one or two groups per lane;
      // elsewhere nontemporal 2 loads as 1
      if (nt == 2 && u == 8 && r == 1) return &launch_one<P, W, SPLIT, OUT, 1, 8, 2, 1>;
      if (nt == 2 && u == 8 && r == 2) return &launch_one<P, W, SPLIT, OUT, 1, 8, 2, 2>;
      if (r == 1) {
        if (u == 4) return &launch_one<P, W, SPLIT, OUT, 1, 4, true, 1>;
        if (u == 8) return &launch_one<P, W, SPLIT, OUT, 1, 8, true, 1>;
        if (u == 16) return &launch_one<P, W, SPLIT, OUT, 1, 16, true, 1>;
      } else if (r == 2) {
        if (u == 4) return &launch_one<P, W, SPLIT, OUT, 1, 4, true, 2>;
        if (u == 8) return &launch_one<P, W, SPLIT, OUT, 1, 8, true, 2>;
      } else if (r == 4) {
        if (u == 4) return &launch_one<P, W, SPLIT, OUT, 1, 4, true, 4>;
        if (u == 8) return &launch_one<P, W, SPLIT, OUT, 1, 8, true, 4>;
      }
      return nullptr;
    }
    if (r != 1) return nullptr;
    switch (kt) {
      case 2: return &launch_one<P, W, SPLIT, OUT, 2, 4, true, 1>;
      case 4: return &launch_one<P, W, SPLIT, OUT, 4, 2, true, 1>;
      case 8: return &launch_one<P, W, SPLIT, OUT, 8, 2, true, 1>;
      case 16: return &launch_one<P, W, SPLIT, OUT, 16, 1, true, 1>;
      case 32: return &launch_one<P, W, SPLIT, OUT, 32, 1, true, 1>;
    }
    return nullptr;
  }
}

template <class P, int W, bool SPLIT>
LaunchFn pick_out(int out, int kt, int u, int nt, int r) {
  if (out == OUT_RUN) return pick_kt<P, W, SPLIT, OUT_RUN>(kt, u, nt, r);
  if (out == OUT_RUNX) return pick_kt<P, W, SPLIT, OUT_RUNX>(kt, u, nt, r);
  if constexpr (std::is_same_v<P, PF64>) return pick_kt<P, W, SPLIT, OUT_RESID>(kt, u, nt, r);
  return nullptr;  // the residual is FP64 only
}

template <class P, int W>
LaunchFn pick_split(bool split, int out, int kt, int u, int nt, int r) {
  return split ? pick_out<P, W, true>(out, kt, u, nt, r) : pick_out<P, W, false>(out, kt, u, nt, r);
}

LaunchFn pick(int prec, int W, bool split, int out, int kt, int u, int nt, int r) {
  // prec: 0 FP64, 1 FP32 compensated, 2 FP32 with FP64 accumulation
  if (prec == 0)
    return W == 64 ? pick_split<PF64, 64>(split, out, kt, u, nt, r)
                   : pick_split<PF64, 32>(split, out, kt, u, nt, r);
  if (W != 64) return nullptr;
  if (prec == 1) return pick_split<PF32C, 64>(split, out, kt, u, nt, r);
  return pick_split<PF32D, 64>(split, out, kt, u, nt, r);
}

using StagedFn = void (*)(const SpmvArgs &, dim3, dim3, size_t, int32_t, int32_t, hipStream_t);

template <class P, bool SPLIT, int OUT, int U, bool PIPE>
void launch_staged(const SpmvArgs &a, dim3 grid, dim3 block, size_t lds, int32_t nwork,
                   int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_staged_kernel<P, SPLIT, OUT, U, PIPE>), grid, block, lds, s, a, nwork, remap);
}

template <class P, bool SPLIT, int OUT>
StagedFn pick_staged_u(int u, bool pipe) {
  if (u == 16) return pipe ? &launch_staged<P, SPLIT, OUT, 16, true> : &launch_staged<P, SPLIT, OUT, 16, false>;
  return pipe ? &launch_staged<P, SPLIT, OUT, 8, true> : &launch_staged<P, SPLIT, OUT, 8, false>;
}

template <class P, bool SPLIT>
StagedFn pick_staged_out(int out, int u, bool pipe) {
  if (out == OUT_RUN) return pick_staged_u<P, SPLIT, OUT_RUN>(u, pipe);
  if (out == OUT_RUNX) return pick_staged_u<P, SPLIT, OUT_RUNX>(u, pipe);
  if constexpr (std::is_same_v<P, PF64>) return pick_staged_u<P, SPLIT, OUT_RESID>(u, pipe);
  return nullptr;
}

template <class P>
StagedFn pick_staged_split(bool split, int out, int u, bool pipe) {
  return split ? pick_staged_out<P, true>(out, u, pipe) : pick_staged_out<P, false>(out, u, pipe);
}

StagedFn pick_staged(int prec, bool split, int out, int u, bool pipe) {
  if (u != 8 && u != 16) return nullptr;
  if (prec == 0) return pick_staged_split<PF64>(split, out, u, pipe);
  if (prec == 1) return pick_staged_split<PF32C>(split, out, u, pipe);
  return pick_staged_split<PF32D>(split, out, u, pipe);
}

// LDS limits of the current device, read once per device (launches are on the
// host's hot path: a kernel of a few microseconds must not wait on queries).
struct LdsLimits {
  int per_cu, per_block;
};

LdsLimits lds_limits() {
  constexpr int kMaxDevices = 64;
  static std::atomic<int64_t> cache[kMaxDevices];  // (per_cu << 32 | per_block) + 1, 0 unset
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev >= 0 && dev < kMaxDevices) {
    const int64_t v = cache[dev].load(std::memory_order_relaxed);
    if (v) return {(int)((v - 1) >> 32), (int)((v - 1) & 0xffffffff)};
  }
  int per_cu = 0, per_block = 0;
  hip_check(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev), "attr");
  hip_check(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attr");
  if (dev >= 0 && dev < kMaxDevices)
    cache[dev].store((((int64_t)per_cu << 32) | (uint32_t)per_block) + 1, std::memory_order_relaxed);
  return {per_cu, per_block};
}

using WinFn = void (*)(const SpmvArgs &, const WinArgs &, dim3, dim3, size_t, int32_t, int32_t,
                       int32_t, hipStream_t);

template <class P, bool SPLIT, int OUT, int U, int KT, int D, int WPW, bool MASK = false>
void launch_win(const SpmvArgs &a, const WinArgs &w, dim3 grid, dim3 block, size_t lds,
                int32_t nwork, int32_t ktiles, int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_win_kernel<P, SPLIT, OUT, U, KT, D, WPW, MASK>), grid, block, lds, s, a,
                     w, nwork, ktiles, remap);
}

// masked last k-tiles: the register-fed shapes of the layout defaults (unroll
// 8, k-tiles 4 and 8, one or two waves per window; run and the residual)
template <class P, bool SPLIT, int OUT>
WinFn pick_win_masked(int u, int kt, int wpw) {
  if (u != 8 || (wpw != 1 && wpw != 2)) return nullptr;
  if (kt == 4)
    return wpw == 1 ? &launch_win<P, SPLIT, OUT, 8, 4, 0, 1, true> : &launch_win<P, SPLIT, OUT, 8, 4, 0, 2, true>;
  if (kt == 8)
    return wpw == 1 ? &launch_win<P, SPLIT, OUT, 8, 8, 0, 1, true> : &launch_win<P, SPLIT, OUT, 8, 8, 0, 2, true>;
  return nullptr;
}

template <class P>
WinFn pick_win_masked_out(bool split, int out, int u, int kt, int wpw) {
  if (out == OUT_RUN)
    return split ? pick_win_masked<P, true, OUT_RUN>(u, kt, wpw) : pick_win_masked<P, false, OUT_RUN>(u, kt, wpw);
  if constexpr (std::is_same_v<P, PF64>)
    if (out == OUT_RESID)
      return split ? pick_win_masked<P, true, OUT_RESID>(u, kt, wpw)
                   : pick_win_masked<P, false, OUT_RESID>(u, kt, wpw);
  return nullptr;
}

WinFn pick_win_mask(int prec, bool split, int out, int u, int kt, int wpw) {
  if (prec == 0) return pick_win_masked_out<PF64>(split, out, u, kt, wpw);
  if (prec == 1) return pick_win_masked_out<PF32C>(split, out, u, kt, wpw);
  return pick_win_masked_out<PF32D>(split, out, u, kt, wpw);
}

// k-tile 1 with unroll 8 or 16; wider k-tiles (2, 4, 8) with unroll 8 and
// two batches in flight (measured best of one, two and four), and 1, 2 or 4
// waves sharing a window where each wave's share is whole 16-byte chunks;
// register-fed (D = 0) also with unroll 16 at k-tiles 4 and 8 (one or two
// waves per window) and, for FP32, down to one right-hand side per wave
template <class P, bool SPLIT, int OUT, int D>
WinFn pick_win_d(int u, int kt, int wpw) {
  constexpr int xsz = (int)sizeof(typename P::X);
  constexpr int D1 = D == 0 ? 0 : 1;  // single right-hand side: one batch in flight
  if (kt == 1) {
    if (wpw != 1) return nullptr;
    if (u == 8) return &launch_win<P, SPLIT, OUT, 8, 1, D1, 1>;
    if (u == 16) return &launch_win<P, SPLIT, OUT, 16, 1, D1, 1>;
    return nullptr;
  }
  if constexpr (D == 0) {
    // register-fed, 16-step batches (two alternating: 32 steps in flight)
    if (u == 16) {
      if (kt == 4 && wpw == 1) return &launch_win<P, SPLIT, OUT, 16, 4, 0, 1>;
      if (kt == 8 && wpw == 1) return &launch_win<P, SPLIT, OUT, 16, 8, 0, 1>;
      if (kt == 8 && wpw == 2) return &launch_win<P, SPLIT, OUT, 16, 8, 0, 2>;
      return nullptr;
    }
  }
  if (u != 8) return nullptr;
  // FP32, register-fed: down to one right-hand side per wave
  constexpr bool fine = xsz == 4 && D == 0;
  if (kt == 2) {
    if (wpw == 1) return &launch_win<P, SPLIT, OUT, 8, 2, D, 1>;
    if constexpr (fine) {
      if (wpw == 2) return &launch_win<P, SPLIT, OUT, 8, 2, D, 2>;
    }
    return nullptr;
  }
  if (kt == 4) {
    if (wpw == 1) return &launch_win<P, SPLIT, OUT, 8, 4, D, 1>;
    if constexpr (2 * xsz >= 16 || fine) {
      if (wpw == 2) return &launch_win<P, SPLIT, OUT, 8, 4, D, 2>;
    }
    if constexpr (fine) {
      if (wpw == 4) return &launch_win<P, SPLIT, OUT, 8, 4, D, 4>;
    }
    return nullptr;
  }
  if (kt == 8) {
    if (wpw == 1) return &launch_win<P, SPLIT, OUT, 8, 8, D, 1>;
    if (wpw == 2) return &launch_win<P, SPLIT, OUT, 8, 8, D, 2>;
    if constexpr (2 * xsz >= 16 || fine) {
      if (wpw == 4) return &launch_win<P, SPLIT, OUT, 8, 8, D, 4>;
    }
    if constexpr (fine) {
      if (wpw == 8) return &launch_win<P, SPLIT, OUT, 8, 8, D, 8>;
    }
    return nullptr;
  }
  return nullptr;
}

// k-tile 1 with unroll 8 or 16; wider k-tiles (2, 4, 8) with two batches in
// flight through the slices (measured best of one, two and four) or, regs,
// loaded straight into registers (the shapes pick_win_d lists)
template <class P, bool SPLIT, int OUT>
WinFn pick_win_u(int u, int kt, int wpw, bool regs) {
  return regs ? pick_win_d<P, SPLIT, OUT, 0>(u, kt, wpw) : pick_win_d<P, SPLIT, OUT, 2>(u, kt, wpw);
}

template <class P, bool SPLIT>
WinFn pick_win_out(int out, int u, int kt, int wpw, bool regs) {
  if (out == OUT_RUN) return pick_win_u<P, SPLIT, OUT_RUN>(u, kt, wpw, regs);
  if (out == OUT_RUNX) return pick_win_u<P, SPLIT, OUT_RUNX>(u, kt, wpw, regs);
  if constexpr (std::is_same_v<P, PF64>) return pick_win_u<P, SPLIT, OUT_RESID>(u, kt, wpw, regs);
  return nullptr;
}

template <class P>
WinFn pick_win_split(bool split, int out, int u, int kt, int wpw, bool regs) {
  return split ? pick_win_out<P, true>(out, u, kt, wpw, regs)
               : pick_win_out<P, false>(out, u, kt, wpw, regs);
}

WinFn pick_win(int prec, bool split, int out, int u, int kt, int wpw, bool regs) {
  if (prec == 0) return pick_win_split<PF64>(split, out, u, kt, wpw, regs);
  if (prec == 1) return pick_win_split<PF32C>(split, out, u, kt, wpw, regs);
  return pick_win_split<PF32D>(split, out, u, kt, wpw, regs);
}

template <class P, bool SPLIT, int OUT, int NB>
void launch_winpk(const SpmvArgs &a, const WinArgs &w, dim3 grid, dim3 block, size_t lds,
                  int32_t nwork, int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_winpk_kernel<P, SPLIT, OUT, NB>), grid, block, lds, s, a, w, nwork,
                     remap);
}

using WinPkFn = void (*)(const SpmvArgs &, const WinArgs &, dim3, dim3, size_t, int32_t, int32_t,
                         hipStream_t);

// nb: 12-step blocks per batch (1 or 2)
template <class P, bool SPLIT, int NB>
WinPkFn pick_winpk_out(int out) {
  if (out == OUT_RUN) return &launch_winpk<P, SPLIT, OUT_RUN, NB>;
  if (out == OUT_RUNX) return &launch_winpk<P, SPLIT, OUT_RUNX, NB>;
  if constexpr (std::is_same_v<P, PF64>) return &launch_winpk<P, SPLIT, OUT_RESID, NB>;
  return nullptr;
}

template <class P, int NB>
WinPkFn pick_winpk_split(bool split, int out) {
  return split ? pick_winpk_out<P, true, NB>(out) : pick_winpk_out<P, false, NB>(out);
}

WinPkFn pick_winpk(int prec, bool split, int out, int nb) {
  if (nb != 1 && nb != 2) return nullptr;
  if (prec == 0) return nb == 1 ? pick_winpk_split<PF64, 1>(split, out) : pick_winpk_split<PF64, 2>(split, out);
  if (prec == 1) return nb == 1 ? pick_winpk_split<PF32C, 1>(split, out) : pick_winpk_split<PF32C, 2>(split, out);
  return nb == 1 ? pick_winpk_split<PF32D, 1>(split, out) : pick_winpk_split<PF32D, 2>(split, out);
}

template <class P, bool SPLIT, int OUT, int NW>
void launch_winnd(const SpmvArgs &a, const WinArgs &w, dim3 grid, dim3 block, size_t lds,
                  int32_t nwork, int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_winnd_kernel<P, SPLIT, OUT, NW>), grid, block, lds, s, a, w, nwork, remap);
}

// run and the residual read entries of three consecutive columns, run_x
// entries of three Kr columns; nw: words (9 steps each) per batch, 2 or 3, or
// 0 (groups of at most 27 steps, no slice)
template <class P, bool SPLIT, int NW>
WinPkFn pick_winnd_out(int out) {
  if (out == OUT_RUN) return &launch_winnd<P, SPLIT, OUT_RUN, NW>;
  if (out == OUT_RUNX) return &launch_winnd<P, SPLIT, OUT_RUNX, NW>;
  if constexpr (std::is_same_v<P, PF64>)
    if (out == OUT_RESID) return &launch_winnd<P, SPLIT, OUT_RESID, NW>;
  return nullptr;
}

template <class P, int NW>
WinPkFn pick_winnd_split(bool split, int out) {
  return split ? pick_winnd_out<P, true, NW>(out) : pick_winnd_out<P, false, NW>(out);
}

template <class P>
WinPkFn pick_winnd_nw(bool split, int out, int nw) {
  if (nw == 0) return pick_winnd_split<P, 0>(split, out);
  if (nw == 2) return pick_winnd_split<P, 2>(split, out);
  if (nw == 3) return pick_winnd_split<P, 3>(split, out);
  return nullptr;
}

WinPkFn pick_winnd(int prec, bool split, int out, int nw) {
  if (prec == 0) return pick_winnd_nw<PF64>(split, out, nw);
  if (prec == 1) return pick_winnd_nw<PF32C>(split, out, nw);
  return pick_winnd_nw<PF32D>(split, out, nw);
}

using WinResFn = void (*)(const SpmvArgs &, const WinArgs &, dim3, dim3, size_t, int32_t, int32_t,
                          int32_t, hipStream_t);

template <class P, bool SPLIT, int OUT, int NB>
void launch_winres(const SpmvArgs &a, const WinArgs &w, dim3 grid, dim3 block, size_t lds,
                   int32_t nwork, int32_t ntiles, int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_winres_kernel<P, SPLIT, OUT, NB>), grid, block, lds, s, a, w, nwork,
                     ntiles, remap);
}

template <class P, bool SPLIT, int NB>
WinResFn pick_winres_out(int out) {
  if (out == OUT_RUN) return &launch_winres<P, SPLIT, OUT_RUN, NB>;
  if (out == OUT_RUNX) return &launch_winres<P, SPLIT, OUT_RUNX, NB>;
  if constexpr (std::is_same_v<P, PF64>) return &launch_winres<P, SPLIT, OUT_RESID, NB>;
  return nullptr;
}

template <class P>
WinResFn pick_winres_p(bool split, int out, int nb) {
  if (nb == 3) return split ? pick_winres_out<P, true, 3>(out) : pick_winres_out<P, false, 3>(out);
  if (nb == 7) return split ? pick_winres_out<P, true, 7>(out) : pick_winres_out<P, false, 7>(out);
  return nullptr;
}

// nb: 12-step blocks held in registers (3 or 7: groups of up to 36 or 84 steps)
WinResFn pick_winres(int prec, bool split, int out, int nb) {
  if (prec == 0) return pick_winres_p<PF64>(split, out, nb);
  if (prec == 1) return pick_winres_p<PF32C>(split, out, nb);
  return pick_winres_p<PF32D>(split, out, nb);
}

size_t staged_lds_per_wave(int prec, int u) {
  return (size_t)u * 64 * ((prec == 0 ? sizeof(double) : sizeof(float)) + sizeof(int32_t));
}

int ceil_pow2(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

}  // namespace

void launch_spmv(const DevLayout &D, int32_t dtype, const wk_launch &cfg, int out, const void *x,
                 int64_t ldx, double *y, int64_t ldy, const double *b, int64_t ldb, int32_t k,
                 hipStream_t stream) {
  if (k < 1) fail(WK_ERR_ARG, "k must be >= 1");
  if (cfg.stage_lds < 0 || cfg.stage_lds > 7) fail(WK_ERR_ARG, "stage_lds must be 0 to 7");
  if (cfg.nontemporal < 0 || cfg.nontemporal > 2) fail(WK_ERR_ARG, "nontemporal must be 0, 1 or 2");
  const int wpb = cfg.waves_per_block;
  if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8) fail(WK_ERR_ARG, "waves_per_block must be 1, 2, 4 or 8");
  if (cfg.k_tile < 1 || cfg.k_tile > 32 || (cfg.k_tile & (cfg.k_tile - 1)))
    fail(WK_ERR_ARG, "k_tile must be a power of two in [1, 32]");
  const int prec = dtype == WK_F64 ? 0 : (cfg.accumulate == WK_ACC_F64 ? 2 : 1);
  if (dtype == WK_F32 && out == OUT_RESID) fail(WK_ERR_ARG, "the residual is computed by FP64 engines");
  const size_t xsz = dtype == WK_F64 ? sizeof(double) : sizeof(float);

  SpmvArgs a{};
  a.vals = D.vals;
  a.cols = (out == OUT_RUNX && D.cols_kr) ? D.cols_kr : D.cols;
  a.row_map = D.row_map;
  a.warp_offsets = D.warp_offsets;
  a.max_nz_rows = D.max_nz_rows;
  a.tpr = D.tpr;
  a.row_off = D.row_off;
  a.ldx = ldx;
  a.ldy = ldy;
  a.ldb = ldb;
  a.ngroups = D.ngroups;
  a.nznrows = D.nznrows;
  a.nrows = D.nrows;
  const bool split = D.variant == 2;
  const int rpl = cfg.rows_per_lane;
  if (rpl != 1 && rpl != 2 && rpl != 4) fail(WK_ERR_ARG, "rows_per_lane must be 1, 2 or 4");
  // dynamic LDS only caps the number of resident blocks per CU
  size_t lds = 0;
  if (cfg.blocks_per_cu > 0) {
    const LdsLimits lim = lds_limits();
    lds = std::min<size_t>((size_t)lim.per_cu / (size_t)cfg.blocks_per_cu, (size_t)lim.per_block);
  }

  int32_t j0 = 0;
  // stage 6: the group's values in registers, one pass per 32-byte k-tile
  // (4 FP64 or 8 FP32 right-hand sides), all passes of a group in one wave;
  // the rest of k (fewer than one tile) continues below
  if (cfg.stage_lds == 6 && D.win_max > 0 && D.win_pk && D.W == 64 && rpl == 1 && cfg.nontemporal) {
    const int ktr = 32 / (int)xsz;
    const int nb = D.pk_longest <= 36 ? 3 : D.pk_longest <= 84 ? 7 : 0;
    const int32_t ntiles = k / ktr;
    const size_t win_bytes = ((size_t)D.win_max * 32 + 15) / 16 * 16;
    WinResFn rf = nb && ntiles ? pick_winres(prec, split, out, nb) : nullptr;
    if (rf && win_bytes * (size_t)wpb <= (size_t)lds_limits().per_block) {
      WinArgs w{};
      w.off = D.win_off;
      w.cols = (out == OUT_RUNX && D.win_cols_kr) ? D.win_cols_kr : D.win_cols;
      w.wave_bytes = (int32_t)win_bytes;
      w.win_bytes = (int32_t)win_bytes;
      w.pk = D.win_pk;
      w.pk_off = D.win_pk_off;
      const int64_t nwork = ((int64_t)D.ngroups + wpb - 1) / wpb;
      const int64_t nf = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);
      if (nwork + nf > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
      a.x = x;
      a.y = y;
      a.b = b;
      a.k = ntiles * ktr;
      if (nwork + nf > 0)
        rf(a, w, dim3((uint32_t)(nwork + nf)), dim3(wpb * 64), std::max(win_bytes * (size_t)wpb, lds),
           (int32_t)nwork, ntiles, cfg.xcd_remap ? 1 : 0, stream);
      hip_check(hipGetLastError(), "spmv launch");
      j0 = ntiles * ktr;
    }
  }
  // k = sum of power-of-two tiles: as many k_tile tiles as fit, then the rest.
  // A rest of 3 or more columns below the tile width goes, where the
  // register-fed windowed kernel takes it (unroll 8, k-tiles 4 and 8, one or
  // two waves per window; run and the residual), as one masked tile of the
  // next power of two (x values of the missing columns read as 0, their y not
  // written): one pass over the values instead of one per power of two (hex
  // elasticity 64^3: k = 7 as 4 + 2 + 1 took 385 us, k = 8 takes 207)
  int kt = D.W == 32 ? 1 : cfg.k_tile;
  bool mask_ok = true;
  while (j0 < k) {
    const int32_t rem = k - j0;
    int32_t kv = 0;  // columns of a launch whose last tile is masked; 0: whole tiles
    if (kt > rem) {
      const int p2 = ceil_pow2(rem);
      if (mask_ok && rem >= 3 && p2 > rem && p2 <= kt) {
        kt = p2;
        kv = rem;
      } else {
        while (kt > rem) kt >>= 1;
      }
    } else if (mask_ok && kt >= 4 && rem % kt >= 3) {
      kv = rem;  // whole tiles and a masked last one in one launch (k = 7 as 4 + 3)
    }
    const int32_t ntiles = kv ? (rem + kt - 1) / kt : rem / kt;
    // rows per lane applies to single right-hand-side tiles of W = 64 layouts
    const int r = (kt == 1 && D.W == 64) ? rpl : 1;
    // LDS staging where it is offered (one right-hand side, W = 64, one group
    // per wave, unroll 8 or 16, nontemporal values); the direct kernel otherwise
    const bool staged = cfg.stage_lds && kt == 1 && D.W == 64 && r == 1 && cfg.nontemporal &&
                        (cfg.unroll == 8 || cfg.unroll == 16);
    // x windows: any k-tile up to 8 (unroll 8 for k-tiles above 1)
    const bool windowed = cfg.stage_lds >= 3 && D.win_max > 0 && D.W == 64 && r == 1 &&
                          cfg.nontemporal &&
                          (kt == 1 ? (cfg.unroll == 8 || cfg.unroll == 16) : kt <= 8);
    // a k-tile-1 rest of a register-fed launch (stage 4, e.g. k = 9 as tiles
    // of 8 and 1) takes node windows where the layout has them
    const int st1 = kt == 1 && k > 1 && cfg.stage_lds == 4 && D.nd_max > 0 ? 7 : cfg.stage_lds;
    if (windowed && st1 == 7 && kt == 1 && D.nd_max > 0 &&
        (out != OUT_RUNX || D.nd_cols_kr)) {
      // node windows: per wave the window of node entries, then a slice of
      // values (unroll 16: 27-step batches, 8: 18-step); groups of at most 27
      // steps take their values straight into registers, no slice (WPK2 at
      // threshold 27, hex elasticity 64^3: 105.8 -> 101.6 us, 32^3: 16.0 ->
      // 14.5 against 18-step batches)
      const size_t vbytes = prec == 0 ? sizeof(double) : sizeof(float);
      const int nw = D.pk_longest <= 27 ? 0 : cfg.unroll == 16 ? 3 : 2;
      const size_t win_bytes = ((size_t)D.nd_max * 3 * xsz + 15) / 16 * 16;
      const size_t wave_bytes = win_bytes + (size_t)9 * nw * 64 * vbytes;
      const size_t need = wave_bytes * (size_t)wpb;
      WinPkFn nf = pick_winnd(prec, split, out, nw);
      if (nf && need <= (size_t)lds_limits().per_block) {
        WinArgs w{};
        w.off = D.win_off;
        w.wave_bytes = (int32_t)wave_bytes;
        w.win_bytes = (int32_t)win_bytes;
        w.nd_cols = D.nd_cols;
        w.nd_cols_kr = D.nd_cols_kr;
        w.nd_cnt = D.nd_cnt;
        w.nd_pos = D.nd_pos;
        w.nd_off = D.nd_off;
        const int64_t nwork = ((int64_t)D.ngroups + wpb - 1) / wpb;
        const int64_t nfill = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);
        if (nwork + nfill > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
        for (int32_t t = 0; t < ntiles; t++) {
          a.x = static_cast<const char *>(x) + (size_t)(j0 + t) * xsz;
          a.y = y + j0 + t;
          a.b = b ? b + j0 + t : nullptr;
          a.k = 1;
          if (nwork + nfill > 0)
            nf(a, w, dim3((uint32_t)(nwork + nfill)), dim3(wpb * 64), std::max(need, lds),
               (int32_t)nwork, cfg.xcd_remap ? 1 : 0, stream);
          hip_check(hipGetLastError(), "spmv launch");
        }
        j0 += ntiles;
        continue;
      }
    }
    if (windowed && st1 >= 5 && kt == 1 && D.win_pk) {  // 5, the rest of 6, 7 elsewhere
      // packed positions: per wave the window, then a 12-step slice of values
      const size_t vbytes = prec == 0 ? sizeof(double) : sizeof(float);
      const size_t win_bytes = ((size_t)D.win_max * xsz + 15) / 16 * 16;
      // unroll 8: 12-step batches (one packed block), unroll 16: 24-step batches
      const int nb = cfg.unroll == 16 ? 2 : 1;
      const size_t wave_bytes = win_bytes + (size_t)12 * nb * 64 * vbytes;
      const size_t need = wave_bytes * (size_t)wpb;
      WinPkFn wf = pick_winpk(prec, split, out, nb);
      if (wf && need <= (size_t)lds_limits().per_block) {
        WinArgs w{};
        w.off = D.win_off;
        w.cols = (out == OUT_RUNX && D.win_cols_kr) ? D.win_cols_kr : D.win_cols;
        w.wave_bytes = (int32_t)wave_bytes;
        w.win_bytes = (int32_t)win_bytes;
        w.pk = D.win_pk;
        w.pk_off = D.win_pk_off;
        const int64_t nwork = ((int64_t)D.ngroups + wpb - 1) / wpb;
        const int64_t nf = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);
        if (nwork + nf > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
        for (int32_t t = 0; t < ntiles; t++) {
          a.x = static_cast<const char *>(x) + (size_t)(j0 + t) * xsz;
          a.y = y + j0 + t;
          a.b = b ? b + j0 + t : nullptr;
          a.k = 1;
          if (nwork + nf > 0)
            wf(a, w, dim3((uint32_t)(nwork + nf)), dim3(wpb * 64), std::max(need, lds), (int32_t)nwork,
               cfg.xcd_remap ? 1 : 0, stream);
          hip_check(hipGetLastError(), "spmv launch");
        }
        j0 += ntiles;
        continue;
      }
    }
    if (windowed) {
      // k-tile 1: per wave, the window (win_max x kt entries, 16-byte
      // multiple), then a U-step slice of values and 16-bit positions, one
      // group per wave.  Wider k-tiles: the block's waves share one group's
      // window, each with its own slice.
      const int u = kt == 1 ? cfg.unroll : ((cfg.stage_lds == 4 || cfg.stage_lds >= 6) && cfg.unroll == 16 ? 16 : 8);
      // waves per window: waves_per_block, at most one per 16-byte chunk of
      // an entry, or (FP32, values in registers) one per right-hand side
      const bool regs = cfg.stage_lds == 4 || cfg.stage_lds >= 6;  // no slices
      const int chunks = regs && xsz == 4 ? kt : std::max(1, kt * (int)xsz / 16);
      const int wpw = kt == 1 ? 1 : std::min(wpb, chunks);
      const int bw = wpw > 1 ? wpw : wpb;  // waves per block
      const size_t vbytes = prec == 0 ? sizeof(double) : sizeof(float);
      const size_t win_bytes = ((size_t)D.win_max * kt * xsz + 15) / 16 * 16;
      const size_t slice_bytes = regs ? 0 : (size_t)u * 64 * (vbytes + sizeof(uint16_t));
      const size_t wave_bytes = win_bytes + slice_bytes;
      const size_t need = wpw > 1 ? win_bytes + slice_bytes * (size_t)wpw : wave_bytes * (size_t)wpb;
      WinFn wf = kv ? (regs ? pick_win_mask(prec, split, out, u, kt, wpw) : nullptr)
                    : pick_win(prec, split, out, u, kt, wpw, regs);
      if (wf && need <= (size_t)lds_limits().per_block) {
        WinArgs w{};
        w.idx = D.win_idx;
        w.off = D.win_off;
        w.cols = (out == OUT_RUNX && D.win_cols_kr) ? D.win_cols_kr : D.win_cols;
        w.wave_bytes = (int32_t)wave_bytes;
        w.win_bytes = (int32_t)win_bytes;
        const int gpb = wpw > 1 ? 1 : wpb;  // groups per block
        const int64_t nb = ((int64_t)D.ngroups + gpb - 1) / gpb;
        const int64_t nf = ((int64_t)D.nrows - D.nznrows + bw * 64 - 1) / (bw * 64);
        const int64_t nwork = nb * ntiles;
        if (nwork + nf > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
        a.x = static_cast<const char *>(x) + (size_t)j0 * xsz;
        a.y = y + j0;
        a.b = b ? b + j0 : nullptr;
        a.k = kv ? kv : ntiles * kt;
        if (nwork + nf > 0)
          wf(a, w, dim3((uint32_t)(nwork + nf)), dim3(bw * 64), std::max(need, lds), (int32_t)nwork,
             ntiles, cfg.xcd_remap ? 1 : 0, stream);
        hip_check(hipGetLastError(), "spmv launch");
        j0 += a.k;
        continue;
      }
      // the window does not fit this block shape: the kernels without windows
    }
    if (kv) {  // only the windowed kernel masks: the rest in whole tiles
      mask_ok = false;
      continue;
    }
    if (staged) {
      StagedFn sf = pick_staged(prec, split, out, cfg.unroll, cfg.stage_lds >= 2);
      if (!sf) fail(WK_ERR_ARG, "unsupported staged configuration");
      const int64_t nb = ((int64_t)D.ngroups + wpb - 1) / wpb;
      const int64_t nf = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);
      const size_t need = staged_lds_per_wave(prec, cfg.unroll) * (size_t)wpb;
      if (need > (size_t)lds_limits().per_block)
        fail(WK_ERR_ARG, "LDS staging: waves_per_block * unroll too large");
      for (int32_t t = 0; t < ntiles; t++) {
        a.x = static_cast<const char *>(x) + (size_t)(j0 + t) * xsz;
        a.y = y + j0 + t;
        a.b = b ? b + j0 + t : nullptr;
        a.k = 1;
        if (nb + nf > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
        if (nb + nf > 0)
          sf(a, dim3((uint32_t)(nb + nf)), dim3(wpb * 64), std::max(need, lds), (int32_t)nb,
             cfg.xcd_remap ? 1 : 0, stream);
        hip_check(hipGetLastError(), "spmv launch");
      }
      j0 += ntiles;
      continue;
    }
    LaunchFn fn = pick(prec, D.W, split, out, kt, cfg.unroll, cfg.nontemporal, r);
    if (!fn) fail(WK_ERR_ARG, "unsupported kernel configuration");
    const int gpw = D.W == 64 ? r : 2;  // groups per wave
    const int64_t waves = ((int64_t)D.ngroups + gpw - 1) / gpw;
    const int64_t ngb = (waves + wpb - 1) / wpb;
    const int64_t nfill = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);
    const int64_t nwork = ngb * ntiles;
    if (nwork + nfill > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
    a.x = static_cast<const char *>(x) + (size_t)j0 * xsz;
    a.y = y + j0;
    a.b = b ? b + j0 : nullptr;
    a.k = ntiles * kt;
    if (nwork + nfill > 0)
      fn(a, dim3((uint32_t)(nwork + nfill)), dim3(wpb * 64), lds, (int32_t)nwork, ntiles,
         cfg.xcd_remap ? 1 : 0, stream);
    hip_check(hipGetLastError(), "spmv launch");
    j0 += ntiles * kt;
  }
}

// The direct kernel with its values not loaded (NT 3), as launch_spmv runs it
// at one right-hand side: the time of the index stream and the x gathers
// alone, the ceiling of a layout whose gathers miss the caches.
template <bool SPLIT>
LaunchFn pick_gather(int u, int r) {
  if (u == 4 && r == 1) return &launch_one<PF64, 64, SPLIT, OUT_RUN, 1, 4, 3, 1>;
  if (u == 8 && r == 1) return &launch_one<PF64, 64, SPLIT, OUT_RUN, 1, 8, 3, 1>;
  if (u == 4 && r == 2) return &launch_one<PF64, 64, SPLIT, OUT_RUN, 1, 4, 3, 2>;
  if (u == 8 && r == 2) return &launch_one<PF64, 64, SPLIT, OUT_RUN, 1, 8, 3, 2>;
  if (u == 4 && r == 4) return &launch_one<PF64, 64, SPLIT, OUT_RUN, 1, 4, 3, 4>;
  if (u == 8 && r == 4) return &launch_one<PF64, 64, SPLIT, OUT_RUN, 1, 8, 3, 4>;
  return nullptr;
}

void launch_gather_ceiling(const DevLayout &D, const wk_launch &cfg, const double *x, double *y,
                           hipStream_t stream) {
  if (D.W != 64) fail(WK_ERR_ARG, "the gather ceiling needs group width 64");
  const int wpb = cfg.waves_per_block;
  if (wpb < 1 || wpb > kMaxThreads / 64) fail(WK_ERR_ARG, "waves_per_block out of range");
  const int r = cfg.rows_per_lane;
  LaunchFn fn = D.variant == 2 ? pick_gather<true>(cfg.unroll, r) : pick_gather<false>(cfg.unroll, r);
  if (!fn) fail(WK_ERR_ARG, "gather ceiling: unroll 4 or 8 and rows_per_lane 1, 2 or 4");
  SpmvArgs a{};
  a.vals = D.vals;
  a.cols = D.cols;
  a.row_map = D.row_map;
  a.warp_offsets = D.warp_offsets;
  a.max_nz_rows = D.max_nz_rows;
  a.tpr = D.tpr;
  a.row_off = D.row_off;
  a.ldx = a.ldy = a.ldb = 1;
  a.ngroups = D.ngroups;
  a.nznrows = D.nznrows;
  a.nrows = D.nrows;
  a.x = x;
  a.y = y;
  a.k = 1;
  const int64_t waves = ((int64_t)D.ngroups + r - 1) / r;
  const int64_t nwork = (waves + wpb - 1) / wpb;
  const int64_t nfill = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);
  if (nwork + nfill > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
  if (nwork + nfill > 0)
    fn(a, dim3((uint32_t)(nwork + nfill)), dim3(wpb * 64), 0, (int32_t)nwork, 1,
       cfg.xcd_remap ? 1 : 0, stream);
  hip_check(hipGetLastError(), "gather ceiling launch");
}

void launch_stream_read(const void *p, int64_t bytes, uint32_t *sink, hipStream_t stream) {
  if (bytes <= 0) return;
  if (bytes % 16 || reinterpret_cast<uintptr_t>(p) % 16)
    fail(WK_ERR_ARG, "stream read: 16-byte aligned buffer and size required");
  hipLaunchKernelGGL(stream_read_kernel, dim3(16384), dim3(256), 0, stream,
                     static_cast<const u32x4 *>(p), bytes / 16, sink);
  hip_check(hipGetLastError(), "stream read launch");
}

// One wave per group: lane l packs its positions of steps 12b .. 12b+11 into
// four words, three 10-bit fields each (steps past the group's length are 0).
__global__ __launch_bounds__(256) void pack_positions_kernel(const uint16_t *__restrict__ idx,
                                                             const int64_t *__restrict__ warp_offsets,
                                                             const uint32_t *__restrict__ max_nz_rows,
                                                             const uint32_t *__restrict__ pk_off,
                                                             int32_t ngroups, u32x4 *__restrict__ out) {
  const int32_t g = (int32_t)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int32_t lane = threadIdx.x & 63;
  if (g >= ngroups) return;
  const int64_t base = warp_offsets[g];
  const uint32_t len = max_nz_rows[g];
  const uint32_t nbat = pk_off[g + 1] - pk_off[g];
  for (uint32_t b = 0; b < nbat; b++) {
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < 12; u++) {
      const uint32_t i = b * 12 + u;
      if (i < len) w[u / 3] |= (uint32_t)idx[base + lane + 64 * (int64_t)i] << (10 * (u % 3));
    }
    out[((int64_t)pk_off[g] + b) * 64 + lane] = u32x4{w[0], w[1], w[2], w[3]};
  }
}

// Node windows from the windows of a layout, one wave per group.  Pass 1
// (nd_pos null) checks that every triple of a lane's steps reads either a node
// (three consecutive window positions holding columns 3n, 3n+1, 3n+2) or one
// position three times, and counts the group's entries: per window position
// in order, a node entry where a node starts there, then a constant entry
// where some triple reads that position three times.  It writes the count, or
// kNdNone when the group does not fit.  Pass 2 (every group fits) writes the
// entries and each lane's entry indices, three 10-bit fields to a word.
constexpr uint32_t kNdNone = 0xffffffffu;
constexpr int kNdWaves = 4;

__global__ __launch_bounds__(kNdWaves * 64) void node_window_kernel(
    const uint16_t *__restrict__ idx, const int64_t *__restrict__ warp_offsets,
    const uint32_t *__restrict__ max_nz_rows, const uint32_t *__restrict__ win_off,
    const int32_t *__restrict__ win_cols, const int32_t *__restrict__ win_cols_kr,
    int32_t ngroups, uint32_t *__restrict__ cnt, const uint32_t *__restrict__ nd_off,
    int32_t *__restrict__ nd_cols, int32_t *__restrict__ nd_cols_kr,
    uint32_t *__restrict__ nd_pos) {
  // per window position: bit 0 a node starts here, bit 1 read as a constant triple
  __shared__ uint8_t kind[kNdWaves][kPackedWindow];
  // per window position: the entry of the node starting there, of the constant
  __shared__ uint16_t entry_of[kNdWaves][kPackedWindow][2];
  const int32_t wib = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int32_t g = (int32_t)blockIdx.x * kNdWaves + wib;
  const bool live = g < ngroups;
  uint32_t n = 0, len = 0, w0 = 0;
  int64_t base = 0;
  if (live) {
    w0 = win_off[g];
    n = win_off[g + 1] - w0;  // at most kPackedWindow (packed windows only)
    len = max_nz_rows[g];
    base = warp_offsets[g];
  }
  uint8_t *kd = kind[wib];
  const int32_t *wc = win_cols + w0;
  for (uint32_t p = (uint32_t)lane; p < n; p += 64) {
    const int32_t c = wc[p];
    const bool st = c % 3 == 0 && p + 2 < n && wc[p + 1] == c + 1 && wc[p + 2] == c + 2;
    kd[p] = st ? 1 : 0;
  }
  __syncthreads();
  bool bad = live && len % 3 != 0;
  if (live && len % 3 == 0) {
    for (uint32_t t = 0; t < len; t += 3) {
      const int64_t s = base + lane + 64 * (int64_t)t;
      const uint32_t p0 = idx[s], p1 = idx[s + 64], p2 = idx[s + 128];
      if (p1 == p0 + 1 && p2 == p0 + 2 && (kd[p0] & 1)) continue;
      if (p0 == p1 && p1 == p2)
        kd[p0] |= 2;  // lanes only set bit 1 here: concurrent writes agree
      else
        bad = true;
    }
  }
  __syncthreads();
  uint32_t total = 0;
  const uint64_t below = (1ull << lane) - 1;
  for (uint32_t p0 = 0; p0 < n; p0 += 64) {
    const uint32_t p = p0 + (uint32_t)lane;
    const uint32_t k = p < n ? kd[p] : 0u;
    const uint64_t ms = __ballot(k & 1u), mc = __ballot((k >> 1) & 1u);
    const uint32_t r = total + (uint32_t)__popcll(ms & below) + (uint32_t)__popcll(mc & below);
    if (k & 1u) {
      entry_of[wib][p][0] = (uint16_t)r;
      if (nd_pos) nd_cols[w0 + r] = wc[p];
      if (nd_pos && nd_cols_kr)
        for (uint32_t d = 0; d < 3; d++) nd_cols_kr[3 * ((int64_t)w0 + r) + d] = win_cols_kr[w0 + p + d];
    }
    if (k & 2u) {
      const uint32_t rc = r + (k & 1u);
      entry_of[wib][p][1] = (uint16_t)rc;
      if (nd_pos) nd_cols[w0 + rc] = -wc[p] - 1;
      if (nd_pos && nd_cols_kr)
        for (uint32_t d = 0; d < 3; d++) nd_cols_kr[3 * ((int64_t)w0 + rc) + d] = win_cols_kr[w0 + p];
    }
    total += (uint32_t)__popcll(ms) + (uint32_t)__popcll(mc);
  }
  bad = __ballot(bad) != 0 || total > n;
  __syncthreads();
  if (!live) return;
  if (!nd_pos) {
    if (lane == 0) cnt[g] = bad ? kNdNone : total;
    return;
  }
  // pass 2: each lane's entry index per triple of steps, 27 steps to 3 words
  const uint32_t nblk = (len + 26) / 27;
  uint32_t *dst = nd_pos + (int64_t)nd_off[g] * 64 + lane;
  for (uint32_t bk = 0; bk < nblk; bk++)
    for (uint32_t q = 0; q < 3; q++) {
      uint32_t word = 0;
      for (uint32_t f = 0; f < 3; f++) {
        const uint32_t t = bk * 27 + q * 9 + f * 3;
        if (t < len) {
          const int64_t s = base + lane + 64 * (int64_t)t;
          const uint32_t p0 = idx[s], p1 = idx[s + 64];
          word |= (uint32_t)entry_of[wib][p0][p1 == p0 + 1 ? 0 : 1] << (10 * f);
        }
      }
      dst[((int64_t)bk * 3 + q) * 64] = word;
    }
}

static void build_node_windows(wk_engine *E, DevLayout &D, hipStream_t stream) {
  D.nd_cols = nullptr;
  D.nd_cols_kr = nullptr;
  D.nd_cnt = nullptr;
  D.nd_pos = nullptr;
  D.nd_off = nullptr;
  D.nd_max = 0;
  const int32_t G = D.ngroups;
  if (!D.win_pk || !D.win_cols || !D.win_off || D.win_total <= 0 || G <= 0) return;
  for (int32_t g = 0; g < G; g++)
    if (D.h_max_nz_rows[g] % 3) return;
  const int64_t blocks = ((int64_t)G + kNdWaves - 1) / kNdWaves;
  uint32_t *cnt = dev_alloc<uint32_t>(E, G);
  hipLaunchKernelGGL(node_window_kernel, dim3((uint32_t)blocks), dim3(kNdWaves * 64), 0, stream,
                     D.win_idx, D.warp_offsets, D.max_nz_rows, D.win_off, D.win_cols, nullptr, G,
                     cnt, nullptr, nullptr, nullptr, nullptr);
  hip_check(hipGetLastError(), "node window launch");
  std::vector<uint32_t> h((size_t)G);
  hip_check(hipMemcpyAsync(h.data(), cnt, sizeof(uint32_t) * (size_t)G, hipMemcpyDeviceToHost, stream),
            "node windows");
  hip_check(hipStreamSynchronize(stream), "node windows");
  uint32_t most = 0;
  std::vector<uint32_t> off((size_t)G + 1, 0);
  for (int32_t g = 0; g < G; g++) {
    if (h[g] == kNdNone) return;
    most = std::max(most, h[g]);
    const uint64_t words = (uint64_t)off[g] + 3ull * ((D.h_max_nz_rows[g] + 26) / 27);
    if (words > UINT32_MAX) return;
    off[g + 1] = (uint32_t)words;
  }
  D.nd_off = dev_upload(E, off.data(), (int64_t)off.size(), stream);
  D.nd_pos = dev_alloc<uint32_t>(E, (int64_t)off.back() * 64);
  D.nd_cols = dev_alloc<int32_t>(E, D.win_total);
  if (D.win_cols_kr) D.nd_cols_kr = dev_alloc<int32_t>(E, 3 * D.win_total);
  hipLaunchKernelGGL(node_window_kernel, dim3((uint32_t)blocks), dim3(kNdWaves * 64), 0, stream,
                     D.win_idx, D.warp_offsets, D.max_nz_rows, D.win_off, D.win_cols,
                     D.win_cols_kr, G, cnt, D.nd_off, D.nd_cols, D.nd_cols_kr, D.nd_pos);
  hip_check(hipGetLastError(), "node window launch");
  hip_check(hipStreamSynchronize(stream), "node windows");
  D.nd_cnt = cnt;
  D.nd_max = (int32_t)most;
}

void build_packed_positions(wk_engine *E, DevLayout &D, hipStream_t stream) {
  D.win_pk = nullptr;
  D.win_pk_off = nullptr;
  if (D.W != 64 || D.win_max <= 0 || D.win_max > kPackedWindow || !D.win_idx || D.ngroups <= 0)
    return;
  std::vector<uint32_t> off((size_t)D.ngroups + 1, 0);
  for (int32_t g = 0; g < D.ngroups; g++) {
    off[g + 1] = off[g] + (D.h_max_nz_rows[g] + 11) / 12;
    D.pk_longest = std::max(D.pk_longest, D.h_max_nz_rows[g]);
  }
  D.win_pk_off = dev_upload(E, off.data(), (int64_t)off.size(), stream);
  D.win_pk = dev_alloc<uint32_t>(E, (int64_t)off.back() * 64 * 4);
  constexpr int wpb = 4;
  const int64_t blocks = ((int64_t)D.ngroups + wpb - 1) / wpb;
  if (D.win_pk)
    hipLaunchKernelGGL(pack_positions_kernel, dim3((uint32_t)blocks), dim3(wpb * 64), 0, stream,
                       D.win_idx, D.warp_offsets, D.max_nz_rows, D.win_pk_off, D.ngroups,
                       reinterpret_cast<u32x4 *>(D.win_pk));
  hip_check(hipGetLastError(), "pack launch");
  hip_check(hipStreamSynchronize(stream), "pack");
  build_node_windows(E, D, stream);
}

void launch_fill_zero(void *p, int64_t bytes, hipStream_t stream) {
  if (bytes <= 0) return;
  if (bytes % 4) fail(WK_ERR_ARG, "fill size must be a multiple of 4 bytes");
  const int64_t n = bytes / 4;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_zero_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream,
                     static_cast<uint32_t *>(p), n);
  hip_check(hipGetLastError(), "fill launch");
}

void launch_scatter_values(const DevLayout &D, int32_t dtype, int64_t nnz, const void *vals_csr,
                           hipStream_t stream) {
  if (nnz <= 0) return;
  if (!D.slot_of_nz) fail(WK_ERR_STATE, "layout has no device scatter map");
  const int64_t blocks = std::min<int64_t>((nnz + 255) / 256, 65536);
  if (dtype == WK_F64)
    hipLaunchKernelGGL(scatter_kernel<double>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       static_cast<const double *>(vals_csr), D.slot_of_nz,
                       static_cast<double *>(D.vals), nnz);
  else
    hipLaunchKernelGGL(scatter_kernel<float>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       static_cast<const float *>(vals_csr), D.slot_of_nz,
                       static_cast<float *>(D.vals), nnz);
  hip_check(hipGetLastError(), "scatter launch");
}

void launch_scatter(const void *src, const uint32_t *index, void *dst, int64_t n, int64_t dst_count,
                    int32_t elem_bytes, hipStream_t stream) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  if (elem_bytes == 8)
    hipLaunchKernelGGL(scatter_bounded_kernel<uint64_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       static_cast<const uint64_t *>(src), index, static_cast<uint64_t *>(dst), n,
                       dst_count);
  else
    hipLaunchKernelGGL(scatter_bounded_kernel<uint32_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       static_cast<const uint32_t *>(src), index, static_cast<uint32_t *>(dst), n,
                       dst_count);
  hip_check(hipGetLastError(), "scatter launch");
}

void launch_dvals(const int32_t *rowptr, const int32_t *colind, const int32_t *tile_row,
                  const int32_t *rowind, int64_t nnz, const double *dy, int64_t lddy,
                  const double *x, int64_t ldx, int32_t k, double *dvals, hipStream_t stream) {
  if (nnz <= 0) return;
  if (k == 1) {
    const int64_t blocks = (nnz + kDvTile - 1) / kDvTile;
    if (blocks > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
    const DvArgs d{rowptr, colind, tile_row, nnz, dy, lddy, x, ldx, k, dvals};
    hipLaunchKernelGGL(dvals_kernel, dim3((uint32_t)blocks), dim3(kDvThreads), 0, stream, d);
  } else {
    const int64_t per = (int64_t)kDvThreads * kCooPer;
    const int64_t blocks = (nnz + per - 1) / per;
    if (blocks > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
    const DvArgs d{rowptr, colind, tile_row, nnz, dy, lddy, x, ldx, k, dvals};
    hipLaunchKernelGGL(dvals_coo_kernel, dim3((uint32_t)blocks), dim3(kDvThreads), 0, stream, d,
                       rowind);
  }
  hip_check(hipGetLastError(), "dvals launch");
}

void launch_backward(const DevLayout &T, const double *dy, double *dx, const int32_t *rowptr,
                     const int32_t *colind, const int32_t *tile_row, int64_t nnz, const double *x,
                     double *dvals, hipStream_t stream) {
  if (T.W != 64) fail(WK_ERR_ARG, "the combined backward needs group width 64");
  constexpr int wpb = kDvThreads / 64;
  SpmvArgs a{};
  a.vals = T.vals;
  a.cols = T.cols;
  a.row_map = T.row_map;
  a.warp_offsets = T.warp_offsets;
  a.max_nz_rows = T.max_nz_rows;
  a.tpr = T.tpr;
  a.row_off = T.row_off;
  a.x = dy;
  a.y = dx;
  a.ldx = a.ldy = a.ldb = 1;
  a.ngroups = T.ngroups;
  a.nznrows = T.nznrows;
  a.nrows = T.nrows;
  a.k = 1;
  const int64_t nwork = ((int64_t)T.ngroups + wpb - 1) / wpb;
  const int64_t nfill = ((int64_t)T.nrows - T.nznrows + kDvThreads - 1) / kDvThreads;
  const int64_t ndv = nnz > 0 ? (nnz + kBwdTiles * kDvTile - 1) / (kBwdTiles * kDvTile) : 0;
  const int64_t total = nwork + nfill + ndv;
  if (total > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
  if (total == 0) return;
  const DvArgs d{rowptr, colind, tile_row, nnz, dy, 1, x, 1, 1, dvals};
  if (T.variant == 2)
    hipLaunchKernelGGL(backward_kernel<true>, dim3((uint32_t)total), dim3(kDvThreads), 0, stream, a,
                       (int32_t)nwork, (int32_t)(nwork + nfill), d, (int32_t)ndv);
  else
    hipLaunchKernelGGL(backward_kernel<false>, dim3((uint32_t)total), dim3(kDvThreads), 0, stream,
                       a, (int32_t)nwork, (int32_t)(nwork + nfill), d, (int32_t)ndv);
  hip_check(hipGetLastError(), "backward launch");
}

int64_t dvals_tiles(int64_t nnz) { return (nnz + kDvTile - 1) / kDvTile; }
int64_t dvals_tile_size() { return kDvTile; }

}  // namespace wk
