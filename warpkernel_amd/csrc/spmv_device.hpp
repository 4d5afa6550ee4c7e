// spmv_device.hpp -- device code shared by the SpMV kernels (kernels.hip)
// and the batched persistent kernel (batch.hip): the per-precision
// arithmetic, the direct batch loops, the LDS-staged group loop and the x
// window (entry layout, fill, one- and two-batch-deep group loops).
//
// Summation order per lane is the reference's (warpkernel.hpp:68-71): the
// rounded first product, then one fma per further slot, trailing pads
// included.  Starting from -0 and adding every slot with fma gives exactly
// that (see PF64::init).  Compiled with -ffp-contract=off: every fusion is an
// explicit fma.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wk {
namespace {

template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// ---------------------------------------------------------------- precision
struct PF64 {
  using V = double;
  using X = double;
  struct Acc {
    double s;
  };
  // fma(v, x, -0) is the rounded product v*x bit for bit (signed zeros
  // included), so starting at -0 and adding every slot with fma reproduces the
  // reference's "first product, then fma" order
  __device__ static __forceinline__ void init(Acc &a) { a.s = -0.0; }
  __device__ static __forceinline__ void first(Acc &a, V v, X x) { a.s = v * x; }
  __device__ static __forceinline__ void step(Acc &a, V v, X x) { a.s = __builtin_fma(v, x, a.s); }
  __device__ static __forceinline__ void add(Acc &a, const Acc &b) { a.s = a.s + b.s; }
  __device__ static __forceinline__ Acc shfl_down(const Acc &a, int o, int width) {
    return Acc{__shfl_down(a.s, o, width)};
  }
  __device__ static __forceinline__ double result(const Acc &a) { return a.s; }
};

// FP32 values and x; error-free transformations (TwoProd by fma, Knuth's
// TwoSum) carry the rounding errors in c (Ogita, Rump, Oishi 2005, Dot2).
struct PF32C {
  using V = float;
  using X = float;
  struct Acc {
    float s, c;
  };
  __device__ static __forceinline__ void init(Acc &a) {
    a.s = -0.0f;
    a.c = 0.0f;
  }
  __device__ static __forceinline__ void first(Acc &a, V v, X x) {
    a.s = v * x;
    a.c = __builtin_fmaf(v, x, -a.s);
  }
  __device__ static __forceinline__ void two_sum(float a, float b, float &s, float &t) {
    s = a + b;
    float bp = s - a;
    float ap = s - bp;
    t = (a - ap) + (b - bp);
  }
  __device__ static __forceinline__ void step(Acc &a, V v, X x) {
    float p = v * x;
    float e = __builtin_fmaf(v, x, -p);
    float s, t;
    two_sum(a.s, p, s, t);
    a.s = s;
    a.c = a.c + (t + e);
  }
  __device__ static __forceinline__ void add(Acc &a, const Acc &b) {
    float s, t;
    two_sum(a.s, b.s, s, t);
    a.s = s;
    a.c = (a.c + b.c) + t;
  }
  __device__ static __forceinline__ Acc shfl_down(const Acc &a, int o, int width) {
    return Acc{__shfl_down(a.s, o, width), __shfl_down(a.c, o, width)};
  }
  __device__ static __forceinline__ double result(const Acc &a) {
    return (double)a.s + (double)a.c;
  }
};

// FP32 values and x; the product of two floats is exact in FP64, so an FP64
// fma chain gives the FP64 CSR sum of the quantised inputs.
struct PF32D {
  using V = float;
  using X = float;
  struct Acc {
    double s;
  };
  __device__ static __forceinline__ void init(Acc &a) { a.s = -0.0; }
  __device__ static __forceinline__ void first(Acc &a, V v, X x) { a.s = (double)v * (double)x; }
  __device__ static __forceinline__ void step(Acc &a, V v, X x) {
    a.s = __builtin_fma((double)v, (double)x, a.s);
  }
  __device__ static __forceinline__ void add(Acc &a, const Acc &b) { a.s = a.s + b.s; }
  __device__ static __forceinline__ Acc shfl_down(const Acc &a, int o, int width) {
    return Acc{__shfl_down(a.s, o, width)};
  }
  __device__ static __forceinline__ double result(const Acc &a) { return a.s; }
};

// Steps [i, end) of R groups in lockstep, U slots per group per batch.  All
// loads of a batch are issued before the first gather waits, so a wave has
// R*U value+index loads in flight per batch.  NT 1: values stream with the
// nontemporal hint (4-byte index loads measured faster without it on banded
// matrices); NT 2: indices too, which leaves more of L2 to scattered x
// gathers; NT 0: plain loads; NT 3: as 2 without the values (each slot
// counts 1.0), which times the index stream and the x gathers alone.
template <class P, int W, int KT, int U, int NT, int R>
__device__ __forceinline__ uint32_t stream_steps(typename P::Acc (&acc)[R][KT],
                                                 const typename P::V *const (&vp)[R],
                                                 const int32_t *const (&cp)[R],
                                                 const typename P::X *xb, int64_t ldx,
                                                 uint32_t i, uint32_t end) {
  using V = typename P::V;
  using X = typename P::X;
  for (; i + U <= end; i += U) {
    V v[R][U];
    int32_t c[R][U];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < U; u++) {
        c[r][u] = ld_stream<(NT >= 2)>(cp[r] + (int64_t)(i + u) * W);
        v[r][u] = NT == 3 ? V(1) : ld_stream<(NT >= 1)>(vp[r] + (int64_t)(i + u) * W);
      }
    X xv[R][U][KT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < U; u++) {
        const X *xr = xb + (int64_t)c[r][u] * ldx;
#pragma unroll
        for (int j = 0; j < KT; j++) xv[r][u][j] = xr[j];
      }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int j = 0; j < KT; j++) P::step(acc[r][j], v[r][u], xv[r][u][j]);
  }
  return i;
}

// The last, partial batch of one group: steps [i, end) with end - i < U, all
// loads issued before the first wait (the guards are wave-uniform for W = 64).
template <class P, int W, int KT, int U, int NT>
__device__ __forceinline__ void stream_tail(typename P::Acc (&acc)[KT], const typename P::V *vp,
                                            const int32_t *cp, const typename P::X *xb,
                                            int64_t ldx, uint32_t i, uint32_t end) {
  using V = typename P::V;
  using X = typename P::X;
  if (i >= end) return;
  int32_t c[U];
  V v[U];
#pragma unroll
  for (int u = 0; u < U - 1; u++)
    if (i + u < end) {
      c[u] = ld_stream<(NT >= 2)>(cp + (int64_t)(i + u) * W);
      v[u] = NT == 3 ? V(1) : ld_stream<(NT >= 1)>(vp + (int64_t)(i + u) * W);
    }
  X xv[U][KT];
#pragma unroll
  for (int u = 0; u < U - 1; u++)
    if (i + u < end) {
      const X *xr = xb + (int64_t)c[u] * ldx;
#pragma unroll
      for (int j = 0; j < KT; j++) xv[u][j] = xr[j];
    }
#pragma unroll
  for (int u = 0; u < U - 1; u++)
    if (i + u < end) {
#pragma unroll
      for (int j = 0; j < KT; j++) P::step(acc[j], v[u], xv[u][j]);
    }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <class P, int U>
constexpr int staged_bytes() {
  return U * 64 * (int)(sizeof(typename P::V) + sizeof(int32_t));
}

// All steps of one group through the wave's LDS slice (U steps per batch,
// the last partial batch by direct loads); PIPE issues the next batch's loads
// before the current batch's FMAs.  Returns the lane's accumulator.
template <class P, int U, bool PIPE>
__device__ __forceinline__ typename P::Acc staged_group(const typename P::V *vbase,
                                                        const int32_t *cbase,
                                                        const typename P::X *__restrict__ xb,
                                                        int64_t ldx, uint32_t len, char *slice,
                                                        int32_t lane) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
  constexpr int W = 64;
  constexpr int NV = U * 64 * (int)sizeof(V) / 1024;  // 16-byte loads per lane, values
  constexpr int NC = U * 64 * 4 / 1024;               // 16-byte loads per lane, indices
  static_assert(NV >= 1 && NC >= 1 && U % 4 == 0, "a batch covers whole 1 KB chunks");
  V *lv = reinterpret_cast<V *>(slice);                                 // [U][64]
  int32_t *lc = reinterpret_cast<int32_t *>(slice + U * 64 * sizeof(V));  // [U][64]
  Acc acc;
  P::init(acc);  // every slot by fma onto -0 (see PF64::init)
  uint32_t i = 0;
  if constexpr (!PIPE) {
    for (; i + U <= len; i += U) {
      const u32x4 *gv = reinterpret_cast<const u32x4 *>(vbase + (int64_t)i * W);
      const u32x4 *gc = reinterpret_cast<const u32x4 *>(cbase + (int64_t)i * W);
      u32x4 tc[NC], tv[NV];
#pragma unroll
      for (int q = 0; q < NC; q++) tc[q] = gc[q * 64 + lane];
#pragma unroll
      for (int q = 0; q < NV; q++) tv[q] = __builtin_nontemporal_load(gv + q * 64 + lane);
#pragma unroll
      for (int q = 0; q < NC; q++) reinterpret_cast<u32x4 *>(lc)[q * 64 + lane] = tc[q];
#pragma unroll
      for (int q = 0; q < NV; q++) reinterpret_cast<u32x4 *>(lv)[q * 64 + lane] = tv[q];
      X xv[U];
#pragma unroll
      for (int u = 0; u < U; u++) xv[u] = xb[(int64_t)lc[u * 64 + lane] * ldx];
#pragma unroll
      for (int u = 0; u < U; u++) P::step(acc, lv[u * 64 + lane], xv[u]);
    }
  } else if (i + U <= len) {
    // Depth-2 pipeline: once a batch sits in LDS its registers are free, so
    // the next batch's 16-byte loads are issued right after this batch's
    // gathers and stay in flight through the FMAs.  The loop body is
    // straight-line and both edges into its header carry the same loads
    // into the same registers.
    u32x4 tc[NC], tv[NV];
    auto issue = [&](uint32_t at) {
      const u32x4 *gv = reinterpret_cast<const u32x4 *>(vbase + (int64_t)at * W);
      const u32x4 *gc = reinterpret_cast<const u32x4 *>(cbase + (int64_t)at * W);
#pragma unroll
      for (int q = 0; q < NC; q++) tc[q] = gc[q * 64 + lane];
#pragma unroll
      for (int q = 0; q < NV; q++) tv[q] = __builtin_nontemporal_load(gv + q * 64 + lane);
    };
    auto stage = [&]() {
#pragma unroll
      for (int q = 0; q < NC; q++) reinterpret_cast<u32x4 *>(lc)[q * 64 + lane] = tc[q];
#pragma unroll
      for (int q = 0; q < NV; q++) reinterpret_cast<u32x4 *>(lv)[q * 64 + lane] = tv[q];
    };
    issue(i);
    for (; i + 2 * U <= len; i += U) {
      stage();
      X xv[U];
#pragma unroll
      for (int u = 0; u < U; u++) xv[u] = xb[(int64_t)lc[u * 64 + lane] * ldx];
      issue(i + U);
      V v[U];
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = lv[u * 64 + lane];
#pragma unroll
      for (int u = 0; u < U; u++) P::step(acc, v[u], xv[u]);
    }
    stage();
    X xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = xb[(int64_t)lc[u * 64 + lane] * ldx];
#pragma unroll
    for (int u = 0; u < U; u++) P::step(acc, lv[u * 64 + lane], xv[u]);
    i += U;
  }
  {
    Acc ar[1] = {acc};
    stream_tail<P, W, 1, U, true>(ar, vbase + lane, cbase + lane, xb, ldx, i, len);
    acc = ar[0];
  }
  return acc;
}

// Bytes of one wave's LDS for the window kernel: the x window (win_max
// entries, rounded up to 16 bytes), then a U-step slice of values and 16-bit
// window positions.
template <class P, int U>
__host__ __device__ constexpr int win_slice_bytes() {
  return U * 64 * (int)(sizeof(typename P::V) + sizeof(uint16_t));
}

// One U-step batch of a group's values and 16-bit window positions: issue()
// loads it with 16-byte loads into registers, stage() copies it into the
// wave's LDS slice, from which each lane reads its own slots.
template <class P, int U>
struct WinBatch {
  static constexpr int NV = U * 64 * (int)sizeof(typename P::V) / 1024;
  static constexpr int NI = U * 64 * 2 / 1024;
  static_assert(NV >= 1 && NI >= 1, "a batch covers whole 1 KB chunks");
  u32x4 tv[NV], ti[NI];
  __device__ __forceinline__ void issue(const typename P::V *vbase, const uint16_t *ibase,
                                        uint32_t at, int32_t lane) {
    const u32x4 *gv = reinterpret_cast<const u32x4 *>(vbase + (int64_t)at * 64);
    const u32x4 *gi = reinterpret_cast<const u32x4 *>(ibase + (int64_t)at * 64);
#pragma unroll
    for (int q = 0; q < NI; q++) ti[q] = gi[q * 64 + lane];
#pragma unroll
    for (int q = 0; q < NV; q++) tv[q] = __builtin_nontemporal_load(gv + q * 64 + lane);
  }
  __device__ __forceinline__ void stage(char *slice, int32_t lane) const {
    u32x4 *lv = reinterpret_cast<u32x4 *>(slice);
    u32x4 *li = reinterpret_cast<u32x4 *>(slice + U * 64 * sizeof(typename P::V));
#pragma unroll
    for (int q = 0; q < NI; q++) li[q * 64 + lane] = ti[q];
#pragma unroll
    for (int q = 0; q < NV; q++) lv[q * 64 + lane] = tv[q];
  }
};

// One window entry: the KT values of x at one column, kBytes = KT*sizeof(X)
// bytes at j*kBytes.  LDS has 32 four-byte banks; with 32- or 64-byte entries
// a 16-byte read at the same chunk of every lane's entry lands on only 16 or
// 8 of them (entry j starts at bank 8j or 16j mod 32), so such entries store
// their 16-byte chunks in an order that depends on j: chunk q sits at
// position q ^ swz(j), and the reads of 64 lanes at scattered entries spread
// over all 32 banks.  Entries of 4 to 16 bytes are stored plainly.
template <class X, int KT>
struct WinEntry {
  static constexpr uint32_t kBytes = KT * (uint32_t)sizeof(X);
  static constexpr uint32_t kChunks = kBytes >= 32 ? kBytes / 16 : 1;
  static constexpr int kPer = 16 / (int)sizeof(X);  // values per chunk
  // entry j's first bank is (kBytes/4 * j) mod 32, which repeats every
  // 8/kChunks entries: the next bits of j pick the chunk order
  static constexpr int kShift = kChunks == 2 ? 2 : kChunks == 4 ? 1 : 0;
  static_assert(kChunks == 1 || kChunks == 2 || kChunks == 4 || kChunks == 8, "entry size");
  __device__ static __forceinline__ uint32_t swz(uint32_t j) {
    return kChunks > 1 ? (j >> kShift) & (kChunks - 1) : 0u;
  }
  __device__ static __forceinline__ void read(const char *xs, uint32_t j, X (&v)[KT]) {
    if constexpr (kChunks == 1) {
      const X *e = reinterpret_cast<const X *>(xs + j * kBytes);
#pragma unroll
      for (int t = 0; t < KT; t++) v[t] = e[t];
    } else {
      const uint32_t base = j * kBytes, s = swz(j);
#pragma unroll
      for (uint32_t q = 0; q < kChunks; q++) {
        const u32x4 c = *reinterpret_cast<const u32x4 *>(xs + base + 16u * (q ^ s));
        __builtin_memcpy(&v[q * kPer], &c, 16);
      }
    }
  }
  // values [t0, t0 + KW) of entry j, t0 a multiple of KW (KW = KT: the whole
  // entry): whole 16-byte chunks, or a 4- or 8-byte part of one chunk
  template <int KW>
  __device__ static __forceinline__ void read_part(const char *xs, uint32_t j, uint32_t t0,
                                                   X (&v)[KW]) {
    if constexpr (KW == KT) {
      read(xs, j, v);
    } else if constexpr (KW * sizeof(X) < 16) {
      constexpr uint32_t pb = KW * (uint32_t)sizeof(X);
      static_assert(pb == 4 || pb == 8, "4- or 8-byte parts");
      const uint32_t b0 = t0 * (uint32_t)sizeof(X);
      const char *p = xs + j * kBytes + 16u * ((b0 >> 4) ^ swz(j)) + (b0 & 15u);
      if constexpr (pb == 8) {
        const uint2 c = *reinterpret_cast<const uint2 *>(p);
        __builtin_memcpy(&v[0], &c, 8);
      } else {
        v[0] = *reinterpret_cast<const X *>(p);
      }
    } else {
      static_assert(KW * sizeof(X) % 16 == 0 && kChunks > 1, "whole chunks of a chunked entry");
      constexpr uint32_t nq = KW * sizeof(X) / 16;
      const uint32_t base = j * kBytes, s = swz(j), q0 = t0 * (uint32_t)sizeof(X) / 16;
#pragma unroll
      for (uint32_t q = 0; q < nq; q++) {
        const u32x4 c = *reinterpret_cast<const u32x4 *>(xs + base + 16u * ((q0 + q) ^ s));
        __builtin_memcpy(&v[q * kPer], &c, 16);
      }
    }
  }
  __device__ static __forceinline__ void write(char *xs, uint32_t j, const X (&v)[KT]) {
    if constexpr (kChunks == 1) {
      X *e = reinterpret_cast<X *>(xs + j * kBytes);
#pragma unroll
      for (int t = 0; t < KT; t++) e[t] = v[t];
    } else {
      const uint32_t base = j * kBytes, s = swz(j);
#pragma unroll
      for (uint32_t q = 0; q < kChunks; q++) {
        u32x4 c;
        __builtin_memcpy(&c, &v[q * kPer], 16);
        *reinterpret_cast<u32x4 *>(xs + base + 16u * (q ^ s)) = c;
      }
    }
  }
};

// The window fill in two halves, so that the column loads go out before a
// group's first value batches and the x gathers right after them: cols()
// loads the first C*S window columns (C per lane, entries first + q*S),
// gather() reads x at them, writes those entries, and repeats for any further
// columns.  S is 64 times the number of waves that fill one window together.
// C keeps one round's gathered values within about 96 registers, and one
// round covers 768 entries (hex elasticity windows hold 646 on average).
// REGS bounds the registers one round of gathered values takes (kernels that
// hold other data in registers pass less).
template <class X, int KT, int S = 64, int REGS = 96>
struct WinFill {
  static constexpr int kRegs = KT * (int)sizeof(X) / 4;  // registers per entry
  static constexpr int kLow = REGS >= 4 * kRegs ? 4 : (REGS / kRegs < 1 ? 1 : REGS / kRegs);
  static constexpr int kByRegs = REGS / kRegs > 12 ? 12 : (REGS / kRegs < kLow ? kLow : REGS / kRegs);
  static constexpr int kByCover = 768 / S > 1 ? 768 / S : 1;
  static constexpr int C = kByRegs < kByCover ? kByRegs : kByCover;
  int32_t c[C];
  __device__ __forceinline__ void cols(const int32_t *__restrict__ wcols, uint32_t n,
                                       int32_t first) {
#pragma unroll
    for (int q = 0; q < C; q++) {
      const uint32_t j = (uint32_t)first + q * S;
      c[q] = j < n ? wcols[j] : 0;
    }
  }
  // MASKED: only kv values of an entry exist in x (a masked last k-tile
  // holds fewer than KT columns); the others are written as 0
  template <bool MASKED = false>
  __device__ __forceinline__ void gather(char *xs, const int32_t *__restrict__ wcols, uint32_t n,
                                         const X *__restrict__ xb, int64_t ldx, int32_t first,
                                         int32_t kv = KT) {
    for (uint32_t j0 = 0;;) {
      X v[C][KT];
#pragma unroll
      for (int q = 0; q < C; q++)
#pragma unroll
        for (int t = 0; t < KT; t++)
          v[q][t] = !MASKED || t < kv ? xb[(int64_t)c[q] * ldx + t] : X(0);
#pragma unroll
      for (int q = 0; q < C; q++) {
        const uint32_t j = j0 + (uint32_t)first + q * S;
        if (j < n) WinEntry<X, KT>::write(xs, j, v[q]);
      }
      j0 += C * S;
      if (j0 >= n) break;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const uint32_t j = j0 + (uint32_t)first + q * S;
        c[q] = j < n ? wcols[j] : 0;
      }
    }
  }
};

// Fills the window in one go: entry j < n holds x[wcols[j] * ldx + t], t < KT.
template <class X, int KT>
__device__ __forceinline__ void fill_window(char *xs, const int32_t *__restrict__ wcols, uint32_t n,
                                            const X *__restrict__ xb, int64_t ldx, int32_t lane) {
  WinFill<X, KT> f;
  f.cols(wcols, n, lane);
  f.gather(xs, wcols, n, xb, ldx, lane);
}

// One group through its x window.  xs holds x at the group's distinct columns
// (filled by fill_window); each slot carries the 16-bit position of its column
// there.  Values and positions stream in U-step batches through the slice,
// the next batch in flight during the current FMAs, and x comes from LDS: the
// same products in the same order as staged_group, without a column index or
// an x gather per slot.  When `issued`, the caller has issued the first batch
// into bt before filling the window.
template <class P, int U, int KT>
__device__ __forceinline__ void window_group(typename P::Acc (&acc)[KT], const typename P::V *vbase,
                                             const uint16_t *ibase, const char *xs, uint32_t len,
                                             WinBatch<P, U> &bt, bool issued, char *slice,
                                             int32_t lane) {
  using V = typename P::V;
  using X = typename P::X;
  using E = WinEntry<X, KT>;
  const V *lv = reinterpret_cast<const V *>(slice);
  const uint16_t *li = reinterpret_cast<const uint16_t *>(slice + U * 64 * sizeof(V));
#pragma unroll
  for (int t = 0; t < KT; t++) P::init(acc[t]);  // every slot by fma onto -0
  uint32_t i = 0;
  if (i + U <= len) {
    if (!issued) bt.issue(vbase, ibase, 0, lane);
    for (; i + 2 * U <= len; i += U) {
      bt.stage(slice, lane);
      uint32_t p[U];
#pragma unroll
      for (int u = 0; u < U; u++) p[u] = li[u * 64 + lane];
      bt.issue(vbase, ibase, i + U, lane);
      V v[U];
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = lv[u * 64 + lane];
#pragma unroll
      for (int u = 0; u < U; u++) {
        X xv[KT];
        E::read(xs, p[u], xv);
#pragma unroll
        for (int t = 0; t < KT; t++) P::step(acc[t], v[u], xv[t]);
      }
    }
    bt.stage(slice, lane);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const V v = lv[u * 64 + lane];
      X xv[KT];
      E::read(xs, li[u * 64 + lane], xv);
#pragma unroll
      for (int t = 0; t < KT; t++) P::step(acc[t], v, xv[t]);
    }
    i += U;
  }
  // the last, partial batch: direct loads, all issued before the first use
  if (i < len) {
    V v[U];
    uint16_t c[U];
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        c[u] = ibase[(int64_t)(i + u) * 64 + lane];
        v[u] = __builtin_nontemporal_load(vbase + (int64_t)(i + u) * 64 + lane);
      }
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        X xv[KT];
        E::read(xs, c[u], xv);
#pragma unroll
        for (int t = 0; t < KT; t++) P::step(acc[t], v[u], xv[t]);
      }
  }
}

// Window positions packed 10 bits each (three per 32-bit word): a 12-step
// block of one lane is one 16-byte word group, loaded straight into registers
// (1.33 bytes per slot instead of 2), while the batch's values are staged
// through the slice as in WinBatch.  A batch is NB blocks (U = 12*NB steps).
// The position loads go out first, so the wait for the values covers them.
template <class P, int NB>
struct WinBatchPk {
  static constexpr int U = 12 * NB;
  static constexpr int NV = U * 64 * (int)sizeof(typename P::V) / 1024;
  static_assert(NV * 1024 == U * 64 * (int)sizeof(typename P::V), "whole 1 KB chunks");
  u32x4 tv[NV], tp[NB];
  __device__ __forceinline__ void issue(const typename P::V *vbase, const u32x4 *pk, uint32_t b,
                                        int32_t lane) {
#pragma unroll
    for (int q = 0; q < NB; q++) tp[q] = pk[((int64_t)b * NB + q) * 64 + lane];
    const u32x4 *gv = reinterpret_cast<const u32x4 *>(vbase + (int64_t)b * U * 64);
#pragma unroll
    for (int q = 0; q < NV; q++) tv[q] = __builtin_nontemporal_load(gv + q * 64 + lane);
  }
  __device__ __forceinline__ void stage(char *slice, int32_t lane) const {
    u32x4 *lv = reinterpret_cast<u32x4 *>(slice);
#pragma unroll
    for (int q = 0; q < NV; q++) lv[q * 64 + lane] = tv[q];
  }
  // position of step u of a block (u is a constant after unrolling)
  __device__ static __forceinline__ uint32_t pos(const u32x4 &w, int u) {
    return (w[u / 3] >> (10 * (u % 3))) & 1023u;
  }
};

// The U steps of one staged batch with packed positions: value from the
// slice, x from the window.
template <class P, int NB>
__device__ __forceinline__ void pk_batch(typename P::Acc &acc, const typename P::V *lv,
                                         const char *xs, const u32x4 (&w)[NB], int32_t lane) {
  constexpr int U = 12 * NB;
  typename P::X xv[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    typename P::X e[1];
    WinEntry<typename P::X, 1>::read(xs, WinBatchPk<P, NB>::pos(w[u / 12], u % 12), e);
    xv[u] = e[0];
  }
#pragma unroll
  for (int u = 0; u < U; u++) P::step(acc, lv[u * 64 + lane], xv[u]);
}

// One group through its x window with packed positions (one right-hand side):
// the same products in the same order as window_group.  When `issued`, the
// caller has issued batch 0 into bt before filling the window.
template <class P, int NB>
__device__ __forceinline__ void window_group_pk(typename P::Acc &acc, const typename P::V *vbase,
                                                const u32x4 *pk, const char *xs, uint32_t len,
                                                WinBatchPk<P, NB> &bt, bool issued, char *slice,
                                                int32_t lane) {
  using V = typename P::V;
  constexpr int U = WinBatchPk<P, NB>::U;
  const V *lv = reinterpret_cast<const V *>(slice);
  P::init(acc);  // every slot by fma onto -0
  uint32_t i = 0, b = 0;
  if (i + U <= len) {
    if (!issued) bt.issue(vbase, pk, 0, lane);
    for (; i + 2 * U <= len; i += U, b++) {
      bt.stage(slice, lane);
      u32x4 w[NB];
#pragma unroll
      for (int q = 0; q < NB; q++) w[q] = bt.tp[q];
      bt.issue(vbase, pk, b + 1, lane);
      pk_batch<P, NB>(acc, lv, xs, w, lane);
    }
    bt.stage(slice, lane);
    u32x4 w[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) w[q] = bt.tp[q];
    pk_batch<P, NB>(acc, lv, xs, w, lane);
    i += U;
    b++;
  }
  // the last, partial batch: its packed blocks (those the group has) and
  // direct value loads
  if (i < len) {
    u32x4 w[NB];
#pragma unroll
    for (int q = 0; q < NB; q++)
      if (i + 12 * q < len) w[q] = pk[((int64_t)b * NB + q) * 64 + lane];
    V v[U];
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) v[u] = __builtin_nontemporal_load(vbase + (int64_t)(i + u) * 64 + lane);
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        typename P::X xv[1];
        WinEntry<typename P::X, 1>::read(xs, WinBatchPk<P, NB>::pos(w[u / 12], u % 12), xv);
        P::step(acc, v[u], xv[0]);
      }
  }
}

// ------------------------------------------------------- node windows
// One right-hand side through a window of node entries (three x values
// each).  A lane's steps go three at a time: steps 3m, 3m+1, 3m+2 read values
// 0, 1, 2 of one entry, whose 10-bit index is field m % 3 of word m / 3 of
// the lane's positions.  A batch of NW words covers 9 * NW steps: at 27
// steps, three words per lane against nine of 16-bit positions or 2.25
// 16-byte blocks of 10-bit ones.

// Batch of 9 * NW steps: the values as 16-byte pieces of the group's
// column-major slots (the last 1 KB chunk may be partial) and NW words.
template <class P, int NW>
struct WinBatchNd {
  static constexpr int U = 9 * NW;
  static constexpr int kPieces = U * 64 * (int)sizeof(typename P::V) / 16;  // per batch
  static constexpr int NV = (kPieces + 63) / 64;
  static_assert(kPieces * 16 == U * 64 * (int)sizeof(typename P::V), "whole 16-byte pieces");
  u32x4 tv[NV];
  uint32_t tp[NW];
  // pw: the group's words plus the lane (word q of the lane at pw[q * 64])
  __device__ __forceinline__ void issue(const typename P::V *vbase, const uint32_t *pw, uint32_t b,
                                        int32_t lane) {
#pragma unroll
    for (int q = 0; q < NW; q++) tp[q] = pw[((int64_t)b * NW + q) * 64];
    const u32x4 *gv = reinterpret_cast<const u32x4 *>(vbase + (int64_t)b * U * 64);
#pragma unroll
    for (int q = 0; q < NV; q++)
      if (q < NV - 1 || q * 64 + lane < kPieces) tv[q] = __builtin_nontemporal_load(gv + q * 64 + lane);
  }
  __device__ __forceinline__ void stage(char *slice, int32_t lane) const {
    u32x4 *lv = reinterpret_cast<u32x4 *>(slice);
#pragma unroll
    for (int q = 0; q < NV; q++)
      if (q < NV - 1 || q * 64 + lane < kPieces) lv[q * 64 + lane] = tv[q];
  }
  // byte offset in the window of the entry that steps 3m .. 3m+2 read
  __device__ static __forceinline__ uint32_t entry(const uint32_t (&p)[NW], int m) {
    return ((p[m / 3] >> (10 * (m % 3))) & 1023u) * 3u * (uint32_t)sizeof(typename P::X);
  }
};

// The steps of one staged batch: value from the slice, x from the window.
template <class P, int NW>
__device__ __forceinline__ void nd_batch(typename P::Acc &acc, const typename P::V *lv,
                                         const char *xs, const uint32_t (&p)[NW], int32_t lane) {
  using X = typename P::X;
  constexpr int U = 9 * NW;
  X xv[U];
#pragma unroll
  for (int m = 0; m < U / 3; m++) {
    const X *e = reinterpret_cast<const X *>(xs + WinBatchNd<P, NW>::entry(p, m));
#pragma unroll
    for (int d = 0; d < 3; d++) xv[3 * m + d] = e[d];
  }
#pragma unroll
  for (int u = 0; u < U; u++) P::step(acc, lv[u * 64 + lane], xv[u]);
}

// Fills a node window: entry j < n (the group's entries) gets its three x
// values.  TRIPLE false: nc[j] = e >= 0 for columns e, e+1, e+2, e < 0 for
// column -e-1 three times; TRIPLE true: the three columns explicitly, at
// nc[3j .. 3j+2] (Kr numbering).  Four entries per lane in flight, twelve
// loads.
template <class X, bool TRIPLE>
struct NdFill {
  static constexpr int C = TRIPLE ? 3 : 1;
  int32_t e[4][C];
  __device__ __forceinline__ void load(const int32_t *__restrict__ nc, uint32_t n, uint32_t j0,
                                       int32_t lane) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t j = j0 + (uint32_t)lane + q * 64;
#pragma unroll
      for (int d = 0; d < C; d++) e[q][d] = j < n ? nc[(int64_t)C * j + d] : (TRIPLE ? 0 : -1);
    }
  }
  __device__ __forceinline__ void cols(const int32_t *__restrict__ nc, uint32_t n, int32_t lane) {
    load(nc, n, 0, lane);
  }
  __device__ static __forceinline__ int64_t column(const int32_t (&c)[C], int d) {
    if constexpr (TRIPLE) return c[d];
    return c[0] >= 0 ? (int64_t)c[0] + d : -(int64_t)c[0] - 1;
  }
  // The first round of gather in two halves, so that other loads can go out
  // between them: issue_x issues the x loads of the lane's first four entries
  // (cols loaded), write_x stores them into the window.  A window of more than
  // 256 entries continues with rest.
  __device__ __forceinline__ void issue_x(X (&v)[4][3], const X *__restrict__ xb, int64_t ldx) const {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int d = 0; d < 3; d++) v[q][d] = xb[column(e[q], d) * ldx];
  }
  __device__ __forceinline__ void write_x(char *xs, const X (&v)[4][3], uint32_t n, int32_t lane) const {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t j = (uint32_t)lane + q * 64;
      if (j < n) {
        X *dst = reinterpret_cast<X *>(xs + j * 3u * (uint32_t)sizeof(X));
#pragma unroll
        for (int d = 0; d < 3; d++) dst[d] = v[q][d];
      }
    }
  }
  __device__ __forceinline__ void rest(char *xs, const int32_t *__restrict__ nc, uint32_t n,
                                       const X *__restrict__ xb, int64_t ldx, int32_t lane) {
    if (n <= 256) return;
    load(nc, n, 256, lane);
    for (uint32_t j0 = 256;;) {
      X v[4][3];
      issue_x(v, xb, ldx);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t j = j0 + (uint32_t)lane + q * 64;
        if (j < n) {
          X *dst = reinterpret_cast<X *>(xs + j * 3u * (uint32_t)sizeof(X));
#pragma unroll
          for (int d = 0; d < 3; d++) dst[d] = v[q][d];
        }
      }
      j0 += 256;
      if (j0 >= n) break;
      load(nc, n, j0, lane);
    }
  }
  __device__ __forceinline__ void gather(char *xs, const int32_t *__restrict__ nc, uint32_t n,
                                         const X *__restrict__ xb, int64_t ldx, int32_t lane) {
    for (uint32_t j0 = 0;;) {
      X v[4][3];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int d = 0; d < 3; d++) v[q][d] = xb[column(e[q], d) * ldx];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t j = j0 + (uint32_t)lane + q * 64;
        if (j < n) {
          X *dst = reinterpret_cast<X *>(xs + j * 3u * (uint32_t)sizeof(X));
#pragma unroll
          for (int d = 0; d < 3; d++) dst[d] = v[q][d];
        }
      }
      j0 += 256;
      if (j0 >= n) break;
      load(nc, n, j0, lane);
    }
  }
};

// One group through its node window: the same products in the same order as
// window_group_pk (every slot by fma onto -0, pads included).  len is a
// multiple of three.  When `issued`, the caller has issued batch 0 into bt
// before filling the window.
template <class P, int NW>
__device__ __forceinline__ void window_group_nd(typename P::Acc &acc, const typename P::V *vbase,
                                                const uint32_t *pw, const char *xs, uint32_t len,
                                                WinBatchNd<P, NW> &bt, bool issued, char *slice,
                                                int32_t lane) {
  using V = typename P::V;
  using X = typename P::X;
  constexpr int U = WinBatchNd<P, NW>::U;
  const V *lv = reinterpret_cast<const V *>(slice);
  P::init(acc);
  uint32_t i = 0, b = 0;
  if (i + U <= len) {
    if (!issued) bt.issue(vbase, pw, 0, lane);
    for (; i + 2 * U <= len; i += U, b++) {
      bt.stage(slice, lane);
      uint32_t p[NW];
#pragma unroll
      for (int q = 0; q < NW; q++) p[q] = bt.tp[q];
      bt.issue(vbase, pw, b + 1, lane);
      nd_batch<P, NW>(acc, lv, xs, p, lane);
    }
    bt.stage(slice, lane);
    uint32_t p[NW];
#pragma unroll
    for (int q = 0; q < NW; q++) p[q] = bt.tp[q];
    nd_batch<P, NW>(acc, lv, xs, p, lane);
    i += U;
    b++;
  }
  // the last, partial batch (fewer than U steps, a multiple of three): the
  // words it has and direct value loads
  if (i < len) {
    uint32_t p[NW];
#pragma unroll
    for (int q = 0; q < NW; q++) p[q] = i + 9 * q < len ? pw[((int64_t)b * NW + q) * 64] : 0u;
    V v[U - 3];
#pragma unroll
    for (int u = 0; u < U - 3; u++)
      if (i + u < len) v[u] = __builtin_nontemporal_load(vbase + (int64_t)(i + u) * 64 + lane);
#pragma unroll
    for (int m = 0; m < U / 3 - 1; m++)
      if (i + 3 * m < len) {
        const X *e = reinterpret_cast<const X *>(xs + WinBatchNd<P, NW>::entry(p, m));
#pragma unroll
        for (int d = 0; d < 3; d++) P::step(acc, v[3 * m + d], e[d]);
      }
  }
}

// The x reads and FMAs of one staged batch: slot u of the lane reads its
// value and window position from the slice, then its KW x values (values
// t0 .. t0+KW-1 of the entry).
template <class P, int U, int KT, int KW>
__device__ __forceinline__ void window_batch(typename P::Acc (&acc)[KW], const char *xs,
                                             const char *slice, uint32_t t0, int32_t lane) {
  using V = typename P::V;
  using X = typename P::X;
  const V *lv = reinterpret_cast<const V *>(slice);
  const uint16_t *li = reinterpret_cast<const uint16_t *>(slice + U * 64 * sizeof(V));
  uint32_t p[U];
  V v[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    p[u] = li[u * 64 + lane];
    v[u] = lv[u * 64 + lane];
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    X xv[KW];
    WinEntry<X, KT>::template read_part<KW>(xs, p[u], t0, xv);
#pragma unroll
    for (int t = 0; t < KW; t++) P::step(acc[t], v[u], xv[t]);
  }
}

// window_group with D batches in flight instead of one.  With KT values per
// slot the window takes KT times the LDS, so fewer waves share a CU, and
// each wave must keep more of the value stream in flight by itself: batch b
// is staged and consumed while batches b+1 .. b+D-1 are loading, and its
// registers take batch b+D right after staging.  Before the window fill the
// caller issues batches 0 .. D-1 (prologue), so they load during the fill.
// Same products, same order.
template <class P, int U, int D>
struct WinPipe {
  WinBatch<P, U> bt[D];
  __device__ __forceinline__ void prologue(const typename P::V *vbase, const uint16_t *ibase,
                                           uint32_t len, int32_t lane) {
    const uint32_t nb = len / U;
#pragma unroll
    for (int d = 0; d < D; d++)
      if ((uint32_t)d < nb) bt[d].issue(vbase, ibase, (uint32_t)d * U, lane);
  }
};

// window_group for wide k-tiles without the slice: each lane loads its own
// values and window positions (8 + 2 bytes per slot, narrow loads) straight
// into registers, two batches alternating so that one loads while the other
// is summed.  With KT values per slot the x reads dominate the LDS traffic,
// and staging the values through LDS would add to it.  Same products, same
// order.
template <class P, int U, int KT, int KW>
__device__ __forceinline__ void window_group_regs(typename P::Acc (&acc)[KW],
                                                  const typename P::V *vbase,
                                                  const uint16_t *ibase, const char *xs,
                                                  uint32_t len, uint32_t t0, int32_t lane) {
  using V = typename P::V;
  using X = typename P::X;
#pragma unroll
  for (int t = 0; t < KW; t++) P::init(acc[t]);  // every slot by fma onto -0
  const V *vp = vbase + lane;
  const uint16_t *ip = ibase + lane;
  auto load = [&](uint32_t at, V (&v)[U], uint16_t (&c)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      c[u] = ip[(int64_t)(at + u) * 64];
      v[u] = __builtin_nontemporal_load(vp + (int64_t)(at + u) * 64);
    }
  };
  auto sum = [&](const V (&v)[U], const uint16_t (&c)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      X xv[KW];
      WinEntry<X, KT>::template read_part<KW>(xs, c[u], t0, xv);
#pragma unroll
      for (int t = 0; t < KW; t++) P::step(acc[t], v[u], xv[t]);
    }
  };
  const uint32_t nb = len / U;
  uint32_t b = 0;
  V va[U], vb[U];
  uint16_t ca[U], cb[U];
  if (nb > 0) load(0, va, ca);
  // two batches per pass: B loads while A is summed, then A while B is
  for (; b + 3 <= nb; b += 2) {
    load((b + 1) * U, vb, cb);
    sum(va, ca);
    load((b + 2) * U, va, ca);
    sum(vb, cb);
  }
  // one or two full batches left (A holds batch b)
  if (b + 2 == nb) {
    load((b + 1) * U, vb, cb);
    sum(va, ca);
    sum(vb, cb);
  } else if (b + 1 == nb) {
    sum(va, ca);
  }
  // the last, partial batch
  const uint32_t i = nb * U;
  if (i < len) {
    V v[U];
    uint16_t c[U];
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        c[u] = ip[(int64_t)(i + u) * 64];
        v[u] = __builtin_nontemporal_load(vp + (int64_t)(i + u) * 64);
      }
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        X xv[KW];
        WinEntry<X, KT>::template read_part<KW>(xs, c[u], t0, xv);
#pragma unroll
        for (int t = 0; t < KW; t++) P::step(acc[t], v[u], xv[t]);
      }
  }
}

template <class P, int U, int KT, int KW, int D>
__device__ __forceinline__ void window_group_deep(typename P::Acc (&acc)[KW],
                                                  const typename P::V *vbase,
                                                  const uint16_t *ibase, const char *xs,
                                                  uint32_t len, WinPipe<P, U, D> &pp, char *slice,
                                                  uint32_t t0, int32_t lane) {
  using V = typename P::V;
  using X = typename P::X;
#pragma unroll
  for (int t = 0; t < KW; t++) P::init(acc[t]);  // every slot by fma onto -0
  const uint32_t nb = len / U;
  uint32_t b = 0;
  // steady state: the batch D ahead of each consumed one exists
  for (; b + 2 * D <= nb; b += D) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      pp.bt[d].stage(slice, lane);
      pp.bt[d].issue(vbase, ibase, (b + D + d) * U, lane);
      window_batch<P, U, KT, KW>(acc, xs, slice, t0, lane);
    }
  }
  // drain: fewer than 2D batches left
#pragma unroll
  for (int d = 0; d < D; d++)
    if (b + d < nb) {
      pp.bt[d].stage(slice, lane);
      if (b + D + d < nb) pp.bt[d].issue(vbase, ibase, (b + D + d) * U, lane);
      window_batch<P, U, KT, KW>(acc, xs, slice, t0, lane);
    }
#pragma unroll
  for (int d = 0; d < D; d++)
    if (b + D + d < nb) {
      pp.bt[d].stage(slice, lane);
      window_batch<P, U, KT, KW>(acc, xs, slice, t0, lane);
    }
  // the last, partial batch: direct loads, all issued before the first use
  const uint32_t i = nb * U;
  if (i < len) {
    V v[U];
    uint16_t c[U];
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        c[u] = ibase[(int64_t)(i + u) * 64 + lane];
        v[u] = __builtin_nontemporal_load(vbase + (int64_t)(i + u) * 64 + lane);
      }
#pragma unroll
    for (int u = 0; u < U - 1; u++)
      if (i + u < len) {
        X xv[KW];
        WinEntry<X, KT>::template read_part<KW>(xs, c[u], t0, xv);
#pragma unroll
        for (int t = 0; t < KW; t++) P::step(acc[t], v[u], xv[t]);
      }
  }
}

}  // namespace
}  // namespace wk
