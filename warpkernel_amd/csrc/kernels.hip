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
#include <type_traits>

#include "wk_internal.hpp"

namespace wk {
namespace {

constexpr int kMaxThreads = 512;

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
// grid.x = nwork + fill blocks.  Work block -> (group block, k-tile), k-tile
// fastest, so blocks running together stream the same groups.  Fill blocks
// write the rows without nonzeros (sorted positions nznrows..nrows-1).
template <class P, int W, bool SPLIT, int OUT, int KT, int U, bool NT>
__global__ __launch_bounds__(kMaxThreads) void spmv_kernel(SpmvArgs a, int32_t nwork,
                                                           int32_t ktiles, int32_t remap) {
  using V = typename P::V;
  using X = typename P::X;
  using Acc = typename P::Acc;
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

  int32_t g, l;
  if constexpr (W == 64) {
    g = wave;
    l = lane;
    if (g >= a.ngroups) return;
  } else {
    g = 2 * wave + (lane >> 5);
    l = lane & 31;
    if (2 * wave >= a.ngroups) return;
  }
  const bool active = g < a.ngroups;
  const int64_t base = active ? a.warp_offsets[g] : 0;
  const uint32_t len = active ? a.max_nz_rows[g] : 0u;
  const uint32_t tpr = SPLIT && active ? a.tpr[g] : 1u;
  const uint32_t row0 = SPLIT ? (active ? a.row_off[g] : 0u) : (uint32_t)g * (uint32_t)W;

  const V *__restrict__ vp = static_cast<const V *>(a.vals) + base + l;
  const int32_t *__restrict__ cp = a.cols + base + l;
  const X *__restrict__ xb = static_cast<const X *>(a.x) + j0;
  const int64_t ldx = a.ldx;

  Acc acc[KT];
  if (len > 0) {
    const V v = ld_stream<NT>(vp);
    const int32_t c = ld_stream<NT>(cp);
    const X *xr = xb + (int64_t)c * ldx;
#pragma unroll
    for (int j = 0; j < KT; j++) P::first(acc[j], v, xr[j]);
  } else {
#pragma unroll
    for (int j = 0; j < KT; j++) P::first(acc[j], V(0), X(0));
  }
  uint32_t i = 1;
  for (; i + U <= len; i += U) {
    V v[U];
    int32_t c[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      v[u] = ld_stream<NT>(vp + (int64_t)(i + u) * W);
      c[u] = ld_stream<NT>(cp + (int64_t)(i + u) * W);
    }
    X xv[U][KT];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const X *xr = xb + (int64_t)c[u] * ldx;
#pragma unroll
      for (int j = 0; j < KT; j++) xv[u][j] = xr[j];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int j = 0; j < KT; j++) P::step(acc[j], v[u], xv[u][j]);
  }
  for (; i < len; i++) {
    const V v = ld_stream<NT>(vp + (int64_t)i * W);
    const int32_t c = ld_stream<NT>(cp + (int64_t)i * W);
    const X *xr = xb + (int64_t)c * ldx;
#pragma unroll
    for (int j = 0; j < KT; j++) P::step(acc[j], v, xr[j]);
  }

  if constexpr (SPLIT) {
    // lockstep pairwise tree of hpp:157-162: lane t adds lane t+o's partial
    for (uint32_t o = 1; o < tpr; o <<= 1) {
#pragma unroll
      for (int j = 0; j < KT; j++) {
        Acc other = P::shfl_down(acc[j], (int)o, W);
        P::add(acc[j], other);
      }
    }
  }
  if (!active) return;
  const uint32_t rowid = row0 + (uint32_t)l / tpr;
  if (((uint32_t)l & (tpr - 1)) != 0 || rowid >= (uint32_t)a.nznrows) return;
  const int64_t dst = OUT == OUT_RUNX ? (int64_t)rowid : (int64_t)a.row_map[rowid];
  double *yr = a.y + dst * a.ldy + j0;
#pragma unroll
  for (int j = 0; j < KT; j++) {
    const double r = P::result(acc[j]);
    if constexpr (OUT == OUT_RESID)
      yr[j] = a.b[dst * a.ldb + j0 + j] - r;
    else
      yr[j] = r;
  }
}

// ------------------------------------------------------------ small kernels
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

// dvals[t] = sum_j dy[row_t][j] * x[col_t][j], CSR order.  Each block takes a
// tile of entries, finds the rows the tile spans, keeps that rowptr slice in
// LDS and binary-searches it per entry.
constexpr int kDvTile = 2048;
constexpr int kDvRows = 2048;

__device__ __forceinline__ int32_t upper_row(const int32_t *rp, int32_t lo, int32_t hi, int64_t e) {
  // largest r in [lo, hi] with rp[r] <= e
  while (lo < hi) {
    int32_t mid = lo + ((hi - lo + 1) >> 1);
    if ((int64_t)rp[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void dvals_kernel(const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colind,
                                                    int32_t nrows, int64_t nnz,
                                                    const double *__restrict__ dy, int64_t lddy,
                                                    const double *__restrict__ x, int64_t ldx,
                                                    int32_t k, double *__restrict__ dv) {
  __shared__ int32_t srp[kDvRows + 1];
  __shared__ int32_t sr[2];
  const int64_t e0 = (int64_t)blockIdx.x * kDvTile;
  const int64_t e1 = e0 + kDvTile < nnz ? e0 + kDvTile : nnz;
  if (threadIdx.x == 0) {
    // rows holding entries e0 and e1-1; skip empty rows by taking the last
    // row whose start is <= e
    sr[0] = upper_row(rowptr, 0, nrows - 1, e0);
    sr[1] = upper_row(rowptr, 0, nrows - 1, e1 - 1);
  }
  __syncthreads();
  const int32_t r0 = sr[0], r1 = sr[1];
  const int32_t span = r1 - r0 + 1;
  const bool in_lds = span + 1 <= kDvRows + 1;
  if (in_lds)
    for (int32_t t = threadIdx.x; t <= span; t += blockDim.x) srp[t] = rowptr[r0 + t];
  __syncthreads();
  for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    int32_t r;
    if (in_lds)
      r = r0 + upper_row(srp, 0, span - 1, e);
    else
      r = upper_row(rowptr, r0, r1, e);
    const double *xr = x + (int64_t)colind[e] * ldx;
    const double *dr = dy + (int64_t)r * lddy;
    double s = dr[0] * xr[0];
    for (int32_t j = 1; j < k; j++) s = __builtin_fma(dr[j], xr[j], s);
    dv[e] = s;
  }
}

// ------------------------------------------------------------- dispatch
using LaunchFn = void (*)(const SpmvArgs &, dim3, dim3, int32_t, int32_t, int32_t, hipStream_t);

template <class P, int W, bool SPLIT, int OUT, int KT, int U, bool NT>
void launch_one(const SpmvArgs &a, dim3 grid, dim3 block, int32_t nwork, int32_t ktiles,
                int32_t remap, hipStream_t s) {
  hipLaunchKernelGGL((spmv_kernel<P, W, SPLIT, OUT, KT, U, NT>), grid, block, 0, s, a, nwork,
                     ktiles, remap);
}

template <class P, int W, bool SPLIT, int OUT>
LaunchFn pick_kt(int kt, int u, bool nt) {
  if constexpr (W == 32) {
    // the reference's own layout width: parity path, one right-hand side per pass
    return kt == 1 ? &launch_one<P, 32, SPLIT, OUT, 1, 8, true> : nullptr;
  } else {
    switch (kt) {
      case 1:
        if (!nt) return &launch_one<P, W, SPLIT, OUT, 1, 8, false>;
        if (u == 4) return &launch_one<P, W, SPLIT, OUT, 1, 4, true>;
        if (u == 16) return &launch_one<P, W, SPLIT, OUT, 1, 16, true>;
        return &launch_one<P, W, SPLIT, OUT, 1, 8, true>;
      case 2: return &launch_one<P, W, SPLIT, OUT, 2, 4, true>;
      case 4: return &launch_one<P, W, SPLIT, OUT, 4, 2, true>;
      case 8: return &launch_one<P, W, SPLIT, OUT, 8, 2, true>;
      case 16: return &launch_one<P, W, SPLIT, OUT, 16, 1, true>;
      case 32: return &launch_one<P, W, SPLIT, OUT, 32, 1, true>;
    }
    return nullptr;
  }
}

template <class P, int W, bool SPLIT>
LaunchFn pick_out(int out, int kt, int u, bool nt) {
  if (out == OUT_RUN) return pick_kt<P, W, SPLIT, OUT_RUN>(kt, u, nt);
  if (out == OUT_RUNX) return pick_kt<P, W, SPLIT, OUT_RUNX>(kt, u, nt);
  if constexpr (std::is_same_v<P, PF64>) return pick_kt<P, W, SPLIT, OUT_RESID>(kt, u, nt);
  return nullptr;  // the residual is FP64 only
}

template <class P, int W>
LaunchFn pick_split(bool split, int out, int kt, int u, bool nt) {
  return split ? pick_out<P, W, true>(out, kt, u, nt) : pick_out<P, W, false>(out, kt, u, nt);
}

LaunchFn pick(int prec, int W, bool split, int out, int kt, int u, bool nt) {
  // prec: 0 FP64, 1 FP32 compensated, 2 FP32 with FP64 accumulation
  if (prec == 0)
    return W == 64 ? pick_split<PF64, 64>(split, out, kt, u, nt)
                   : pick_split<PF64, 32>(split, out, kt, u, nt);
  if (W != 64) return nullptr;
  if (prec == 1) return pick_split<PF32C, 64>(split, out, kt, u, nt);
  return pick_split<PF32D, 64>(split, out, kt, u, nt);
}

int floor_pow2(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}

}  // namespace

void launch_spmv(const DevLayout &D, int32_t dtype, const wk_launch &cfg, int out, const void *x,
                 int64_t ldx, double *y, int64_t ldy, const double *b, int64_t ldb, int32_t k,
                 hipStream_t stream) {
  if (k < 1) fail(WK_ERR_ARG, "k must be >= 1");
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
  const int gpw = D.W == 64 ? 1 : 2;
  const int64_t waves = ((int64_t)D.ngroups + gpw - 1) / gpw;
  const int64_t ngb = (waves + wpb - 1) / wpb;
  const int64_t nfill = ((int64_t)D.nrows - D.nznrows + wpb * 64 - 1) / (wpb * 64);

  // k = sum of power-of-two tiles: as many k_tile tiles as fit, then the rest
  int32_t j0 = 0;
  int kt = std::min(cfg.k_tile, floor_pow2(k));
  if (D.W == 32) kt = 1;
  while (j0 < k) {
    const int32_t rem = k - j0;
    while (kt > rem) kt >>= 1;
    const int32_t ntiles = rem / kt;
    LaunchFn fn = pick(prec, D.W, split, out, kt, cfg.unroll, cfg.nontemporal != 0);
    if (!fn) fail(WK_ERR_ARG, "unsupported kernel configuration");
    const int64_t nwork = ngb * ntiles;
    if (nwork + nfill > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
    a.x = static_cast<const char *>(x) + (size_t)j0 * xsz;
    a.y = y + j0;
    a.b = b ? b + j0 : nullptr;
    a.k = ntiles * kt;
    if (nwork + nfill > 0)
      fn(a, dim3((uint32_t)(nwork + nfill)), dim3(wpb * 64), (int32_t)nwork, ntiles,
         cfg.xcd_remap ? 1 : 0, stream);
    hip_check(hipGetLastError(), "spmv launch");
    j0 += ntiles * kt;
  }
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

void launch_dvals(const int32_t *rowptr, const int32_t *colind, int32_t nrows, int64_t nnz,
                  const double *dy, int64_t lddy, const double *x, int64_t ldx, int32_t k,
                  double *dvals, hipStream_t stream) {
  if (nnz <= 0) return;
  const int64_t blocks = (nnz + kDvTile - 1) / kDvTile;
  if (blocks > INT32_MAX) fail(WK_ERR_LIMIT, "grid too large");
  hipLaunchKernelGGL(dvals_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, rowptr, colind,
                     nrows, nnz, dy, lddy, x, ldx, k, dvals);
  hip_check(hipGetLastError(), "dvals launch");
}

}  // namespace wk
