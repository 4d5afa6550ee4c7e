/*
 * wk_oracle.h -- CPU oracle for the WarpKernel (WPK1/WPK2) SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * host scan / process steps (warpkernel.hpp:215-771) and of the summation order
 * of its kernels (warpkernel.hpp:53-211), plus an FP64 CSR SpMV in original row
 * order.  Only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it; the product library never links or calls it.
 *
 * Every function names the reference lines it follows.  Generalisations over the
 * reference, all reducing to it at W = 32 and max_tpr = 32:
 *   - the group width W (the reference's WARP_SIZE) is a parameter;
 *   - WPK2's literal thmax_per_row = 32 (warpkernel.hpp:643) is max_tpr;
 *   - warp_offsets are int64 (the reference uses int).
 * Parity status of each function is recorded beside it ("pinned by ...").
 */
#ifndef WK_ORACLE_H
#define WK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wko_layout {
  int32_t variant;        /* 1 = structure, 2 = structure2 */
  int32_t W;              /* group width (WARP_SIZE) */
  int32_t threshold;      /* nmax_per_thread (variant 2) */
  int32_t max_tpr;        /* thmax_per_row (variant 2) */
  int32_t nrows;
  int32_t nznrows;
  int32_t nwarps;         /* number of groups */
  int64_t nz;
  int64_t allocate_nz;
  uint32_t min_nz, max_nz;          /* variant 2 only */
  int32_t *row_map;                 /* [nrows]    y[row_map[i]] = y'[i] */
  int32_t *row_map_inv;             /* [nrows] */
  uint32_t *reorder_rows;           /* [nz]       new_values[reorder_rows[i]] = A[i] */
  int64_t *warp_offsets;            /* [nwarps+1] */
  uint32_t *max_nz_rows;            /* [nwarps] */
  uint32_t *threads_per_row;        /* [nwarps]   variant 2, else NULL */
  uint32_t *row_offset_warp;        /* [nwarps+1] variant 2, else NULL */
} wko_layout;

/* structure::scan (warpkernel.hpp:309-365).  Returns NULL on allocation failure. */
wko_layout *wko_scan1(int32_t W, int32_t nrows, int64_t nz, const int32_t *rowptr);
/* structure2::scan (warpkernel.hpp:578-736). */
wko_layout *wko_scan2(int32_t W, int32_t threshold, int32_t max_tpr, int32_t nrows,
                      int64_t nz, const int32_t *rowptr);
void wko_free(wko_layout *L);

/* structure::reorder / process (warpkernel.hpp:379-398, 532-548): zero-filled
 * padded arrays of allocate_nz entries; new[reorder_rows[i]] = A[i]. */
void wko_process(const wko_layout *L, const double *vals, const int32_t *colind,
                 double *pvals, int32_t *pcols);
/* reorder_columns_coalesced (warpkernel.hpp:446-457). */
void wko_reorder_columns_coalesced(const wko_layout *L, const int32_t *colind, int32_t *pcols_kr);
/* reorder_x (warpkernel.hpp:517-530): xout[row_map_inv[r]] = x[r]. */
void wko_reorder_x(const wko_layout *L, const double *x, double *xout);

/* reorder_columns_rowsort (warpkernel.hpp:414-444): within each row, the
 * entries ordered by permuted column row_map_inv[col] ascending;
 * cols_out[n] = that permuted column, reorder_A_rows[n] = the CSR index the
 * entry came from.  The reference's std::sort leaves the order of equal
 * columns (duplicates within a row) unspecified; here they keep CSR order.
 * Pinned by hand-worked cases (tests/golden/kat_wpk.json) and by the
 * permutation / sortedness properties. */
void wko_reorder_columns_rowsort(const wko_layout *L, const int32_t *rowptr,
                                 const int32_t *colind, int32_t *cols_out,
                                 int32_t *reorder_A_rows);
/* reorder_values_coalesced_rowsort (warpkernel.hpp:505-515): zero-filled
 * padded values, pvals[reorder_rows[i]] = vals[reorder_A_rows[i]]. */
void wko_reorder_values_rowsort(const wko_layout *L, const double *vals,
                                const int32_t *reorder_A_rows, double *pvals);

/* FP64 CSR SpMV in original row order, k right-hand sides (x is [ncols][ldx],
 * y is [nrows][ldy], row-major).  fused = 1: first product rounded, then fma
 * (the order of warpKernel, warpkernel.hpp:68-71, with contraction as nvcc does
 * by default); fused = 0: separate multiply and add.  Empty rows get 0. */
void wko_csr_spmv(int32_t nrows, const int32_t *rowptr, const int32_t *colind,
                  const double *vals, const double *x, double *y, int32_t k,
                  int64_t ldx, int64_t ldy, int32_t fused);
/* Same summation as wko_csr_spmv(fused = 1), rows split over nthreads. */
void wko_csr_spmv_omp(int32_t nrows, const int32_t *rowptr, const int32_t *colind,
                      const double *vals, const double *x, double *y, int32_t nthreads);

/* Emulates the WPK kernels lane by lane on the padded arrays, pads included:
 * warpKernel / warpKernel_reorderx (warpkernel.hpp:53-75, 103-125) for variant 1,
 * warpKernel2 / warpKernel2_reorderx (warpkernel.hpp:127-211) for variant 2.
 * kr = 0 stores y[row_map[row]] (run), kr = 1 stores y[row] (run_x).
 * Empty rows are written as 0. */
void wko_wpk_spmv(const wko_layout *L, const double *pvals, const int32_t *pcols,
                  const double *x, double *y, int32_t k, int64_t ldx, int64_t ldy,
                  int32_t kr);

/* CSR -> CSR of the transpose, entries of each column in increasing row order.
 * perm[t] = CSR index of transposed entry t. */
void wko_transpose(int32_t nrows, int32_t ncols, const int32_t *rowptr, const int32_t *colind,
                   int32_t *colptr, int32_t *rowind, int64_t *perm);

/* dvals[j] = dy[row_j] * x[col_j] (the pattern gradient of y = A(vals) x). */
void wko_dvals(int32_t nrows, const int32_t *rowptr, const int32_t *colind,
               const double *dy, const double *x, double *dvals, int32_t k,
               int64_t ldx, int64_t lddy);

#ifdef __cplusplus
}
#endif
#endif
