/*
 * wk_oracle.c -- CPU oracle for the WarpKernel (WPK1/WPK2) SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm,
 * used by the tests, the smoke check and the CPU-baseline leg of the benchmark
 * as the checker.  The product library never links or calls this file.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so that a*b+c is never fused
 * behind our back; fusion is spelled out with fma()).
 *
 * Parity pinning (what checks each function against something other than
 * itself; the tests named live under tests/):
 *   wko_scan1           hand-worked known-answer vector (golden/kat_wpk.json),
 *                       closed-form allocate_nz on the FE configurations,
 *                       brute-force decode invariants on random inputs.
 *   wko_scan2           three hand-worked vectors, one per branch of the split
 *                       (golden/kat_wpk.json), plus decode invariants.
 *   wko_process / _reorder_*   scatter through the pinned maps; pads checked.
 *   wko_csr_spmv        textbook CSR product; checked against a dense
 *                       matrix-vector product and against closed forms
 *                       (A*1 = 0 on Neumann Laplacians, rigid modes).
 *   wko_wpk_spmv        bit-exact against wko_csr_spmv(fused=1) for variant 1
 *                       (pads add +-0 only); for variant 2 against an
 *                       independent chunk+pairwise-tree sum.
 *   wko_transpose       against a dense transpose.
 *   wko_dvals           against central finite differences of y = A(v) x.
 */
#include "wk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* (row, length) pair; checkPairOrder (warpkernel.hpp:222-225) orders by length
 * descending, then row ascending.  This is a strict total order on pairs with
 * distinct rows, so any correct sort gives the same permutation. */
typedef struct {
  uint32_t row, len;
} wko_pair;

static int pair_cmp(const void *a, const void *b) {
  const wko_pair *p = (const wko_pair *)a, *q = (const wko_pair *)b;
  if (p->len != q->len) return p->len > q->len ? -1 : 1;
  if (p->row != q->row) return p->row < q->row ? -1 : 1;
  return 0;
}

/* structure::sort (warpkernel.hpp:232-250): pairs of (row, rowsize), sorted;
 * counts the non-empty rows. */
static wko_pair *sort_rows(int32_t nrows, const int32_t *rowptr, int32_t *nznrows) {
  wko_pair *p = (wko_pair *)malloc(sizeof(wko_pair) * (size_t)(nrows > 0 ? nrows : 1));
  int32_t nz = 0;
  for (int32_t r = 0; r < nrows; r++) {
    p[r].row = (uint32_t)r;
    p[r].len = (uint32_t)(rowptr[r + 1] - rowptr[r]);
    if (p[r].len > 0) nz++;
  }
  qsort(p, (size_t)nrows, sizeof(wko_pair), pair_cmp);
  *nznrows = nz;
  return p;
}

/* structure::storemap (warpkernel.hpp:288-296). */
static void store_map(wko_layout *L, const wko_pair *p) {
  L->row_map = (int32_t *)malloc(sizeof(int32_t) * (size_t)(L->nrows > 0 ? L->nrows : 1));
  L->row_map_inv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(L->nrows > 0 ? L->nrows : 1));
  for (int32_t r = 0; r < L->nrows; r++) L->row_map[r] = (int32_t)p[r].row;
  for (int32_t i = 0; i < L->nrows; i++) L->row_map_inv[L->row_map[i]] = i;
}

wko_layout *wko_scan1(int32_t W, int32_t nrows, int64_t nz, const int32_t *rowptr) {
  wko_layout *L = (wko_layout *)calloc(1, sizeof(wko_layout));
  L->variant = 1;
  L->W = W;
  L->nrows = nrows;
  L->nz = nz;
  wko_pair *p = sort_rows(nrows, rowptr, &L->nznrows);
  store_map(L, p);
  /* warpkernel.hpp:324-340: groups of W sorted non-empty rows; A_w = longest. */
  L->nwarps = (L->nznrows + W - 1) / W;
  int32_t G = L->nwarps;
  L->max_nz_rows = (uint32_t *)calloc((size_t)(G > 0 ? G : 1), sizeof(uint32_t));
  for (int32_t w = 0; w < G; w++) {
    uint32_t mx = 0;
    for (int32_t r = W * w; r < L->nznrows && r < W * (w + 1); r++)
      if (p[r].len > mx) mx = p[r].len;
    L->max_nz_rows[w] = mx;
  }
  /* warpkernel.hpp:343-349 */
  L->warp_offsets = (int64_t *)malloc(sizeof(int64_t) * (size_t)(G + 1));
  L->warp_offsets[0] = 0;
  for (int32_t w = 0; w < G; w++)
    L->warp_offsets[w + 1] = L->warp_offsets[w] + (int64_t)L->max_nz_rows[w] * W;
  L->allocate_nz = L->warp_offsets[G];
  /* warpkernel.hpp:352-364: lane (r_s % W) of group w takes the row; its i-th
   * nonzero lands W slots after its (i-1)-th. */
  L->reorder_rows = (uint32_t *)calloc((size_t)(nz > 0 ? nz : 1), sizeof(uint32_t));
  for (int32_t w = 0; w < G; w++) {
    for (int32_t rs = W * w; rs < L->nznrows && rs < W * (w + 1); rs++) {
      int32_t r = (int32_t)p[rs].row;
      int32_t len = (int32_t)p[rs].len;
      for (int32_t i = 0; i < len; i++)
        L->reorder_rows[rowptr[r] + i] =
            (uint32_t)(L->warp_offsets[w] + (rs % W) + (int64_t)i * W);
    }
  }
  free(p);
  return L;
}

/* Growable uint32 array with insert-before, standing in for std::vector<uint>
 * and its insert() in warpkernel.hpp:670-672. */
typedef struct {
  uint32_t *v;
  int64_t n, cap;
} uvec;

static void uvec_push(uvec *a, uint32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? a->cap * 2 : 16;
    a->v = (uint32_t *)realloc(a->v, sizeof(uint32_t) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}

static void uvec_insert(uvec *a, int64_t at, uint32_t x) {
  uvec_push(a, 0);
  memmove(a->v + at + 1, a->v + at, sizeof(uint32_t) * (size_t)(a->n - 1 - at));
  a->v[at] = x;
}

wko_layout *wko_scan2(int32_t W, int32_t threshold, int32_t max_tpr, int32_t nrows,
                      int64_t nz, const int32_t *rowptr) {
  wko_layout *L = (wko_layout *)calloc(1, sizeof(wko_layout));
  L->variant = 2;
  L->W = W;
  L->threshold = threshold;
  L->max_tpr = max_tpr;
  L->nrows = nrows;
  L->nz = nz;
  wko_pair *p = sort_rows(nrows, rowptr, &L->nznrows);
  store_map(L, p);
  int32_t G0 = (L->nznrows + W - 1) / W;

  /* warpkernel.hpp:599-628: per group the longest row (A_w), the nominal rows
   * per group (W) and the actual row count (rows_w); global min/max length. */
  uvec A_w = {0}, rows_per_warp = {0}, rows_w = {0};
  L->min_nz = (uint32_t)nrows;
  L->max_nz = 0;
  for (int32_t w = 0; w < G0; w++) {
    uint32_t mx = 0, cnt = 0;
    for (int32_t r = W * w; r < L->nznrows && r < W * (w + 1); r++) {
      uint32_t len = p[r].len;
      if (len > mx) mx = len;
      cnt++;
      if (len > 0) {
        if (len < L->min_nz) L->min_nz = len;
        if (len > L->max_nz) L->max_nz = len;
      }
    }
    uvec_push(&A_w, mx);
    uvec_push(&rows_per_warp, (uint32_t)W);
    uvec_push(&rows_w, cnt);
  }

  /* warpkernel.hpp:633-691, with the three iterators as indices.  A group whose
   * longest row exceeds the threshold gets i lanes per row (smallest power of
   * two with A_w <= cap*i), each lane a contiguous chunk of ceil(A_w/i); a full
   * group becomes i sub-groups of W/i rows, the tail group (rows_w*i < W) only
   * changes its nominal rows per group.  The tail branch advances past the end,
   * which ends the loop as in the reference. */
  {
    int64_t wi = 0, ri = 0, ci = 0;
    const uint32_t T = (uint32_t)threshold;
    for (; wi < A_w.n; wi++) {
      if (A_w.v[wi] > T) {
        uint32_t cap = T;
        if (A_w.v[wi] > T * (uint32_t)max_tpr)
          cap = (A_w.v[wi] + (uint32_t)max_tpr - 1) / (uint32_t)max_tpr;
        uint32_t temp = A_w.v[wi];
        for (int32_t i = 2; i <= max_tpr; i *= 2) {
          if (temp <= cap * (uint32_t)i) {
            uint32_t nzpt = (temp + (uint32_t)i - 1) / (uint32_t)i;
            A_w.v[wi] = nzpt;
            if (rows_w.v[ci] * (uint32_t)i < (uint32_t)W) {
              rows_per_warp.v[ri] = (uint32_t)(W / i);
            } else {
              rows_per_warp.v[ri] = (uint32_t)(W / i);
              rows_w.v[ci] = (uint32_t)(W / i);
              for (int32_t j = 0; j < i - 1; j++) {
                uvec_insert(&A_w, wi, nzpt);
                uvec_insert(&rows_per_warp, ri, (uint32_t)(W / i));
                uvec_insert(&rows_w, ci, (uint32_t)(W / i));
              }
            }
            wi += i - 1;
            ri += i - 1;
            ci += i - 1;
            break;
          }
        }
      }
      ci += 1;
      ri += 1;
    }
  }
  int32_t G = (int32_t)A_w.n;
  L->nwarps = G;
  L->max_nz_rows = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(G > 0 ? G : 1));
  if (G) memcpy(L->max_nz_rows, A_w.v, sizeof(uint32_t) * (size_t)G);

  /* warpkernel.hpp:696-702 */
  L->warp_offsets = (int64_t *)malloc(sizeof(int64_t) * (size_t)(G + 1));
  L->warp_offsets[0] = 0;
  for (int32_t w = 0; w < G; w++)
    L->warp_offsets[w + 1] = L->warp_offsets[w] + (int64_t)A_w.v[w] * W;
  L->allocate_nz = L->warp_offsets[G];

  /* warpkernel.hpp:704-707 */
  L->threads_per_row = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(G > 0 ? G : 1));
  for (int32_t w = 0; w < G; w++) L->threads_per_row[w] = (uint32_t)W / rows_per_warp.v[w];

  /* warpkernel.hpp:709-717: cumulative actual rows; the sentinel is nrows. */
  L->row_offset_warp = (uint32_t *)calloc((size_t)(G + 1), sizeof(uint32_t));
  for (int32_t w = 0; w < G; w++) L->row_offset_warp[w + 1] = L->row_offset_warp[w] + rows_w.v[w];
  L->row_offset_warp[G] = (uint32_t)nrows;

  /* warpkernel.hpp:719-735: lane (tid + r*tpr) of group w holds chunk tid of
   * sorted row row_offset_warp[w] + r, in CSR order. */
  L->reorder_rows = (uint32_t *)calloc((size_t)(nz > 0 ? nz : 1), sizeof(uint32_t));
  for (int32_t w = 0; w < G; w++) {
    uint32_t tpr = L->threads_per_row[w];
    uint32_t aw = A_w.v[w];
    for (uint32_t r = 0; r < rows_per_warp.v[w] && L->row_offset_warp[w] + r < (uint32_t)L->nznrows;
         r++) {
      uint32_t rs = L->row_offset_warp[w] + r;
      int32_t row = (int32_t)p[rs].row;
      uint32_t len = p[rs].len;
      for (uint32_t tid = 0; tid < tpr; tid++)
        for (uint32_t i = 0; i + tid * aw < len && i < aw; i++)
          L->reorder_rows[rowptr[row] + i + tid * aw] =
              (uint32_t)(L->warp_offsets[w] + (tid + r * tpr) + (int64_t)i * W);
    }
  }
  free(A_w.v);
  free(rows_per_warp.v);
  free(rows_w.v);
  free(p);
  return L;
}

void wko_free(wko_layout *L) {
  if (!L) return;
  free(L->row_map);
  free(L->row_map_inv);
  free(L->reorder_rows);
  free(L->warp_offsets);
  free(L->max_nz_rows);
  free(L->threads_per_row);
  free(L->row_offset_warp);
  free(L);
}

void wko_process(const wko_layout *L, const double *vals, const int32_t *colind,
                 double *pvals, int32_t *pcols) {
  memset(pvals, 0, sizeof(double) * (size_t)L->allocate_nz);
  memset(pcols, 0, sizeof(int32_t) * (size_t)L->allocate_nz);
  for (int64_t i = 0; i < L->nz; i++) {
    pvals[L->reorder_rows[i]] = vals[i];
    pcols[L->reorder_rows[i]] = colind[i];
  }
}

void wko_reorder_columns_coalesced(const wko_layout *L, const int32_t *colind, int32_t *pcols_kr) {
  memset(pcols_kr, 0, sizeof(int32_t) * (size_t)L->allocate_nz);
  for (int64_t i = 0; i < L->nz; i++) pcols_kr[L->reorder_rows[i]] = L->row_map_inv[colind[i]];
}

void wko_reorder_x(const wko_layout *L, const double *x, double *xout) {
  for (int32_t r = 0; r < L->nrows; r++) xout[L->row_map_inv[r]] = x[r];
}

typedef struct {
  int32_t key; /* permuted column */
  int32_t at;  /* position within the row */
} wko_colkey;

static int colkey_cmp(const void *a, const void *b) {
  const wko_colkey *p = (const wko_colkey *)a, *q = (const wko_colkey *)b;
  if (p->key != q->key) return p->key < q->key ? -1 : 1;
  return p->at < q->at ? -1 : (p->at > q->at);
}

/* warpkernel.hpp:414-444: pairs (n - rows[r], row_map_inv[colinds[n]]) sorted by
 * the second member (checkPairInOrder, hpp:228-230), ties by position. */
void wko_reorder_columns_rowsort(const wko_layout *L, const int32_t *rowptr,
                                 const int32_t *colind, int32_t *cols_out,
                                 int32_t *reorder_A_rows) {
  int32_t maxlen = 0;
  for (int32_t r = 0; r < L->nrows; r++)
    if (rowptr[r + 1] - rowptr[r] > maxlen) maxlen = rowptr[r + 1] - rowptr[r];
  wko_colkey *tmp = (wko_colkey *)malloc(sizeof(wko_colkey) * (size_t)(maxlen ? maxlen : 1));
  if (!tmp) return;
  for (int32_t r = 0; r < L->nrows; r++) {
    const int32_t b = rowptr[r], n = rowptr[r + 1] - rowptr[r];
    for (int32_t i = 0; i < n; i++) {
      tmp[i].key = L->row_map_inv[colind[b + i]];
      tmp[i].at = i;
    }
    qsort(tmp, (size_t)n, sizeof(wko_colkey), colkey_cmp);
    for (int32_t i = 0; i < n; i++) {
      cols_out[b + i] = tmp[i].key;
      reorder_A_rows[b + i] = b + tmp[i].at;
    }
  }
  free(tmp);
}

/* warpkernel.hpp:505-515 */
void wko_reorder_values_rowsort(const wko_layout *L, const double *vals,
                                const int32_t *reorder_A_rows, double *pvals) {
  memset(pvals, 0, sizeof(double) * (size_t)L->allocate_nz);
  for (int64_t i = 0; i < L->nz; i++) pvals[L->reorder_rows[i]] = vals[reorder_A_rows[i]];
}

static void csr_rows(int32_t r0, int32_t r1, const int32_t *rowptr, const int32_t *colind,
                     const double *vals, const double *x, double *y, int32_t k, int64_t ldx,
                     int64_t ldy, int32_t fused) {
  for (int32_t r = r0; r < r1; r++) {
    int32_t b = rowptr[r], e = rowptr[r + 1];
    for (int32_t j = 0; j < k; j++) {
      double s = 0.0;
      if (e > b) {
        s = vals[b] * x[(int64_t)colind[b] * ldx + j];
        for (int32_t t = b + 1; t < e; t++) {
          double xv = x[(int64_t)colind[t] * ldx + j];
          if (fused)
            s = fma(vals[t], xv, s);
          else {
            double prod = vals[t] * xv;
            s = s + prod;
          }
        }
      }
      y[(int64_t)r * ldy + j] = s;
    }
  }
}

void wko_csr_spmv(int32_t nrows, const int32_t *rowptr, const int32_t *colind,
                  const double *vals, const double *x, double *y, int32_t k, int64_t ldx,
                  int64_t ldy, int32_t fused) {
  csr_rows(0, nrows, rowptr, colind, vals, x, y, k, ldx, ldy, fused);
}

void wko_csr_spmv_omp(int32_t nrows, const int32_t *rowptr, const int32_t *colind,
                      const double *vals, const double *x, double *y, int32_t nthreads) {
#ifdef _OPENMP
  int32_t nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    int32_t t = omp_get_thread_num(), n = omp_get_num_threads();
    int32_t chunk = (nrows + n - 1) / n;
    int32_t r0 = t * chunk, r1 = r0 + chunk < nrows ? r0 + chunk : nrows;
    if (r0 < r1) csr_rows(r0, r1, rowptr, colind, vals, x, y, 1, 1, 1, 1);
  }
#else
  (void)nthreads;
  csr_rows(0, nrows, rowptr, colind, vals, x, y, 1, 1, 1, 1);
#endif
}

void wko_wpk_spmv(const wko_layout *L, const double *pvals, const int32_t *pcols,
                  const double *x, double *y, int32_t k, int64_t ldx, int64_t ldy, int32_t kr) {
  const int32_t W = L->W;
  double *lane = (double *)malloc(sizeof(double) * (size_t)W);
  for (int32_t pos = L->nznrows; pos < L->nrows; pos++) {
    int64_t dst = kr ? pos : L->row_map[pos];
    for (int32_t j = 0; j < k; j++) y[dst * ldy + j] = 0.0;
  }
  for (int32_t g = 0; g < L->nwarps; g++) {
    uint32_t tpr = L->variant == 2 ? L->threads_per_row[g] : 1u;
    uint32_t row0 = L->variant == 2 ? L->row_offset_warp[g] : (uint32_t)g * (uint32_t)W;
    uint32_t aw = L->max_nz_rows[g];
    for (int32_t j = 0; j < k; j++) {
      /* each lane: first product rounded, then fma along its slots
       * (warpkernel.hpp:68-71, 149-153) */
      for (int32_t l = 0; l < W; l++) {
        int64_t s0 = L->warp_offsets[g] + l;
        double s = pvals[s0] * x[(int64_t)pcols[s0] * ldx + j];
        for (uint32_t i = 1; i < aw; i++) {
          int64_t si = s0 + (int64_t)i * W;
          s = fma(pvals[si], x[(int64_t)pcols[si] * ldx + j], s);
        }
        lane[l] = s;
      }
      /* lockstep tree (warpkernel.hpp:157-162): every lane reads its partner's
       * value from the previous step, then all write. */
      for (uint32_t o = 1; o < tpr; o <<= 1) {
        for (int32_t l = 0; l + (int32_t)o < W; l++)
          if ((l & (int32_t)(2 * o - 1)) == 0) lane[l] = lane[l] + lane[l + o];
      }
      for (int32_t l = 0; l < W; l += (int32_t)tpr) {
        uint32_t rowid = row0 + (uint32_t)l / tpr;
        if (rowid < (uint32_t)L->nznrows) {
          int64_t dst = kr ? (int64_t)rowid : (int64_t)L->row_map[rowid];
          y[dst * ldy + j] = lane[l];
        }
      }
    }
  }
  free(lane);
}

void wko_transpose(int32_t nrows, int32_t ncols, const int32_t *rowptr, const int32_t *colind,
                   int32_t *colptr, int32_t *rowind, int64_t *perm) {
  memset(colptr, 0, sizeof(int32_t) * (size_t)(ncols + 1));
  for (int32_t r = 0; r < nrows; r++)
    for (int32_t t = rowptr[r]; t < rowptr[r + 1]; t++) colptr[colind[t] + 1]++;
  for (int32_t c = 0; c < ncols; c++) colptr[c + 1] += colptr[c];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ncols > 0 ? ncols : 1));
  for (int32_t c = 0; c < ncols; c++) fill[c] = colptr[c];
  for (int32_t r = 0; r < nrows; r++)
    for (int32_t t = rowptr[r]; t < rowptr[r + 1]; t++) {
      int32_t d = fill[colind[t]]++;
      rowind[d] = r;
      perm[d] = t;
    }
  free(fill);
}

void wko_dvals(int32_t nrows, const int32_t *rowptr, const int32_t *colind, const double *dy,
               const double *x, double *dvals, int32_t k, int64_t ldx, int64_t lddy) {
  for (int32_t r = 0; r < nrows; r++)
    for (int32_t t = rowptr[r]; t < rowptr[r + 1]; t++) {
      const double *xr = x + (int64_t)colind[t] * ldx;
      const double *dr = dy + (int64_t)r * lddy;
      double s = dr[0] * xr[0];
      for (int32_t j = 1; j < k; j++) s = fma(dr[j], xr[j], s);
      dvals[t] = s;
    }
}
