// fegen.cpp -- finite-element matrices on structured grids, assembled straight
// into CSR (sorted columns, structural zeros kept).  These are the synthetic
// workloads the SpMV is measured on; the WarpKernel reference reads external
// MatrixMarket files instead (wpk1.cu:43-44).
//
// Matrices (nodes numbered x-fastest, dof interleaved: row = ndof*node + d):
//   q1_laplace_2d   bilinear Q1 Laplacian, 9-point, per-element coefficient
//   hex_laplace_3d  trilinear Q1 diffusion, 27-point
//   hex_elastic_3d  trilinear Q1 isotropic linear elasticity, 3 dof, <= 81/row
//   tet_poisson_3d  P1 Laplacian on the Kuhn subdivision (6 tets per cube),
//                   with a per-element keep mask
// All element matrices come from Gauss quadrature (2 points per axis) on the
// reference square / cube, or from barycentric gradients for tets; elements are
// unit-sized, so values are O(coefficient).
//
// Every generator is two-phase: *_rowptr fills rowptr[nrows+1] (returns nnz),
// *_fill writes colind/vals given that rowptr.  Rows are independent, so the
// fill phase runs on several host threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Grid {
  int dim;          // 2 or 3
  int ex, ey, ez;   // elements per axis (ez = 1 for dim 2)
  int nx, ny, nz;   // nodes per axis (nz = 1 for dim 2)
  int64_t nodes() const { return (int64_t)nx * ny * nz; }
  int64_t elem(int a, int b, int c) const { return a + (int64_t)ex * (b + (int64_t)ey * c); }
};

Grid make_grid(int dim, int ex, int ey, int ez) {
  Grid g;
  g.dim = dim;
  g.ex = ex;
  g.ey = ey;
  g.ez = dim == 3 ? ez : 1;
  g.nx = ex + 1;
  g.ny = ey + 1;
  g.nz = dim == 3 ? ez + 1 : 1;
  return g;
}

// Q1 shape functions on [0,1]^dim, local node l = lx + 2 ly (+ 4 lz).
void q1_grad(int dim, int l, const double *p, double *grad) {
  double f[3], df[3];
  for (int a = 0; a < dim; a++) {
    int bit = (l >> a) & 1;
    f[a] = bit ? p[a] : 1.0 - p[a];
    df[a] = bit ? 1.0 : -1.0;
  }
  for (int a = 0; a < dim; a++) {
    double v = df[a];
    for (int b = 0; b < dim; b++)
      if (b != a) v *= f[b];
    grad[a] = v;
  }
}

// Reference element matrix on the unit square/cube, 2-point Gauss per axis.
// ndof = 1: Laplacian; ndof = 3: isotropic elasticity with E = 1.
std::vector<double> q1_element(int dim, int ndof, double nu) {
  const int nloc = 1 << dim, n = nloc * ndof;
  std::vector<double> K((size_t)n * n, 0.0);
  const double gq[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  const int ngp = 1 << dim;
  const double w = 1.0 / ngp;
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  for (int q = 0; q < ngp; q++) {
    double p[3] = {gq[q & 1], gq[(q >> 1) & 1], gq[(q >> 2) & 1]};
    double G[8][3];
    for (int l = 0; l < nloc; l++) q1_grad(dim, l, p, G[l]);
    if (ndof == 1) {
      for (int a = 0; a < nloc; a++)
        for (int b = 0; b < nloc; b++) {
          double s = 0;
          for (int c = 0; c < dim; c++) s += G[a][c] * G[b][c];
          K[(size_t)a * n + b] += w * s;
        }
    } else {
      // K_(a,i),(b,j) = lam * dNa/dxi dNb/dxj + mu * (dNa/dxj dNb/dxi + delta_ij grad Na . grad Nb)
      for (int a = 0; a < nloc; a++)
        for (int b = 0; b < nloc; b++) {
          double gg = 0;
          for (int c = 0; c < 3; c++) gg += G[a][c] * G[b][c];
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
              double v = lam * G[a][i] * G[b][j] + mu * G[a][j] * G[b][i];
              if (i == j) v += mu * gg;
              K[(size_t)(a * 3 + i) * n + (b * 3 + j)] += w * v;
            }
        }
    }
  }
  return K;
}

// Kuhn subdivision: tet t follows the axis order perm[t] from corner 0 to 7.
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

void kuhn_vertices(int t, int v[4]) {
  v[0] = 0;
  v[1] = 1 << kPerm[t][0];
  v[2] = v[1] | (1 << kPerm[t][1]);
  v[3] = 7;
}

// P1 Laplacian of each Kuhn tet of the unit cube: vol * grad(lambda_a).grad(lambda_b).
void tet_elements(double K[6][4][4]) {
  for (int t = 0; t < 6; t++) {
    int v[4];
    kuhn_vertices(t, v);
    double P[4][3];
    for (int a = 0; a < 4; a++)
      for (int c = 0; c < 3; c++) P[a][c] = (v[a] >> c) & 1;
    double M[3][3];  // rows: p_a - p_0
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) M[a][c] = P[a + 1][c] - P[0][c];
    double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                 M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                 M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    // inverse of M; grad lambda_{a+1} = column a of M^{-1}
    double I[3][3];
    I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    double g[4][3];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) g[a + 1][c] = I[c][a];
    for (int c = 0; c < 3; c++) g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
    double vol = std::fabs(det) / 6.0;
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        double s = 0;
        for (int c = 0; c < 3; c++) s += g[a][c] * g[b][c];
        K[t][a][b] = vol * s;
      }
  }
}

// One matrix row: columns ascending with their values.
struct RowOut {
  int32_t cols[81];
  double vals[81];
  int n;
};

// Q1 row (hex or quad): node (i,j,k), dof d.  Neighbours in increasing node
// index, each contributing ndof columns.  Element contributions are summed in
// lexicographic element order.
void q1_row(const Grid &g, int ndof, const std::vector<double> &K, const double *coef,
            int64_t node, int d, RowOut &out) {
  const int nloc = 1 << g.dim, n = nloc * ndof;
  int i = (int)(node % g.nx), j = (int)((node / g.nx) % g.ny), k = (int)(node / ((int64_t)g.nx * g.ny));
  out.n = 0;
  const int kr = g.dim == 3 ? 1 : 0;
  for (int dk = -kr; dk <= kr; dk++)
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        int bi = i + di, bj = j + dj, bk = k + dk;
        if (bi < 0 || bj < 0 || bk < 0 || bi >= g.nx || bj >= g.ny || bk >= g.nz) continue;
        double acc[3] = {0, 0, 0};
        bool shared = false;
        // elements containing both nodes: [max-1, min] per axis, clipped
        const int ek0 = g.dim == 3 ? std::max(std::max(k, bk) - 1, 0) : 0;
        const int ek1 = g.dim == 3 ? std::min(std::min(k, bk), g.ez - 1) : 0;
        for (int ek = ek0; ek <= ek1; ek++) {
          for (int ej = std::max(j, bj) - 1; ej <= std::min(j, bj); ej++) {
            if (ej < 0 || ej >= g.ey) continue;
            for (int ei = std::max(i, bi) - 1; ei <= std::min(i, bi); ei++) {
              if (ei < 0 || ei >= g.ex) continue;
              int la = (i - ei) + 2 * (j - ej) + 4 * (k - ek);
              int lb = (bi - ei) + 2 * (bj - ej) + 4 * (bk - ek);
              double c = coef[g.elem(ei, ej, ek)];
              for (int e = 0; e < ndof; e++) acc[e] += c * K[(size_t)(la * ndof + d) * n + lb * ndof + e];
              shared = true;
            }
          }
        }
        if (!shared) continue;
        int64_t b = bi + (int64_t)g.nx * (bj + (int64_t)g.ny * bk);
        for (int e = 0; e < ndof; e++) {
          out.cols[out.n] = (int32_t)(b * ndof + e);
          out.vals[out.n] = acc[e];
          out.n++;
        }
      }
}

// P1 Kuhn-tet row: node (i,j,k).  keep[e] == 0 drops element e = t + 6*cube.
void tet_row(const Grid &g, const double K[6][4][4], const double *coef, const uint8_t *keep,
             int64_t node, RowOut &out) {
  int i = (int)(node % g.nx), j = (int)((node / g.nx) % g.ny), k = (int)(node / ((int64_t)g.nx * g.ny));
  double acc[27];
  bool present[27];
  for (int o = 0; o < 27; o++) {
    acc[o] = 0;
    present[o] = false;
  }
  for (int ck = k - 1; ck <= k; ck++) {
    if (ck < 0 || ck >= g.ez) continue;
    for (int cj = j - 1; cj <= j; cj++) {
      if (cj < 0 || cj >= g.ey) continue;
      for (int ci = i - 1; ci <= i; ci++) {
        if (ci < 0 || ci >= g.ex) continue;
        int corner = (i - ci) | ((j - cj) << 1) | ((k - ck) << 2);
        int64_t cube = g.elem(ci, cj, ck);
        for (int t = 0; t < 6; t++) {
          int64_t e = 6 * cube + t;
          if (keep && !keep[e]) continue;
          int v[4];
          kuhn_vertices(t, v);
          int la = -1;
          for (int a = 0; a < 4; a++)
            if (v[a] == corner) la = a;
          if (la < 0) continue;
          for (int b = 0; b < 4; b++) {
            int dx = (v[b] & 1) - (corner & 1), dy = ((v[b] >> 1) & 1) - ((corner >> 1) & 1),
                dz = ((v[b] >> 2) & 1) - ((corner >> 2) & 1);
            int o = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
            acc[o] += coef[e] * K[t][la][b];
            present[o] = true;
          }
        }
      }
    }
  }
  out.n = 0;
  for (int o = 0; o < 27; o++) {
    if (!present[o]) continue;
    int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
    int64_t b = (i + dx) + (int64_t)g.nx * ((j + dy) + (int64_t)g.ny * (k + dz));
    out.cols[out.n] = (int32_t)b;
    out.vals[out.n] = acc[o];
    out.n++;
  }
}

template <class RowFn>
int64_t count_rows(int64_t nrows, RowFn fn, int32_t *rowptr) {
  rowptr[0] = 0;
  RowOut o;
  int64_t s = 0;
  for (int64_t r = 0; r < nrows; r++) {
    fn(r, o);
    s += o.n;
    rowptr[r + 1] = (int32_t)s;
  }
  return s;
}

template <class RowFn>
void fill_rows(int64_t nrows, RowFn fn, const int32_t *rowptr, int32_t *colind, double *vals) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      RowOut o;
      int64_t r0 = nrows * t / nt, r1 = nrows * (t + 1) / nt;
      for (int64_t r = r0; r < r1; r++) {
        fn(r, o);
        std::memcpy(colind + rowptr[r], o.cols, sizeof(int32_t) * o.n);
        std::memcpy(vals + rowptr[r], o.vals, sizeof(double) * o.n);
      }
    });
  for (auto &x : th) x.join();
}

struct Q1Spec {
  Grid g;
  int ndof;
  std::vector<double> K;
};

Q1Spec q1_spec(int dim, int ex, int ey, int ez, int ndof, double nu) {
  Q1Spec s;
  s.g = make_grid(dim, ex, ey, ez);
  s.ndof = ndof;
  s.K = q1_element(dim, ndof, nu);
  return s;
}

}  // namespace

extern "C" {

// Q1 family: dim 2 (ez ignored) or 3; ndof 1 (Laplacian/diffusion) or 3
// (elasticity, dim 3 only).  coef[elements] scales each element (kappa or E).
int64_t fe_q1_rowptr(int dim, int ex, int ey, int ez, int ndof, int32_t *rowptr) {
  Q1Spec s = q1_spec(dim, ex, ey, ez, ndof, 0.3);
  std::vector<double> ones((size_t)s.g.ex * s.g.ey * s.g.ez, 1.0);
  return count_rows(
      s.g.nodes() * ndof,
      [&](int64_t r, RowOut &o) { q1_row(s.g, ndof, s.K, ones.data(), r / ndof, (int)(r % ndof), o); },
      rowptr);
}

void fe_q1_fill(int dim, int ex, int ey, int ez, int ndof, double nu, const double *coef,
                const int32_t *rowptr, int32_t *colind, double *vals) {
  Q1Spec s = q1_spec(dim, ex, ey, ez, ndof, nu);
  fill_rows(
      s.g.nodes() * ndof,
      [&](int64_t r, RowOut &o) { q1_row(s.g, ndof, s.K, coef, r / ndof, (int)(r % ndof), o); },
      rowptr, colind, vals);
}

// Reference element matrix (for tests): (2^dim*ndof)^2 doubles, row-major.
void fe_q1_element(int dim, int ndof, double nu, double *out) {
  std::vector<double> K = q1_element(dim, ndof, nu);
  std::memcpy(out, K.data(), sizeof(double) * K.size());
}

// P1 Kuhn tets on an ex x ey x ez cube grid; keep[6*cubes] (NULL keeps all).
int64_t fe_tet_rowptr(int ex, int ey, int ez, const uint8_t *keep, int32_t *rowptr) {
  Grid g = make_grid(3, ex, ey, ez);
  double K[6][4][4];
  tet_elements(K);
  std::vector<double> ones((size_t)6 * ex * ey * ez, 1.0);
  return count_rows(
      g.nodes(), [&](int64_t r, RowOut &o) { tet_row(g, K, ones.data(), keep, r, o); }, rowptr);
}

void fe_tet_fill(int ex, int ey, int ez, const uint8_t *keep, const double *coef,
                 const int32_t *rowptr, int32_t *colind, double *vals) {
  Grid g = make_grid(3, ex, ey, ez);
  double K[6][4][4];
  tet_elements(K);
  fill_rows(g.nodes(), [&](int64_t r, RowOut &o) { tet_row(g, K, coef, keep, r, o); }, rowptr,
            colind, vals);
}

void fe_tet_element(double *out) {
  double K[6][4][4];
  tet_elements(K);
  std::memcpy(out, K, sizeof(K));
}

// Symmetric permutation B = P A P^T with B[perm[i], perm[j]] = A[i, j]; columns
// of each row of B sorted ascending.  rowptr_out must be filled first by
// fe_permute_rowptr.
void fe_permute_rowptr(int32_t nrows, const int32_t *rowptr, const int32_t *perm,
                       int32_t *rowptr_out) {
  std::vector<int32_t> inv(nrows);
  for (int32_t i = 0; i < nrows; i++) inv[perm[i]] = i;
  rowptr_out[0] = 0;
  for (int32_t r = 0; r < nrows; r++) {
    int32_t o = inv[r];
    rowptr_out[r + 1] = rowptr_out[r] + (rowptr[o + 1] - rowptr[o]);
  }
}

void fe_permute_fill(int32_t nrows, const int32_t *rowptr, const int32_t *colind,
                     const double *vals, const int32_t *perm, const int32_t *rowptr_out,
                     int32_t *colind_out, double *vals_out) {
  std::vector<int32_t> inv(nrows);
  for (int32_t i = 0; i < nrows; i++) inv[perm[i]] = i;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      std::vector<std::pair<int32_t, double>> tmp;
      int64_t r0 = (int64_t)nrows * t / nt, r1 = (int64_t)nrows * (t + 1) / nt;
      for (int64_t r = r0; r < r1; r++) {
        int32_t o = inv[r];
        tmp.clear();
        for (int32_t q = rowptr[o]; q < rowptr[o + 1]; q++) tmp.emplace_back(perm[colind[q]], vals[q]);
        std::sort(tmp.begin(), tmp.end(),
                  [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) {
                    return a.first < b.first;
                  });
        for (size_t q = 0; q < tmp.size(); q++) {
          colind_out[rowptr_out[r] + q] = tmp[q].first;
          vals_out[rowptr_out[r] + q] = tmp[q].second;
        }
      }
    });
  for (auto &x : th) x.join();
}

}  // extern "C"
