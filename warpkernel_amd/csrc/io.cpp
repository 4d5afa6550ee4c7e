// io.cpp -- files: MatrixMarket matrices (the drivers' input,
// cusp::io::read_matrix_market_file at wpk1.cu:44, wpk2.cu:55) and WPK layouts
// on disk (the save / load of serialize.cu:135-139 over the fields that
// structure::serialize lists, warpkernel.hpp:252-270).
//
// MatrixMarket.  CUSP is not part of the reference tree; its reader is
// restated from the format's published definition and CUSP's documented
// behaviour: "coordinate" matrices with field real, double, integer or pattern
// (pattern entries are 1) and symmetry general, symmetric, skew-symmetric or
// hermitian (real: as symmetric); 1-based indices; a symmetric file's
// off-diagonal entries are mirrored (negated for skew-symmetric); the entries
// are then ordered by (row, column), stably, and duplicates are kept.  The CSR
// conversion (`A = B`, wpk1.cu:48) is the row-offset count of that order.
//
// Layout files: little-endian binary, "WKLAYOUT" + version, the scalars, then
// each array with its length, then an FNV-1a 64 checksum of everything before
// it.  load() checks every invariant the kernels index by (offsets, the row
// permutation, lanes per row), so a damaged or hand-made file is refused
// rather than handed to the GPU.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>

#include "wk_internal.hpp"

struct wk_csr {
  int32_t nrows = 0, ncols = 0;
  std::vector<int32_t> rowptr, colind;
  std::vector<double> vals;
};

namespace wk {
namespace {

std::string lower(std::string s) {
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// ------------------------------------------------------------ MatrixMarket
struct MMEntry {
  int32_t r, c;
  double v;
};

void mm_read(const char *path, wk_csr &A) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(WK_ERR_ARG, std::string("cannot open ") + path);
  std::string line;
  if (!std::getline(f, line)) fail(WK_ERR_CSR, "empty MatrixMarket file");
  std::istringstream hs(line);
  std::string banner, object, format, field, symmetry;
  hs >> banner >> object >> format >> field >> symmetry;
  if (banner != "%%MatrixMarket") fail(WK_ERR_CSR, "missing %%MatrixMarket banner");
  object = lower(object), format = lower(format), field = lower(field), symmetry = lower(symmetry);
  if (object != "matrix") fail(WK_ERR_CSR, "MatrixMarket object must be 'matrix'");
  if (format != "coordinate") fail(WK_ERR_CSR, "only the coordinate format is supported");
  const bool pattern = field == "pattern";
  if (!pattern && field != "real" && field != "double" && field != "integer")
    fail(WK_ERR_CSR, "unsupported MatrixMarket field '" + field + "'");
  const bool skew = symmetry == "skew-symmetric";
  const bool mirror = skew || symmetry == "symmetric" || symmetry == "hermitian";
  if (!mirror && symmetry != "general")
    fail(WK_ERR_CSR, "unsupported MatrixMarket symmetry '" + symmetry + "'");
  // size line: the first line that is neither a comment nor blank
  int64_t M = -1, N = -1, L = -1;
  while (std::getline(f, line)) {
    const size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos || line[p] == '%') continue;
    std::istringstream ss(line);
    if (!(ss >> M >> N >> L)) fail(WK_ERR_CSR, "bad MatrixMarket size line");
    break;
  }
  if (M < 0 || N < 0 || L < 0) fail(WK_ERR_CSR, "missing MatrixMarket size line");
  if (M > INT32_MAX || N > INT32_MAX) fail(WK_ERR_LIMIT, "matrix dimension exceeds 2^31");
  if ((mirror ? 2 * L : L) > INT32_MAX) fail(WK_ERR_LIMIT, "nnz exceeds 2^31");
  std::vector<MMEntry> e;
  e.reserve((size_t)(mirror ? 2 * L : L));
  int64_t read = 0;
  while (read < L && std::getline(f, line)) {
    const char *s = line.c_str(), *end = s + line.size();
    while (s < end && (*s == ' ' || *s == '\t')) s++;
    if (s == end || *s == '%' || *s == '\r') continue;
    int64_t i = 0, j = 0;
    auto r1 = std::from_chars(s, end, i);
    if (r1.ec != std::errc()) fail(WK_ERR_CSR, "bad entry line " + std::to_string(read + 1));
    s = r1.ptr;
    while (s < end && (*s == ' ' || *s == '\t')) s++;
    auto r2 = std::from_chars(s, end, j);
    if (r2.ec != std::errc()) fail(WK_ERR_CSR, "bad entry line " + std::to_string(read + 1));
    s = r2.ptr;
    double v = 1.0;
    if (!pattern) {
      char *q = nullptr;
      v = std::strtod(s, &q);
      if (q == s) fail(WK_ERR_CSR, "missing value on entry line " + std::to_string(read + 1));
    }
    if (i < 1 || i > M || j < 1 || j > N)
      fail(WK_ERR_CSR, "entry index out of range on entry line " + std::to_string(read + 1));
    e.push_back({(int32_t)(i - 1), (int32_t)(j - 1), v});
    read++;
  }
  if (read != L) fail(WK_ERR_CSR, "MatrixMarket file ends before its entries do");
  if (mirror) {
    const size_t n0 = e.size();
    for (size_t t = 0; t < n0; t++)
      if (e[t].r != e[t].c) e.push_back({e[t].c, e[t].r, skew ? -e[t].v : e[t].v});
  }
  std::stable_sort(e.begin(), e.end(), [](const MMEntry &a, const MMEntry &b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  A.nrows = (int32_t)M;
  A.ncols = (int32_t)N;
  A.rowptr.assign((size_t)M + 1, 0);
  A.colind.resize(e.size());
  A.vals.resize(e.size());
  for (size_t t = 0; t < e.size(); t++) {
    A.rowptr[(size_t)e[t].r + 1]++;
    A.colind[t] = e[t].c;
    A.vals[t] = e[t].v;
  }
  for (int64_t r = 0; r < M; r++) A.rowptr[r + 1] += A.rowptr[r];
}

void mm_write(const char *path, int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rowptr,
              const int32_t *colind, const double *vals) {
  validate_csr(nrows, ncols, nnz, rowptr, colind);
  if (nnz && !vals) fail(WK_ERR_ARG, "vals is NULL");
  std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "wb"), &std::fclose);
  if (!f) fail(WK_ERR_ARG, std::string("cannot create ") + path);
  std::fprintf(f.get(), "%%%%MatrixMarket matrix coordinate real general\n");
  std::fprintf(f.get(), "%d %d %lld\n", nrows, ncols, (long long)nnz);
  // shortest decimal that reads back to the same double
  char buf[96];
  for (int32_t r = 0; r < nrows; r++)
    for (int32_t t = rowptr[r]; t < rowptr[r + 1]; t++) {
      char *p = buf + std::snprintf(buf, sizeof buf, "%d %d ", r + 1, colind[t] + 1);
      auto res = std::to_chars(p, buf + sizeof buf - 1, vals[t]);
      *res.ptr++ = '\n';
      if (std::fwrite(buf, 1, (size_t)(res.ptr - buf), f.get()) != (size_t)(res.ptr - buf))
        fail(WK_ERR_ARG, "write failed");
    }
  if (std::fflush(f.get()) != 0) fail(WK_ERR_ARG, "write failed");
}

// ------------------------------------------------------------ layout files
constexpr char kMagic[8] = {'W', 'K', 'L', 'A', 'Y', 'O', 'U', 'T'};
constexpr uint32_t kVersion = 1;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  }
};

struct Writer {
  std::ofstream f;
  Fnv sum;
  void raw(const void *p, size_t n) {
    f.write(static_cast<const char *>(p), (std::streamsize)n);
    sum.add(p, n);
  }
  template <class T>
  void put(T v) {
    raw(&v, sizeof v);
  }
  template <class T>
  void arr(const std::vector<T> &a) {
    put<int64_t>((int64_t)a.size());
    if (!a.empty()) raw(a.data(), sizeof(T) * a.size());
  }
};

struct Reader {
  std::ifstream f;
  Fnv sum;
  int64_t left;
  void raw(void *p, size_t n) {
    if ((int64_t)n > left || !f.read(static_cast<char *>(p), (std::streamsize)n))
      fail(WK_ERR_CSR, "layout file is truncated");
    left -= (int64_t)n;
    sum.add(p, n);
  }
  template <class T>
  T get() {
    T v;
    raw(&v, sizeof v);
    return v;
  }
  template <class T>
  void arr(std::vector<T> &a) {
    const int64_t n = get<int64_t>();
    if (n < 0 || n > left / (int64_t)sizeof(T)) fail(WK_ERR_CSR, "layout file is truncated");
    a.resize((size_t)n);
    if (n) raw(a.data(), sizeof(T) * (size_t)n);
  }
};

void layout_save(const wk_layout &L, const char *path, const char *driver_name) {
  Writer w;
  w.f.open(path, std::ios::binary | std::ios::trunc);
  if (!w.f) fail(WK_ERR_ARG, std::string("cannot create ") + path);
  w.raw(kMagic, sizeof kMagic);
  w.put<uint32_t>(kVersion);
  w.put<uint32_t>(0x01020304u);
  for (int32_t v : {L.variant, L.W, L.threshold, L.max_tpr, L.nrows, L.ncols, L.nznrows, L.nwarps})
    w.put<int32_t>(v);
  w.put<int64_t>(L.nz);
  w.put<int64_t>(L.allocate_nz);
  w.put<uint32_t>(L.min_nz);
  w.put<uint32_t>(L.max_nz);
  const std::string name = driver_name ? driver_name : "";
  w.put<int64_t>((int64_t)name.size());
  w.raw(name.data(), name.size());
  w.arr(L.row_map);
  w.arr(L.row_map_inv);
  w.arr(L.reorder_rows);
  w.arr(L.warp_offsets);
  w.arr(L.max_nz_rows);
  w.arr(L.threads_per_row);
  w.arr(L.row_offset_warp);
  w.arr(L.reorder_A_rows);
  w.arr(L.rowptr);
  const uint64_t h = w.sum.h;
  w.f.write(reinterpret_cast<const char *>(&h), sizeof h);
  w.f.flush();
  if (!w.f) fail(WK_ERR_ARG, std::string("write failed: ") + path);
}

}  // namespace

// Every invariant the kernels and the host helpers index by.
void check_layout(const wk_layout &L, const char *what) {
  auto bad = [what](const std::string &m) { fail(WK_ERR_CSR, std::string(what) + ": " + m); };
  if (L.variant != 1 && L.variant != 2) bad("variant");
  if (L.W != 32 && L.W != 64) bad("group width");
  if (L.nrows < 0 || L.ncols < 0 || L.nznrows < 0 || L.nznrows > L.nrows || L.nwarps < 0) bad("sizes");
  if (L.nz < 0 || L.nz > INT32_MAX || L.allocate_nz < 0 || L.allocate_nz > (int64_t)UINT32_MAX) bad("sizes");
  if ((int64_t)L.row_map.size() != L.nrows || (int64_t)L.row_map_inv.size() != L.nrows) bad("row_map");
  for (int32_t i = 0; i < L.nrows; i++) {
    const int32_t r = L.row_map[i];
    if (r < 0 || r >= L.nrows || L.row_map_inv[r] != i) bad("row_map is not a permutation");
  }
  if ((int64_t)L.reorder_rows.size() != L.nz) bad("reorder_rows");
  for (uint32_t s : L.reorder_rows)
    if ((int64_t)s >= L.allocate_nz) bad("reorder_rows outside the padded arrays");
  if ((int64_t)L.warp_offsets.size() != (int64_t)L.nwarps + 1 ||
      (int64_t)L.max_nz_rows.size() != L.nwarps)
    bad("group arrays");
  if (L.warp_offsets[0] != 0) bad("warp_offsets[0]");
  for (int32_t g = 0; g < L.nwarps; g++)
    if (L.warp_offsets[g + 1] - L.warp_offsets[g] != (int64_t)L.max_nz_rows[g] * L.W)
      bad("warp_offsets do not match max_nz_rows");
  if (L.warp_offsets[L.nwarps] != L.allocate_nz) bad("allocate_nz");
  if (L.variant == 2) {
    if ((int64_t)L.threads_per_row.size() != L.nwarps ||
        (int64_t)L.row_offset_warp.size() != (int64_t)L.nwarps + 1)
      bad("structure2 arrays");
    for (uint32_t t : L.threads_per_row)
      if (t == 0 || t > (uint32_t)L.W || (t & (t - 1))) bad("threads_per_row");
  } else if (!L.threads_per_row.empty() || !L.row_offset_warp.empty()) {
    bad("structure2 arrays in a structure file");
  } else if ((int64_t)L.nwarps != ((int64_t)L.nznrows + L.W - 1) / L.W) {
    bad("group count");
  }
  if ((int64_t)L.rowptr.size() != (int64_t)L.nrows + 1 || L.rowptr[0] != 0 ||
      (int64_t)L.rowptr[L.nrows] != L.nz)
    bad("rowptr");
  for (int32_t r = 0; r < L.nrows; r++)
    if (L.rowptr[r + 1] < L.rowptr[r]) bad("rowptr decreases");
  if (!L.reorder_A_rows.empty()) {
    if ((int64_t)L.reorder_A_rows.size() != L.nz) bad("reorder_A_rows");
    for (int32_t s : L.reorder_A_rows)
      if (s < 0 || (int64_t)s >= L.nz) bad("reorder_A_rows outside the matrix");
  }
}

namespace {

void layout_load(const char *path, wk_layout &L, std::string &driver_name) {
  Reader r;
  r.f.open(path, std::ios::binary);
  if (!r.f) fail(WK_ERR_ARG, std::string("cannot open ") + path);
  r.f.seekg(0, std::ios::end);
  const int64_t size = (int64_t)r.f.tellg();
  r.f.seekg(0);
  if (size < (int64_t)sizeof(kMagic) + 8) fail(WK_ERR_CSR, "not a layout file");
  r.left = size - 8;  // the checksum is not part of the sum
  char magic[8];
  r.raw(magic, sizeof magic);
  if (std::memcmp(magic, kMagic, sizeof magic)) fail(WK_ERR_CSR, "not a layout file");
  if (r.get<uint32_t>() != kVersion) fail(WK_ERR_CSR, "unsupported layout file version");
  if (r.get<uint32_t>() != 0x01020304u) fail(WK_ERR_CSR, "layout file has the other byte order");
  L.variant = r.get<int32_t>();
  L.W = r.get<int32_t>();
  L.threshold = r.get<int32_t>();
  L.max_tpr = r.get<int32_t>();
  L.nrows = r.get<int32_t>();
  L.ncols = r.get<int32_t>();
  L.nznrows = r.get<int32_t>();
  L.nwarps = r.get<int32_t>();
  L.nz = r.get<int64_t>();
  L.allocate_nz = r.get<int64_t>();
  L.min_nz = r.get<uint32_t>();
  L.max_nz = r.get<uint32_t>();
  const int64_t nlen = r.get<int64_t>();
  if (nlen < 0 || nlen > r.left) fail(WK_ERR_CSR, "layout file is truncated");
  driver_name.resize((size_t)nlen);
  if (nlen) r.raw(&driver_name[0], (size_t)nlen);
  r.arr(L.row_map);
  r.arr(L.row_map_inv);
  r.arr(L.reorder_rows);
  r.arr(L.warp_offsets);
  r.arr(L.max_nz_rows);
  r.arr(L.threads_per_row);
  r.arr(L.row_offset_warp);
  r.arr(L.reorder_A_rows);
  r.arr(L.rowptr);
  if (r.left != 0) fail(WK_ERR_CSR, "layout file has trailing data");
  uint64_t h = 0;
  if (!r.f.read(reinterpret_cast<char *>(&h), sizeof h) || h != r.sum.h)
    fail(WK_ERR_CSR, "layout file checksum mismatch");
  check_layout(L, "layout file");
}

}  // namespace
}  // namespace wk

using namespace wk;

extern "C" {

int wk_mm_read(const char *path, wk_csr **out) {
  return guarded([&] {
    if (!path || !out) fail(WK_ERR_ARG, "NULL argument");
    *out = nullptr;
    auto A = std::make_unique<wk_csr>();
    mm_read(path, *A);
    *out = A.release();
  });
}

int wk_csr_get(const wk_csr *A, int32_t *nrows, int32_t *ncols, int64_t *nnz,
               const int32_t **rowptr, const int32_t **colind, const double **vals) {
  return guarded([&] {
    if (!A) fail(WK_ERR_ARG, "csr is NULL");
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = (int64_t)A->colind.size();
    if (rowptr) *rowptr = A->rowptr.data();
    if (colind) *colind = A->colind.data();
    if (vals) *vals = A->vals.data();
  });
}

void wk_csr_destroy(wk_csr *A) { delete A; }

int wk_mm_write(const char *path, int32_t nrows, int32_t ncols, int64_t nnz,
                const int32_t *rowptr_host, const int32_t *colind_host, const double *vals_host) {
  return guarded([&] {
    if (!path) fail(WK_ERR_ARG, "path is NULL");
    mm_write(path, nrows, ncols, nnz, rowptr_host, colind_host, vals_host);
  });
}

int wk_layout_save(const wk_layout *L, const char *path, const char *driver_name) {
  return guarded([&] {
    if (!L || !path) fail(WK_ERR_ARG, "NULL argument");
    layout_save(*L, path, driver_name);
  });
}

int wk_layout_load(const char *path, wk_layout **out, char *driver_name, int64_t name_capacity) {
  return guarded([&] {
    if (!path || !out) fail(WK_ERR_ARG, "NULL argument");
    *out = nullptr;
    auto L = std::make_unique<wk_layout>();
    std::string name;
    layout_load(path, *L, name);
    if (driver_name && name_capacity > 0) {
      const size_t n = std::min<size_t>(name.size(), (size_t)name_capacity - 1);
      std::memcpy(driver_name, name.data(), n);
      driver_name[n] = '\0';
    }
    *out = L.release();
  });
}

}  // extern "C"
