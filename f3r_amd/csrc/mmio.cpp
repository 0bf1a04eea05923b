// MatrixMarket input of the solve: a coordinate matrix read into CSR and the
// <name>_b.mtx right-hand side, with the reference reader's semantics
// (sparse/matrix/io.hpp:104-217, io::CSR(coo) io.hpp:42-64):
//   * the header's first five whitespace-separated tokens are read; the
//     matrix is symmetric only when the fifth is exactly "symmetric"
//     (io.hpp:133-137), so skew-symmetric and hermitian files are read as
//     general (the stored triangle only), as the reference reads them;
//   * lines that start with '%' at the top of the file are comments;
//   * an explicit zero off the diagonal is dropped, a zero diagonal entry is
//     kept (io.hpp:168-171);
//   * an off-diagonal entry of a symmetric file is mirrored (io.hpp:172-177);
//   * entries are placed by row in file order and each row is sorted by column
//     (io.hpp:49-62).  A repeated (row, column) entry, which the reference
//     keeps as two entries of one row, is refused (F3R_ERR_UNSUPPORTED): this
//     library's CSR has one entry per position.
// Where the reference calls exit(1) (unreadable header, a right-hand side that
// is not "array ... general" with one column or not of the matrix's size),
// these return F3R_ERR_ARG.  Additionally, entries outside the declared shape
// and truncated files are errors (so is a header that declares more entries
// than the file has bytes for, before anything is allocated), and
// "pattern" / "complex" files, which the reference's real reader misparses,
// are F3R_ERR_UNSUPPORTED.
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace f3r {
namespace {

std::string slurp(const char *path) {
  std::ifstream f(path, std::ios::in | std::ios::binary);
  F3R_CHECK(f.is_open(), F3R_ERR_ARG, std::string("mm: cannot open ") + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// the first five whitespace-separated tokens (sscanf "%s %s %s %s %s")
std::vector<std::string> header_tokens(const std::string &s, const char *path) {
  std::istringstream in(s);
  std::vector<std::string> t(5);
  for (auto &x : t) F3R_CHECK(static_cast<bool>(in >> x), F3R_ERR_ARG, std::string("mm: short header in ") + path);
  return t;
}

// position after the comment lines at the top (lines starting with '%')
size_t skip_comments(const std::string &s) {
  size_t p = 0;
  while (p < s.size() && s[p] == '%') {
    const size_t nl = s.find('\n', p);
    if (nl == std::string::npos) return s.size();
    p = nl + 1;
  }
  return p;
}

// one number as the reference reads it (strtod, then one separator skipped)
struct Cursor {
  const char *p, *end;
  const char *path;
  double next() {
    F3R_CHECK(p < end, F3R_ERR_ARG, std::string("mm: truncated file ") + path);
    char *e = nullptr;
    const double v = std::strtod(p, &e);
    F3R_CHECK(e != p, F3R_ERR_ARG, std::string("mm: expected a number in ") + path);
    p = e < end ? e + 1 : end;
    return v;
  }
  int64_t index(int64_t lo, int64_t hi) {
    const double v = next();
    F3R_CHECK(v == std::floor(v) && v >= (double)lo && v <= (double)hi, F3R_ERR_ARG,
              std::string("mm: index out of range in ") + path);
    return (int64_t)v;
  }
};

struct MMMatrix {
  int32_t nrows = 0, ncols = 0;
  std::vector<int32_t> rptr, col;
  std::vector<double> val;
};

MMMatrix read_matrix(const char *path) {
  const std::string s = slurp(path);
  const auto tok = header_tokens(s, path);
  F3R_CHECK(tok[0] == "%%MatrixMarket" && tok[2] == "coordinate", F3R_ERR_ARG,
            std::string("mm: not a MatrixMarket coordinate file: ") + path);
  F3R_CHECK(tok[3] != "pattern" && tok[3] != "complex", F3R_ERR_UNSUPPORTED,
            std::string("mm: only real and integer matrices are read: ") + path);
  const bool sym = tok[4] == "symmetric";
  Cursor c{s.data() + skip_comments(s), s.data() + s.size(), path};
  const int64_t nr = c.index(0, INT32_MAX - 1), nc = c.index(0, INT32_MAX), nl = c.index(0, INT64_MAX / 4);
  // an entry takes at least 6 bytes ("1 1 1\n")
  F3R_CHECK(nl <= (int64_t)s.size() / 6, F3R_ERR_ARG, std::string("mm: truncated file ") + path);
  std::vector<int32_t> r, cl;
  std::vector<double> v;
  r.reserve(sym ? 2 * nl : nl);
  cl.reserve(sym ? 2 * nl : nl);
  v.reserve(sym ? 2 * nl : nl);
  for (int64_t k = 0; k < nl; ++k) {
    const int32_t i = (int32_t)(c.index(1, nr) - 1);
    const int32_t j = (int32_t)(c.index(1, nc) - 1);
    const double x = c.next();
    if (x == 0.0 && i != j) continue;
    r.push_back(i);
    cl.push_back(j);
    v.push_back(x);
    if (sym && i != j) {
      F3R_CHECK(j < nr && i < nc, F3R_ERR_ARG, std::string("mm: symmetric entry outside a square shape in ") + path);
      r.push_back(j);
      cl.push_back(i);
      v.push_back(x);
    }
  }
  MMMatrix M;
  M.nrows = (int32_t)nr;
  M.ncols = (int32_t)nc;
  F3R_CHECK(r.size() < (size_t)INT32_MAX, F3R_ERR_RANGE, "mm: more than 2^31 entries");
  M.rptr.assign(nr + 1, 0);
  for (int32_t i : r) M.rptr[i + 1]++;
  for (int64_t i = 0; i < nr; ++i) M.rptr[i + 1] += M.rptr[i];
  M.col.resize(r.size());
  M.val.resize(r.size());
  std::vector<int32_t> fill(M.rptr.begin(), M.rptr.end() - 1);
  for (size_t k = 0; k < r.size(); ++k) {
    const int32_t p = fill[r[k]]++;
    M.col[p] = cl[k];
    M.val[p] = v[k];
  }
  std::vector<std::pair<int32_t, double>> row;
  for (int64_t i = 0; i < nr; ++i) {
    row.clear();
    for (int32_t k = M.rptr[i]; k < M.rptr[i + 1]; ++k) row.push_back({M.col[k], M.val[k]});
    std::stable_sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    for (size_t q = 0; q < row.size(); ++q) {
      F3R_CHECK(q == 0 || row[q].first != row[q - 1].first, F3R_ERR_UNSUPPORTED,
                std::string("mm: a repeated entry (row ") + std::to_string(i + 1) + ", column " +
                    std::to_string(row[q].first + 1) + ") in " + path);
      M.col[M.rptr[i] + q] = row[q].first;
      M.val[M.rptr[i] + q] = row[q].second;
    }
  }
  return M;
}

std::vector<double> read_vector(const char *path) {
  const std::string s = slurp(path);
  const auto tok = header_tokens(s, path);
  F3R_CHECK(tok[2] == "array" && tok[4] == "general", F3R_ERR_ARG,
            std::string("mm: a right-hand side must be \"array ... general\": ") + path);
  Cursor c{s.data() + skip_comments(s), s.data() + s.size(), path};
  const int64_t nr = c.index(0, INT32_MAX), nc = c.index(0, INT32_MAX);
  F3R_CHECK(nc == 1, F3R_ERR_ARG, std::string("mm: a right-hand side has one column: ") + path);
  F3R_CHECK(nr <= (int64_t)s.size() / 2, F3R_ERR_ARG, std::string("mm: truncated file ") + path);  // "1\n" each
  std::vector<double> b(nr);
  for (auto &x : b) x = c.next();
  return b;
}

}  // namespace
}  // namespace f3r

using namespace f3r;

extern "C" {

int f3r_mm_read(f3r_ctx ctx, const char *path, f3r_mat *A) {
  F3R_API_BEGIN
  F3R_CHECK(path && A, F3R_ERR_ARG, "mm_read: null argument");
  *A = nullptr;
  MMMatrix M = read_matrix(path);
  const int st = f3r_mat_create_csr(ctx, M.nrows, M.ncols, (int64_t)M.col.size(), M.rptr.data(), M.col.data(),
                                    M.val.data(), A);
  if (st != F3R_OK) return st;
  F3R_API_END
}

int f3r_mm_read_vector(const char *path, int64_t cap, double *out, int64_t *n) {
  F3R_API_BEGIN
  F3R_CHECK(path && n, F3R_ERR_ARG, "mm_read_vector: null argument");
  std::vector<double> b = read_vector(path);
  *n = (int64_t)b.size();
  if (out) {
    F3R_CHECK(cap >= (int64_t)b.size(), F3R_ERR_ARG, "mm_read_vector: output shorter than the vector");
    std::copy(b.begin(), b.end(), out);
  }
  F3R_API_END
}

int f3r_mm_rhs_path(const char *path, char *out, int64_t cap) {
  F3R_API_BEGIN
  F3R_CHECK(path && out, F3R_ERR_ARG, "mm_rhs_path: null argument");
  std::string p = path;
  F3R_CHECK(p.size() >= 4, F3R_ERR_ARG, "mm_rhs_path: a matrix path ends in .mtx");
  p.insert(p.size() - 4, "_b");  // io.hpp:108-109
  F3R_CHECK((int64_t)p.size() < cap, F3R_ERR_ARG, "mm_rhs_path: output too short");
  std::memcpy(out, p.c_str(), p.size() + 1);
  F3R_API_END
}

int f3r_gen_uniform_mt(int64_t n, double lo, double hi, uint64_t seed, double *out) {
  F3R_API_BEGIN
  F3R_CHECK(n >= 0 && (out || n == 0), F3R_ERR_ARG, "gen_uniform_mt: bad argument");
  std::mt19937_64 eng(seed);
  std::uniform_real_distribution<double> dist(lo, hi);
  for (int64_t i = 0; i < n; ++i) out[i] = dist(eng);
  F3R_API_END
}

int f3r_mat_shape(f3r_mat A, int32_t *nrows, int32_t *ncols) {
  F3R_API_BEGIN
  F3R_CHECK(A && nrows && ncols, F3R_ERR_ARG, "mat_shape: null argument");
  const std::shared_ptr<Mat> m = mat_of(A);
  F3R_CHECK(m && m->host, F3R_ERR_ARG, "mat_shape: a matrix carrying its host CSR");
  *nrows = m->host->n;
  *ncols = m->host->ncols;
  F3R_API_END
}

int f3r_mat_csr_get(f3r_mat A, int32_t *rptr, int32_t *col, double *val) {
  F3R_API_BEGIN
  F3R_CHECK(A, F3R_ERR_ARG, "mat_csr_get: null matrix");
  const std::shared_ptr<Mat> m = mat_of(A);
  F3R_CHECK(m && m->host, F3R_ERR_ARG, "mat_csr_get: a matrix carrying its host CSR");
  const HostCSR &h = *m->host;
  if (rptr) std::copy(h.rptr.begin(), h.rptr.end(), rptr);
  if (col) std::copy(h.idx.begin(), h.idx.end(), col);
  if (val) std::copy(h.val.begin(), h.val.end(), val);
  F3R_API_END
}

}  // extern "C"
