// Host-side set-up of the solve: diagonal scaling, SELL-C-sigma layout,
// SDAINV and ILU(0) factorizations, and the synthetic input generators.
// These run once before the timed solve (the reference also runs them on the
// host, outside its timer: work/src/f3r-gpu.cpp:65-74).
//
// Compiled with -ffp-contract=off: the factorizations must round exactly as
// the reference's host code does (one rounding per operation).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "common.hpp"

namespace f3r {

// fn(begin, end) over [0, n) split into contiguous ranges, one per hardware
// thread (host set-up only; every range writes disjoint outputs)
template <class Fn> static void parallel_ranges(int64_t n, Fn fn) {
  const int64_t hw = std::max<int64_t>(1, std::thread::hardware_concurrency());
  const int64_t nt = std::min<int64_t>(hw, std::max<int64_t>(1, n / 256));
  if (nt <= 1) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
  for (auto &x : th) x.join();
}

// algorithm::scaling (sparse/matrix/algorithm.hpp:45-59): every entry divided
// by sqrt|a_ii| * sqrt|a_jj|; rows without a diagonal use 1.
void csr_scaling(int32_t n, const int32_t *rptr, const int32_t *idx, double *val) {
  std::vector<double> d(n, 1.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = rptr[i]; k < rptr[i + 1]; ++k)
      if (idx[k] == i) d[i] = std::sqrt(std::fabs(val[k]));
  for (int32_t i = 0; i < n; ++i) {
    const double di = d[i];
    for (int32_t k = rptr[i]; k < rptr[i + 1]; ++k) val[k] /= di * d[idx[k]];
  }
}

// Rows sorted by descending length inside each window of sigma rows
// (csr_to_sell_sigma, sell.hpp:314-327), ties kept in row order.
std::vector<int32_t> sigma_perm(int32_t n, const int32_t *rptr, int32_t sigma) {
  std::vector<int32_t> p(n);
  std::iota(p.begin(), p.end(), 0);
  if (sigma <= 1) return p;
  for (int64_t w0 = 0; w0 < n; w0 += sigma) {
    auto b = p.begin() + w0, e = p.begin() + std::min<int64_t>(n, w0 + sigma);
    std::stable_sort(b, e, [rptr](int32_t a, int32_t c) {
      return (rptr[a + 1] - rptr[a]) > (rptr[c + 1] - rptr[c]);
    });
  }
  return p;
}

static inline int32_t round_up(int32_t v, int32_t m) { return (v + m - 1) / m * m; }

int64_t sell_nslots(int32_t n, const int32_t *rptr, int32_t C, int32_t sigma, int32_t K) {
  auto p = sigma_perm(n, rptr, sigma);
  int64_t tot = 0;
  for (int64_t s0 = 0; s0 < n; s0 += C) {
    int32_t tc = (int32_t)std::min<int64_t>(C, n - s0);
    int32_t mx = 0;
    for (int32_t j = 0; j < tc; ++j) {
      int32_t o = p[s0 + j];
      mx = std::max(mx, rptr[o + 1] - rptr[o]);
    }
    tot += (int64_t)round_up(mx, K) * tc;
  }
  return tot;
}

HostSell build_sell(const HostCSR &A, int32_t C, int32_t sigma, int32_t K,
                    const std::vector<int32_t> *perm_override, bool renumber, bool offsets) {
  F3R_CHECK(C >= 1 && K >= 1 && sigma >= 1, F3R_ERR_ARG, "SELL: C, K and sigma must be >= 1");
  HostSell S;
  const int32_t n = A.n;
  S.n = n;
  S.C = C;
  S.K = K;
  S.perm = perm_override ? *perm_override : sigma_perm(n, A.rptr.data(), sigma);
  std::vector<int32_t> inv;
  if (renumber) {
    inv.resize(n);
    for (int32_t r = 0; r < n; ++r) inv[S.perm[r]] = r;
  }
  auto col = [&](int32_t k) { return renumber ? inv[A.idx[k]] : A.idx[k]; };
  // a wavefront of 64 rows must lie in one chunk for the offset format, and
  // the x gathers address (n + offset range) elements with 32-bit byte offsets
  offsets = offsets && C >= 64 && n < (1 << 27);
  S.nchunks = (n + C - 1) / C;
  S.cptr.assign(S.nchunks + 1, 0);
  S.cwidth.assign(S.nchunks, 0);
  std::vector<std::vector<int32_t>> uni(offsets ? S.nchunks : 0);
  // pass 1 (parallel over chunks): the widest row and, for the offset
  // format, the union of col - row
  parallel_ranges(S.nchunks, [&](int64_t c0, int64_t c1) {
    std::vector<int32_t> u, first;
    for (int64_t s = c0; s < c1; ++s) {
      int64_t r0 = s * C;
      int32_t tc = (int32_t)std::min<int64_t>(C, n - r0);
      int32_t mx = 0;
      for (int32_t j = 0; j < tc; ++j) {
        int32_t o = S.perm[r0 + j];
        mx = std::max(mx, A.rptr[o + 1] - A.rptr[o]);
      }
      int32_t width = mx;
      if (offsets && mx > 0) {
        // eligible when every row's columns ascend (CSR order = offset
        // order), every row + offset is a row of the matrix, and the union
        // is at most 5/4 of the widest row.  Rows repeating the previous
        // row's offsets (the stencil case) skip the merge.
        u.clear();
        first.clear();
        bool ok = true;
        for (int32_t j = 0; j < tc && ok; ++j) {
          const int32_t o = S.perm[r0 + j], r = (int32_t)(r0 + j);
          const int32_t len = A.rptr[o + 1] - A.rptr[o];
          bool same = j > 0 && len == (int32_t)first.size();
          int32_t prev = INT32_MIN;
          for (int32_t q = 0; q < len; ++q) {
            const int32_t c = col(A.rptr[o] + q);
            if (c <= prev) ok = false;
            prev = c;
            if (same && first[q] != c - r) same = false;
          }
          if (same) continue;
          first.clear();
          for (int32_t q = 0; q < len; ++q) first.push_back(col(A.rptr[o] + q) - r);
          u.insert(u.end(), first.begin(), first.end());
        }
        if (ok) {
          std::sort(u.begin(), u.end());
          u.erase(std::unique(u.begin(), u.end()), u.end());
          ok = (int64_t)u.size() * 4 <= (int64_t)mx * 5 && r0 + u.front() >= 0 && r0 + tc - 1 + u.back() < n;
        }
        if (ok) {
          uni[s] = u;
          S.cwidth[s] = offset_chunk_width((int32_t)u.size(), K);
          continue;
        }
      }
      S.cwidth[s] = round_up(width, K) / K;
    }
  });
  int64_t tot = 0;
  for (int32_t s = 0; s < S.nchunks; ++s) {
    int32_t tc = (int32_t)std::min<int64_t>(C, n - (int64_t)s * C);
    tot += (int64_t)chunk_slots(S.cwidth[s], K) * tc;
    F3R_CHECK(tot < INT32_MAX, F3R_ERR_RANGE, "SELL: more than 2^31 slots");
    S.cptr[s + 1] = (int32_t)tot;
  }
  S.val.assign(tot, 0.0);
  int64_t itot = 0;
  if (offsets) {
    S.iptr.assign(S.nchunks, 0);
    for (int32_t s = 0; s < S.nchunks; ++s) {
      const int64_t w = chunk_slots(S.cwidth[s], K);
      if (!uni[s].empty()) {
        if (s > 0 && uni[s] == uni[s - 1]) {
          // the previous chunk's offsets: share its table, so the tables the
          // wavefronts read stay few and cached (one for a stencil's interior)
          S.iptr[s] = S.iptr[s - 1];
        } else {
          S.iptr[s] = -(int32_t)S.coff.size() - 1;
          S.off_bias = std::max(S.off_bias, -uni[s].front());
          for (int64_t t = 0; t < w; ++t) S.coff.push_back(t < (int64_t)uni[s].size() ? uni[s][t] : 0);
        }
        ++S.offset_chunks;
      } else {
        S.iptr[s] = (int32_t)itot;
        itot += (S.cptr[s + 1] - S.cptr[s]);
      }
    }
  } else {
    itot = tot;
  }
  S.idx.assign(itot, 0);
  // pass 3 (parallel over chunks): values and indices; every chunk writes
  // its own range
  std::vector<int64_t> off_slots(S.nchunks, 0), expl_nnz(S.nchunks, 0);
  parallel_ranges(S.nchunks, [&](int64_t c0, int64_t c1) {
  for (int64_t s = c0; s < c1; ++s) {
    int64_t r0 = s * C;
    int32_t tc = (int32_t)std::min<int64_t>(C, n - r0);
    int32_t width = chunk_slots(S.cwidth[s], K);
    if (offsets && !uni[s].empty()) {
      // offset chunk: the value of column r + uni[t] goes to slot t, the
      // other slots stay 0 (their columns r + uni[t] are in range)
      const std::vector<int32_t> &us = uni[s];
      off_slots[s] = (int64_t)width * tc;
      for (int32_t j = 0; j < tc; ++j) {
        const int32_t o = S.perm[r0 + j], r = (int32_t)(r0 + j);
        int32_t t = 0;
        for (int32_t k = A.rptr[o]; k < A.rptr[o + 1]; ++k) {
          const int32_t d = col(k) - r;
          while (us[t] < d) ++t;  // columns ascend, so the slot only moves forward
          S.val[chunk_slot_pos(S.cptr[s], S.cwidth[s], tc, j, t, K)] = A.val[k];
        }
      }
      continue;
    }
    const int64_t ibase = offsets ? (int64_t)S.iptr[s] - S.cptr[s] : 0;
    for (int32_t j = 0; j < tc; ++j) {
      int32_t o = S.perm[r0 + j];
      int32_t len = A.rptr[o + 1] - A.rptr[o];
      expl_nnz[s] += len;
      int32_t last = 0;
      for (int32_t q = 0; q < width; ++q) {
        int64_t pos = S.cptr[s] + ((int64_t)(q / K) * tc + j) * K + (q % K);
        if (q < len) {
          last = col(A.rptr[o] + q);
          S.idx[pos + ibase] = last;
          S.val[pos] = A.val[A.rptr[o] + q];
        } else {  // padding: value 0, the previous slot's column (sell.hpp:281-285)
          S.idx[pos + ibase] = last;
          S.val[pos] = 0.0;
        }
      }
    }
  }
  });
  for (int32_t s = 0; s < S.nchunks; ++s) {
    S.offset_slots += off_slots[s];
    S.explicit_nnz += expl_nnz[s];
  }
  return S;
}

// ----------------------------------------------------------------- SDAINV --
// Sparse biconjugation restricted to A's pattern (left_sdainv,
// ainv.hpp:268-329).  Each z_i / w_i is a list of (column, value) sorted by
// column; updates drop entries with |v| < tol (spvec.hpp:66-92) and never
// keep exact zeros (spvec.hpp:15-18).
namespace {
struct SparseRow {
  std::vector<int32_t> c;
  std::vector<double> v;
};

// sum over matching columns of r and (cols, vals), in column order
double merge_dot(const SparseRow &r, const int32_t *cols, const double *vals, int32_t len) {
  double s = 0.0;
  size_t a = 0;
  int32_t b = 0;
  while (a < r.c.size() && b < len) {
    if (r.c[a] < cols[b]) {
      ++a;
    } else if (r.c[a] > cols[b]) {
      ++b;
    } else {
      s += r.v[a] * vals[b];
      ++a;
      ++b;
    }
  }
  return s;
}

// out = y + alpha * x with the absolute drop rule
void merge_axpy(double alpha, const SparseRow &x, const SparseRow &y, double tol,
                SparseRow &out) {
  out.c.clear();
  out.v.clear();
  size_t a = 0, b = 0;
  while (a < x.c.size() || b < y.c.size()) {
    int64_t cx = a < x.c.size() ? x.c[a] : INT64_MAX;
    int64_t cy = b < y.c.size() ? y.c[b] : INT64_MAX;
    int32_t col;
    double v;
    if (cx < cy) {
      col = (int32_t)cx;
      v = x.v[a] * alpha;
      ++a;
    } else if (cx == cy) {
      col = (int32_t)cy;
      v = y.v[b] + alpha * x.v[a];
      ++a;
      ++b;
    } else {
      col = (int32_t)cy;
      v = y.v[b];
      ++b;
    }
    if (std::fabs(v) >= tol && v != 0.0) {
      out.c.push_back(col);
      out.v.push_back(v);
    }
  }
}

}  // namespace

HostCSR csr_transpose(const HostCSR &A) {
  HostCSR T;
  T.n = A.ncols ? A.ncols : A.n;
  T.ncols = A.n;
  T.rptr.assign(T.n + 1, 0);
  for (int32_t k : A.idx) T.rptr[k + 1]++;
  for (int32_t i = 0; i < T.n; ++i) T.rptr[i + 1] += T.rptr[i];
  T.idx.resize(A.nnz());
  T.val.resize(A.nnz());
  std::vector<int32_t> fill(T.rptr.begin(), T.rptr.end() - 1);
  for (int32_t i = 0; i < A.n; ++i)
    for (int32_t k = A.rptr[i]; k < A.rptr[i + 1]; ++k) {
      int32_t p = fill[A.idx[k]]++;
      T.idx[p] = i;
      T.val[p] = A.val[k];
    }
  return T;
}

namespace {
HostCSR rows_to_csr(const std::vector<SparseRow> &rows) {
  HostCSR R;
  R.n = R.ncols = (int32_t)rows.size();
  R.rptr.assign(R.n + 1, 0);
  for (int32_t i = 0; i < R.n; ++i) R.rptr[i + 1] = R.rptr[i] + (int32_t)rows[i].c.size();
  R.idx.reserve(R.rptr[R.n]);
  R.val.reserve(R.rptr[R.n]);
  for (const auto &r : rows) {
    R.idx.insert(R.idx.end(), r.c.begin(), r.c.end());
    R.val.insert(R.val.end(), r.v.begin(), r.v.end());
  }
  return R;
}
}  // namespace

HostFactors sdainv_host(const HostCSR &A, double tol, double acc) {
  const int32_t N = A.n;
  HostCSR At = csr_transpose(A);
  std::vector<double> d(N, 0.0);
  std::vector<SparseRow> z(N), w(N);
  for (int32_t i = 0; i < N; ++i) {
    z[i].c = {i};
    z[i].v = {1.0};
    w[i].c = {i};
    w[i].v = {1.0};
  }
  SparseRow tmp;
  for (int32_t i = 0; i < N; ++i) {
    const int32_t *arow = A.idx.data() + A.rptr[i];
    const double *aval = A.val.data() + A.rptr[i];
    const int32_t alen = A.rptr[i + 1] - A.rptr[i];
    const int32_t *acol = At.idx.data() + At.rptr[i];
    const double *acv = At.val.data() + At.rptr[i];
    const int32_t clen = At.rptr[i + 1] - At.rptr[i];
    // w_i against the z of the lower part of row i
    for (int32_t k = 0; k < alen && arow[k] < i; ++k) {
      const int32_t j = arow[k];
      const double alpha = -merge_dot(z[j], arow, aval, alen) / d[j];
      merge_axpy(alpha, w[j], w[i], tol, tmp);
      std::swap(w[i], tmp);
    }
    // z_i against the w of the upper part of column i
    for (int32_t k = 0; k < clen && acol[k] < i; ++k) {
      const int32_t j = acol[k];
      const double alpha = -merge_dot(w[j], acol, acv, clen) / d[j];
      merge_axpy(alpha, z[j], z[i], tol, tmp);
      std::swap(z[i], tmp);
    }
    // pivot with the diagonal weight acc on w_i's last entry (ainv.hpp:310-316)
    if (!w[i].v.empty()) {
      double &last = w[i].v.back();
      last *= acc;
      d[i] = merge_dot(w[i], acol, acv, clen);
      last /= acc;
    }
    if (d[i] == 0.0) d[i] = 1.0;
  }
  HostFactors F;
  // Z = (rows z_i / d_i)^T, W^T = rows w_i (ainv.hpp:57-64)
  for (int32_t i = 0; i < N; ++i)
    for (double &v : z[i].v) v /= d[i];
  F.f[0] = csr_transpose(rows_to_csr(z));
  F.f[1] = rows_to_csr(w);
  return F;
}

// ----------------------------------------------------------------- ILU(0) --
// IKJ ILU(0) on the diagonal blocks of `blocks` equal row ranges, diagonal
// multiplied by acc before elimination (ilu.hpp:128-164,
// csr.hpp:109-155), then split into unit L (1 stored last in each row) and
// U with its diagonal first (algorithm.hpp:8-43).
HostCSR ilu0_restrict(const HostCSR &A, int32_t blocks) {
  const int32_t n = A.n;
  F3R_CHECK(blocks >= 1 && blocks <= n, F3R_ERR_ARG, "ILU(0): blocks must be in [1, n]");
  const int32_t bs = (n + blocks - 1) / blocks;
  F3R_CHECK((int64_t)bs * (blocks - 1) <= n, F3R_ERR_ARG,
            "ILU(0): blocks leave an empty or negative last block");
  std::vector<int32_t> bnd(blocks + 1);
  for (int32_t b = 0; b < blocks; ++b) bnd[b] = b * bs;
  bnd[blocks] = n;
  HostCSR R;
  R.n = R.ncols = n;
  R.rptr.assign(n + 1, 0);
  for (int32_t b = 0; b < blocks; ++b)
    for (int32_t i = bnd[b]; i < bnd[b + 1]; ++i) {
      for (int32_t k = A.rptr[i]; k < A.rptr[i + 1]; ++k)
        if (A.idx[k] >= bnd[b] && A.idx[k] < bnd[b + 1]) {
          R.idx.push_back(A.idx[k]);
          R.val.push_back(A.val[k]);
        }
      R.rptr[i + 1] = (int32_t)R.idx.size();
    }
  return R;
}

std::vector<int32_t> ilu0_diag_pos(const HostCSR &R) {
  std::vector<int32_t> dpos(R.n, -1);
  for (int32_t i = 0; i < R.n; ++i)
    for (int32_t k = R.rptr[i]; k < R.rptr[i + 1]; ++k)
      if (R.idx[k] == i) dpos[i] = k;
  for (int32_t i = 0; i < R.n; ++i)
    F3R_CHECK(dpos[i] >= 0, F3R_ERR_ARG, "ILU(0): a row has no diagonal entry");
  return dpos;
}

HostFactors ilu0_split(const HostCSR &R) { return ilu0_split(R, R.val.data()); }

HostFactors ilu0_split(const HostCSR &R, const double *v) {
  const int32_t n = R.n;
  const auto &c = R.idx;
  const auto &p = R.rptr;
  HostFactors F;
  HostCSR &L = F.f[0], &U = F.f[1];
  L.n = U.n = L.ncols = U.ncols = n;
  L.rptr.assign(n + 1, 0);
  U.rptr.assign(n + 1, 0);
  const size_t half = c.size() / 2 + (size_t)n;
  L.idx.reserve(half);
  L.val.reserve(half);
  U.idx.reserve(half);
  U.val.reserve(half);
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t k = p[i]; k < p[i + 1]; ++k) {
      if (c[k] < i) {
        L.idx.push_back(c[k]);
        L.val.push_back(v[k]);
      } else if (c[k] == i) {
        L.idx.push_back(i);
        L.val.push_back(1.0);
      }
      if (c[k] >= i) {
        U.idx.push_back(c[k]);
        U.val.push_back(v[k]);
      }
    }
    L.rptr[i + 1] = (int32_t)L.idx.size();
    U.rptr[i + 1] = (int32_t)U.idx.size();
  }
  return F;
}

HostFactors ilu0_host(const HostCSR &A, int32_t blocks, double acc) {
  HostCSR R = ilu0_restrict(A, blocks);
  const int32_t n = R.n;
  auto &v = R.val;
  const auto &c = R.idx;
  const auto &p = R.rptr;
  const std::vector<int32_t> dpos = ilu0_diag_pos(R);
  for (int32_t i = 0; i < n; ++i) {
    v[dpos[i]] *= acc;
    for (int32_t q = p[i]; q < p[i + 1] && c[q] < i; ++q) {
      const int32_t k = c[q];
      int32_t r = dpos[k];
      v[q] = v[q] / v[r];
      int32_t t = q + 1;
      while (t < p[i + 1] && r < p[k + 1]) {
        if (c[r] < c[t]) {
          ++r;
        } else if (c[r] > c[t]) {
          ++t;
        } else {
          v[t] -= v[q] * v[r];
          ++t;
          ++r;
        }
      }
    }
  }
  return ilu0_split(R);
}

int32_t factor_levels(const HostCSR &A, bool both, std::vector<int32_t> &lptr, std::vector<int32_t> &rows) {
  // rows are sorted: the entries before column i, then (diagonal and) after
  const int32_t n = A.n;
  std::vector<int32_t> lev(n), need(both ? n : 0, 0);
  const int32_t *rp = A.rptr.data(), *ix = A.idx.data();
  int32_t top = -1;
  for (int32_t i = 0; i < n; ++i) {
    int32_t l = both ? need[i] : 0;
    const int32_t *p = ix + rp[i], *e = ix + rp[i + 1];
    for (; p < e && *p < i; ++p) {
      const int32_t c = lev[*p] + 1;
      l = c > l ? c : l;
    }
    lev[i] = l;
    top = l > top ? l : top;
    if (both) {
      const int32_t c = l + 1;
      for (; p < e; ++p) {
        int32_t &d = need[*p];
        d = (*p != i && c > d) ? c : d;
      }
    }
  }
  const int32_t nl = top + 1;
  lptr.assign(nl + 1, 0);
  for (int32_t i = 0; i < n; ++i) lptr[lev[i] + 1]++;
  for (int32_t l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
  rows.resize(n);
  std::vector<int32_t> fill(lptr.begin(), lptr.end() - 1);
  for (int32_t i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
  return nl;
}

uint16_t d2h_bits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  const uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
  const uint64_t mag = u & 0x7fffffffffffffffULL;
  if (mag >= 0x7ff0000000000000ULL)  // inf / nan
    return sign | 0x7c00u | (mag > 0x7ff0000000000000ULL ? 0x0200u : 0u);
  const int e = (int)(mag >> 52) - 1023;
  // fp64 subnormals and |d| < 2^-25 (half the smallest fp16 subnormal) -> 0
  if ((mag >> 52) == 0 || e < -25) return sign;
  if (e > 15) return sign | 0x7c00u;
  const uint64_t m = (mag & 0xfffffffffffffULL) | (1ULL << 52);
  // keep 11 significant bits for normals, fewer for subnormals
  const int drop = e >= -14 ? 42 : 42 + (-14 - e);
  uint64_t q = m >> drop;
  const uint64_t rest = m & ((1ULL << drop) - 1), half = 1ULL << (drop - 1);
  q += (rest > half || (rest == half && (q & 1))) ? 1 : 0;
  if (e >= -14) {
    uint32_t ex = (uint32_t)(e + 15);
    if (q >> 11) {  // mantissa overflow into the next binade
      q >>= 1;
      ++ex;
    }
    if (ex >= 31) return sign | 0x7c00u;
    return (uint16_t)(sign | (ex << 10) | (q & 0x3ffu));
  }
  return (uint16_t)(sign | q);
}

// ---------------------------------------------------------- generators ----
namespace {
inline uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline double unit53(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

// Gaussian random field of unit variance: white noise smoothed by a
// separable Gaussian kernel of standard deviation `corr` cells, then
// standardised to mean 0 and variance 1.
std::vector<double> gaussian_field(int32_t nx, int32_t ny, int32_t nz, double corr,
                                   uint64_t seed) {
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<double> f(n);
  uint64_t s = seed * 0x2545f4914f6cdd1dULL + 1;
  for (int64_t i = 0; i < n; i += 2) {
    double u1 = 1.0 - unit53(s), u2 = unit53(s);
    double r = std::sqrt(-2.0 * std::log(u1));
    f[i] = r * std::cos(6.283185307179586 * u2);
    if (i + 1 < n) f[i + 1] = r * std::sin(6.283185307179586 * u2);
  }
  if (corr > 0) {
    const int rad = std::max(1, (int)std::ceil(3.0 * corr));
    std::vector<double> ker(2 * rad + 1);
    for (int t = -rad; t <= rad; ++t) ker[t + rad] = std::exp(-0.5 * t * t / (corr * corr));
    std::vector<double> g(n);
    const int64_t dims[3] = {nx, ny, nz};
    const int64_t strides[3] = {1, nx, (int64_t)nx * ny};
    for (int ax = 0; ax < 3; ++ax) {
      const int64_t L = dims[ax], st = strides[ax];
      for (int64_t i = 0; i < n; ++i) {
        const int64_t pos = (i / st) % L;
        double acc = 0.0;
        for (int t = -rad; t <= rad; ++t) {
          int64_t q = pos + t;
          if (q < 0 || q >= L) continue;  // zero padding
          acc += ker[t + rad] * f[i + t * st];
        }
        g[i] = acc;
      }
      f.swap(g);
    }
  }
  double mean = 0.0;
  for (double v : f) mean += v;
  mean /= (double)n;
  double var = 0.0;
  for (double v : f) var += (v - mean) * (v - mean);
  var /= (double)n;
  const double inv = var > 0 ? 1.0 / std::sqrt(var) : 1.0;
  for (double &v : f) v = (v - mean) * inv;
  return f;
}
}  // namespace

}  // namespace f3r

// ---------------------------------------------------------------- C ABI ----
using namespace f3r;

extern "C" {

int f3r_gen_stencil27_nnz(int32_t nx, int32_t ny, int32_t nz, int64_t *nnz) {
  if (nx < 1 || ny < 1 || nz < 1 || !nnz) return F3R_ERR_ARG;
  *nnz = (int64_t)(3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
  return F3R_OK;
}

int f3r_gen_stencil27(int32_t nx, int32_t ny, int32_t nz, double beta, int32_t *rptr,
                      int32_t *col, double *val) {
  if (nx < 1 || ny < 1 || nz < 1 || !rptr || !col || !val) return F3R_ERR_ARG;
  if ((int64_t)nx * ny * nz >= INT32_MAX) return F3R_ERR_RANGE;
  int64_t cnt = 0;
  rptr[0] = 0;
  for (int32_t k = 0; k < nz; ++k)
    for (int32_t j = 0; j < ny; ++j)
      for (int32_t i = 0; i < nx; ++i) {
        const int64_t row = ((int64_t)k * ny + j) * nx + i;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              if (k + dz < 0 || k + dz >= nz || j + dy < 0 || j + dy >= ny || i + dx < 0 ||
                  i + dx >= nx)
                continue;
              col[cnt] = (int32_t)(((int64_t)(k + dz) * ny + (j + dy)) * nx + (i + dx));
              val[cnt] = (dz == 0 && dy == 0 && dx == 0) ? 26.0 : -1.0 + beta * dx;
              ++cnt;
            }
        rptr[row + 1] = (int32_t)cnt;
      }
  return F3R_OK;
}

int f3r_gen_convdiff7_nnz(int32_t nx, int32_t ny, int32_t nz, int64_t *nnz) {
  if (nx < 1 || ny < 1 || nz < 1 || !nnz) return F3R_ERR_ARG;
  const int64_t n = (int64_t)nx * ny * nz;
  *nnz = n + 2 * ((int64_t)(nx - 1) * ny * nz + (int64_t)nx * (ny - 1) * nz +
                  (int64_t)nx * ny * (nz - 1));
  return F3R_OK;
}

int f3r_gen_convdiff7(int32_t nx, int32_t ny, int32_t nz, double sigma, double corr,
                      double peclet, uint64_t seed, int32_t *rptr, int32_t *col, double *val) {
  if (nx < 1 || ny < 1 || nz < 1 || !rptr || !col || !val) return F3R_ERR_ARG;
  if ((int64_t)nx * ny * nz >= INT32_MAX) return F3R_ERR_RANGE;
  try {
    std::vector<double> g = gaussian_field(nx, ny, nz, corr, seed);
    for (double &v : g) v = std::exp(sigma * v);  // lognormal diffusion
    auto face = [](double a, double b) { return 2.0 * a * b / (a + b); };
    int64_t cnt = 0;
    rptr[0] = 0;
    for (int32_t k = 0; k < nz; ++k)
      for (int32_t j = 0; j < ny; ++j)
        for (int32_t i = 0; i < nx; ++i) {
          const int64_t row = ((int64_t)k * ny + j) * nx + i;
          const double kc = g[row];
          // neighbours in column order: -z, -y, -x, (diag), +x, +y, +z
          const int64_t nb[6] = {row - (int64_t)nx * ny, row - nx, row - 1,
                                 row + 1, row + nx, row + (int64_t)nx * ny};
          const bool ok[6] = {k > 0, j > 0, i > 0, i < nx - 1, j < ny - 1, k < nz - 1};
          double diag = 0.0, off[6];
          for (int t = 0; t < 6; ++t) {
            // Dirichlet faces on the boundary: half-cell distance
            const double kf = ok[t] ? face(kc, g[nb[t]]) : 2.0 * kc;
            diag += kf;
            off[t] = -kf;
          }
          // first-order upwind convection along +x: inflow from -x
          diag += peclet;
          off[2] -= peclet;
          int64_t start = cnt;
          for (int t = 0; t < 3; ++t)
            if (ok[t]) {
              col[cnt] = (int32_t)nb[t];
              val[cnt++] = off[t];
            }
          col[cnt] = (int32_t)row;
          val[cnt++] = diag;
          for (int t = 3; t < 6; ++t)
            if (ok[t]) {
              col[cnt] = (int32_t)nb[t];
              val[cnt++] = off[t];
            }
          (void)start;
          rptr[row + 1] = (int32_t)cnt;
        }
  } catch (...) {
    return F3R_ERR_ALLOC;
  }
  return F3R_OK;
}

int f3r_gen_dominant_nnz(int32_t npts, int32_t nx, int32_t ny, int32_t nz, int64_t *nnz) {
  if (npts == 27) return f3r_gen_stencil27_nnz(nx, ny, nz, nnz);
  if (npts == 7) return f3r_gen_convdiff7_nnz(nx, ny, nz, nnz);
  return F3R_ERR_ARG;
}

int f3r_gen_dominant(int32_t npts, int32_t nx, int32_t ny, int32_t nz, double margin,
                     int32_t *rptr, int32_t *col, double *val) {
  if (npts != 7 && npts != 27) return F3R_ERR_ARG;
  if (nx < 1 || ny < 1 || nz < 1 || !rptr || !col || !val) return F3R_ERR_ARG;
  int64_t cnt = 0;
  rptr[0] = 0;
  for (int32_t k = 0; k < nz; ++k)
    for (int32_t j = 0; j < ny; ++j)
      for (int32_t i = 0; i < nx; ++i) {
        const int64_t row = ((int64_t)k * ny + j) * nx + i;
        const int64_t start = cnt;
        int64_t dpos = -1;
        int noff = 0;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              const int man = std::abs(dz) + std::abs(dy) + std::abs(dx);
              if (npts == 7 && man > 1) continue;
              if (k + dz < 0 || k + dz >= nz || j + dy < 0 || j + dy >= ny || i + dx < 0 ||
                  i + dx >= nx)
                continue;
              col[cnt] = (int32_t)(((int64_t)(k + dz) * ny + (j + dy)) * nx + (i + dx));
              if (man == 0) {
                dpos = cnt;
              } else {
                val[cnt] = -1.0;
                ++noff;
              }
              ++cnt;
            }
        val[dpos] = (1.0 + margin) * noff;
        (void)start;
        rptr[row + 1] = (int32_t)cnt;
      }
  return F3R_OK;
}

int f3r_gen_uniform(int64_t n, double lo, double hi, uint64_t seed, double *out) {
  if (n < 0 || !out) return F3R_ERR_ARG;
  uint64_t s = seed;
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * unit53(s);
  return F3R_OK;
}

int f3r_csr_scaling(int32_t n, const int32_t *rptr, const int32_t *col, double *val) {
  if (n < 0 || !rptr || !col || !val) return F3R_ERR_ARG;
  csr_scaling(n, rptr, col, val);
  return F3R_OK;
}

int f3r_sell_nslots(int32_t n, const int32_t *rptr, int32_t C, int32_t sigma, int32_t K,
                    int64_t *nslots) {
  if (n < 0 || !rptr || !nslots || C < 1 || sigma < 1 || K < 1) return F3R_ERR_ARG;
  *nslots = sell_nslots(n, rptr, C, sigma, K);
  return F3R_OK;
}

}  // extern "C"
