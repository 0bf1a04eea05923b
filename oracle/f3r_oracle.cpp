// f3r_oracle.cpp -- CPU oracle (test infrastructure only; see f3r_oracle.h).
//
// Restates the reference's hot-path arithmetic in plain C++17 so that the HIP
// product can be checked against it.  Every function cites the reference
// file:line it follows.  Built with -ffp-contract=off: a multiply-add is fused
// only where `madd` is called in FMA mode, so the oracle can match either the
// reference built without contraction (fma off) or this build's GPU kernels,
// which use explicit fused multiply-adds (fma on).
#include "f3r_oracle.h"

#include <omp.h>
#include <immintrin.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <limits>
#include <memory>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------- fp16 ----
struct h16 {
  uint16_t b;
};

inline float h2f(h16 h) { return _cvtsh_ss(h.b); }
inline h16 f2h(float f) { return h16{(uint16_t)_cvtss_sh(f, _MM_FROUND_TO_NEAREST_INT)}; }

// Correctly rounded fp64 -> fp16 (round to nearest even), no detour via fp32.
uint16_t d2h_bits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
  uint64_t a = u & 0x7fffffffffffffffULL;
  if (a >= 0x7ff0000000000000ULL)
    return sign | 0x7c00u | (a > 0x7ff0000000000000ULL ? 0x200u : 0u);
  int bexp = (int)(a >> 52);
  if (bexp == 0) return sign;  // fp64 subnormal: far below fp16 range
  int e = bexp - 1023;
  if (e > 15) return sign | 0x7c00u;
  uint64_t mant = (a & 0xfffffffffffffULL) | (1ULL << 52);
  int shift;
  uint32_t he;
  if (e >= -14) {
    shift = 42;
    he = (uint32_t)(e + 15);
  } else {
    shift = 42 + (-14 - e);
    he = 0;
    if (shift > 63) return sign;
  }
  uint64_t q = mant >> shift;
  uint64_t rem = mant & ((1ULL << shift) - 1);
  uint64_t halfway = 1ULL << (shift - 1);
  if (rem > halfway || (rem == halfway && (q & 1))) q++;
  if (he) {
    if (q == 2048) {
      q = 1024;
      he++;
    }
    if (he >= 31) return sign | 0x7c00u;
    return (uint16_t)(sign | (he << 10) | (q & 0x3ffu));
  }
  return (uint16_t)(sign | q);  // q == 1024 is the smallest normal
}
inline h16 d2h(double d) { return h16{d2h_bits(d)}; }

// ------------------------------------------------------------ bfloat16 ----
// 1 sign, 8 exponent, 7 stored mantissa bits: the top half of an fp32.
struct b16 {
  uint16_t b;
};
inline float b2f(b16 h) {
  uint32_t u = (uint32_t)h.b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// fp32 -> bf16, round to nearest even: the value lies between the two bf16
// numbers whose bits are the top 16 of u and that plus one; the discarded
// low half decides, ties to the even one
inline b16 f2b(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return b16{(uint16_t)((u >> 16) | 0x40u)};
  uint32_t hi = u >> 16, lo = u & 0xffffu;
  if (lo > 0x8000u || (lo == 0x8000u && (hi & 1u))) ++hi;  // carries into the exponent as it should
  return b16{(uint16_t)hi};
}
// fp64 -> bf16, correctly rounded: from the fp64 value's position between
// consecutive bf16 numbers (exponent e >= -126: spacing 2^(e-7); below:
// subnormal spacing 2^-133)
uint16_t d2b_bits(double d) {
  if (std::isnan(d)) return std::signbit(d) ? 0xffc0u : 0x7fc0u;
  const uint16_t sign = std::signbit(d) ? 0x8000u : 0u;
  double a = std::fabs(d);
  if (a == 0.0) return sign;
  int e;
  std::frexp(a, &e);  // a = f * 2^e, f in [0.5, 1): unbiased exponent e - 1
  const int ue = e - 1;
  const int q = ue >= -126 ? ue - 7 : -133;  // log2 of the spacing at a
  const double scaled = std::ldexp(a, -q);    // exact: a power-of-two scaling
  double k = std::floor(scaled);
  const double r = scaled - k;                // exact
  if (r > 0.5 || (r == 0.5 && std::fmod(k, 2.0) != 0.0)) k += 1.0;
  const double v = std::ldexp(k, q);
  if (v >= 0x1p128) return sign | 0x7f80u;   // beyond the largest finite bf16
  float fv = (float)v;                        // exact: v has at most 8 significant bits
  uint32_t u;
  std::memcpy(&u, &fv, 4);
  return (uint16_t)(sign | (u >> 16));
}
inline b16 d2b(double d) { return b16{d2b_bits(d)}; }

// ------------------------------------------------------------ traits ------
// load: value as it enters arithmetic; cvt<T>: rounding to storage type T.
inline double ld(double x) { return x; }
inline float ld(float x) { return x; }
inline float ld(h16 x) { return h2f(x); }
inline float ld(b16 x) { return b2f(x); }

template <class T> struct cv;
template <> struct cv<double> {
  static double from(double x) { return x; }
  static double from(float x) { return x; }
};
template <> struct cv<float> {
  static float from(double x) { return (float)x; }
  static float from(float x) { return x; }
};
template <> struct cv<h16> {
  static h16 from(double x) { return d2h(x); }
  static h16 from(float x) { return f2h(x); }
};
template <> struct cv<b16> {
  static b16 from(double x) { return d2b(x); }
  static b16 from(float x) { return f2b(x); }
};
template <class T, class S> inline T cvt(S x) { return cv<T>::from(x); }

bool g_fma = true;  // fused multiply-adds where the GPU build fuses them
inline double madd(double a, double b, double c) { return g_fma ? std::fma(a, b, c) : a * b + c; }
inline float madd(float a, float b, float c) { return g_fma ? std::fmaf(a, b, c) : a * b + c; }

// Deterministic reduction: fixed blocks of 4096 summed in order, then the
// partials in order; a single block (sequential order) with one thread.
template <class TA, class F> TA reduce_sum(int64_t n, F f) {
  int nt = omp_get_max_threads();
  if (nt == 1) {
    TA s = 0;
    for (int64_t i = 0; i < n; ++i) s += f(i);
    return s;
  }
  const int64_t B = 4096;
  int64_t nb = (n + B - 1) / B;
  std::vector<TA> part(nb);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    TA s = 0;
    int64_t e = std::min(n, (b + 1) * B);
    for (int64_t i = b * B; i < e; ++i) s += f(i);
    part[b] = s;
  }
  TA s = 0;
  for (int64_t b = 0; b < nb; ++b) s += part[b];
  return s;
}

// ------------------------------------------------------------- CSR --------
struct CSRd {
  int32_t n = 0;
  std::vector<int32_t> rptr, idx;
  std::vector<double> val;
  int64_t nnz() const { return (int64_t)idx.size(); }
};

template <class V> struct CSRv {  // values rounded to V, pattern shared
  int32_t n = 0;
  const int32_t *rptr = nullptr, *idx = nullptr;
  std::vector<V> val;
  CSRv() = default;
  explicit CSRv(const CSRd &A) : n(A.n), rptr(A.rptr.data()), idx(A.idx.data()), val(A.val.size()) {
    for (size_t k = 0; k < A.val.size(); ++k) val[k] = cvt<V>(A.val[k]);
  }
};

CSRd make_csr(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val) {
  CSRd A;
  A.n = n;
  A.rptr.assign(rptr, rptr + n + 1);
  int64_t nnz = rptr[n];
  A.idx.assign(idx, idx + nnz);
  A.val.assign(val, val + nnz);
  return A;
}

// csr_to_csc: counting sort by column (sparse/preconditioner/ainv.hpp:331-357)
CSRd csr_to_csc(const CSRd &A) {
  int32_t N = A.n;
  CSRd R;
  R.n = N;
  R.rptr.assign(N + 1, 0);
  R.idx.resize(A.nnz());
  R.val.resize(A.nnz());
  std::vector<int32_t> num(N, 0);
  for (int32_t i = 0; i < N; i++)
    for (int32_t j = A.rptr[i]; j < A.rptr[i + 1]; j++) num[A.idx[j]]++;
  for (int32_t i = 0; i < N; i++) {
    R.rptr[i + 1] = R.rptr[i] + num[i];
    num[i] = 0;
  }
  for (int32_t i = 0; i < N; i++)
    for (int32_t j = A.rptr[i]; j < A.rptr[i + 1]; j++) {
      int32_t c = A.idx[j];
      int32_t p = R.rptr[c] + num[c]++;
      R.idx[p] = i;
      R.val[p] = A.val[j];
    }
  return R;
}

// ------------------------------------------------------------ SpMV --------
// Per-row accumulation in slot order (sparse/matrix/sell.hpp:63-69; the SELL
// slots of a row hold the CSR entries in order, padding adds +0).
template <class TA, class V, class X, class Y>
void spmv_t(const CSRv<V> &A, const X *x, Y *y) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < A.n; ++i) {
    TA t = 0;
    for (int32_t j = A.rptr[i]; j < A.rptr[i + 1]; ++j)
      t = madd((TA)ld(A.val[j]), (TA)ld(x[A.idx[j]]), t);
    y[i] = cvt<Y>(t);
  }
}
template <class TA, class V, class B, class X, class Y>
void resid_t(const CSRv<V> &A, const B *b, const X *x, Y *y) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < A.n; ++i) {
    TA t = 0;
    for (int32_t j = A.rptr[i]; j < A.rptr[i + 1]; ++j)
      t = madd((TA)ld(A.val[j]), (TA)ld(x[A.idx[j]]), t);
    y[i] = cvt<Y>((TA)ld(b[i]) - t);
  }
}
// fp16 accumulation of the reference's __half build: one fp16 rounding per
// statement `t += val * x` (products of two halves are exact in fp32).
void spmv_h16acc(const CSRv<h16> &A, const h16 *x, h16 *y, const h16 *b) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < A.n; ++i) {
    h16 t = f2h(0.f);
    for (int32_t j = A.rptr[i]; j < A.rptr[i + 1]; ++j)
      t = f2h(h2f(t) + h2f(A.val[j]) * h2f(x[A.idx[j]]));
    y[i] = b ? f2h(h2f(b[i]) - h2f(t)) : t;
  }
}

// ------------------------------------------------------------ SDAINV ------
// SpVec: sorted (value, index) list (sparse/spvec.hpp:12-92)
struct SpVec {
  std::vector<std::pair<double, int32_t>> e;
  void append(std::pair<double, int32_t> x) {
    if (x.first != 0) e.push_back(x);
  }
  double dot(const double *v, const int32_t *ix, int32_t len) const {
    double r = 0;
    size_t p = 0;
    int32_t q = 0;
    while (p < e.size() && q < len) {
      int32_t a = e[p].second, b = ix[q];
      if (a < b)
        p++;
      else if (a == b) {
        r += e[p].first * v[q];
        p++;
        q++;
      } else
        q++;
    }
    return r;
  }
  // this = y + alpha * x with absolute drop |entry| >= tol (spvec.hpp:66-92)
  void set_axpy(double alpha, const SpVec &x, const SpVec &y, double tol) {
    size_t xp = 0, yp = 0;
    e.clear();
    while (xp < x.e.size() || yp < y.e.size()) {
      int64_t xt = xp < x.e.size() ? x.e[xp].second : INT_MAX;
      int64_t yt = yp < y.e.size() ? y.e[yp].second : INT_MAX;
      std::pair<double, int32_t> el;
      if (xt < yt) {
        el = x.e[xp];
        el.first *= alpha;
        xp++;
      } else if (xt == yt) {
        el = y.e[yp];
        el.first += alpha * x.e[xp].first;
        xp++;
        yp++;
      } else {
        el = y.e[yp];
        yp++;
      }
      if (std::abs(el.first) >= tol) append(el);
    }
  }
};

struct Factors {
  CSRd f[2];
};

// left_sdainv + sdainv (sparse/preconditioner/ainv.hpp:27-65, 268-329)
Factors sdainv(const CSRd &A, double tol, double acc) {
  CSRd At = csr_to_csc(A);
  int32_t N = A.n;
  std::vector<double> d(N);
  std::vector<SpVec> z(N), w(N);
  for (int32_t i = 0; i < N; i++) {
    z[i].append({1.0, i});
    w[i].append({1.0, i});
  }
  SpVec temp;
  for (int32_t i = 0; i < N; i++) {
    int32_t a_len = A.rptr[i + 1] - A.rptr[i], off = A.rptr[i];
    int32_t at_len = At.rptr[i + 1] - At.rptr[i], c_off = At.rptr[i];
    int32_t z_len = 0, w_len = 0;
    while (w_len < a_len && A.idx[off + w_len] < i) w_len++;
    while (z_len < at_len && At.idx[c_off + z_len] < i) z_len++;
    for (int32_t j = 0; j < w_len; j++) {
      int32_t col = A.idx[off + j];
      double valu = z[col].dot(&A.val[off], &A.idx[off], a_len);
      double alpha = -valu / d[col];
      temp.set_axpy(alpha, w[col], w[i], tol);
      w[i].e = temp.e;
    }
    for (int32_t j = 0; j < z_len; j++) {
      int32_t col = At.idx[c_off + j];
      double valu = w[col].dot(&At.val[c_off], &At.idx[c_off], at_len);
      double alpha = -valu / d[col];
      temp.set_axpy(alpha, z[col], z[i], tol);
      z[i].e = temp.e;
    }
    if (!w[i].e.empty()) {
      size_t end = w[i].e.size() - 1;
      w[i].e[end].first *= acc;
      d[i] = w[i].dot(&At.val[c_off], &At.idx[c_off], at_len);
      w[i].e[end].first /= acc;
    } else {
      d[i] = 0;
    }
    if (d[i] == 0) d[i] = 1;
  }
  // convert_to_csc (ainv.hpp:359-375): row i of Zt / Wt holds z_i / w_i
  auto to_rows = [N](const std::vector<SpVec> &v) {
    CSRd R;
    R.n = N;
    R.rptr.assign(N + 1, 0);
    for (int32_t i = 0; i < N; i++) R.rptr[i + 1] = R.rptr[i] + (int32_t)v[i].e.size();
    R.idx.resize(R.rptr[N]);
    R.val.resize(R.rptr[N]);
    for (int32_t i = 0; i < N; i++)
      for (size_t j = 0; j < v[i].e.size(); j++) {
        R.val[R.rptr[i] + j] = v[i].e[j].first;
        R.idx[R.rptr[i] + j] = v[i].e[j].second;
      }
    return R;
  };
  CSRd Zt = to_rows(z), Wt = to_rows(w);
  for (int32_t i = 0; i < N; i++)
    for (int32_t j = Zt.rptr[i]; j < Zt.rptr[i + 1]; ++j) Zt.val[j] /= d[i];
  Factors F;
  F.f[0] = csr_to_csc(Zt);
  F.f[1] = std::move(Wt);
  return F;
}

// ------------------------------------------------------------ ILU(0) ------
// duplicate_block_diagonal (csr.hpp:109-155) -> block bounds + filtered CSR
std::vector<int32_t> block_bounds(int32_t n, int32_t block_num) {
  int32_t bs = (n + block_num - 1) / block_num;
  std::vector<int32_t> b(block_num + 1);
  b[0] = 0;
  for (int32_t i = 0; i < block_num; ++i) b[i + 1] = (i != block_num - 1) ? b[i] + bs : n;
  return b;
}
CSRd block_diagonal(const CSRd &A, const std::vector<int32_t> &bb) {
  CSRd R;
  R.n = A.n;
  R.rptr.assign(A.n + 1, 0);
  for (size_t bi = 0; bi + 1 < bb.size(); bi++) {
    int32_t l = bb[bi], r = bb[bi + 1];
    for (int32_t i = l; i < r; i++) {
      for (int32_t j = A.rptr[i]; j < A.rptr[i + 1]; ++j)
        if (l <= A.idx[j] && A.idx[j] < r) {
          R.idx.push_back(A.idx[j]);
          R.val.push_back(A.val[j]);
        }
      R.rptr[i + 1] = (int32_t)R.idx.size();
    }
  }
  return R;
}
// _ilu0 (ilu.hpp:128-164) then split (algorithm.hpp:8-43)
Factors ilu0(const CSRd &A, int32_t block_num, double acc) {
  auto bb = block_bounds(A.n, block_num);
  CSRd R = block_diagonal(A, bb);
  auto &val = R.val;
  auto &cind = R.idx;
  auto &rptr = R.rptr;
  if (!val.empty()) val[0] *= acc;
  for (int32_t id = 0; id < block_num; id++) {
    int32_t s = id == 0 ? 1 : bb[id];
    int32_t e = bb[id + 1];
    for (int32_t i = s; i < e; ++i) {
      int32_t d = rptr[i];
      while (cind[d] != i) ++d;
      val[d] *= acc;
      for (int32_t pvt = rptr[i]; pvt < rptr[i + 1] && cind[pvt] < i; ++pvt) {
        int32_t k = cind[pvt];
        int32_t ref = rptr[k];
        while (cind[ref] != cind[pvt]) ++ref;
        val[pvt] = val[pvt] / val[ref];
        int32_t base = pvt + 1;
        while (base < rptr[i + 1] && ref < rptr[k + 1]) {
          if (cind[ref] < cind[base])
            ++ref;
          else if (cind[ref] > cind[base])
            ++base;
          else
            val[base++] -= val[pvt] * val[ref++];
        }
      }
    }
  }
  Factors F;
  CSRd &L = F.f[0], &U = F.f[1];
  L.n = U.n = R.n;
  L.rptr.assign(R.n + 1, 0);
  U.rptr.assign(R.n + 1, 0);
  for (int32_t i = 0; i < R.n; i++) {
    for (int32_t j = rptr[i]; j < rptr[i + 1]; ++j) {
      if (cind[j] <= i) {
        L.val.push_back(cind[j] < i ? val[j] : 1.0);
        L.idx.push_back(cind[j]);
      }
      if (cind[j] >= i) {
        U.val.push_back(val[j]);
        U.idx.push_back(cind[j]);
      }
    }
    L.rptr[i + 1] = (int32_t)L.idx.size();
    U.rptr[i + 1] = (int32_t)U.idx.size();
  }
  return F;
}

// BJ triangular solves (trsv/bj.hpp:58-98).  L: unit diagonal stored last.
// U: d_upper rotated by 180 degrees with the inverted diagonal
// (csr.hpp:60-87): row r is processed from its largest column down and the
// sum is multiplied by 1/u_rr (inverted in fp64 before rounding to storage).
template <class S> struct BJ {
  int32_t n;
  std::vector<int32_t> lrptr, lidx, urptr, uidx;
  std::vector<S> lval, uval;  // uval: off-diagonals; uinv: inverted diagonal
  std::vector<S> uinv;
  std::vector<int32_t> bb;
  BJ(const Factors &F, int32_t block_num) {
    const CSRd &L = F.f[0], &U = F.f[1];
    n = L.n;
    lrptr = L.rptr;
    lidx = L.idx;
    lval.resize(L.val.size());
    for (size_t k = 0; k < L.val.size(); ++k) lval[k] = cvt<S>(L.val[k]);
    urptr = U.rptr;
    uidx = U.idx;
    uval.resize(U.val.size());
    for (size_t k = 0; k < U.val.size(); ++k) uval[k] = cvt<S>(U.val[k]);
    uinv.resize(n);
    for (int32_t r = 0; r < n; ++r) uinv[r] = cvt<S>(1. / U.val[U.rptr[r]]);
    bb = block_bounds(n, block_num);
  }
  // Arithmetic: AT = accumulator; `stmt_round` rounds every statement to S
  // (reference _Float16 semantics) instead of accumulating in AT.
  template <class AT, bool stmt_round> void solve(const S *in, S *out, S *tmp) const {
    int32_t nb = (int32_t)bb.size() - 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t b = 0; b < nb; ++b)
      for (int32_t i = bb[b]; i < bb[b + 1]; ++i) {
        if (stmt_round) {
          S t = in[i];
          for (int32_t j = lrptr[i]; j < lrptr[i + 1] - 1; ++j)
            t = cvt<S>((AT)ld(t) - (AT)ld(lval[j]) * (AT)ld(tmp[lidx[j]]));
          tmp[i] = t;
        } else {
          AT t = (AT)ld(in[i]);
          for (int32_t j = lrptr[i]; j < lrptr[i + 1] - 1; ++j)
            t = madd(-(AT)ld(lval[j]), (AT)ld(tmp[lidx[j]]), t);
          tmp[i] = cvt<S>(t);
        }
      }
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t b = 0; b < nb; ++b)
      for (int32_t r = bb[b + 1] - 1; r >= bb[b]; --r) {
        if (stmt_round) {
          S t = tmp[r];
          for (int32_t q = urptr[r + 1] - 1; q > urptr[r]; --q)
            t = cvt<S>((AT)ld(t) - (AT)ld(uval[q]) * (AT)ld(out[uidx[q]]));
          out[r] = cvt<S>((AT)ld(t) * (AT)ld(uinv[r]));
        } else {
          AT t = (AT)ld(tmp[r]);
          for (int32_t q = urptr[r + 1] - 1; q > urptr[r]; --q)
            t = madd(-(AT)ld(uval[q]), (AT)ld(out[uidx[q]]), t);
          out[r] = cvt<S>(t * (AT)ld(uinv[r]));
        }
      }
  }
};

// ------------------------------------------------------------ FGMRES ------
struct SolveRes {
  bool is_solved;
  int32_t res_iter;
  double res_nrm2;
  void operator+=(const SolveRes &o) {  // sparse/models.hpp:195-199
    is_solved = is_solved || o.is_solved;
    res_iter += o.res_iter;
    res_nrm2 = o.res_nrm2;
  }
};
using Cond = std::function<bool(int, double)>;

template <class T> struct OpT {
  virtual ~OpT() = default;
  virtual void apply(const T *in, T *out) = 0;
};

// Blas1::full_grot / final_grot / trsv (sparse/solver/gmres.hpp:12-53)
template <class T> void full_grot(T *c, T *s, T *H, int k) {
  for (int i = 0; i < k; i++) {
    T t = H[i];
    T h1 = H[i + 1];
    H[i] = madd(c[i], t, s[i] * h1);
    H[i + 1] = madd(-s[i], t, c[i] * h1);
  }
  T t = std::sqrt(madd(H[k], H[k], H[k + 1] * H[k + 1]));
  c[k] = H[k] / t;
  s[k] = H[k + 1] / t;
  H[k] = t;
  H[k + 1] = 0;
}
template <class T> void final_grot(const T *c, const T *s, T *e, int k) {
  e[k + 1] = -s[k] * e[k];
  e[k] = c[k] * e[k];
}
template <class T> void hess_trsv(const T *U, const T *b, T *x, int h, int w) {
  for (int i = w - 1; i >= 0; i--) {
    T t = b[i];
    for (int j = w - 1; j > i; j--) t = madd(-U[j * h + i], x[j], t);
    x[i] = t / U[i * h + i];
  }
}

// FGMRES<T, L> with matrix values V (sparse/solver/gmres.hpp:56-135).
template <class T, class V> struct FGMRES {
  const CSRv<V> *A;
  OpT<T> *M;
  int m;
  bool normalize;
  int64_t n;
  std::vector<T> Vb, Zb, H, c, s, y, e;
  FGMRES(const CSRv<V> *A_, OpT<T> *M_, int m_, bool norm)
      : A(A_), M(M_), m(m_), normalize(norm), n(A_->n), Vb((size_t)n * (m_ + 1)),
        Zb((size_t)n * m_), H((size_t)(m_ + 1) * m_), c(m_), s(m_), y(m_), e(m_ + 1) {}
  T *Vc(int j) { return Vb.data() + (size_t)n * j; }
  T *Zc(int j) { return Zb.data() + (size_t)n * j; }

  template <bool with_init, class TI, class TO>
  SolveRes kernel(const TI *rhs, TO *x, const Cond &cond) {
    bool flag = false;
    T h;
    if (normalize) {
      h = std::sqrt(reduce_sum<T>(n, [&](int64_t i) { T v = (T)ld(rhs[i]); return v * v; }));
    } else {
      h = 1;
    }
    e[0] = h;
    T hn = h;  // h_nrm_r
    if (normalize) {
      T inv = (T)(1. / h);
      T *v0 = Vc(0);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) v0[i] = inv * (T)ld(rhs[i]);
    } else {
      T *v0 = Vc(0);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) v0[i] = (T)ld(rhs[i]);
    }
    if (cond && cond(0, std::abs((double)hn))) {
      if (with_init)
        for (int64_t i = 0; i < n; ++i) x[i] = cvt<TO>((T)0);
      return SolveRes{true, 0, std::abs((double)hn)};
    }
    int j = 0;
    for (; j < m; j++) {
      M->apply(Vc(j), Zc(j));
      spmv_t<T>(*A, Zc(j), Vc(j + 1));
      T *w = Vc(j + 1);
      T *Hj = H.data() + (size_t)j * (m + 1);
      for (int k = 0; k <= j; k++) {
        const T *vk = Vc(k);
        Hj[k] = reduce_sum<T>(n, [&](int64_t i) { return w[i] * vk[i]; });
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        T r = Vb[i] * Hj[0];
        for (int k = 1; k <= j; k++) r = madd(Vb[(size_t)n * k + i], Hj[k], r);
        w[i] = w[i] - r;
      }
      T nv = std::sqrt(reduce_sum<T>(n, [&](int64_t i) { return w[i] * w[i]; }));
      Hj[j + 1] = nv;
      T inv = (T)(1. / nv);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) w[i] *= inv;
      full_grot(c.data(), s.data(), Hj, j);
      final_grot(c.data(), s.data(), e.data(), j);
      hn = e[j + 1];
      if (cond && cond(j + 1, std::abs((double)hn))) {
        j++;
        flag = true;
        break;
      }
    }
    hess_trsv(H.data(), e.data(), y.data(), m + 1, j);
    if (j > 0) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        T r = Zb[i] * y[0];
        for (int k = 1; k < j; k++) r = madd(Zb[(size_t)n * k + i], y[k], r);
        if (with_init)
          x[i] = cvt<TO>(r);
        else
          x[i] = cvt<TO>((T)ld(x[i]) + r);
      }
    } else if (with_init) {
      for (int64_t i = 0; i < n; ++i) x[i] = cvt<TO>((T)0);
    }
    return flag ? SolveRes{true, j, std::abs((double)hn)} : SolveRes{false, m, std::abs((double)hn)};
  }
  template <class TI, class TO> void operate(const TI *in, TO *out) { kernel<true>(in, out, nullptr); }
};

// Adapter: an FGMRES level of basis T2 called with T1 vectors.
template <class T1, class T2, class V> struct LevelOp : OpT<T1> {
  FGMRES<T2, V> *F;
  explicit LevelOp(FGMRES<T2, V> *f) : F(f) {}
  void apply(const T1 *in, T1 *out) override { F->operate(in, out); }
};

// ---------------------------------------------------- adaptive Richardson --
// Shared state of the adaptive weights (work/src/f3r-gpu.cpp:106-107,224-227)
// With a static weight w > 0 there are no probes and every step uses w
// (work/src/static.cpp:99-119).
struct RichState {
  int c = 64;
  int cnt = 64;
  int den = 0;
  bool fixed = false;
  std::vector<double> sum, ome;
  void reset(int m4, int c_, double w = 0.0) {
    c = c_;
    cnt = c_;
    den = 0;
    fixed = w > 0;
    sum.assign(m4, 1.0);
    ome.assign(m4, fixed ? w : 1.0);
  }
  bool probe() const { return !fixed && cnt % c == 0; }
};

// Preconditioner of the innermost level in storage type S.
template <class S> struct Prec {
  int kind = 0;  // 0 SDAINV, 1 BJ-ILU(0)
  CSRd Wd, Zd;   // factor patterns, owned here (CSRv points into them)
  CSRv<S> W, Z;
  std::unique_ptr<BJ<S>> bj;
  std::vector<S> tmp;
  // AT: accumulator precision (fp16 storage uses AT=float), ref16: fp16 sums
  template <class AT, bool ref16> void apply(const S *in, S *out) {
    if (kind == 0) {
      if constexpr (ref16) {
        spmv_h16acc(reinterpret_cast<const CSRv<h16> &>(W), reinterpret_cast<const h16 *>(in),
                    reinterpret_cast<h16 *>(tmp.data()), nullptr);
        spmv_h16acc(reinterpret_cast<const CSRv<h16> &>(Z), reinterpret_cast<const h16 *>(tmp.data()),
                    reinterpret_cast<h16 *>(out), nullptr);
      } else {
        spmv_t<AT>(W, in, tmp.data());
        spmv_t<AT>(Z, tmp.data(), out);
      }
    } else {
      bj->template solve<AT, ref16>(in, out, tmp.data());
    }
  }
};

// Richardson level of the DOUBLE / FLOAT builds (work/src/f3r-gpu.cpp:123-152,
// 174-203; work/src/f3r.cpp:110-139,159-188): vectors and weights in T, the
// preconditioner in T, and the SpMVs with the fp64 matrix A accumulated in
// fp64 (decltype(double * T)), also in the FLOAT build.
template <class T> struct RichT : OpT<T> {
  const CSRv<double> *A;
  Prec<T> *P;
  RichState *st;
  int m4;
  std::vector<T> tmp, tmp2, hr;
  int64_t calls = 0;
  RichT(const CSRv<double> *A_, Prec<T> *P_, RichState *s, int m4_)
      : A(A_), P(P_), st(s), m4(m4_), tmp(A_->n), tmp2(A_->n), hr(A_->n) {}
  void apply(const T *in, T *out) override {
    int64_t n = A->n;
    ++calls;
    ++st->cnt;
    bool probe = st->probe();
    P->template apply<T, false>(in, out);
    if (probe) {
      spmv_t<double>(*A, out, tmp.data());
      T dar = reduce_sum<T>(n, [&](int64_t i) { return tmp[i] * tmp[i]; });
      T dr = reduce_sum<T>(n, [&](int64_t i) { return tmp[i] * in[i]; });
      T al = dr / dar;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) out[i] *= al;
      st->sum[0] = (T)(st->sum[0] + al);
      st->den = st->cnt / st->c;
      st->ome[0] = (T)(st->sum[0] / st->den);
    } else {
      T om = (T)st->ome[0];
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) out[i] *= om;
    }
    for (int k = 1; k < m4; k++) {
      resid_t<double>(*A, in, out, hr.data());
      P->template apply<T, false>(hr.data(), tmp2.data());
      if (probe) {
        spmv_t<double>(*A, tmp2.data(), tmp.data());
        T dar = reduce_sum<T>(n, [&](int64_t i) { return tmp[i] * tmp[i]; });
        T dr = reduce_sum<T>(n, [&](int64_t i) { return tmp[i] * hr[i]; });
        T al = dr / dar;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) out[i] = madd(al, tmp2[i], out[i]);
        st->sum[k] = (T)(st->sum[k] + al);
        st->ome[k] = (T)(st->sum[k] / st->den);
      } else {
        T om = (T)st->ome[k];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) out[i] = madd(om, tmp2[i], out[i]);
      }
    }
  }
};

// Richardson level of the HALF build (work/src/f3r-gpu.cpp:210-260):
// vectors stored in fp16, input/output fp32.
//  ref16 = true : the reference's arithmetic: fp16 SpMV sums, one fp16
//                 rounding per statement, weights stored in fp16.
//  ref16 = false: this build: fp32 sums of fp16 data, fp32 weights, and a
//                 power-of-two range factor on the input.
template <bool ref16> struct RichH : OpT<float> {
  const CSRv<h16> *A;
  Prec<h16> *P;
  RichState *st;
  int m4;
  float scale;
  std::vector<h16> hin, hout, tmp, tmp2, hr;
  int64_t calls = 0;
  RichH(const CSRv<h16> *A_, Prec<h16> *P_, RichState *s, int m4_, float scale_)
      : A(A_), P(P_), st(s), m4(m4_), scale(scale_), hin(A_->n), hout(A_->n), tmp(A_->n),
        tmp2(A_->n), hr(A_->n) {}
  void spmv16(const h16 *x, h16 *y) {
    if (ref16)
      spmv_h16acc(*A, x, y, nullptr);
    else
      spmv_t<float>(*A, x, y);
  }
  void apply(const float *in, float *out) override {
    int64_t n = A->n;
    ++calls;
    ++st->cnt;
    bool probe = st->probe();
    const float sc = ref16 ? 1.f : scale;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) hin[i] = f2h(in[i] * sc);
    P->template apply<float, ref16>(hin.data(), hout.data());
    if (probe) {
      spmv16(hout.data(), tmp.data());
      float dar = reduce_sum<float>(n, [&](int64_t i) { float t = h2f(tmp[i]); return t * t; });
      float dr = reduce_sum<float>(n, [&](int64_t i) { return h2f(tmp[i]) * h2f(hin[i]); });
      float al = dr / dar;
      st->den = st->cnt / st->c;
      if (ref16) {
        float alh = h2f(f2h(al));
        for (int64_t i = 0; i < n; ++i) hout[i] = f2h(h2f(hout[i]) * alh);
        st->sum[0] = (float)st->sum[0] + al;
        st->ome[0] = h2f(f2h((float)st->sum[0] / (float)st->den));
      } else {
        for (int64_t i = 0; i < n; ++i) hout[i] = f2h(h2f(hout[i]) * al);
        st->sum[0] = (float)st->sum[0] + al;
        st->ome[0] = (float)st->sum[0] / (float)st->den;
      }
    } else {
      float om = (float)st->ome[0];
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) hout[i] = f2h(h2f(hout[i]) * om);
    }
    for (int k = 1; k < m4; k++) {
      if (ref16)
        spmv_h16acc(*A, hout.data(), hr.data(), hin.data());
      else
        resid_t<float>(*A, hin.data(), hout.data(), hr.data());
      P->template apply<float, ref16>(hr.data(), tmp2.data());
      if (probe) {
        spmv16(tmp2.data(), tmp.data());
        float dar = reduce_sum<float>(n, [&](int64_t i) { float t = h2f(tmp[i]); return t * t; });
        float dr = reduce_sum<float>(n, [&](int64_t i) { return h2f(tmp[i]) * h2f(hr[i]); });
        float al = dr / dar;
        if (ref16) {
          for (int64_t i = 0; i < n; ++i)
            hout[i] = f2h(h2f(hout[i]) + h2f(f2h(al * h2f(tmp2[i]))));
          st->sum[k] = (float)st->sum[k] + al;
          st->ome[k] = h2f(f2h((float)st->sum[k] / (float)st->den));
        } else {
          for (int64_t i = 0; i < n; ++i) hout[i] = f2h(std::fmaf(al, h2f(tmp2[i]), h2f(hout[i])));
          st->sum[k] = (float)st->sum[k] + al;
          st->ome[k] = (float)st->sum[k] / (float)st->den;
        }
      } else {
        float om = (float)st->ome[k];
        if (ref16) {
#pragma omp parallel for schedule(static)
          for (int64_t i = 0; i < n; ++i)
            hout[i] = f2h(h2f(hout[i]) + h2f(f2h(om * h2f(tmp2[i]))));
        } else {
#pragma omp parallel for schedule(static)
          for (int64_t i = 0; i < n; ++i) hout[i] = f2h(std::fmaf(om, h2f(tmp2[i]), h2f(hout[i])));
        }
      }
    }
    const float inv = 1.f / sc;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = h2f(hout[i]) * inv;
  }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class S> void setup_prec(Prec<S> &P, const CSRd &A, const orc_f3r_opts &o) {
  P.kind = o.precond;
  P.tmp.assign(A.n, S{});
  if (o.precond == 0) {
    Factors F = sdainv(A, o.tol, o.acc);
    P.Wd = std::move(F.f[1]);
    P.Zd = std::move(F.f[0]);
    P.W = CSRv<S>(P.Wd);
    P.Z = CSRv<S>(P.Zd);
  } else {
    Factors F = ilu0(A, o.bj_blocks, o.acc);
    P.bj.reset(new BJ<S>(F, o.bj_blocks));
  }
}

// The nested solve of one build variant.  Restarted<double>(FGMRES<double>
// (A, L2, {m1, true}), {restart}) as composed in work/src/f3r-gpu.cpp:261-265.
struct F3R {
  virtual ~F3R() = default;
  virtual SolveRes solve(const double *b, double *x, const Cond &cond) = 0;
  virtual void reset() = 0;
  virtual int64_t calls() const = 0;
  virtual void level_apply(int level, const double *in, double *out) = 0;
};

template <class T2, class V3, class S, bool ref16> struct F3RImpl : F3R {
  // level matrices
  CSRv<double> A64;
  CSRv<T2> A2;
  CSRv<V3> A3;
  CSRv<S> A4;
  Prec<S> P;
  RichState st;
  std::unique_ptr<OpT<T2>> R4;
  std::unique_ptr<FGMRES<T2, V3>> L3;
  std::unique_ptr<LevelOp<T2, T2, V3>> L3op;
  std::unique_ptr<FGMRES<T2, T2>> L2;
  std::unique_ptr<LevelOp<double, T2, T2>> L2op;
  std::unique_ptr<FGMRES<double, double>> L1;
  orc_f3r_opts o;
  std::vector<double> r;

  F3RImpl(const CSRd &A, const orc_f3r_opts &opts)
      : A64(A), A2(A), A3(A), A4(A), o(opts), r(A.n) {
    setup_prec(P, A, o);
    st.reset(o.m4, o.c);
    if constexpr (std::is_same_v<S, h16>)
      R4.reset(new RichH<ref16>(&A4, &P, &st, o.m4, (float)o.range_scale));
    else
      R4.reset(new RichT<S>(&A64, &P, &st, o.m4));
    L3.reset(new FGMRES<T2, V3>(&A3, R4.get(), o.m3, false));
    L3op.reset(new LevelOp<T2, T2, V3>(L3.get()));
    L2.reset(new FGMRES<T2, T2>(&A2, L3op.get(), o.m2, false));
    L2op.reset(new LevelOp<double, T2, T2>(L2.get()));
    L1.reset(new FGMRES<double, double>(&A64, L2op.get(), o.m1, true));
  }
  void reset() override {
    // the reference keeps the static weight in fp16 (static.cpp:84-85)
    double w = o.static_weight;
    if (w > 0 && ref16) w = h2f(f2h((float)w));
    st.reset(o.m4, o.c, w);
  }
  int64_t calls() const override {
    if constexpr (std::is_same_v<S, h16>)
      return static_cast<RichH<ref16> *>(R4.get())->calls;
    else
      return static_cast<RichT<S> *>(R4.get())->calls;
  }
  SolveRes solve(const double *b, double *x, const Cond &cond) override {
    // FGMRES::solve: act = b - A x, then the normalized kernel (gmres.hpp:84-89)
    auto one = [&]() {
      resid_t<double>(A64, b, x, r.data());
      return L1->template kernel<false>(r.data(), x, cond);
    };
    SolveRes flag = one();
    int cnt = 1;
    while (!flag.is_solved && cnt < o.restart) {
      flag += one();
      cnt++;
    }
    return flag;
  }
  void level_apply(int level, const double *in, double *out) override {
    int64_t n = A64.n;
    if (level == 2) {
      L2->operate(in, out);
    } else {
      std::vector<T2> a(n), b2(n);
      for (int64_t i = 0; i < n; ++i) a[i] = (T2)in[i];
      if (level == 3)
        L3->operate(a.data(), b2.data());
      else
        R4->apply(a.data(), b2.data());
      for (int64_t i = 0; i < n; ++i) out[i] = (double)b2[i];
    }
  }
};

// ------------------------------------------------------ depth ablations --
// The nests of work/src/depth.cpp:122-167: FGMRES levels (gmres2.hpp:56-150,
// normalize = false inside) down to the preconditioner applied directly.
// This build's arithmetic: an fp16 level stores V and Z in fp16 and computes
// in fp32 (sums, Hessenberg, Givens), with its vectors kept scaled by the
// power-of-two range factor s (stored = s * value; dots and norms take
// 1/s^2, the output 1/s).  The reference computes such a level in host
// _Float16 arithmetic.

// FGMRES level with fp16 storage (F2H, F3H, the innermost level of F4)
template <class V> struct FGMRESH {
  const CSRv<V> *A;
  OpT<h16> *M;
  int m;
  int64_t n;
  float s, inv_s, inv_s2;
  std::vector<h16> Vb, Zb;
  std::vector<float> H, c, sn, y, e;
  FGMRESH(const CSRv<V> *A_, OpT<h16> *M_, int m_, float s_)
      : A(A_), M(M_), m(m_), n(A_->n), s(s_), inv_s(1.f / s_), inv_s2(1.f / (s_ * s_)),
        Vb((size_t)n * (m_ + 1)), Zb((size_t)n * m_), H((size_t)(m_ + 1) * m_), c(m_), sn(m_), y(m_),
        e(m_ + 1) {}
  h16 *Vc(int j) { return Vb.data() + (size_t)n * j; }
  h16 *Zc(int j) { return Zb.data() + (size_t)n * j; }
  template <class TI, class TO> void operate(const TI *in, TO *out) {
    h16 *v0 = Vc(0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) v0[i] = cvt<h16>(in[i] * (TI)s);
    e[0] = 1;
    for (int j = 0; j < m; j++) {
      M->apply(Vc(j), Zc(j));
      spmv_t<float>(*A, Zc(j), Vc(j + 1));
      h16 *w = Vc(j + 1);
      float *Hj = H.data() + (size_t)j * (m + 1);
      for (int k = 0; k <= j; k++) {
        const h16 *vk = Vc(k);
        Hj[k] = reduce_sum<float>(n, [&](int64_t i) { return h2f(w[i]) * h2f(vk[i]); }) * inv_s2;
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        float r = h2f(Vb[i]) * Hj[0];
        for (int k = 1; k <= j; k++) r = madd(h2f(Vb[(size_t)n * k + i]), Hj[k], r);
        w[i] = f2h(h2f(w[i]) - r);
      }
      const float nv =
          std::sqrt(reduce_sum<float>(n, [&](int64_t i) { float t = h2f(w[i]); return t * t; }) * inv_s2);
      Hj[j + 1] = nv;
      const float inv = (float)(1. / nv);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) w[i] = f2h(h2f(w[i]) * inv);
      full_grot(c.data(), sn.data(), Hj, j);
      final_grot(c.data(), sn.data(), e.data(), j);
    }
    hess_trsv(H.data(), e.data(), y.data(), m + 1, m);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      float r = h2f(Zb[i]) * y[0];
      for (int k = 1; k < m; k++) r = madd(h2f(Zb[(size_t)n * k + i]), y[k], r);
      out[i] = cvt<TO>(r * inv_s);
    }
  }
};
template <class T1, class V> struct LevelOpH : OpT<T1> {
  FGMRESH<V> *F;
  explicit LevelOpH(FGMRESH<V> *f) : F(f) {}
  void apply(const T1 *in, T1 *out) override { F->operate(in, out); }
};

// M applied directly, from an fp32 level: fp16 copy scaled by s, fp16
// factors with fp32 sums, back to fp32
struct MDirectF : OpT<float> {
  Prec<h16> *P;
  int64_t n;
  float s;
  std::vector<h16> hin, hout;
  int64_t calls = 0;
  MDirectF(Prec<h16> *P_, int64_t n_, float s_) : P(P_), n(n_), s(s_), hin(n_), hout(n_) {}
  void apply(const float *in, float *out) override {
    ++calls;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) hin[i] = f2h(in[i] * s);
    P->template apply<float, false>(hin.data(), hout.data());
    const float inv = 1.f / s;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = h2f(hout[i]) * inv;
  }
};
// from an fp16 level: its stored (scaled) vectors go straight in
struct MDirectH : OpT<h16> {
  Prec<h16> *P;
  int64_t calls = 0;
  explicit MDirectH(Prec<h16> *P_) : P(P_) {}
  void apply(const h16 *in, h16 *out) override {
    ++calls;
    P->template apply<float, false>(in, out);
  }
};

struct DepthImpl : F3R {
  CSRv<double> A64;
  CSRv<float> A32;
  CSRv<h16> A16;
  Prec<h16> P;
  orc_f3r_opts o;
  std::unique_ptr<MDirectF> mf;
  std::unique_ptr<MDirectH> mh;
  std::vector<std::shared_ptr<void>> own;  // the levels and their adapters
  OpT<double> *top = nullptr;
  std::unique_ptr<FGMRES<double, double>> L1;
  std::vector<double> r;

  template <class X> X *keep(X *p) {
    own.emplace_back(std::shared_ptr<X>(p));
    return p;
  }
  DepthImpl(const CSRd &A, const orc_f3r_opts &opts) : A64(A), A32(A), A16(A), o(opts), r(A.n) {
    setup_prec(P, A, o);
    const float s = (float)o.range_scale;
    const int64_t n = A.n;
    switch (o.depth) {
      case 1: {  // F2: FGMRES<float>(A32, M, 64)
        mf.reset(new MDirectF(&P, n, s));
        auto *l2 = keep(new FGMRES<float, float>(&A32, mf.get(), 64, false));
        top = keep(new LevelOp<double, float, float>(l2));
        break;
      }
      case 2: {  // F2H: FGMRES<half>(A16, M, 64)
        mh.reset(new MDirectH(&P));
        auto *l2 = keep(new FGMRESH<h16>(&A16, mh.get(), 64, s));
        top = keep(new LevelOpH<double, h16>(l2));
        break;
      }
      case 3: {  // F3: FGMRES<float>(A32, FGMRES<float>(A16, M, 8), 8)
        mf.reset(new MDirectF(&P, n, s));
        auto *l3 = keep(new FGMRES<float, h16>(&A16, mf.get(), 8, false));
        auto *l3op = keep(new LevelOp<float, float, h16>(l3));
        auto *l2 = keep(new FGMRES<float, float>(&A32, l3op, 8, false));
        top = keep(new LevelOp<double, float, float>(l2));
        break;
      }
      case 4: {  // F3H: FGMRES<float>(A32, FGMRES<half>(A16, M, 8), 8)
        mh.reset(new MDirectH(&P));
        auto *l3 = keep(new FGMRESH<h16>(&A16, mh.get(), 8, s));
        auto *l3op = keep(new LevelOpH<float, h16>(l3));
        auto *l2 = keep(new FGMRES<float, float>(&A32, l3op, 8, false));
        top = keep(new LevelOp<double, float, float>(l2));
        break;
      }
      case 5: {  // F4: FGMRES<float>(A32, FGMRES<float>(A16, FGMRES<half>(A16, M, 2), 4), 8)
        mh.reset(new MDirectH(&P));
        auto *l4 = keep(new FGMRESH<h16>(&A16, mh.get(), 2, s));
        auto *l4op = keep(new LevelOpH<float, h16>(l4));
        auto *l3 = keep(new FGMRES<float, h16>(&A16, l4op, 4, false));
        auto *l3op = keep(new LevelOp<float, float, h16>(l3));
        auto *l2 = keep(new FGMRES<float, float>(&A32, l3op, 8, false));
        top = keep(new LevelOp<double, float, float>(l2));
        break;
      }
      default:
        throw std::runtime_error("depth must be 1..5");
    }
    L1.reset(new FGMRES<double, double>(&A64, top, o.m1, true));
  }
  void reset() override {}
  int64_t calls() const override { return mf ? mf->calls : mh->calls; }
  SolveRes solve(const double *b, double *x, const Cond &cond) override {
    auto one = [&]() {
      resid_t<double>(A64, b, x, r.data());
      return L1->template kernel<false>(r.data(), x, cond);
    };
    SolveRes flag = one();
    int cnt = 1;
    while (!flag.is_solved && cnt < o.restart) {
      flag += one();
      cnt++;
    }
    return flag;
  }
  void level_apply(int, const double *in, double *out) override { top->apply(in, out); }
};

std::unique_ptr<F3R> make_f3r(const CSRd &A, const orc_f3r_opts &o) {
  if (o.depth != 0) return std::unique_ptr<F3R>(new DepthImpl(A, o));
  if (o.variant == 2) return std::unique_ptr<F3R>(new F3RImpl<double, double, double, false>(A, o));
  if (o.variant == 1) return std::unique_ptr<F3R>(new F3RImpl<float, float, float, false>(A, o));
  if (o.ref_arith == 1) return std::unique_ptr<F3R>(new F3RImpl<float, h16, h16, true>(A, o));
  return std::unique_ptr<F3R>(new F3RImpl<float, h16, h16, false>(A, o));
}

struct ThreadScope {
  int saved;
  explicit ThreadScope(int n) : saved(omp_get_max_threads()) {
    if (n > 0) omp_set_num_threads(n);
  }
  ~ThreadScope() { omp_set_num_threads(saved); }
};

// One step of an MLP layer: y = relu?(W x + b), W row-major [out][in].
void dense(int rows, int nin, int nout, const double *x, const double *W, const double *b,
           double *y, bool relu) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r)
    for (int o = 0; o < nout; ++o) {
      double s = b[o];
      for (int i = 0; i < nin; ++i) s += W[(size_t)o * nin + i] * x[(size_t)r * nin + i];
      y[(size_t)r * nout + o] = relu ? (s > 0 ? s : 0) : s;
    }
}

}  // namespace

// ================================================================ C ABI ===
extern "C" {

uint16_t orc_f2h(float f) { return f2h(f).b; }
float orc_h2f(uint16_t h) { return h2f(h16{h}); }
uint16_t orc_d2h(double d) { return d2h_bits(d); }
uint16_t orc_d2b(double d) { return d2b_bits(d); }
uint16_t orc_f2b(float f) { return f2b(f).b; }
float orc_b2f(uint16_t b) { return b2f(b16{b}); }
void orc_d2h_array(int64_t n, const double *in, uint16_t *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = d2h_bits(in[i]);
}
void orc_f2h_array(int64_t n, const float *in, uint16_t *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = f2h(in[i]).b;
}
void orc_h2f_array(int64_t n, const uint16_t *in, float *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = h2f(h16{in[i]});
}

int32_t orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_fma(int32_t on) { g_fma = on != 0; }

// scaling (sparse/matrix/algorithm.hpp:45-59): a_ij /= sqrt|a_ii| * sqrt|a_jj|
void orc_scaling(int32_t n, const int32_t *rptr, const int32_t *idx, double *val) {
  std::vector<double> diag(n, 1.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = rptr[i]; j < rptr[i + 1]; ++j)
      if (idx[j] == i) diag[i] = std::sqrt(std::abs(val[j]));
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = rptr[i]; j < rptr[i + 1]; ++j) val[j] /= diag[i] * diag[idx[j]];
}

// Row order of SELL-C-sigma: rows sorted by descending length inside windows
// of sigma rows (csr_to_sell_sigma, sell.hpp:309-327).  The reference sorts
// with its own quicksort (core/sort.hpp), which is not stable; this oracle
// uses a stable sort, and the product must do the same.
static std::vector<int32_t> sigma_perm(int32_t n, const int32_t *rptr, int32_t sigma) {
  std::vector<int32_t> p(n);
  for (int32_t i = 0; i < n; ++i) p[i] = i;
  if (sigma > 1)
    for (int32_t w = 0; w < n; w += sigma) {
      int32_t e = std::min(n, w + sigma);
      std::stable_sort(p.begin() + w, p.begin() + e, [&](int32_t a, int32_t b) {
        return rptr[a + 1] - rptr[a] > rptr[b + 1] - rptr[b];
      });
    }
  return p;
}

int64_t orc_sell_nslots(int32_t n, const int32_t *rptr, int32_t C, int32_t sigma) {
  auto p = sigma_perm(n, rptr, sigma);
  int64_t tot = 0;
  for (int32_t s = 0; s < n; s += C) {
    int32_t tc = std::min(C, n - s);
    int32_t mx = 0;
    for (int32_t j = 0; j < tc; ++j) mx = std::max(mx, rptr[p[s + j] + 1] - rptr[p[s + j]]);
    tot += (int64_t)mx * tc;
  }
  return tot;
}

// csr_to_sell (sell.hpp:247-287, left aligned): chunk s of t_chunk rows;
// slot k of row r sits at cptr[s] + k*t_chunk + r; padding value 0 with the
// previous slot's column (0 for an empty row).  cptr[s] is the element offset
// (the reference stores sptr[s] = offset / w; the last chunk may be partial).
void orc_csr_to_sell(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
                     int32_t C, int32_t sigma, int32_t *cptr, int32_t *sidx, double *sval,
                     int32_t *perm) {
  auto p = sigma_perm(n, rptr, sigma);
  for (int32_t i = 0; i < n; ++i) perm[i] = p[i];
  int32_t nch = (n + C - 1) / C;
  cptr[0] = 0;
  for (int32_t s = 0; s < nch; ++s) {
    int32_t r0 = s * C, tc = std::min(C, n - r0);
    int32_t mx = 0;
    for (int32_t j = 0; j < tc; ++j) mx = std::max(mx, rptr[p[r0 + j] + 1] - rptr[p[r0 + j]]);
    int32_t off = cptr[s];
    for (int32_t j = 0; j < tc; ++j) {
      int32_t row = p[r0 + j], cnt = 0;
      for (int32_t k = rptr[row]; k < rptr[row + 1]; ++k, ++cnt) {
        sval[off + cnt * tc + j] = val[k];
        sidx[off + cnt * tc + j] = idx[k];
      }
      for (; cnt < mx; ++cnt) {
        sval[off + cnt * tc + j] = 0.;
        sidx[off + cnt * tc + j] = cnt == 0 ? 0 : sidx[off + (cnt - 1) * tc + j];
      }
    }
    cptr[s + 1] = off + mx * tc;
  }
}

void orc_spmv(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
              int32_t val_prec, const void *x, int32_t x_prec, void *y, int32_t y_prec,
              int32_t acc_prec, int32_t residual, const void *b, int32_t b_prec) {
  CSRd A = make_csr(n, rptr, idx, val);
  // The combinations exercised by the solve (precision codes 0/1/2):
#define ORC_CASE(VP, XP, YP, AP, V, X, Y, TA)                                     \
  if (val_prec == VP && x_prec == XP && y_prec == YP && acc_prec == AP) {         \
    CSRv<V> Av(A);                                                                \
    if (residual) {                                                               \
      if (b_prec != YP) return;                                                   \
      resid_t<TA>(Av, (const Y *)b, (const X *)x, (Y *)y);                       \
    } else                                                                        \
      spmv_t<TA>(Av, (const X *)x, (Y *)y);                                       \
    return;                                                                       \
  }
  ORC_CASE(0, 0, 0, 0, double, double, double, double)
  ORC_CASE(1, 1, 1, 1, float, float, float, float)
  ORC_CASE(2, 1, 1, 1, h16, float, float, float)
  ORC_CASE(2, 2, 2, 1, h16, h16, h16, float)
  ORC_CASE(2, 2, 1, 1, h16, h16, float, float)
  ORC_CASE(3, 3, 3, 1, b16, b16, b16, float)
  ORC_CASE(3, 1, 1, 1, b16, float, float, float)
#undef ORC_CASE
  if (val_prec == 2 && x_prec == 2 && y_prec == 2 && acc_prec == 2) {
    CSRv<h16> Av(A);
    spmv_h16acc(Av, (const h16 *)x, (h16 *)y, residual ? (const h16 *)b : nullptr);
  }
}

void orc_spmv_ld(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
                 const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    long double t = 0;
    for (int32_t j = rptr[i]; j < rptr[i + 1]; ++j) t += (long double)val[j] * (long double)x[idx[j]];
    y[i] = (double)t;
  }
}

void *orc_sdainv(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
                 double tol, double acc) {
  CSRd A = make_csr(n, rptr, idx, val);
  return new Factors(sdainv(A, tol, acc));
}
void *orc_ilu0(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
               int32_t block_num, double acc) {
  CSRd A = make_csr(n, rptr, idx, val);
  auto *F = new Factors(ilu0(A, block_num, acc));
  // keep the block count in an unused slot: f[0].n is n for both factors
  return F;
}
int64_t orc_factor_nnz(void *h, int32_t which) { return ((Factors *)h)->f[which].nnz(); }
void orc_factor_get(void *h, int32_t which, int32_t *rptr, int32_t *idx, double *val) {
  const CSRd &F = ((Factors *)h)->f[which];
  std::memcpy(rptr, F.rptr.data(), sizeof(int32_t) * (F.n + 1));
  std::memcpy(idx, F.idx.data(), sizeof(int32_t) * F.nnz());
  std::memcpy(val, F.val.data(), sizeof(double) * F.nnz());
}
void orc_factor_free(void *h) { delete (Factors *)h; }

void orc_bj_solve(void *ilu, const double *in, double *out, double *tmp) {
  const Factors &F = *(Factors *)ilu;
  BJ<double> bj(F, 1);  // block bounds only steer the parallel split
  bj.solve<double, false>(in, out, tmp);
}

void orc_bj_solve16(void *ilu, const uint16_t *in, uint16_t *out, uint16_t *tmp, int32_t ref_arith) {
  const Factors &F = *(Factors *)ilu;
  BJ<h16> bj(F, 1);
  if (ref_arith)
    bj.solve<float, true>((const h16 *)in, (h16 *)out, (h16 *)tmp);
  else
    bj.solve<float, false>((const h16 *)in, (h16 *)out, (h16 *)tmp);
}

int32_t orc_f3r_solve(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
                      const double *b, double *x, const orc_f3r_opts *opts,
                      orc_f3r_result *res, double *hist) {
  if (opts->depth < 0 || opts->depth > 5 || (opts->depth != 0 && opts->ref_arith != 0)) return -1;
  ThreadScope ts(opts->nthreads);
  const bool saved_fma = g_fma;
  g_fma = opts->ref_arith == 1 ? false : true;
  CSRd A = make_csr(n, rptr, idx, val);
  double t0 = now_s();
  auto S = make_f3r(A, *opts);
  double t1 = now_s();
  S->reset();
  double nrm_b = std::sqrt(reduce_sum<double>(n, [&](int64_t i) { return b[i] * b[i]; }));
  for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
  int steps = 0;
  const double eps = opts->eps;
  const int maxo = opts->max_outer;
  Cond cond = [&](int, double e) {
    if (hist) hist[steps] = e / nrm_b;
    ++steps;
    if (maxo > 0 && steps > maxo) return true;
    return e / nrm_b < eps;
  };
  double t2 = now_s();
  SolveRes fl = S->solve(b, x, cond);
  double t3 = now_s();
  std::vector<double> r(n);
  orc_spmv_ld(n, rptr, idx, val, x, r.data());
  double rr = 0;
  for (int32_t i = 0; i < n; ++i) rr += (b[i] - r[i]) * (b[i] - r[i]);
  res->is_solved = fl.is_solved;
  res->res_iter = fl.res_iter;
  res->res_nrm2 = fl.res_nrm2;
  res->nrm_b = nrm_b;
  res->expl_res = std::sqrt(rr) / nrm_b;
  res->solve_s = t3 - t2;
  res->setup_s = t1 - t0;
  res->precond_calls = S->calls();
  res->nthreads = omp_get_max_threads();
  g_fma = saved_fma;
  return 0;
}

int32_t orc_level_apply(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
                        const orc_f3r_opts *opts, int32_t level, const double *in, double *out) {
  if (opts->depth < 0 || opts->depth > 5 || (opts->depth != 0 && opts->ref_arith != 0)) return -1;
  ThreadScope ts(opts->nthreads);
  const bool saved_fma = g_fma;
  g_fma = opts->ref_arith == 1 ? false : true;
  CSRd A = make_csr(n, rptr, idx, val);
  auto S = make_f3r(A, *opts);
  S->reset();
  S->level_apply(level, in, out);
  g_fma = saved_fma;
  return 0;
}

// Restarted GMRES(m), fp64, no preconditioner (same Arnoldi/Givens as FGMRES
// with M = I); the accuracy reference of the precision study.
int32_t orc_gmres(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
                  const double *b, double *x, int32_t m, int32_t max_iter, double tol,
                  double *rel_res) {
  CSRd A = make_csr(n, rptr, idx, val);
  CSRv<double> Av(A);
  struct Ident : OpT<double> {
    int64_t n;
    void apply(const double *in, double *out) override { std::memcpy(out, in, sizeof(double) * n); }
  } I;
  I.n = n;
  FGMRES<double, double> G(&Av, &I, m, true);
  double nb = std::sqrt(reduce_sum<double>(n, [&](int64_t i) { return b[i] * b[i]; }));
  std::vector<double> r(n);
  int total = 0;
  Cond cond = [&](int, double e) { return e / nb < tol; };
  while (total < max_iter) {
    resid_t<double>(Av, b, x, r.data());
    SolveRes s = G.kernel<false>(r.data(), x, cond);
    total += s.res_iter;
    if (s.is_solved) break;
  }
  std::vector<double> ax(n);
  orc_spmv_ld(n, rptr, idx, val, x, ax.data());
  double rr = 0;
  for (int32_t i = 0; i < n; ++i) rr += (b[i] - ax[i]) * (b[i] - ax[i]);
  if (rel_res) *rel_res = std::sqrt(rr) / nb;
  return total;
}

void orc_mlp_forward(int32_t rows, int32_t nin, int32_t h1, int32_t h2, const double *x,
                     const double *w1, const double *b1, const double *w2, const double *b2,
                     const double *w3, const double *b3, double *y) {
  std::vector<double> a1((size_t)rows * h1), a2((size_t)rows * h2);
  dense(rows, nin, h1, x, w1, b1, a1.data(), true);
  dense(rows, h1, h2, a1.data(), w2, b2, a2.data(), true);
  dense(rows, h2, 1, a2.data(), w3, b3, y, false);
}

void orc_mlp_backward(int32_t rows, int32_t nin, int32_t h1, int32_t h2, const double *x,
                      const double *w1, const double *b1, const double *w2, const double *b2,
                      const double *w3, const double *b3, const double *g, double *gw1,
                      double *gb1, double *gw2, double *gb2, double *gw3, double *gb3) {
  std::vector<double> a1((size_t)rows * h1), a2((size_t)rows * h2);
  dense(rows, nin, h1, x, w1, b1, a1.data(), true);
  dense(rows, h1, h2, a1.data(), w2, b2, a2.data(), true);
  std::fill(gw1, gw1 + (size_t)h1 * nin, 0.0);
  std::fill(gb1, gb1 + h1, 0.0);
  std::fill(gw2, gw2 + (size_t)h2 * h1, 0.0);
  std::fill(gb2, gb2 + h2, 0.0);
  std::fill(gw3, gw3 + h2, 0.0);
  gb3[0] = 0;
  std::vector<double> d2(h2), d1(h1);
  for (int r = 0; r < rows; ++r) {
    const double gr = g[r];
    const double *xr = x + (size_t)r * nin;
    const double *a1r = a1.data() + (size_t)r * h1;
    const double *a2r = a2.data() + (size_t)r * h2;
    gb3[0] += gr;
    for (int o = 0; o < h2; ++o) {
      gw3[o] += gr * a2r[o];
      d2[o] = a2r[o] > 0 ? gr * w3[o] : 0.0;
    }
    for (int o = 0; o < h2; ++o) {
      gb2[o] += d2[o];
      for (int i = 0; i < h1; ++i) gw2[(size_t)o * h1 + i] += d2[o] * a1r[i];
    }
    for (int i = 0; i < h1; ++i) {
      double s = 0;
      for (int o = 0; o < h2; ++o) s += w2[(size_t)o * h1 + i] * d2[o];
      d1[i] = a1r[i] > 0 ? s : 0.0;
    }
    for (int o = 0; o < h1; ++o) {
      gb1[o] += d1[o];
      for (int i = 0; i < nin; ++i) gw1[(size_t)o * nin + i] += d1[o] * xr[i];
    }
  }
  (void)b1;
  (void)b2;
  (void)b3;
}

}  // extern "C"
