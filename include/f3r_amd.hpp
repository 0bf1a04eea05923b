/*
 * f3r_amd.hpp -- header-only C++ layer over the C ABI of f3r_amd.h with the
 * class shapes of the reference's library (suzuki-hpc/F3R,
 * library/umineko-sparse), so that a driver written like
 * work/src/f3r-gpu.cpp:210-266 composes the same nest:
 *
 *   auto data = f3r::io::mm::read_system(path).first;       // io.hpp:104-127
 *   f3r::algorithm::scaling(data);                           // algorithm.hpp:45-59
 *   auto A   = f3r::SELL<64, double>(data);                  // sell.hpp:35-97
 *   auto A32 = f3r::SELL<64, float>(A);                      // sell.hpp:51-53
 *   auto A16 = f3r::SELL<64, f3r::half>(A);
 *   auto M   = f3r::SDAINV<f3r::half>(data, {0.1, 1.0});     // ainv.hpp:129-147
 *   auto R   = f3r::AutoRichardson<f3r::half>(A16, M, {2, 64});   // richardson.hpp:96-156
 *   auto L3  = f3r::FGMRES<float>(A16, R, {4, false});       // gmres.hpp:56-135
 *   auto L2  = f3r::FGMRES<float>(A32, L3, {8, false});
 *   auto S   = f3r::Restarted<double>(A, f3r::FGMRES<double>(A, L2, {100, true}), {20});
 *   auto res = S.solve(b, x, [&](int, double e) { return e / nrm_b < 1e-8; });
 *
 * Objects are cheap shared handles, as the reference's are (core/share.hpp).
 * Device objects live on a context; constructors without one use
 * f3r::Context::current(), a process-wide context on GPU 0 (the reference's
 * implicit default stream).  The innermost operator of a nest may be a
 * preconditioner, the adaptive Richardson or an Operator(shape, lambda); the
 * lambda receives device vector views and may call any operation of this
 * layer on them.  Every failing call throws f3r::error.
 */
#ifndef F3R_AMD_HPP
#define F3R_AMD_HPP

#include <cstdint>
#include <cstdio>
#include <exception>
#include <functional>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <utility>
#include <vector>

#include "f3r_amd.h"

namespace f3r {

// ------------------------------------------------------------- errors ----
struct error : std::runtime_error {
  int status;
  error(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};
inline void check(int st, const char *where) {
  if (st != F3R_OK)
    throw error(st, std::string(where) + ": " + f3r_strerror(st) + " (" + f3r_last_error() + ")");
}

// element types: the precisions of kmm::vector<T, L>.  `half` is a tag for
// fp16 (IEEE binary16) storage; host code never computes with it.
struct half;
template <class T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr int value = F3R_F64; };
template <> struct dtype_of<float> { static constexpr int value = F3R_F32; };
template <> struct dtype_of<half> { static constexpr int value = F3R_F16; };

struct solve_res_t {  // sparse/models.hpp:191-200
  bool is_solved;
  int32_t res_iter;
  double res_nrm2;
};

// ------------------------------------------------------------ context ----
class Context {
 public:
  explicit Context(int device = 0) {
    f3r_ctx c = nullptr;
    check(f3r_ctx_create(device, &c), "ctx_create");
    h_ = std::shared_ptr<f3r_ctx_s>(c, [](f3r_ctx p) { f3r_ctx_destroy(p); });
  }
  f3r_ctx get() const { return h_.get(); }
  void sync() const { check(f3r_ctx_sync(get()), "ctx_sync"); }
  // the process-wide context of the constructors that take none
  static Context &current() {
    static Context c(0);
    return c;
  }

 private:
  std::shared_ptr<f3r_ctx_s> h_;
};

// ------------------------------------------------------------ vectors ----
// kmm::vector<T, device> (core/tensor.hpp:204-316) of runtime dtype: an
// owning vector or a non-owning view (the operands an Operator receives)
class view {
 public:
  view() = default;
  view(f3r_vec v, std::shared_ptr<f3r_vec_s> keep) : v_(v), keep_(std::move(keep)) {}
  f3r_vec get() const { return v_; }
  int dtype() const {
    int dt;
    check(f3r_vec_info(v_, &dt, nullptr, nullptr), "vec_info");
    return dt;
  }
  int64_t size() const {
    int64_t n;
    check(f3r_vec_info(v_, nullptr, &n, nullptr), "vec_info");
    return n;
  }
  void fill(double a) const { check(f3r_vec_fill(v_, a), "vec_fill"); }
  void copy(const view &src) const { check(f3r_vec_copy(src.get(), v_), "vec_copy"); }
  void axpy(double alpha, const view &x) const { check(f3r_vec_axpy(alpha, x.get(), v_), "vec_axpy"); }
  void scale(double alpha) const { check(f3r_vec_scale(v_, alpha), "vec_scale"); }

 protected:
  f3r_vec v_ = nullptr;
  std::shared_ptr<f3r_vec_s> keep_;
};
inline double dot(const view &a, const view &b) {
  double d;
  check(f3r_vec_dot(a.get(), b.get(), &d), "vec_dot");
  return d;
}
inline double nrm(const view &a) {
  double d;
  check(f3r_vec_nrm2(a.get(), &d), "vec_nrm2");
  return d;
}

template <class T> class vector : public view {
 public:
  explicit vector(int64_t n, Context &ctx = Context::current()) {
    f3r_vec v = nullptr;
    check(f3r_vec_create(ctx.get(), dtype_of<T>::value, n, &v), "vec_create");
    keep_ = std::shared_ptr<f3r_vec_s>(v, [](f3r_vec p) { f3r_vec_destroy(p); });
    v_ = v;
  }
  // host data of T's precision (double or float), exactly size() entries
  void upload(const std::vector<T> &h) const {
    if ((int64_t)h.size() != size()) throw error(F3R_ERR_ARG, "vector::upload: host data of another size");
    check(f3r_vec_upload(v_, h.data()), "vec_upload");
  }
  std::vector<T> download() const {
    std::vector<T> h(size());
    check(f3r_vec_download(v_, h.data()), "vec_download");
    return h;
  }
};

// ---------------------------------------------------------- host CSR ----
// CSR<double, host> (sparse/matrix/csr.hpp:12-50): int32 row pointers and
// column indices (sorted within a row), fp64 values
struct CSR {
  int32_t n = 0, ncols = 0;
  std::vector<int32_t> rptr, idx;
  std::vector<double> val;
  int64_t nnz() const { return (int64_t)idx.size(); }
  int32_t nrows() const { return n; }
  // the matrix on a context (created on first use, dropped when values change)
  f3r_mat handle(Context &ctx) const {
    if (!h_ || hctx_ != ctx.get()) {
      f3r_mat m = nullptr;
      check(f3r_mat_create_csr(ctx.get(), n, ncols, nnz(), rptr.data(), idx.data(), val.data(), &m), "mat_create_csr");
      h_ = std::shared_ptr<f3r_mat_s>(m, [](f3r_mat p) { f3r_mat_destroy(p); });
      hctx_ = ctx.get();
    }
    return h_.get();
  }
  void invalidate() { h_.reset(); }

 private:
  mutable std::shared_ptr<f3r_mat_s> h_;
  mutable f3r_ctx hctx_ = nullptr;
};

namespace algorithm {
// a_ij /= sqrt|a_ii| sqrt|a_jj| in place (sparse/matrix/algorithm.hpp:45-59)
inline void scaling(CSR &A) {
  check(f3r_csr_scaling(A.n, A.rptr.data(), A.idx.data(), A.val.data()), "csr_scaling");
  A.invalidate();
}
}  // namespace algorithm

namespace io {
struct mm {  // sparse/matrix/io.hpp:104-217
  static CSR read_matrix(const std::string &path) {
    f3r_mat m = nullptr;
    check(f3r_mm_read(nullptr, path.c_str(), &m), "mm_read");
    std::shared_ptr<f3r_mat_s> keep(m, [](f3r_mat p) { f3r_mat_destroy(p); });
    CSR A;
    int64_t nnz;
    check(f3r_mat_shape(m, &A.n, &A.ncols), "mat_shape");
    check(f3r_mat_info(m, nullptr, &nnz, nullptr, nullptr, nullptr), "mat_info");
    A.rptr.resize(A.n + 1);
    A.idx.resize(nnz);
    A.val.resize(nnz);
    check(f3r_mat_csr_get(m, A.rptr.data(), A.idx.data(), A.val.data()), "mat_csr_get");
    return A;
  }
  // (A, b): b from <name>_b.mtx, else U[0, 1) of std::mt19937_64(seed)
  static std::pair<CSR, std::vector<double>> read_system(const std::string &path, uint64_t seed = 0) {
    CSR A = read_matrix(path);
    std::vector<char> bp(path.size() + 8);
    check(f3r_mm_rhs_path(path.c_str(), bp.data(), (int64_t)bp.size()), "mm_rhs_path");
    std::vector<double> b;
    int64_t nb = 0;
    // an existing _b.mtx must be read; only a missing one falls back
    std::FILE *bf = std::fopen(bp.data(), "rb");
    if (bf) {
      std::fclose(bf);
      check(f3r_mm_read_vector(bp.data(), 0, nullptr, &nb), "mm_read_vector");
      b.resize(nb);
      check(f3r_mm_read_vector(bp.data(), nb, b.data(), &nb), "mm_read_vector");
      if (nb != A.n) throw error(F3R_ERR_ARG, "mm::read_system: the right-hand side's size differs from the matrix's");
    } else {
      b.resize(A.n);
      check(f3r_gen_uniform_mt(A.n, 0.0, 1.0, seed, b.data()), "gen_uniform_mt");
    }
    return {std::move(A), std::move(b)};
  }
};
}  // namespace io

// -------------------------------------------------------------- SELL ----
// SELL<w, T, device> (sparse/matrix/sell.hpp:35-97): chunks of w rows
// (a power of two >= 16), values in T; the cross-precision constructor shares
// the index arrays (sell.hpp:51-53).  Keeps the CSR it was built from, which
// a nest needs for its own layouts.
template <int w, class T> class SELL {
 public:
  static_assert(w >= 16 && (w & (w - 1)) == 0, "SELL<w>: w a power of two >= 16");
  using value_type = T;
  static constexpr int chunk = w;
  explicit SELL(const CSR &A, Context &ctx = Context::current()) : ctx_(&ctx), csr_(std::make_shared<CSR>(A)) {
    f3r_mat m = nullptr;
    check(f3r_mat_to_sell(ctx.get(), csr_->handle(ctx), dtype_of<T>::value, w, 1, 1, &m), "mat_to_sell");
    h_ = std::shared_ptr<f3r_mat_s>(m, [](f3r_mat p) { f3r_mat_destroy(p); });
  }
  template <class T2> explicit SELL(const SELL<w, T2> &o) : ctx_(&o.context()), csr_(o.csr()) {
    f3r_mat m = nullptr;
    check(f3r_mat_cast(o.get(), dtype_of<T>::value, &m), "mat_cast");
    h_ = std::shared_ptr<f3r_mat_s>(m, [](f3r_mat p) { f3r_mat_destroy(p); });
  }
  int32_t nrows() const { return csr_->n; }
  // y = A x (sell.hpp:55-71) and r = b - A x (sell.hpp:72-89)
  void operate(const view &x, const view &y) const { check(f3r_spmv(get(), x.get(), y.get()), "spmv"); }
  void compute_residual(const view &b, const view &x, const view &r) const {
    check(f3r_residual(get(), b.get(), x.get(), r.get()), "residual");
  }
  f3r_mat get() const { return h_.get(); }
  Context &context() const { return *ctx_; }
  const std::shared_ptr<CSR> &csr() const { return csr_; }

 private:
  Context *ctx_;
  std::shared_ptr<CSR> csr_;
  std::shared_ptr<f3r_mat_s> h_;
};

// ---------------------------------------------------- preconditioners ----
class Preconditioner {
 public:
  void operate(const view &in, const view &out) const { check(f3r_prec_apply(get(), in.get(), out.get()), "prec_apply"); }
  f3r_prec get() const { return h_.get(); }

 protected:
  void own(f3r_prec p) { h_ = std::shared_ptr<f3r_prec_s>(p, [](f3r_prec q) { f3r_prec_destroy(q); }); }
  std::shared_ptr<f3r_prec_s> h_;
};
// SDAINV<T, device>(A, {tol, acc}) (sparse/preconditioner/ainv.hpp:129-147)
template <class T> class SDAINV : public Preconditioner {
 public:
  struct params {
    double tol, acc;
  };
  SDAINV(const CSR &A, params p, Context &ctx = Context::current()) {
    f3r_prec h = nullptr;
    check(f3r_sdainv_create(ctx.get(), A.handle(ctx), p.tol, p.acc, dtype_of<T>::value, &h), "sdainv_create");
    own(h);
  }
};
// ILU(0) on `blocks` diagonal blocks (ilu.hpp:128-164 + algorithm::split),
// solved block by block as trsv::BJ does (trsv/bj.hpp:58-98)
template <class T> class ILU0 : public Preconditioner {
 public:
  struct params {
    int32_t blocks;
    double acc;
  };
  ILU0(const CSR &A, params p, Context &ctx = Context::current()) {
    f3r_prec h = nullptr;
    check(f3r_ilu0_create(ctx.get(), A.handle(ctx), p.blocks, p.acc, dtype_of<T>::value, &h), "ilu0_create");
    own(h);
  }
};

// ---------------------------------------------------------- Operator ----
// Operator<L>(shape, lambda) (sparse/models.hpp:46-126): lambda(in, out) on
// device views of `dtype` (F3R_F32 by default; see f3r_op_create)
class Operator {
 public:
  using fn_t = std::function<void(const view &, const view &)>;
  Operator(int64_t n, fn_t f, int dtype = F3R_F32, Context &ctx = Context::current())
      : st_(std::make_shared<State>()) {
    st_->f = std::move(f);
    f3r_op op = nullptr;
    check(f3r_op_create(ctx.get(), (int32_t)n, dtype, &Operator::trampoline, st_.get(), &op), "op_create");
    st_->op = std::shared_ptr<f3r_op_s>(op, [](f3r_op p) { f3r_op_destroy(p); });
  }
  f3r_op get() const { return st_->op.get(); }

 private:
  struct State {
    fn_t f;
    std::shared_ptr<f3r_op_s> op;
  };
  static int trampoline(void *user, f3r_vec in, f3r_vec out) {
    try {
      static_cast<State *>(user)->f(view(in, nullptr), view(out, nullptr));
      return F3R_OK;
    } catch (const error &e) {
      return e.status;
    } catch (...) {
      return F3R_ERR_ARG;
    }
  }
  std::shared_ptr<State> st_;
};

// ------------------------------------------------------- solver nests ----
// What a nest is made of, innermost first: the Krylov levels below the outer
// one and the innermost operator.  Each solver object below adds itself.
struct nest_desc {
  std::vector<f3r_level> levels;  // outermost inner level first
  f3r_inner inner{};
  std::vector<std::shared_ptr<void>> keep;  // objects the nest refers to
};

// the adaptive Richardson of work/src/f3r-gpu.cpp:229-260 in its library
// form AutoRichardson<T, L>(A, M, {max_iter, period})
// (sparse/solver/richardson.hpp:96-156), on the fp16 matrix with an fp16
// preconditioner; a static weight (work/src/static.cpp:99-119) replaces the
// probes when > 0
template <class T> class AutoRichardson {
 public:
  struct params {
    int32_t max_iter;
    int32_t period;
    double static_weight = 0.0;
  };
  template <int w, class V>
  AutoRichardson(const SELL<w, V> &A, const Preconditioner &M, params p) : M_(M), p_(p) {
    static_assert(std::is_same<T, half>::value && std::is_same<V, half>::value,
                  "AutoRichardson runs on fp16 data (the HALF build)");
    (void)A;
  }
  void describe(nest_desc &d) const {
    d.inner.kind = F3R_INNER_RICHARDSON;
    d.inner.P = M_.get();
    d.inner.m4 = p_.max_iter;
    d.inner.c = p_.period;
    d.inner.static_weight = p_.static_weight;
    d.keep.push_back(std::make_shared<Preconditioner>(M_));
  }

 private:
  Preconditioner M_;
  params p_;
};

namespace detail {
inline void describe_inner(nest_desc &d, const Preconditioner &M) {
  d.inner.kind = F3R_INNER_PREC;
  d.inner.P = M.get();
  d.keep.push_back(std::make_shared<Preconditioner>(M));
}
inline void describe_inner(nest_desc &d, const Operator &op) {
  d.inner.kind = F3R_INNER_OP;
  d.inner.op = op.get();
  d.keep.push_back(std::make_shared<Operator>(op));
}
template <class X> inline auto describe_inner(nest_desc &d, const X &x) -> decltype(x.describe(d)) {
  return x.describe(d);
}
}  // namespace detail

// FGMRES<T, L>(A, M, {max_iter, normalize}) (sparse/solver/gmres.hpp:56-135):
// T = float (fp32 basis) or half (fp16 storage, fp32 arithmetic) for an
// inner level; T = double for the outer level under Restarted.  M is another
// FGMRES, an AutoRichardson, a preconditioner or an Operator.
template <class T> class FGMRES {
 public:
  struct params {
    int32_t max_iter;
    bool normalize;
  };
  template <int w, class V, class M>
  FGMRES(const SELL<w, V> &A, const M &inner, params p)
      : d_(std::make_shared<nest_desc>()), m_(p.max_iter), normalize_(p.normalize), C_(w), csr_(A.csr()),
        ctx_(&A.context()) {
    if (!std::is_same<T, double>::value) {
      if (p.normalize) throw error(F3R_ERR_UNSUPPORTED, "FGMRES: inner levels run with normalize = false");
      d_->levels.push_back({p.max_iter, dtype_of<T>::value, dtype_of<V>::value});
    }
    nest_desc in;
    detail::describe_inner(in, inner);
    d_->levels.insert(d_->levels.end(), in.levels.begin(), in.levels.end());
    d_->inner = in.inner;
    d_->keep = in.keep;
  }
  void describe(nest_desc &d) const {
    d.levels.insert(d.levels.end(), d_->levels.begin(), d_->levels.end());
    d.inner = d_->inner;
    d.keep.insert(d.keep.end(), d_->keep.begin(), d_->keep.end());
  }
  int32_t max_iter() const { return m_; }
  bool normalize() const { return normalize_; }
  int chunk() const { return C_; }
  const std::shared_ptr<CSR> &csr() const { return csr_; }
  Context &context() const { return *ctx_; }
  const nest_desc &desc() const { return *d_; }

 private:
  std::shared_ptr<nest_desc> d_;
  int32_t m_;
  bool normalize_;
  int C_;
  std::shared_ptr<CSR> csr_;
  Context *ctx_;
};

// Restarted<T, L>(A, S, {max_iter}) (sparse/solver/restart.hpp:8-37) over
// the fp64 outer FGMRES(A, ..., {m1, true}): the whole nest, built on the
// device at construction (f3r_nest_create)
template <class T> class Restarted {
 public:
  static_assert(std::is_same<T, double>::value, "Restarted<double>: the outer level is fp64");
  struct params {
    int32_t max_iter;
  };
  template <int w>
  Restarted(const SELL<w, double> &A, const FGMRES<double> &S, params p) : st_(std::make_shared<State>()) {
    if (!S.normalize()) throw error(F3R_ERR_UNSUPPORTED, "Restarted: the outer FGMRES normalizes");
    f3r_params prm;
    check(f3r_params_default(&prm), "params_default");
    prm.m1 = S.max_iter();
    prm.restart = p.max_iter;
    prm.C = w;
    const nest_desc &d = S.desc();
    st_->keep = d.keep;
    st_->csr = A.csr();
    f3r_solver h = nullptr;
    check(f3r_nest_create(A.context().get(), A.csr()->handle(A.context()), d.levels.data(), (int32_t)d.levels.size(),
                          &d.inner, &prm, &h),
          "nest_create");
    st_->s = std::shared_ptr<f3r_solver_s>(h, [](f3r_solver q) { f3r_solver_destroy(q); });
  }
  // Solver::solve(b, x, cond) (sparse/models.hpp:191-267): cond(i, |e|).
  // An exception from cond stops the solve and is rethrown here.
  solve_res_t solve(const view &b, const view &x, std::function<bool(int, double)> cond) const {
    f3r_result r{};
    Cond c{std::move(cond), nullptr};
    const int st = f3r_solve_cond(get(), b.get(), x.get(), &Restarted::trampoline, &c, &r);
    if (c.err) std::rethrow_exception(c.err);
    check(st, "solve_cond");
    last_ = r;
    return {r.is_solved != 0, r.res_iter, r.res_nrm2};
  }
  // the library's full result of the last solve (time, preconditioner calls)
  const f3r_result &last() const { return last_; }
  f3r_solver get() const { return st_->s.get(); }

 private:
  struct State {
    std::shared_ptr<f3r_solver_s> s;
    std::vector<std::shared_ptr<void>> keep;
    std::shared_ptr<CSR> csr;
  };
  struct Cond {
    std::function<bool(int, double)> f;
    std::exception_ptr err;
  };
  // no exception crosses the C ABI: it is kept, and the solve is stopped
  static int trampoline(void *user, int32_t i, double e) {
    Cond *c = static_cast<Cond *>(user);
    try {
      return c->f(i, e) ? 1 : 0;
    } catch (...) {
      c->err = std::current_exception();
      return 1;
    }
  }
  std::shared_ptr<State> st_;
  mutable f3r_result last_{};
};

}  // namespace f3r

#endif
