// ref_driver.cpp -- runs the reference's own host-side library code
// (suzuki-hpc/F3R, library/umineko-{core,sparse}, included from where it lies
// under the reference checkout) on small inputs and writes the results as
// binary files.  Test infrastructure only: it produces the golden vectors that
// pin the oracle.  Nothing from the reference is copied here; this file only
// calls the reference's headers.
//
// Binary CSR file: int32 n, int64 nnz, int32 rptr[n+1], int32 idx[nnz],
// double val[nnz].  Binary vector file: int64 len, then len elements.
//
// Commands
//   scale   A.bin out.bin                      algorithm::scaling
//   sell32  A.bin out.bin                      SELL<32,double,host>(CSR)
//   sdainv  A.bin tol acc Z.bin W.bin          algorithm::sdainv -> {Z, Wt}
//   ilu0    A.bin blocks acc L.bin U.bin       ilup(duplicate_block_diagonal) + split
//   spmv    A.bin x.bin mode y.bin             SELL<32,T,host>::operate, mode:
//            64 (fp64), 32 (fp32), 1616 (fp16 values, fp16 x, fp16 sums),
//            1632 (fp16 values, fp32 x, fp32 sums)
//   mm      A.mtx out.bin [b.bin]              io::mm::read_system + io::CSR(coo);
//            b.bin gets the right-hand side (<name>_b.mtx, else U[0,1) of
//            std::mt19937_64(0), as work/src/f3r.cpp:39 asks for it)
//   f3r     A.bin b.bin prec m2 m3 m4 c restart acc x.bin hist.bin
//            nested solve composed from the reference's FGMRES / Restarted /
//            SELL<32,T,host> / sdainv, with the adaptive Richardson level of
//            work/src/f3r-gpu.cpp:229-260 on the host tag; prec = 64|32|16.
#include "umineko-sparse/matrix/algorithm.hpp"
#include "umineko-sparse/matrix/csr.hpp"
#include "umineko-sparse/matrix/io.hpp"
#include "umineko-sparse/matrix/sell.hpp"
#include "umineko-sparse/preconditioner/ainv.hpp"
#include "umineko-sparse/preconditioner/ilu.hpp"
#include "umineko-sparse/solver/gmres.hpp"
#include "umineko-sparse/solver/restart.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace kmm;

static CSR<double, host> read_csr(const char *path) {
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    std::exit(2);
  }
  int32_t n;
  int64_t nnz;
  if (std::fread(&n, 4, 1, f) != 1 || std::fread(&nnz, 8, 1, f) != 1) std::exit(2);
  auto A = CSR<double, host>(impl::shape{n, n}, (idx_t)nnz);
  if (std::fread(A.rptr.raw(), 4, n + 1, f) != (size_t)(n + 1)) std::exit(2);
  if (std::fread(A.idx.raw(), 4, nnz, f) != (size_t)nnz) std::exit(2);
  if (std::fread(A.val.raw(), 8, nnz, f) != (size_t)nnz) std::exit(2);
  std::fclose(f);
  return A;
}

static void write_csr(const char *path, const CSR<double, host> &A) {
  FILE *f = std::fopen(path, "wb");
  int32_t n = A.nrows();
  int64_t nnz = A.nnz();
  std::fwrite(&n, 4, 1, f);
  std::fwrite(&nnz, 8, 1, f);
  std::fwrite(A.rptr.raw(), 4, n + 1, f);
  std::fwrite(A.idx.raw(), 4, nnz, f);
  std::fwrite(A.val.raw(), 8, nnz, f);
  std::fclose(f);
}

template <class T> static std::vector<T> read_vec(const char *path) {
  FILE *f = std::fopen(path, "rb");
  int64_t len;
  if (!f || std::fread(&len, 8, 1, f) != 1) std::exit(2);
  std::vector<T> v(len);
  if (std::fread(v.data(), sizeof(T), len, f) != (size_t)len) std::exit(2);
  std::fclose(f);
  return v;
}

template <class T> static void write_raw(const char *path, const T *p, int64_t len) {
  FILE *f = std::fopen(path, "wb");
  std::fwrite(&len, 8, 1, f);
  std::fwrite(p, sizeof(T), len, f);
  std::fclose(f);
}

template <class P> static int run_f3r(int argc, char **argv);

int main(int argc, char **argv) {
  if (argc < 2) return 1;
  std::string cmd = argv[1];
  if (cmd == "scale") {
    auto A = read_csr(argv[2]);
    algorithm::scaling(A);
    write_csr(argv[3], A);
  } else if (cmd == "sell32") {
    auto A = read_csr(argv[2]);
    auto S = SELL<32, double, host>(A);
    int64_t ns = S.sptr.size(0);
    write_raw(argv[3], S.sptr.raw(), ns);
    std::string p = argv[3];
    write_raw((p + ".idx").c_str(), S.idx.raw(), (int64_t)S.idx.size(0));
    write_raw((p + ".val").c_str(), S.val.raw(), (int64_t)S.val.size(0));
  } else if (cmd == "sdainv") {
    auto A = read_csr(argv[2]);
    auto [Z, W] = algorithm::sdainv(A, std::atof(argv[3]), std::atof(argv[4]));
    write_csr(argv[5], Z);
    write_csr(argv[6], W);
  } else if (cmd == "ilu0") {
    auto A = read_csr(argv[2]);
    int blocks = std::atoi(argv[3]);
    double acc = std::atof(argv[4]);
    auto bd = A.duplicate_block_diagonal(blocks, 1);
    auto [l, u] = algorithm::split(algorithm::ilup(bd, 0, acc, blocks, 1));
    write_csr(argv[5], l);
    write_csr(argv[6], u);
  } else if (cmd == "spmv") {
    auto A = read_csr(argv[2]);
    int mode = std::atoi(argv[4]);
    auto S = SELL<32, double, host>(A);
    int64_t n = A.nrows();
    if (mode == 64) {
      auto xv = read_vec<double>(argv[3]);
      auto x = vector<double, host>(xv.data(), impl::_shape<1>(n));
      auto y = vector<double, host>(n);
      S.operate(x, y);
      write_raw(argv[5], y.raw(), n);
    } else if (mode == 32) {
      auto S32 = SELL<32, float, host>(S);
      auto xv = read_vec<float>(argv[3]);
      auto x = vector<float, host>(xv.data(), impl::_shape<1>(n));
      auto y = vector<float, host>(n);
      S32.operate(x, y);
      write_raw(argv[5], y.raw(), n);
    } else if (mode == 1616) {
      auto S16 = SELL<32, _Float16, host>(S);
      auto xv = read_vec<_Float16>(argv[3]);
      auto x = vector<_Float16, host>(xv.data(), impl::_shape<1>(n));
      auto y = vector<_Float16, host>(n);
      S16.operate(x, y);
      write_raw(argv[5], y.raw(), n);
    } else if (mode == 1632) {
      auto S16 = SELL<32, _Float16, host>(S);
      auto xv = read_vec<float>(argv[3]);
      auto x = vector<float, host>(xv.data(), impl::_shape<1>(n));
      auto y = vector<float, host>(n);
      S16.operate(x, y);
      write_raw(argv[5], y.raw(), n);
    }
  } else if (cmd == "mm") {
    auto [coo, rhs] = io::mm::read_system(argv[2], [](vector<double, host> v) {
      v.random<0, 1>(std::mt19937_64(0));
    });
    auto C = io::CSR<double>(coo);
    FILE *f = std::fopen(argv[3], "wb");
    int32_t n = C.nrows(), nc = C.ncols();
    int64_t nnz = C.nnz();
    std::fwrite(&n, 4, 1, f);
    std::fwrite(&nc, 4, 1, f);
    std::fwrite(&nnz, 8, 1, f);
    std::fwrite(C.rptr.raw(), 4, n + 1, f);
    std::fwrite(C.idx.raw(), 4, nnz, f);
    std::fwrite(C.val.raw(), 8, nnz, f);
    std::fclose(f);
    if (argc > 4) write_raw(argv[4], rhs.raw(), (int64_t)rhs.size(0));
  } else if (cmd == "f3r") {
    int prec = std::atoi(argv[4]);
    if (prec == 64) return run_f3r<double>(argc, argv);
    if (prec == 32) return run_f3r<float>(argc, argv);
    return run_f3r<_Float16>(argc, argv);
  }
  return 0;
}

// P: storage type of the innermost level.  For _Float16 the levels are
// FGMRES<float>(A16) and FGMRES<float>(A32); otherwise FGMRES<P>(A_P) twice,
// the DOUBLE / FLOAT / HALF compositions of the reference's GPU driver.
template <class P> static int run_f3r(int, char **argv) {
  using tag = host;
  using W_t = std::conditional_t<std::is_same_v<P, _Float16>, float, P>;
  auto data = read_csr(argv[2]);
  auto bv = read_vec<double>(argv[3]);
  int m2 = std::atoi(argv[5]), m3 = std::atoi(argv[6]), m4 = std::atoi(argv[7]);
  int c = std::atoi(argv[8]), restart = std::atoi(argv[9]);
  double acc = std::atof(argv[10]);
  int n = data.nrows();

  auto A = SELL<32, double, tag>(data);
  auto b = vector<double, tag>(n);
  for (int i = 0; i < n; ++i) b[i] = bv[i];
  auto x = vector<double, tag>(n);
  impl::pool<double, tag>::init(n);
  impl::pool<float, tag>::init(n);
  auto [z, w] = algorithm::sdainv(data, 0.1, acc);

  // level-3 matrix in P; the Richardson level uses A16 in the HALF
  // composition and the fp64 A in the DOUBLE and FLOAT ones
  auto AP = SELL<32, P, tag>(A);
  auto A32 = SELL<32, W_t, tag>(A);
  using R_t = std::conditional_t<std::is_same_v<P, _Float16>, _Float16, double>;
  auto AR = SELL<32, R_t, tag>(A);
  auto Wm = SELL<32, P, tag>(w);
  auto Zm = SELL<32, P, tag>(z);
  auto t_in = vector<P, tag>(n), t_out = vector<P, tag>(n), t1 = vector<P, tag>(n),
       t2 = vector<P, tag>(n), t_r = vector<P, tag>(n);
  std::vector<scalar<W_t, tag>> sum(m4);
  std::vector<scalar<P, tag>> ome(m4);
  scalar<W_t, tag> dot_ar, dot_r;
  scalar<int, tag> den;
  int cnt = c;
  for (int k = 0; k < m4; ++k) {
    sum[k] = 1.0;
    ome[k] = 1.0;
  }
  auto f32 = [](auto v) { return static_cast<W_t>(v); };
  // Adaptive Richardson on the innermost level: weights probed every c-th
  // call by minimising the residual, then averaged (including the initial 1).
  auto R = Operator<tag>(data.get_shape(), [&](auto in, auto out) {
    cnt += 1;
    const bool probe = (cnt % c) == 0;
    t_in.copy(in);
    Wm.operate(t_in, t1);
    Zm.operate(t1, t_out);
    if (probe) {
      AR.operate(t_out, t1);
      dot_ar = dot(wildcard(t1, f32), wildcard(t1, f32));
      dot_r = dot(wildcard(t1, f32), wildcard(t_in, f32));
      t_out *= dot_r / dot_ar;
      sum[0] += dot_r / dot_ar;
      ome[0] = sum[0] / (den = cnt / c);
    } else {
      t_out *= ome[0];
    }
    for (int k = 1; k < m4; ++k) {
      AR.compute_residual(t_in, t_out, t_r);
      Wm.operate(t_r, t1);
      Zm.operate(t1, t2);
      if (probe) {
        AR.operate(t2, t1);
        dot_ar = dot(wildcard(t1, f32), wildcard(t1, f32));
        dot_r = dot(wildcard(t1, f32), wildcard(t_r, f32));
        t_out += dot_r / dot_ar * t2;
        sum[k] += dot_r / dot_ar;
        ome[k] = sum[k] / den;
      } else {
        t_out += ome[k] * t2;
      }
    }
    out.copy(t_out);
  });
  auto L3 = FGMRES<W_t, tag>(AP, R, {m3, false});
  auto L2 = FGMRES<W_t, tag>(A32, L3, {m2, false});
  auto inner = Solver<tag>(FGMRES<double, tag>(A, L2, {100, true}));
  auto solver = Restarted<double, tag>(A, inner, {restart});

  x.fill(0.0);
  scalar<double, host> nrm_b;
  nrm_b = nrm(b);
  std::vector<double> hist;
  auto cond = [&](int, double e) {
    hist.push_back(e / nrm_b[0]);
    return e / nrm_b[0] < 1.e-8;
  };
  auto flag = solver.solve(b, x, cond);
  auto r = vector<double, tag>(n);
  A.operate(x, r);
  r = b - r;
  scalar<double, host> nrm_r;
  nrm_r = nrm(r);
  std::printf("%d,%d,%.17e,%.17e\n", (int)flag.is_solved, flag.res_iter,
              flag.res_nrm2 / nrm_b[0], nrm_r[0] / nrm_b[0]);
  write_raw(argv[11], x.raw(), n);
  write_raw(argv[12], hist.data(), (int64_t)hist.size());
  return 0;
}
