// f3r_gpu -- the reference driver's command line on the C ABI of libf3r_amd.
//
//   f3r_gpu <matrix.mtx> <acc> <reps> <m2> <m3> <m4> <c> <restart> [ilu0]
//   f3r_gpu gen:27:<nx>[:<beta>] ...   synthetic 27-point stencil instead of a file
//
// Same arguments and the same CSV result line as work/src/f3r-gpu.cpp:45-54,
// 268-272, 97-102:
//   Problem,Method,Prec,M2,M3,M4,W,Precond,ACC,Time,Iter,ImplRes,ExplRes
// The matrix is read by the library's MatrixMarket reader (f3r_mm_read:
// symmetric entries expanded, explicit off-diagonal zeros dropped, columns
// sorted per row; sparse/matrix/io.hpp:42-64,129-192), the right-hand side
// from <name>_b.mtx when present, else U[0,1) from std::mt19937_64(0)
// (work/src/f3r-gpu.cpp:60).  The matrix is scaled (algorithm::scaling) and
// the solve is timed alone, averaged over <reps> repetitions.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "f3r_amd.h"

namespace {

void die(int st, const char *where) {
  std::fprintf(stderr, "%s: %s (%s)\n", where, f3r_strerror(st), f3r_last_error());
  std::exit(1);
}
#define CK(call)                  \
  do {                            \
    int _s = (call);              \
    if (_s != F3R_OK) die(_s, #call); \
  } while (0)

struct Csr {
  int32_t n = 0;
  std::vector<int32_t> rptr, idx;
  std::vector<double> val;
};

// the library's MatrixMarket reader (f3r_mm_read: io::mm::read_matrix +
// io::CSR of sparse/matrix/io.hpp) into host arrays
int read_mm(const std::string &path, Csr &A) {
  f3r_mat M;
  int st = f3r_mm_read(nullptr, path.c_str(), &M);
  if (st != F3R_OK) return st;
  int32_t nr, nc;
  int64_t nnz;
  if ((st = f3r_mat_shape(M, &nr, &nc)) == F3R_OK && (st = f3r_mat_info(M, nullptr, &nnz, nullptr, nullptr, nullptr)) == F3R_OK) {
    if (nr != nc) {
      st = F3R_ERR_ARG;
    } else {
      A.n = nr;
      A.rptr.resize(nr + 1);
      A.idx.resize(nnz);
      A.val.resize(nnz);
      st = f3r_mat_csr_get(M, A.rptr.data(), A.idx.data(), A.val.data());
    }
  }
  f3r_mat_destroy(M);
  return st;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: %s <matrix.mtx|gen:27:<nx>[:beta]> <acc> <reps> <m2> <m3> <m4> <c> <restart> [ilu0]\n",
                 argv[0]);
    return 2;
  }
  std::string src = argv[1];
  const double acc = std::atof(argv[2]);
  const int reps = std::max(1, std::atoi(argv[3]));
  const int m2 = std::atoi(argv[4]), m3 = std::atoi(argv[5]), m4 = std::atoi(argv[6]);
  const int c = std::atoi(argv[7]), restart = std::atoi(argv[8]);
  const bool ilu = argc > 9 && std::string(argv[9]) == "ilu0";

  Csr A;
  std::vector<double> b;
  std::string name;
  if (src.rfind("gen:27:", 0) == 0) {
    int nx = std::atoi(src.c_str() + 7);
    const char *bp = std::strchr(src.c_str() + 7, ':');
    double beta = bp ? std::atof(bp + 1) : 0.0;
    int64_t nnz;
    CK(f3r_gen_stencil27_nnz(nx, nx, nx, &nnz));
    A.n = nx * nx * nx;
    A.rptr.resize(A.n + 1);
    A.idx.resize(nnz);
    A.val.resize(nnz);
    CK(f3r_gen_stencil27(nx, nx, nx, beta, A.rptr.data(), A.idx.data(), A.val.data()));
    name = "stencil27_" + std::to_string(nx);
  } else {
    CK(read_mm(src, A));
    std::string base = src.substr(src.find_last_of('/') + 1);
    name = base.substr(0, base.size() - 4);
    std::vector<char> bpath(src.size() + 8);
    CK(f3r_mm_rhs_path(src.c_str(), bpath.data(), (int64_t)bpath.size()));
    int64_t nb = 0;
    // a _b.mtx that exists must be read (io.hpp:108-126); only a missing one
    // falls back to the generated right-hand side
    if (std::FILE *bf = std::fopen(bpath.data(), "rb")) {
      std::fclose(bf);
      CK(f3r_mm_read_vector(bpath.data(), 0, nullptr, &nb));
      b.resize(nb);
      CK(f3r_mm_read_vector(bpath.data(), nb, b.data(), &nb));
      if ((int32_t)b.size() != A.n) {
        std::fprintf(stderr, "%s: %lld entries for %d rows\n", bpath.data(), (long long)nb, A.n);
        return 1;
      }
    }
  }
  if (b.empty()) {  // U[0,1) from mt19937_64(0), the reference drivers' fallback
    b.resize(A.n);
    CK(f3r_gen_uniform_mt(A.n, 0.0, 1.0, 0, b.data()));
  }

  CK(f3r_csr_scaling(A.n, A.rptr.data(), A.idx.data(), A.val.data()));
  f3r_ctx ctx;
  CK(f3r_ctx_create(0, &ctx));
  f3r_mat M;
  CK(f3r_mat_create_csr(ctx, A.n, A.n, (int64_t)A.idx.size(), A.rptr.data(), A.idx.data(), A.val.data(), &M));
  f3r_prec P;
  if (ilu)
    CK(f3r_ilu0_create(ctx, M, 112, acc, F3R_F16, &P));
  else
    CK(f3r_sdainv_create(ctx, M, 0.1, acc, F3R_F16, &P));
  f3r_params prm;
  f3r_params_default(&prm);
  prm.m2 = m2;
  prm.m3 = m3;
  prm.m4 = m4;
  prm.c = c;
  prm.restart = restart;
  f3r_solver S;
  CK(f3r_solver_create(ctx, M, P, &prm, &S));
  f3r_vec bv, xv, rv;
  CK(f3r_vec_create(ctx, F3R_F64, A.n, &bv));
  CK(f3r_vec_create(ctx, F3R_F64, A.n, &xv));
  CK(f3r_vec_create(ctx, F3R_F64, A.n, &rv));
  CK(f3r_vec_upload(bv, b.data()));
  std::printf("%s,F3R,__half,%d,%d,%d,%d,%s,%s,", name.c_str(), m2, m3, m4, c, ilu ? "ILU0" : "SDAINV", argv[2]);
  double tsum = 0;
  int itr = 0;
  f3r_result res{};
  for (int r = 0; r < reps; ++r) {
    CK(f3r_vec_fill(xv, 0.0));
    CK(f3r_solve(S, bv, xv, &res));
    tsum += res.solve_s;
    itr += res.res_iter;
  }
  // explicit residual ||b - A x|| / ||b|| in fp64
  CK(f3r_residual(M, bv, xv, rv));
  double nr, nb;
  CK(f3r_vec_nrm2(rv, &nr));
  CK(f3r_vec_nrm2(bv, &nb));
  std::printf("%e,%d,%e,%e\n", tsum / reps, itr * m2 * m3 * m4 / reps, res.res_nrm2 / nb, nr / nb);
  f3r_vec_destroy(bv);
  f3r_vec_destroy(xv);
  f3r_vec_destroy(rv);
  f3r_solver_destroy(S);
  f3r_prec_destroy(P);
  f3r_mat_destroy(M);
  f3r_ctx_destroy(ctx);
  return res.is_solved ? 0 : 3;
}
