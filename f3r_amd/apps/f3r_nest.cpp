// f3r_nest -- the reference driver's F3R solve composed from objects with the
// C++ layer (include/f3r_amd.hpp), in the structure of work/src/f3r-gpu.cpp:
// read and scale the system, build SELL copies in three precisions and the
// SDAINV preconditioner, compose Restarted(FGMRES<double>(FGMRES<float>(
// FGMRES<float>(AutoRichardson)))) and solve with the caller's stopping rule.
//
//   f3r_nest <matrix.mtx|gen:27:<nx>[:<beta>]> <acc> <reps> <m2> <m3> <m4> <c> <restart> [check]
//
// Prints the reference's CSV line (work/src/f3r-gpu.cpp:268-272):
//   Problem,Method,Prec,M2,M3,M4,W,Precond,ACC,Time,Iter,ImplRes,ExplRes
// With `check` it also solves through f3r_solver_create and compares, bit for
// bit, and composes two more nests (M applied directly under two levels, and
// the same with M inside an Operator lambda), printing one line per check;
// the exit status is the number of failed checks.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "f3r_amd.hpp"

namespace {

f3r::CSR stencil27(int nx, double beta) {
  int64_t nnz;
  f3r::check(f3r_gen_stencil27_nnz(nx, nx, nx, &nnz), "gen_stencil27_nnz");
  f3r::CSR A;
  A.n = A.ncols = nx * nx * nx;
  A.rptr.resize(A.n + 1);
  A.idx.resize(nnz);
  A.val.resize(nnz);
  f3r::check(f3r_gen_stencil27(nx, nx, nx, beta, A.rptr.data(), A.idx.data(), A.val.data()), "gen_stencil27");
  return A;
}

bool same_bits(const std::vector<double> &a, const std::vector<double> &b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(double) * a.size()) == 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: %s <matrix.mtx|gen:27:<nx>[:beta]> <acc> <reps> <m2> <m3> <m4> <c> <restart> [check]\n",
                 argv[0]);
    return 2;
  }
  const std::string src = argv[1];
  const double acc = std::atof(argv[2]);
  const int reps = std::max(1, std::atoi(argv[3]));
  const int m2 = std::atoi(argv[4]), m3 = std::atoi(argv[5]), m4 = std::atoi(argv[6]);
  const int c = std::atoi(argv[7]), restart = std::atoi(argv[8]);
  const bool checks = argc > 9 && std::string(argv[9]) == "check";
  const double eps = 1.e-8;
  try {
    // the system: a MatrixMarket file with its _b.mtx (else U[0,1) of
    // mt19937_64(0)), or a generated 27-point stencil with that right-hand side
    f3r::CSR data;
    std::vector<double> rhs;
    std::string name;
    if (src.rfind("gen:27:", 0) == 0) {
      const int nx = std::atoi(src.c_str() + 7);
      const char *bp = std::strchr(src.c_str() + 7, ':');
      data = stencil27(nx, bp ? std::atof(bp + 1) : 0.0);
      rhs.resize(data.n);
      f3r::check(f3r_gen_uniform_mt(data.n, 0.0, 1.0, 0, rhs.data()), "gen_uniform_mt");
      name = "stencil27_" + std::to_string(nx);
    } else {
      auto sys = f3r::io::mm::read_system(src);
      data = std::move(sys.first);
      rhs = std::move(sys.second);
      const std::string base = src.substr(src.find_last_of('/') + 1);
      name = base.substr(0, base.size() - 4);
    }
    f3r::algorithm::scaling(data);

    auto A = f3r::SELL<64, double>(data);
    auto b = f3r::vector<double>(data.n);
    b.upload(rhs);
    auto x = f3r::vector<double>(data.n);
    auto r = f3r::vector<double>(data.n);
    auto A32 = f3r::SELL<64, float>(A);
    auto A16 = f3r::SELL<64, f3r::half>(A);
    auto M = f3r::SDAINV<f3r::half>(data, {0.1, acc});
    auto R = f3r::AutoRichardson<f3r::half>(A16, M, {m4, c});
    auto L3 = f3r::FGMRES<float>(A16, R, {m3, false});
    auto L2 = f3r::FGMRES<float>(A32, L3, {m2, false});
    auto solver = f3r::Restarted<double>(A, f3r::FGMRES<double>(A, L2, {100, true}), {restart});

    const double nrm_b = f3r::nrm(b);
    auto cond = [&](int, double e) { return e / nrm_b < eps; };
    std::printf("%s,F3R,__half,%d,%d,%d,%d,SDAINV,%s,", name.c_str(), m2, m3, m4, c, argv[2]);
    double tsum = 0;
    int itr = 0;
    f3r::solve_res_t flag{};
    for (int i = 0; i < reps; ++i) {
      x.fill(0.0);
      flag = solver.solve(b, x, cond);
      tsum += solver.last().solve_s;
      itr += flag.res_iter;
    }
    A.compute_residual(b, x, r);
    std::printf("%e,%d,%e,%e\n", tsum / reps, itr * m2 * m3 * m4 / reps, flag.res_nrm2 / nrm_b, f3r::nrm(r) / nrm_b);
    if (!checks) return flag.is_solved ? 0 : 3;

    int failed = 0;
    auto report = [&](const char *what, bool ok, const std::string &detail) {
      std::printf("check %s: %s%s\n", what, ok ? "ok" : "FAILED", detail.c_str());
      failed += ok ? 0 : 1;
    };
    const std::vector<double> x_nest = x.download();

    // 1. the same nest through f3r_solver_create (the C ABI's F3R solver)
    {
      f3r_params p;
      f3r::check(f3r_params_default(&p), "params_default");
      p.m2 = m2;
      p.m3 = m3;
      p.m4 = m4;
      p.c = c;
      p.restart = restart;
      p.C = 64;
      f3r_solver S;
      auto &ctx = f3r::Context::current();
      f3r::check(f3r_solver_create(ctx.get(), data.handle(ctx), M.get(), &p, &S), "solver_create");
      auto x2 = f3r::vector<double>(data.n);
      x2.fill(0.0);
      f3r_result res;
      f3r::check(f3r_solve(S, b.get(), x2.get(), &res), "solve");
      f3r_solver_destroy(S);
      const bool ok = same_bits(x_nest, x2.download()) && res.res_iter == flag.res_iter &&
                      res.res_nrm2 == flag.res_nrm2;
      report("composed F3R nest = f3r_solver_create", ok,
             " (" + std::to_string(flag.res_iter) + " vs " + std::to_string(res.res_iter) + " outer steps)");
    }
    // 2. M applied directly under two levels (the F3 nest of work/src/depth.cpp:133-141)
    //    against the depth ablation solver of the C ABI
    std::vector<double> x_f3;
    {
      auto F = f3r::FGMRES<float>(A16, M, {8, false});
      auto inner = f3r::FGMRES<float>(A32, F, {8, false});
      auto S3 = f3r::Restarted<double>(A, f3r::FGMRES<double>(A, inner, {100, true}), {restart});
      x.fill(0.0);
      auto f3 = S3.solve(b, x, cond);
      x_f3 = x.download();
      f3r_params p;
      f3r::check(f3r_params_default(&p), "params_default");
      p.depth = 3;
      p.restart = restart;
      p.C = 64;
      f3r_solver S;
      auto &ctx = f3r::Context::current();
      f3r::check(f3r_solver_create(ctx.get(), data.handle(ctx), M.get(), &p, &S), "solver_create");
      auto x2 = f3r::vector<double>(data.n);
      x2.fill(0.0);
      f3r_result res;
      f3r::check(f3r_solve(S, b.get(), x2.get(), &res), "solve");
      f3r_solver_destroy(S);
      report("composed F3 nest = depth ablation F3", same_bits(x_f3, x2.download()) && res.res_iter == f3.res_iter,
             " (" + std::to_string(f3.res_iter) + " outer steps)");
    }
    // 3. the same nest with M inside an Operator(shape, lambda)
    {
      int calls = 0;
      auto Mop = f3r::Operator(
          data.n,
          [&](const f3r::view &in, const f3r::view &out) {
            ++calls;
            M.operate(in, out);
          },
          F3R_F16);
      auto F = f3r::FGMRES<float>(A16, Mop, {8, false});
      auto inner = f3r::FGMRES<float>(A32, F, {8, false});
      auto S3 = f3r::Restarted<double>(A, f3r::FGMRES<double>(A, inner, {100, true}), {restart});
      x.fill(0.0);
      auto f3 = S3.solve(b, x, cond);
      report("Operator lambda innermost = M applied directly", same_bits(x_f3, x.download()) && calls > 0,
             " (" + std::to_string(f3.res_iter) + " outer steps, " + std::to_string(calls) + " lambda calls)");
    }
    // 4. a deeper nest with the Richardson innermost, and a stopping rule of
    //    the caller's that ends after two outer steps
    {
      auto L4 = f3r::FGMRES<float>(A16, R, {2, false});
      auto L3b = f3r::FGMRES<float>(A16, L4, {2, false});
      auto L2b = f3r::FGMRES<float>(A32, L3b, {4, false});
      auto S4 = f3r::Restarted<double>(A, f3r::FGMRES<double>(A, L2b, {100, true}), {restart});
      x.fill(0.0);
      auto f4 = S4.solve(b, x, cond);
      A.compute_residual(b, x, r);
      const double rel = f3r::nrm(r) / nrm_b;
      report("three levels over the Richardson solve to 1e-8", f4.is_solved && rel <= eps,
             " (" + std::to_string(f4.res_iter) + " outer steps, true residual " + std::to_string(rel) + ")");
      x.fill(0.0);
      auto two = S4.solve(b, x, [](int i, double) { return i >= 2; });
      report("the caller's stopping rule", two.is_solved && two.res_iter == 2,
             " (" + std::to_string(two.res_iter) + " outer steps)");
      // 5. an exception thrown by the stopping rule stops the solve and
      //    reaches the caller unchanged
      bool caught = false;
      try {
        x.fill(0.0);
        S4.solve(b, x, [](int i, double) -> bool {
          if (i >= 1) throw std::runtime_error("stop rule failed");
          return false;
        });
      } catch (const std::runtime_error &e) {
        caught = std::string(e.what()) == "stop rule failed";
      }
      report("an exception in the stopping rule reaches the caller", caught, "");
    }
    return failed;
  } catch (const f3r::error &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
