/*
 * f3r_oracle.h -- CPU oracle for the F3R nested mixed-precision solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product library (f3r_amd/) never calls it.
 *
 * It restates, in plain C++17 compiled by g++, the arithmetic of the
 * reference (suzuki-hpc/F3R, umineko-sparse) on the hot path:
 *   scaling                sparse/matrix/algorithm.hpp:45-59
 *   csr_to_sell(_sigma)    sparse/matrix/sell.hpp:247-343
 *   SELL/CSR operate/res   sparse/matrix/sell.hpp:55-89, csr.hpp:32-50
 *   SDAINV                 sparse/preconditioner/ainv.hpp:27-65,268-375,
 *                          sparse/spvec.hpp:40-92
 *   ILU(0) + split + BJ    sparse/preconditioner/ilu.hpp:128-164,
 *                          sparse/matrix/algorithm.hpp:8-43,
 *                          sparse/matrix/csr.hpp:60-155, trsv/bj.hpp:58-98
 *   FGMRES + Givens        sparse/solver/gmres.hpp:12-135
 *   Restarted              sparse/solver/restart.hpp:8-37
 *   adaptive Richardson    work/src/f3r-gpu.cpp:229-260, work/src/f3r.cpp:212-243
 *   static Richardson      work/src/static.cpp:99-119
 *   nesting / driver       work/src/f3r-gpu.cpp:77-104,210-266
 *   depth ablations        work/src/depth.cpp:122-167, solver/gmres2.hpp:56-150
 *
 * Precision codes used throughout: 0 = fp64, 1 = fp32, 2 = fp16.
 *
 * What pins each part (tests/test_oracle*.py, fixtures in tests/golden/ written
 * by the reference's own code, built from its sources by oracle/ref_harness):
 *   fp16 conversions      every fp16 value against numpy's; round to nearest
 *                         even at ties; fp64 -> fp16 against exact rounding
 *   bf16 conversions      exact rational rounding of random and boundary values
 *   scaling, SELL layout  the reference's arrays, bit for bit
 *   SpMV (all precisions) the reference's results, bit for bit; a brute-force
 *                         dense product; long-double sums within 1e-14
 *   SDAINV, ILU(0), BJ    the reference's factors, bit for bit; exact inverse
 *                         and LU of tridiagonal matrices (closed forms)
 *   FGMRES / Richardson / the reference's nests (DOUBLE, FLOAT, HALF) run on the
 *   Restarted / driver    host: solution vectors and residual histories bit for
 *                         bit (reference arithmetic); its CPU driver's CSV line
 *   static weights        the reference's static.cpp CSV lines (+-1 outer step)
 *   depth ablations       the reference's depth.cpp CSV lines (F4, F3, F2 within
 *                         one step; F3H, F2H no more steps), F3 = F3R with one
 *                         static step bit for bit, fp16 levels against their
 *                         fp32 twins
 *   MLP forward/backward  numpy's fp64 matrix products (forward) and finite
 *                         differences of the forward (backward)
 * Nothing here is left unpinned.
 */
#ifndef F3R_ORACLE_H
#define F3R_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- fp16 conversions (IEEE binary16, round to nearest even) ---- */
uint16_t orc_f2h(float f);
float orc_h2f(uint16_t h);
uint16_t orc_d2h(double d);
/* bfloat16 (precision code 3 of orc_spmv): fp64 and fp32 to bf16 correctly
 * rounded (nearest even), bf16 to fp32 exactly */
uint16_t orc_d2b(double d);
uint16_t orc_f2b(float f);
float orc_b2f(uint16_t b);
void orc_d2h_array(int64_t n, const double *in, uint16_t *out);
void orc_f2h_array(int64_t n, const float *in, uint16_t *out);
void orc_h2f_array(int64_t n, const uint16_t *in, float *out);

/* ---- formats ---- */
void orc_scaling(int32_t n, const int32_t *rptr, const int32_t *idx, double *val);
/* number of padded slots of SELL-C-sigma (C rows per chunk, rows sorted by
 * descending length inside windows of sigma rows; sigma=1 is plain SELL-C) */
int64_t orc_sell_nslots(int32_t n, const int32_t *rptr, int32_t C, int32_t sigma);
void orc_csr_to_sell(int32_t n, const int32_t *rptr, const int32_t *idx,
                     const double *val, int32_t C, int32_t sigma, int32_t *cptr,
                     int32_t *sidx, double *sval, int32_t *perm);

/* ---- SpMV: y = A x (residual=0) or y = b - A x (residual=1).
 * val_prec: precision the matrix values are rounded to before use.
 * acc_prec: accumulator precision (fp16 accumulation rounds every product
 *           and every sum to fp16, as the reference's __half build does).
 * Every multiply-add of a fp32/fp64 accumulator is a fused multiply-add.  */
void orc_spmv(int32_t n, const int32_t *rptr, const int32_t *idx, const double *val,
              int32_t val_prec, const void *x, int32_t x_prec, void *y,
              int32_t y_prec, int32_t acc_prec, int32_t residual, const void *b,
              int32_t b_prec);
/* fp64 SpMV with long-double accumulation: accuracy reference */
void orc_spmv_ld(int32_t n, const int32_t *rptr, const int32_t *idx,
                 const double *val, const double *x, double *y);

/* ---- factorizations: results are held in an opaque handle with two CSR
 * factors.  SDAINV: factor 0 = Z, factor 1 = Wt (ainv.hpp:57-64).
 * ILU(0)+split: factor 0 = L (unit lower, 1.0 stored last in each row),
 * factor 1 = U (diagonal first in each row).                               */
void *orc_sdainv(int32_t n, const int32_t *rptr, const int32_t *idx,
                 const double *val, double tol, double acc);
void *orc_ilu0(int32_t n, const int32_t *rptr, const int32_t *idx,
               const double *val, int32_t block_num, double acc);
int64_t orc_factor_nnz(void *h, int32_t which);
void orc_factor_get(void *h, int32_t which, int32_t *rptr, int32_t *idx, double *val);
void orc_factor_free(void *h);
/* block-Jacobi triangular solves of trsv/bj.hpp on the split ILU(0) factors,
 * in fp64 */
void orc_bj_solve(void *ilu, const double *in, double *out, double *tmp);
/* the same with fp16 factors and vectors: fp32 sums (ref_arith = 0) or one
 * fp16 rounding per statement (ref_arith = 1) */
void orc_bj_solve16(void *ilu, const uint16_t *in, uint16_t *out, uint16_t *tmp,
                    int32_t ref_arith);

/* ---- nested solve ---- */
typedef struct {
  int32_t m1, m2, m3, m4;  /* outer FGMRES(m1), level-2 m2, level-3 m3, Richardson m4 */
  int32_t c;               /* probe period of the adaptive weights */
  int32_t restart;         /* cycles of the outer FGMRES */
  double eps;              /* stop when |e_j| / ||b|| < eps */
  int32_t variant;         /* 0 = HALF, 1 = FLOAT, 2 = DOUBLE (work/MakefileGPU.mk:41-48) */
  int32_t ref_arith;       /* 0 = this build's arithmetic: fused multiply-adds,
                              and in the HALF variant fp32 sums of fp16 data;
                              1 = the reference's arithmetic: no contraction,
                              and in HALF fp16 sums and fp16 weights */
  double range_scale;      /* HALF build mode: power-of-two factor applied to
                              the fp16 level's input (1 = none) */
  int32_t precond;         /* 0 = SDAINV, 1 = block-Jacobi ILU(0) */
  double tol, acc;         /* SDAINV drop tolerance and diagonal weight */
  int32_t bj_blocks;       /* ILU(0) diagonal blocks */
  int32_t max_outer;       /* >0: stop after this many outer steps (timing samples) */
  int32_t nthreads;        /* OpenMP threads, 0 = default */
  double static_weight;    /* >0: Richardson with this fixed weight for every
                              step and no probes (work/src/static.cpp:99-119) */
  int32_t depth;           /* 0: the F3R nest above; 1..5: the nesting-depth
                              ablations F2, F2H, F3, F3H, F4 of
                              work/src/depth.cpp:122-167 (this build's
                              arithmetic only) */
} orc_f3r_opts;

typedef struct {
  int32_t is_solved;
  int32_t res_iter;        /* outer Arnoldi steps summed over restarts */
  double res_nrm2;         /* |e| of the last outer step (implicit residual) */
  double nrm_b;
  double expl_res;         /* ||b - A x|| / ||b|| in fp64 */
  double solve_s;          /* wall time of solve() only */
  double setup_s;          /* preconditioner + format setup */
  int64_t precond_calls;   /* innermost operator applications */
  int32_t nthreads;
} orc_f3r_result;

/* Solves A x = b with the matrix as given (the caller applies scaling).
 * Returns 0, or -1 for a depth outside 0..5 or a depth with ref_arith = 1.
 * hist (optional, length >= m1*restart+1) receives |e_j|/||b|| per outer step.*/
int32_t orc_f3r_solve(int32_t n, const int32_t *rptr, const int32_t *idx,
                      const double *val, const double *b, double *x,
                      const orc_f3r_opts *opts, orc_f3r_result *res, double *hist);

/* Plain restarted GMRES(m) in fp64 without preconditioner (oracle for the
 * precision study); returns the number of iterations. */
int32_t orc_gmres(int32_t n, const int32_t *rptr, const int32_t *idx,
                  const double *val, const double *b, double *x, int32_t m,
                  int32_t max_iter, double tol, double *rel_res);

/* One application of a single F3R level as an operator, used to measure the
 * error of each level against fp64: level 2 (fp32 FGMRES), 3 (fp16-value
 * FGMRES) or 4 (Richardson) applied to `in`, written to `out` (fp64). */
int32_t orc_level_apply(int32_t n, const int32_t *rptr, const int32_t *idx,
                        const double *val, const orc_f3r_opts *opts,
                        int32_t level, const double *in, double *out);

/* ---- learned correction MLP (fp64 reference): 12 -> H1 -> H2 -> 1, ReLU.
 * Weights are row-major [out][in].  Forward writes y (rows).  Backward
 * accumulates dL/dparams for L = sum_r g[r] * y[r].                        */
void orc_mlp_forward(int32_t rows, int32_t nin, int32_t h1, int32_t h2,
                     const double *x, const double *w1, const double *b1,
                     const double *w2, const double *b2, const double *w3,
                     const double *b3, double *y);
void orc_mlp_backward(int32_t rows, int32_t nin, int32_t h1, int32_t h2,
                      const double *x, const double *w1, const double *b1,
                      const double *w2, const double *b2, const double *w3,
                      const double *b3, const double *g, double *gw1, double *gb1,
                      double *gw2, double *gb2, double *gw3, double *gb3);

int32_t orc_num_threads(void);
/* 1 (default): fused multiply-adds as the GPU build computes them;
 * 0: one rounding per operation, as the reference built without contraction */
void orc_set_fma(int32_t on);

#ifdef __cplusplus
}
#endif

#endif
