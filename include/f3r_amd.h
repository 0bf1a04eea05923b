/*
 * f3r_amd.h -- C ABI of the MI355X-native F3R nested mixed-precision solver.
 *
 * Plain pointers, sizes and opaque handles only.  Every call returns an int
 * status (F3R_OK = 0, negative on error; see f3r_strerror).  Non-convergence
 * is not an error: it is reported in f3r_result.is_solved.
 *
 * Each entry point names the reference interface it replaces
 * (suzuki-hpc/F3R, library/umineko-{core,sparse}, work/src).
 */
#ifndef F3R_AMD_H
#define F3R_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- status ---- */
enum {
  F3R_OK = 0,
  F3R_ERR_ARG = -1,        /* invalid argument or shape mismatch */
  F3R_ERR_HIP = -2,        /* a HIP runtime call failed */
  F3R_ERR_ALLOC = -3,      /* device or host allocation failed */
  F3R_ERR_UNSUPPORTED = -4,/* dtype / format combination not built */
  F3R_ERR_NODEVICE = -5,   /* no usable GPU */
  F3R_ERR_STATE = -6,      /* object used in a state that does not allow it */
  F3R_ERR_RANGE = -7       /* size beyond what the format can index */
};
/* replaces: nothing (the reference exits on errors, sparse/matrix/io.hpp:123) */
const char *f3r_strerror(int status);
const char *f3r_last_error(void); /* detail of the last error on this thread */

/* element types: the precisions of kmm::vector<T, device>, and bfloat16
 * (8-bit exponent, 8 significant bits; fp32 arithmetic like fp16) */
enum { F3R_F64 = 0, F3R_F32 = 1, F3R_F16 = 2, F3R_I32 = 3, F3R_BF16 = 4 };

typedef struct f3r_ctx_s *f3r_ctx;
typedef struct f3r_vec_s *f3r_vec;
typedef struct f3r_mat_s *f3r_mat;
typedef struct f3r_prec_s *f3r_prec;
typedef struct f3r_solver_s *f3r_solver;

/* ------------------------------------------------------------ context ---- */
/* One HIP stream per context; calls on one context are serialised.
 * replaces: the implicit default CUDA stream of exec<device>
 * (core/loop.hpp:142-181) and the global scratch pool (core/tensor.hpp:492-507) */
int f3r_ctx_create(int device, f3r_ctx *ctx);
int f3r_ctx_destroy(f3r_ctx ctx);
int f3r_ctx_sync(f3r_ctx ctx);
int f3r_device_count(int *count);

/* ------------------------------------------------------------ vectors ---- */
/* replaces: kmm::vector<T, device> (core/tensor.hpp:204-316) and
 * memory<device>::{alloc, copy} (core/memory.hpp:38-54) */
int f3r_vec_create(f3r_ctx ctx, int dtype, int64_t n, f3r_vec *v);
/* non-owning view of n elements starting at `offset` (vector(T*, shape),
 * core/tensor.hpp:229-230; matrix column views, tensor.hpp:347-349) */
int f3r_vec_view(f3r_vec base, int64_t offset, int64_t n, f3r_vec *v);
int f3r_vec_destroy(f3r_vec v);
int f3r_vec_upload(f3r_vec v, const void *host);   /* host array of v's dtype */
int f3r_vec_download(f3r_vec v, void *host);
int f3r_vec_info(f3r_vec v, int *dtype, int64_t *n, void **device_ptr);
/* vector::fill (tensor.hpp:247) */
int f3r_vec_fill(f3r_vec v, double value);
/* precision-converting copy dst = T(src), round to nearest even
 * (vector::copy, tensor.hpp:272-277) */
int f3r_vec_copy(f3r_vec src, f3r_vec dst);
/* y += alpha * x  and  x *= alpha (vector expressions, tensor.hpp:302-311) */
int f3r_vec_axpy(double alpha, f3r_vec x, f3r_vec y);
int f3r_vec_scale(f3r_vec x, double alpha);
/* dot / nrm of DOT / NRM (tensor.hpp:509-571): accumulated in the wider of
 * the two element types, fp16 data in fp32; a single-pass grid reduction. */
int f3r_vec_dot(f3r_vec a, f3r_vec b, double *out);
int f3r_vec_nrm2(f3r_vec a, double *out);
/* f3r_vec_dot that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  out[0] (an fp64 vector of >= 1 element on the same
 * context) receives <a, b> on the device, stream-ordered, without a host
 * synchronisation; the bits equal f3r_vec_dot's.  The reduction's counters and
 * partial sums live in a static device array of 73,984 bytes per GPU (the
 * counters of one reduction and the partials of up to 1024 blocks; any n fits,
 * the grid is capped), held until the process ends.  Calls that overlap on
 * one GPU (two contexts or threads at once) race on that array; calls on one
 * context are ordered by its stream. */
int f3r_vec_dot_noalloc(f3r_vec a, f3r_vec b, f3r_vec out);

/* ------------------------------------------------------- host set-up ---- */
/* CPU routines; no GPU is needed.  CSR arrays are int32 row pointers and
 * column indices (columns sorted within a row) and fp64 values.           */
/* algorithm::scaling (sparse/matrix/algorithm.hpp:45-59), in place */
int f3r_csr_scaling(int32_t n, const int32_t *rptr, const int32_t *col, double *val);
/* SELL-C-sigma slot count and layout (sparse/matrix/sell.hpp:247-343):
 * rows sorted by descending length inside windows of `sigma` rows, chunks of
 * C rows, slots interleaved by K; see f3r_mat_to_sell. */
int f3r_sell_nslots(int32_t n, const int32_t *rptr, int32_t C, int32_t sigma, int32_t K,
                    int64_t *nslots);
/* Synthetic inputs (write-only outputs sized by the *_nnz query):
 * 27-point stencil on nx*ny*nz, diagonal 26, neighbours -1 + beta*sign(dx)
 * (matrix/hpcg.cpp:13-57 at beta = 0). */
int f3r_gen_stencil27_nnz(int32_t nx, int32_t ny, int32_t nz, int64_t *nnz);
int f3r_gen_stencil27(int32_t nx, int32_t ny, int32_t nz, double beta, int32_t *rptr,
                      int32_t *col, double *val);
/* 7-point variable-coefficient convection-diffusion: lognormal diffusion
 * exp(sigma * g) of a seeded unit-variance Gaussian field g with correlation
 * length `corr` cells, harmonic face averages, first-order upwind convection
 * of cell Peclet number `peclet` along +x (0 = pure diffusion). */
int f3r_gen_convdiff7_nnz(int32_t nx, int32_t ny, int32_t nz, int64_t *nnz);
int f3r_gen_convdiff7(int32_t nx, int32_t ny, int32_t nz, double sigma, double corr,
                      double peclet, uint64_t seed, int32_t *rptr, int32_t *col,
                      double *val);
/* 7- or 27-point stencil with unit off-diagonals and diagonal
 * (1 + margin) * sum|off|, for the batched systems */
int f3r_gen_dominant_nnz(int32_t npts, int32_t nx, int32_t ny, int32_t nz, int64_t *nnz);
int f3r_gen_dominant(int32_t npts, int32_t nx, int32_t ny, int32_t nz, double margin,
                     int32_t *rptr, int32_t *col, double *val);
/* seeded uniform [lo, hi) vector (splitmix64 -> 53-bit mantissa) */
int f3r_gen_uniform(int64_t n, double lo, double hi, uint64_t seed, double *out);

/* ----------------------------------------------------------- matrices ---- */
/* CSR<double, host> copied to the device (csr.hpp:12-50); the host copy is
 * kept for set-up (SELL build, SDAINV, ILU(0)).  ctx may be NULL: a
 * host-only matrix, usable for preconditioner set-up but not for SpMV. */
int f3r_mat_create_csr(f3r_ctx ctx, int32_t nrows, int32_t ncols, int64_t nnz,
                       const int32_t *rptr, const int32_t *col, const double *val,
                       f3r_mat *A);
/* host copy of a matrix's CSR (rptr[nrows+1], col[nnz], val[nnz]; sizes from
 * f3r_mat_info); any output may be NULL.  replaces: direct access to
 * CSR<double, host>::{rptr, idx, val} (csr.hpp:12-30) */
int f3r_mat_csr_get(f3r_mat A, int32_t *rptr, int32_t *col, double *val);
/* rows and columns of a matrix carrying its CSR (spmat::get_shape,
 * sparse/matrix/base.hpp) */
int f3r_mat_shape(f3r_mat A, int32_t *nrows, int32_t *ncols);

/* ------------------------------------------------------ MatrixMarket ---- */
/* io::mm::read_matrix + io::CSR(coo) (sparse/matrix/io.hpp:42-64,129-192):
 * a coordinate file into a CSR matrix (ctx may be NULL: host-only).  The
 * matrix is symmetric only when the header's fifth token is exactly
 * "symmetric" (io.hpp:133-137; skew-symmetric files are read as general);
 * off-diagonal explicit zeros are dropped; symmetric entries are mirrored;
 * rows sorted by column.  Errors instead of exit(1): F3R_ERR_ARG for an
 * unreadable file, header or entry outside the declared shape;
 * F3R_ERR_UNSUPPORTED for pattern and complex files. */
int f3r_mm_read(f3r_ctx ctx, const char *path, f3r_mat *A);
/* io::mm::_read_matrix_market_b (io.hpp:193-217): an "array ... general"
 * file of one column.  *n = its length; values copied when out != NULL
 * (cap >= *n).  F3R_ERR_ARG when the file is missing or not of that form. */
int f3r_mm_read_vector(const char *path, int64_t cap, double *out, int64_t *n);
/* the right-hand side's path: "_b" inserted before the last 4 characters
 * (read_system, io.hpp:108-109) */
int f3r_mm_rhs_path(const char *path, char *out, int64_t cap);
/* the reference drivers' right-hand side when <name>_b.mtx is absent:
 * std::uniform_real_distribution(lo, hi) over std::mt19937_64(seed)
 * (vector::random<0,1>(std::mt19937_64(0)), work/src/f3r-gpu.cpp:60,
 * core/tensor.hpp:253-262) */
int f3r_gen_uniform_mt(int64_t n, double lo, double hi, uint64_t seed, double *out);

/* SELL<w, T, device>(CSR) (sell.hpp:35-97) and SELLsig (sell.hpp:99-166):
 * C rows per chunk (multiple of 16), sigma = sort window (1 = unsorted),
 * K = slots interleaved per row (1, 2, 4 or 8; K = 1 is the reference layout),
 * and with F3R_SELL_TPR (f3r_mat_to_sell_ex) also the threads per row.
 * Rows of the result are in sorted order; y of f3r_spmv is written back to
 * the original row order. */
int f3r_mat_to_sell(f3r_ctx ctx, f3r_mat csr, int dtype, int32_t C, int32_t sigma,
                    int32_t K, f3r_mat *sell);
/* the same with flags; bit 0: chunk-offset index format (C >= 64): a chunk
 * whose rows share a few column offsets col - row (stencil matrices) stores
 * them once per chunk instead of a column index per slot.  The sums keep the
 * CSR order of every row.  No counterpart in the reference. */
#define F3R_SELL_OFFSETS 1
/* flag bit 1: build the layout on the device from the CSR's device copy
 * (identical arrays to the host build; sigma up to 4096, wider windows and
 * chunks of more than 4096 entries in the offset format are built on the
 * host) */
#define F3R_SELL_DEVICE 2
/* flag bit 2: LDS-staged x gathers (int32-index layout built on the host):
 * every block of 256 rows whose slots read a column window of at most 32 KB
 * of x copies that window into LDS with coalesced loads and gathers from
 * there; wider blocks gather from global memory.  Same sums, same bits.  For
 * banded matrices; f3r_mat_lds_blocks tells how many blocks are staged. */
#define F3R_SELL_LDS 4
/* flag bit 3: K threads per row (K in {2, 4, 8}; int32-index layout, no LDS
 * staging): lane k of a row sums slot k of each of its groups and the K
 * lanes add their sums.  The same layout; the sum is reassociated (not the
 * CSR order).  For long rows: the threads-per-row dimension of the SELL
 * SpMV, chosen by the autotuner with C and sigma. */
#define F3R_SELL_TPR 8
int f3r_mat_to_sell_ex(f3r_ctx ctx, f3r_mat csr, int dtype, int32_t C, int32_t sigma,
                       int32_t K, int32_t flags, f3r_mat *sell);
/* host-only SELL-C-sigma layout of a CSR matrix (csr_to_sell /
 * csr_to_sell_sigma, sell.hpp:247-343): cptr[nchunks+1] element offsets,
 * cwidth[nchunks] slot groups, sidx / sval [nslots], perm[n] (sorted row ->
 * original row).  Sizes from f3r_sell_nslots. */
int f3r_sell_layout(int32_t n, const int32_t *rptr, const int32_t *col, const double *val,
                    int32_t C, int32_t sigma, int32_t K, int32_t *cptr, int32_t *cwidth,
                    int32_t *sidx, double *sval, int32_t *perm);
/* host-only chunk-offset layout (F3R_SELL_OFFSETS, sigma = 1).  A first call
 * with the array outputs NULL returns the sizes: value slots, int32 column
 * indices (of the explicit chunks) and offset-table entries.  Then cptr[nchunks+1],
 * cwidth[nchunks], iptr[nchunks] (>= 0: start of an explicit chunk in idx;
 * < 0: offset chunk with table coff[-iptr-1 ...]), idx, coff and sval.
 * cwidth of an offset chunk counts its full groups of K slots; bit 30 set
 * adds a half group of K / 2 slots per row after them, stored row by row:
 * slot G*K + k of row j at cptr[s] + G*K*tc + j*(K/2) + k (G full groups,
 * tc rows in the chunk). */
int f3r_sell_offset_layout(int32_t n, const int32_t *rptr, const int32_t *col, const double *val,
                           int32_t C, int32_t K, int64_t *nslots, int64_t *nidx, int64_t *ncoff,
                           int32_t *cptr, int32_t *cwidth, int32_t *iptr, int32_t *idx,
                           int32_t *coff, double *sval);
/* cross-precision copy sharing the index arrays: a SELL (sell.hpp:51-53) or a
 * CSR on a context (CSR<T, device>(CSR<double>), csr.hpp:12-30) */
int f3r_mat_cast(f3r_mat sell, int dtype, f3r_mat *out);
/* lanes per row of a CSR matrix's SpMV: a power of two in [1, 64], or 0 for
 * the largest power of two not above half the mean row length.  With 1 each row is
 * summed in CSR order (bit-identical to the SELL kernels); with more, the
 * lanes' strided partial sums are added by a butterfly (coalesced loads).
 * No counterpart in the reference (one OpenMP iteration per row, csr.hpp:32-50). */
int f3r_mat_set_threads_per_row(f3r_mat A, int32_t tpr);
int f3r_mat_destroy(f3r_mat A);
int f3r_mat_info(f3r_mat A, int32_t *nrows, int64_t *nnz, int64_t *nslots, int *dtype,
                 int *format);
/* SELL with F3R_SELL_LDS: blocks of 256 rows, and blocks whose x window fits the LDS */
int f3r_mat_lds_blocks(f3r_mat A, int32_t *blocks, int32_t *staged);
/* SELL: chunks in total and chunks stored in the chunk-offset format */
int f3r_mat_chunks(f3r_mat A, int32_t *chunks, int32_t *offset_chunks);
/* SELL: *where = 0: layout built on the host; 1: on the device
 * (F3R_SELL_DEVICE, when it could run there).  No counterpart in the
 * reference. */
int f3r_mat_setup_where(f3r_mat A, int *where);
/* y = A x  (SELL::operate / CSR::operate, sell.hpp:55-71, csr.hpp:32-40).
 * Accumulates in fp64 for fp64 data, else in fp32 (fp16 x fp16 too). */
int f3r_spmv(f3r_mat A, f3r_vec x, f3r_vec y);
/* r = b - A x  (SELL::compute_residual, sell.hpp:72-89; csr.hpp:41-50) */
int f3r_residual(f3r_mat A, f3r_vec b, f3r_vec x, f3r_vec r);
/* mean device time of `reps` SpMVs on the context's stream (HIP events) */
int f3r_spmv_time(f3r_mat A, f3r_vec x, f3r_vec y, int32_t reps, double *ms);

/* ----------------------------------------------------- preconditioners ---- */
/* SDAINV<T, device>(A, {tol, acc}) (ainv.hpp:27-65,129-147): M = Z D^-1 W^T,
 * applied as two SpMVs (W^T then Z) in `dtype` with fp32 sums.  The factor
 * set-up runs on the host (ainv.hpp:268-329).  With ctx = NULL only the host
 * factors are built (f3r_prec_factor_get works, f3r_prec_apply does not). */
int f3r_sdainv_create(f3r_ctx ctx, f3r_mat csr, double tol, double acc, int dtype,
                      f3r_prec *P);
/* ILU(0) on `blocks` diagonal blocks (ilu.hpp:128-164 + algorithm::split),
 * applied by level-scheduled forward / backward solves (the block-Jacobi
 * solves of trsv/bj.hpp:58-98, same per-row order). */
int f3r_ilu0_create(f3r_ctx ctx, f3r_mat csr, int32_t blocks, double acc, int dtype,
                    f3r_prec *P);
/* The same with set-up flags.  F3R_SETUP_DEVICE: factorise on the device
 * (needs ctx): rows in level sets, one wavefront per row, factors bit-identical
 * to the host set-up; when a row outgrows the device's row capacity (SDAINV:
 * 1024 entries of z_i or w_i; ILU(0): 512 entries per row) the host set-up
 * runs instead.  f3r_prec_setup_where tells which ran. */
#define F3R_SETUP_DEVICE 1
int f3r_sdainv_create_ex(f3r_ctx ctx, f3r_mat csr, double tol, double acc, int dtype, int flags,
                         f3r_prec *P);
int f3r_ilu0_create_ex(f3r_ctx ctx, f3r_mat csr, int32_t blocks, double acc, int dtype, int flags,
                       f3r_prec *P);
/* The level sets the device factorisations run in: level[i] for every row
 * (row i waits for the rows j < i with A[i][j] != 0 and, when `both`, with
 * A[j][i] != 0; SDAINV uses both, ILU(0) its lower triangle) and the number
 * of levels (= kernel launches).  Host only. */
int f3r_factor_levels(f3r_mat csr, int both, int32_t *level, int32_t *nlevels);
/* *where = 0: factors built on the host; 1: on the device; 2: the device
 * build was asked for and handed the rows back to the host. */
int f3r_prec_setup_where(f3r_prec P, int *where);
/* out = M in (IZW::operate ainv.hpp:110-114; ILU::operate ilu.hpp:57-61) */
int f3r_prec_apply(f3r_prec P, f3r_vec in, f3r_vec out);
/* factor sizes and a host copy of the factors (for tests):
 * SDAINV: which 0 = Z, 1 = W^T.  ILU(0): 0 = L (unit, 1 stored last), 1 = U. */
int f3r_prec_factor_nnz(f3r_prec P, int which, int64_t *nnz);
int f3r_prec_factor_get(f3r_prec P, int which, int32_t *rptr, int32_t *col, double *val);
int f3r_prec_destroy(f3r_prec P);

/* --------------------------------------------------------- the solver ---- */
typedef struct {
  int32_t m1;          /* outer FGMRES(m1), fp64 (work/src/f3r-gpu.cpp:264: 100) */
  int32_t m2, m3, m4;  /* level-2 / level-3 FGMRES steps, Richardson steps */
  int32_t c;           /* probe period of the adaptive weights */
  int32_t restart;     /* outer cycles (Restarted, restart.hpp:8-37) */
  double eps;          /* stop when |e_j| / ||b|| < eps */
  int32_t variant;     /* 0 HALF (fp16 innermost), 1 FLOAT, 2 DOUBLE */
  double range_scale;  /* power-of-two factor on the fp16 level's input; 0 = auto */
  int32_t C, sigma, K; /* SELL layout of the solver's matrices (0 = default: C 64, sigma 1,
                          K 8 for rows of 16 or more entries on average, else 4) */
  int32_t use_graph;   /* capture the inner levels in HIP graphs */
  int32_t flags;       /* bit 0: record the outer residual history; bit 1: run the SpMV
                          kernels with fused reductions on the capped grid too; bit 2:
                          build the matrix layouts on the device (as F3R_SELL_DEVICE; needs
                          A created with a context); bit 5: every SpMV with K threads per
                          row (F3R_SELL_TPR; needs index 0 and K >= 2; the sums are
                          reassociated, so level 4 is no longer bit-exact to CSR order) */
  int32_t grid;        /* blocks of the fused SpMV + reduction kernels; 0 = 8 per CU */
  int32_t index;       /* 0: a column index per slot (SELL); 1: chunks whose rows share their
                          column offsets store the offsets once (default; needs C >= 64) */
  double static_weight; /* > 0: Richardson with this fixed weight in every step and no probes
                           (work/src/static.cpp:99-119); 0 = adaptive weights (default) */
  int32_t depth;       /* 0: the F3R nest (default).  1..5: the nesting-depth ablations of
                          work/src/depth.cpp:122-167, in the HALF build, M applied directly
                          under FGMRES levels (basis storage / matrix values):
                          1 F2  fp32/fp32(64)
                          2 F2H fp16/fp16(64)
                          3 F3  fp32/fp32(8) > fp32/fp16(8)
                          4 F3H fp32/fp32(8) > fp16/fp16(8)
                          5 F4  fp32/fp32(8) > fp32/fp16(4) > fp16/fp16(2)
                          fp16 levels compute in fp32 with their vectors scaled by the
                          range factor; m2, m3, m4, c are not used */
} f3r_params;

typedef struct {
  int32_t is_solved;   /* solve_res_t (sparse/models.hpp:191-200) */
  int32_t res_iter;    /* outer Arnoldi steps summed over restarts */
  double res_nrm2;     /* |e| of the last outer step */
  double nrm_b;
  double solve_s;      /* wall time of the solve only (src/f3r-gpu.cpp:87-89) */
  int64_t precond_calls;
  int32_t outer_cycles;
} f3r_result;

/* fills the reference defaults: m = (100; 8, 4, 2), c = 64, restart 20,
 * eps 1e-8 (work/suite-gpu.py:49, work/src/f3r-gpu.cpp:18,264-265) */
int f3r_params_default(f3r_params *p);
/* Restarted<double>(A, FGMRES<double>(A, FGMRES<float>(A32, FGMRES<float>(
 * A16, Richardson(M))))) of work/src/f3r-gpu.cpp:210-266.  `A` is the CSR
 * matrix as solved (the caller applies f3r_csr_scaling first, as the
 * reference driver does).  P is the innermost preconditioner (SDAINV or
 * ILU(0)), built in the precision of the variant's innermost level. */
int f3r_solver_create(f3r_ctx ctx, f3r_mat A, f3r_prec P, const f3r_params *params,
                      f3r_solver *S);
/* solve A x = b from the given x (Restarted::solve, restart.hpp:26-36);
 * resets the adaptive weights first, as the reference driver does per
 * repetition (src/f3r-gpu.cpp:276-281). b and x are fp64 device vectors. */
int f3r_solve(f3r_solver S, f3r_vec b, f3r_vec x, f3r_result *res);
/* |e_j| / ||b|| of each outer step of the last solve (flags bit 0) */
int f3r_solver_history(f3r_solver S, double *hist, int32_t cap, int32_t *len);
/* *where = 0: the solver's matrix layouts were built on the host; 1: on the
 * device (flags bit 2, when it could run there).  No counterpart in the
 * reference. */
int f3r_solver_setup_where(f3r_solver S, int *where);
/* one application of level 2, 3 or 4 as an operator on fp64 vectors, for the
 * per-level error study */
int f3r_solver_level_apply(f3r_solver S, int32_t level, f3r_vec in, f3r_vec out);
/* mean device time per launch of one of the solver's kernels (which: SpMV
 * with 0 fp16 values and vectors, 1 fp16 values / fp32 vectors, 2 fp32, 3 fp64;
 * 4 the level-3 SpMV + CGS dots and 5 the CGS update of its last step; the
 * solve's own mixes of level-3 Arnoldi steps j = 0 .. m3 - 1: 6 the SpMV + CGS
 * dots, 7 the CGS update + norm + Givens; with diagonal SDAINV factors, 8 the
 * fused Richardson step and 9 the innermost entry), measured with HIP events
 * on the solver's stream over `reps` rounds of back-to-back launches; `bytes` =
 * the algorithmic bytes of one launch (each nonzero's value once, padding
 * excluded; an int32 index only for the nonzeros of int32-index chunks, since
 * chunk-offset chunks read none; each vector once), averaged over the mix.
 * which | F3R_TIME_COLD: caches cold, each launch timed alone after a pass over
 * 512 MiB that evicts the operands from the L2s and the last-level cache. */
#define F3R_TIME_COLD 0x100
int f3r_solver_time_kernel(f3r_solver S, int32_t which, int32_t reps, double *ms,
                           double *bytes);
int f3r_solver_destroy(f3r_solver S);

/* --------------------------------------------------- composed nests ---- */
/* The reference composes its solvers from objects: FGMRES<T>(A, M, {m,
 * normalize}) levels whose M is another level, a preconditioner, or an
 * Operator(shape, lambda) (sparse/models.hpp:46-126, solver/gmres.hpp:56-135),
 * under Restarted<double>(A, FGMRES<double>(A, ..., {m1, true}), {restart})
 * (solver/restart.hpp:8-37).  f3r_nest_create builds such a chain.
 *
 * A caller's operator: out = Op(in) on non-owning views of n elements of
 * `dtype` on the context's stream (the callback may call any f3r_* function
 * of that context on them, e.g. f3r_spmv, f3r_prec_apply, f3r_vec_axpy).  An
 * F3R_F16 operator receives vectors multiplied by the solver's power-of-two
 * range factor (f3r_params.range_scale), as the fp16 preconditioner does, and
 * must be linear; F3R_F32 / F3R_F64 operators see the level's values.
 * Returns F3R_OK or an error status, which the solve returns.
 * replaces: Operator<L>(shape, lambda) (sparse/models.hpp:46-126) */
typedef int (*f3r_apply_fn)(void *user, f3r_vec in, f3r_vec out);
typedef struct f3r_op_s *f3r_op;
int f3r_op_create(f3r_ctx ctx, int32_t n, int dtype, f3r_apply_fn fn, void *user, f3r_op *op);
int f3r_op_destroy(f3r_op op);

/* one FGMRES<T>(A_level, M, {m, false}) level: basis storage F3R_F32
 * (FGMRES<float>) or F3R_F16 (FGMRES<_Float16>: fp16 storage, fp32
 * arithmetic), the level's matrix values F3R_F32 or F3R_F16 */
typedef struct {
  int32_t m;
  int32_t basis;
  int32_t matrix;
} f3r_level;
/* the innermost operator of a nest */
enum { F3R_INNER_PREC = 0,        /* M applied directly (the nests of work/src/depth.cpp) */
       F3R_INNER_RICHARDSON = 1,  /* the adaptive (or static-weight) fp16 Richardson of
                                     work/src/f3r-gpu.cpp:229-260 with M, m4 steps, probe
                                     period c; the level above it must have an fp32 basis */
       F3R_INNER_OP = 2 };        /* the caller's operator */
typedef struct {
  int32_t kind;
  f3r_prec P;            /* PREC, RICHARDSON */
  f3r_op op;             /* OP */
  int32_t m4, c;         /* RICHARDSON */
  double static_weight;  /* RICHARDSON: > 0 fixed weight, no probes */
} f3r_inner;
/* Restarted<double>(A, FGMRES<double>(A, L_0(L_1(... L_{k-1}(inner))), {m1,
 * true}), {restart}) with k = nlevels in [1, 4] levels of m in [1, 64]; from
 * params: m1, restart, eps, range_scale and the layout (C, sigma, K, index,
 * flags).  The composition of f3r_solver_create (levels {m2, F32, F32},
 * {m3, F32, F16} over the Richardson) runs f3r_solver_create's solver, bit for
 * bit; any other composition runs the generic levels (no HIP graphs). */
int f3r_nest_create(f3r_ctx ctx, f3r_mat A, const f3r_level *levels, int32_t nlevels, const f3r_inner *inner,
                    const f3r_params *params, f3r_solver *S);
/* Solver::solve(b, x, cond) (sparse/models.hpp:191-267): stop when
 * cond(user, i, |e_i|) returns nonzero, called at the start of every outer
 * cycle (i = 0) and after every outer step i, instead of |e| / ||b|| < eps */
typedef int (*f3r_cond_fn)(void *user, int32_t iter, double res_nrm);
int f3r_solve_cond(f3r_solver S, f3r_vec b, f3r_vec x, f3r_cond_fn cond, void *user, f3r_result *res);

/* ----------------------------------------------------- batched solves ---- */
/* Many independent systems (each right-hand side its own segment) solved in
 * lockstep: every launch covers all segments.  Segment s is A[s] with the
 * SDAINV preconditioner P[s] (fp16, any context: only its host factors are
 * read); a system solved for k right-hand sides appears k times.  The
 * segments are laid out as one block-diagonal matrix with each segment
 * padded to whole 2048-row tiles; all reductions, Hessenberg/Givens updates
 * and Richardson weights are per segment.  With use_graph each outer step
 * of all segments is one HIP graph launch, after which the host reads one
 * array of residual norms.  The nest and its per-segment semantics
 * are f3r_solver_create's HALF build (params: m1..m4, c, restart, eps, C, K,
 * index, range_scale; 0 = the per-segment automatic range factor; flags bit
 * 4: the SpMVs with K threads per row, which needs index 0 and K >= 2).  The
 * reference solves one system per run (work/src/f3r-gpu.cpp:44-288); this
 * has no counterpart there. */
typedef struct f3r_batch_s *f3r_batch;
int f3r_batch_create(f3r_ctx ctx, int32_t nseg, const f3r_mat *A, const f3r_prec *P, const f3r_params *params,
                     f3r_batch *B);
/* b[s], x[s]: fp64 vectors of segment s's size (x is the initial guess and
 * the solution); res[s]: per-segment result, solve_s = the whole batch */
int f3r_batch_solve(f3r_batch B, const f3r_vec *b, f3r_vec *x, f3r_result *res);
/* padded rows of all segments, and the segment count */
int f3r_batch_info(f3r_batch B, int64_t *rows, int32_t *segments);
int f3r_batch_destroy(f3r_batch B);

/* record copies of the innermost level's input vectors during later solves:
 * every `every`-th call, up to `max_vectors` (0 stops and frees).  Level 2
 * runs without HIP graphs while vectors are being recorded.  These are the
 * probe vectors of the learned correction below; the reference has no
 * counterpart (its Richardson operator receives the same vectors, src/f3r-gpu.cpp:229-231). */
int f3r_solver_capture(f3r_solver S, int32_t max_vectors, int32_t every);
/* the recorded vectors as fp32, count x n, in the matrix's row order; `out`
 * may be NULL to query the count */
int f3r_solver_captured(f3r_solver S, int32_t *count, float *out);

/* ---- learned correction of SDAINV (no counterpart in the reference) ----
 * A per-row MLP 12 -> 64 -> 64 -> 1 (ReLU; bf16 weights, fp32 accumulation
 * on MFMA) maps row features to y_i; the corrected preconditioner is
 * M = Z diag(s) W^T with s_i = exp(clamp(y_i, -3, 3)), SDAINV's factors
 * (ainv.hpp:57-64) with the columns of Z rescaled.  Training minimises a
 * residual loss over probe vectors (f3r_lsys_loss). */
#define F3R_MLP_NIN 16   /* feature columns per row; columns 12..15 are zero */
#define F3R_MLP_NFEAT 12
typedef struct f3r_mlp_s *f3r_mlp;
typedef struct f3r_lsys_s *f3r_lsys;
int f3r_mlp_nparams(int64_t *n);
/* He-uniform hidden layers from `seed`, last layer zero (M = SDAINV) */
int f3r_mlp_create(f3r_ctx ctx, uint64_t seed, f3r_mlp *m);
int f3r_mlp_destroy(f3r_mlp m);
/* fp32 master parameters: W1 [64][16], b1 [64], W2 [64][64], b2 [64],
 * w3 [64], b3 (row-major, out x in); set also clears the Adam moments */
int f3r_mlp_get(f3r_mlp m, float *params);
int f3r_mlp_set(f3r_mlp m, const float *params);
/* y = MLP(feat) for n rows of F3R_MLP_NIN host floats (rounded to bf16) */
int f3r_mlp_forward(f3r_mlp m, int64_t n, const float *feat, float *y);
int f3r_mlp_grad_zero(f3r_mlp m);
/* gradient += d(sum_r gy_r y_r)/d(params) */
int f3r_mlp_backward(f3r_mlp m, int64_t n, const float *feat, const float *gy);
int f3r_mlp_grad_get(f3r_mlp m, float *grad);
/* one Adam step on the accumulated gradient (the gradient is kept) */
int f3r_mlp_adam(f3r_mlp m, float lr, float beta1, float beta2, float eps);
/* the 12 features of every row (n x F3R_MLP_NIN) from the scaled CSR matrix,
 * its SDAINV preconditioner and, optionally, the unscaled diagonal */
int f3r_learn_features(f3r_mat A, f3r_prec P, const double *diag, float *feat);
/* a training system: A, A^T, Z, Z^T, W^T in fp32 SELL and bf16 features */
int f3r_lsys_create(f3r_ctx ctx, f3r_mat A, f3r_prec P, const float *feat, f3r_lsys *L);
int f3r_lsys_destroy(f3r_lsys L);
/* k probe vectors, k x n host floats */
int f3r_lsys_set_probes(f3r_lsys L, int32_t k, const float *probes);
/* the mean over the probes z of, with q = A M z,
 *   kind 0: ||z - q||^2 / ||z||^2
 *   kind 1: min_w ||z - w q||^2 / ||z||^2 = 1 - <z,q>^2 / (<z,z><q,q>)
 * (kind 1 ignores a global scale of M, which the adaptive Richardson weights
 * absorb); with grad != 0 its gradient is added to the MLP's */
int f3r_lsys_loss(f3r_lsys L, f3r_mlp m, int32_t kind, int32_t grad, double *loss);
/* s = exp(clamp(MLP(features))) for every row; with normalize != 0 divided
 * by its geometric mean (for models trained with the scale-blind loss) */
int f3r_lsys_scales(f3r_lsys L, f3r_mlp m, int32_t normalize, float *s);
/* a new SDAINV preconditioner with Z's column i scaled by s_i (s > 0) */
int f3r_sdainv_rescale(f3r_ctx ctx, f3r_prec P, const float *s, f3r_prec *out);

#ifdef __cplusplus
}
#endif

#endif
