// Declarations shared between the translation units of libf3r_amd.
#pragma once

#include "common.hpp"
#include "sell_kernel.hpp"

namespace f3r {

// Innermost preconditioner: SDAINV (M = Z D^-1 W^T as two SELL SpMVs, or two
// diagonal scalings when both factors are diagonal) or ILU(0) (level-scheduled
// triangular solves).  Values in `dtype`, sums in fp32 for fp16 / fp32 data.
struct Prec {
  std::shared_ptr<Context> ctx;
  int kind = 0;  // 0 SDAINV, 1 ILU(0)
  int dtype = F3R_F16;
  double tol = 0.1, acc = 1.0;
  int32_t n = 0;
  int where = 0;  // factors built: 0 host, 1 device, 2 device asked, host ran
  HostFactors host;
  DevPtr tmp;
  // SDAINV
  std::shared_ptr<Mat> W, Z;  // W holds W^T (rows w_i), Z holds Z D^-1
  bool diag = false;
  DevPtr dw, dz;              // diagonals when `diag`
  DevPtr dwz;                 // fp16 only: (dw[i], dz[i]) interleaved, one 4-byte load per row
  // ILU(0)
  std::vector<int32_t> l_lptr, u_lptr;  // level pointers into *_rows
  DevPtr l_rows, u_rows, l_rptr, l_idx, l_val, u_rptr, u_idx, u_val, u_dinv;
  // block-Jacobi split with at least kBjBlockKernel blocks: rows grouped by
  // block, one workgroup per block (bj_groups of them)
  int32_t bj_blocks = 0, bj_groups = 0;
  DevPtr lb_rows, lb_ptr, lb_start, ub_rows, ub_ptr, ub_start;
  // fused L+U solve with the block's vector in LDS (bj_R entries per record; 0: not used)
  int32_t bj_R = 0, bj_threads = 1024;
  size_t bj_lds = 0;
  DevPtr lr_row, lr_cnt, lr_col, lr_val, ur_row, ur_cnt, ur_col, ur_val, ur_dinv, bj_b0, bj_bn;
};
// LDS the fused block-Jacobi kernel may take for a block's vector
constexpr size_t kBjLdsBytes = 64 * 1024;
// fewer diagonal blocks than this: one launch per level instead (a block
// per workgroup would leave most of the GPU idle)
constexpr int32_t kBjBlockKernel = 32;

std::shared_ptr<Context> ctx_create(int device);
double read_scalar(Context *c, const double *dev);
SellArgs sell_args(const SellShared &sh, bool with_perm);
DevPtr upload_values(Context *ctx, const std::vector<double> &v, int dtype);
HostCSR permute_csr(const HostCSR &M, const std::vector<int32_t> &perm);
void prec_build_device(Prec *P, const std::vector<int32_t> *perm, int32_t C, int32_t K, bool offsets = true);
void prec_apply_raw(Prec *P, const void *in, void *out, hipStream_t st);
void vec_copy(Vec *s, Vec *d);
// dot product of two device arrays (fp16 data summed in fp32), host result
double dot_raw(Context *c, int da, const void *a, int db, const void *b, int64_t n);

#define F3R_API_BEGIN try {
#define F3R_API_END                                   \
  }                                                   \
  catch (const f3r::Error &e) {                       \
    f3r::set_error(e.msg);                            \
    return e.status;                                  \
  }                                                   \
  catch (const std::bad_alloc &) {                    \
    f3r::set_error("host allocation failed");         \
    return F3R_ERR_ALLOC;                             \
  }                                                   \
  catch (const std::exception &e) {                   \
    f3r::set_error(e.what());                         \
    return F3R_ERR_ARG;                               \
  }                                                   \
  return F3R_OK;

}  // namespace f3r

struct f3r_ctx_s;
struct f3r_vec_s;
struct f3r_mat_s;
struct f3r_prec_s;
namespace f3r {
std::shared_ptr<Context> ctx_of(f3r_ctx c);
Vec *vec_of(f3r_vec v);
std::shared_ptr<Mat> mat_of(f3r_mat m);
std::shared_ptr<Prec> prec_of(f3r_prec p);
f3r_prec wrap_prec(std::shared_ptr<Prec> p);  // new handle owning a reference
f3r_vec wrap_vec(const Vec &v);                // new handle of a vector (view)
void release_vec(f3r_vec v);
// device build of a SELL-C-sigma layout from a device CSR (build.hip), the
// columns renumbered into the sorted order when `renumber`; false (nothing
// built) when a chunk or a sigma window is too long for the on-chip sorts
bool build_sell_device(Context *c, const HostCSR &A, const int32_t *d_rptr, const int32_t *d_col,
                       const double *d_val, int32_t C, int32_t sigma, int32_t K, bool offsets, bool renumber,
                       std::shared_ptr<SellShared> &sh, DevPtr &sval);
// device factorisations (factor.hip), bit-identical to sdainv_host /
// ilu0_host; false (F untouched) when a row outgrows the device capacity
bool sdainv_device(Context *c, const HostCSR &A, const int32_t *d_rptr, const int32_t *d_col, const double *d_val,
                   double tol, double acc, HostFactors &F);
bool ilu0_device(Context *c, const HostCSR &A, const int32_t *d_rptr, const int32_t *d_col, const double *d_val,
                 int32_t blocks, double acc, HostFactors &F);
// fp64 slot values on the device -> dtype (correctly rounded)
DevPtr cast_values_device(Context *c, const DevPtr &v64, int64_t n, int dtype);
}  // namespace f3r
