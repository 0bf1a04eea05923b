// Internal declarations shared by the HIP kernels, the host set-up code and
// the C ABI of libf3r_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "f3r_amd.h"

namespace f3r {

// ------------------------------------------------------------- errors -----
void set_error(const std::string &msg);
struct Error {
  int status;
  std::string msg;
};

#define F3R_HIP(call)                                                               \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess)                                                           \
      throw ::f3r::Error{F3R_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)}; \
  } while (0)
#define F3R_CHECK(cond, status, msg)                                                \
  do {                                                                              \
    if (!(cond)) throw ::f3r::Error{status, msg};                                   \
  } while (0)

inline size_t dtype_size(int dt) {
  switch (dt) {
    case F3R_F64: return 8;
    case F3R_F32: return 4;
    case F3R_F16: return 2;
    case F3R_I32: return 4;
    case F3R_BF16: return 2;
  }
  return 0;
}


// ------------------------------------------------------ device memory -----
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  void alloc(size_t b);
  ~DevBuf();
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  template <class T> T *as() const { return static_cast<T *>(p); }
};
using DevPtr = std::shared_ptr<DevBuf>;
inline DevPtr dev_alloc(size_t bytes) { return std::make_shared<DevBuf>(bytes); }

// Scratch of the single-pass grid reductions: per-block partial sums and
// arrival counters.  Counters are reset by the last block of each launch, so
// no memset is ever recorded (also under graph capture).
struct Reducer {
  static constexpr int64_t kPartialDoubles = 1 << 20;  // 8 MiB of partial sums
  DevPtr partial;
  static constexpr int kSlots = 24;                     // reduction slots of kCounterSlot counters
  DevPtr counters;
  DevPtr scalars;   // 64 doubles of results
};

struct Context {
  ~Context();
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 256;
  Reducer red;
  double *pinned = nullptr;  // host-pinned landing buffer for small reads
  void *na_scratch = nullptr;  // device address of f3r_vec_dot_noalloc's scratch on this device
};

struct Vec {
  std::shared_ptr<Context> ctx;
  int dtype = F3R_F64;
  int64_t n = 0;
  DevPtr buf;         // owner (shared with views)
  int64_t offset = 0; // element offset into buf
  void *ptr() const { return static_cast<char *>(buf->p) + offset * dtype_size(dtype); }
};

// Host CSR (fp64 values), int32 indices.
struct HostCSR {
  int32_t n = 0, ncols = 0;
  std::vector<int32_t> rptr, idx;
  std::vector<double> val;
  int64_t nnz() const { return (int64_t)idx.size(); }
};

// SELL-C-sigma with K-slot interleave on the device.  Chunk s holds rows
// [s*C, s*C + tc) of the sorted order (tc = C except for the last chunk);
// slot group g of row j of the chunk occupies elements
//   cptr[s] + (g * tc + j) * K + k,  k < K
// (K = 1 gives the reference's column-major layout, sell.hpp:272-287).
// A chunk-offset chunk (below) may end in a half group: with bit kHalfTail
// of cwidth[s] set, K / 2 more slots follow its G full groups, K / 2 per
// row: slot G * K + k of row j at cptr[s] + G * K * tc + j * (K / 2) + k.
// A 27-entry stencil row then takes 28 slots with K = 8, not 32.
// Row r of the sorted order is row perm[r] of the original matrix; the
// column indices stay in the original numbering unless `internal` is set, in
// which case they are renumbered into the sorted order.
#ifndef F3R_HALF_GROUPS
#define F3R_HALF_GROUPS 1
#endif
constexpr int32_t kHalfTail = 1 << 30;
__host__ __device__ inline int32_t chunk_groups(int32_t cw) { return cw & (kHalfTail - 1); }
__host__ __device__ inline int32_t chunk_slots(int32_t cw, int32_t K) {
  return chunk_groups(cw) * K + ((cw & kHalfTail) ? K / 2 : 0);
}
// width entry of an offset chunk with `width` slots in use: full groups, and
// a half group when the rest fits in K / 2 slots
__host__ __device__ inline int32_t offset_chunk_width(int32_t width, int32_t K) {
  const int32_t g = width / K, rest = width - g * K;
  if (rest == 0) return g;
  if (F3R_HALF_GROUPS && K >= 2 && rest <= K / 2) return g | kHalfTail;
  return g + 1;
}
// element of slot t of row j in a chunk starting at c0 with tc rows
__host__ __device__ inline int64_t chunk_slot_pos(int64_t c0, int32_t cw, int32_t tc, int32_t j, int32_t t, int32_t K) {
  const int32_t g = chunk_groups(cw);
  if (t < g * K) return c0 + ((int64_t)(t / K) * tc + j) * K + (t % K);
  return c0 + (int64_t)g * K * tc + (int64_t)j * (K / 2) + (t - g * K);
}

struct SellShared {  // index data shared by all precisions of one matrix
  int32_t n = 0, C = 64, sigma = 1, K = 1, nchunks = 0;
  // lanes per row of the standalone SpMV: 1 (every slot of a row on one
  // lane, CSR order) or K (F3R_SELL_TPR: lane k of a row takes slot k of
  // every group, then the K lanes add their sums)
  int32_t lanes = 1;
  int64_t nslots = 0, nnz = 0;
  bool internal = false;
  DevPtr cptr, cwidth, idx, perm;  // perm null when identity
  DevPtr iptr, coff;               // chunk-offset format (null when not used)
  int32_t offset_chunks = 0;
  int64_t offset_slots = 0, explicit_nnz = 0;
  int32_t off_bias = 0;
  // LDS-staged x gathers (F3R_SELL_LDS): per block of kBlock sorted rows, the
  // column window (first column, length) its slots read; null when not used
  DevPtr bwin;
  int32_t lds_cap = 0, lds_blocks = 0;  // window capacity in elements of x; blocks that fit
  int32_t lds_fit[4] = {0, 0, 0, 0};    // by log2(sizeof x) - 1: the widest window that fits
  // algorithmic bytes of one SpMV: every nonzero's value once (no padding),
  // an int32 index only for the nonzeros of int32-index chunks (offset
  // chunks read none), x and y once
  double algo_bytes(int sv, int sx, int sy) const {
    return (double)(nnz - explicit_nnz) * sv + (double)explicit_nnz * (sv + 4) + (double)n * (sx + sy);
  }
  std::vector<int32_t> h_perm;     // host copy of the permutation
};

struct Mat {
  std::shared_ptr<Context> ctx;
  int format = 0;  // 0 = CSR (fp64), 1 = SELL
  int dtype = F3R_F64;
  // CSR
  std::shared_ptr<HostCSR> host;
  DevPtr d_rptr, d_idx, d_val;  // d_val: fp64 (set-up reads it)
  DevPtr c_val;                 // the values in `dtype` for SpMV when dtype is not fp64
  int32_t tpr = 0;              // CSR SpMV lanes per row (0 = by the mean row length)
  // SELL
  std::shared_ptr<SellShared> sh;
  DevPtr s_val;
  std::shared_ptr<std::vector<double>> sell_host_val;  // fp64 slot values, for casts
  DevPtr sell_dev_val64;  // fp64 slot values on the device (layouts built there)
};

// ----------------------------------------------------- host set-up API ----
void csr_scaling(int32_t n, const int32_t *rptr, const int32_t *idx, double *val);
std::vector<int32_t> sigma_perm(int32_t n, const int32_t *rptr, int32_t sigma);
int64_t sell_nslots(int32_t n, const int32_t *rptr, int32_t C, int32_t sigma, int32_t K);
struct HostSell {
  int32_t n, C, K, nchunks;
  std::vector<int32_t> cptr, cwidth, idx, perm;
  std::vector<double> val;
  // chunk-offset index format (empty when not used): iptr[s] >= 0 is the
  // first column index of explicit chunk s in `idx`; iptr[s] < 0 marks an
  // offset chunk whose slot t of row r reads column r + coff[-iptr[s] - 1 + t]
  std::vector<int32_t> iptr, coff;
  int32_t offset_chunks = 0;
  int64_t offset_slots = 0;  // value slots of offset chunks (read in full)
  int32_t off_bias = 0;      // -(smallest offset of any offset chunk), >= 0
  int64_t explicit_nnz = 0;  // nonzeros of int32-index chunks
};
// Builds the layout; when `renumber` the columns are mapped through the
// inverse permutation (the matrix P A P^T in sorted order).  With `offsets`
// (and C >= 64) every chunk whose rows share a small set of column offsets
// col - row, all in range and in CSR order, stores that set once instead of
// a column index per slot.
HostSell build_sell(const HostCSR &A, int32_t C, int32_t sigma, int32_t K,
                    const std::vector<int32_t> *perm_override, bool renumber, bool offsets = false);
struct HostFactors {
  HostCSR f[2];
};
HostFactors sdainv_host(const HostCSR &A, double tol, double acc);
HostFactors ilu0_host(const HostCSR &A, int32_t blocks, double acc);
HostCSR csr_transpose(const HostCSR &A);
// pieces of ilu0_host: the block-diagonal part of A, the position of each
// row's diagonal entry in it, and the split of the eliminated rows into L, U
HostCSR ilu0_restrict(const HostCSR &A, int32_t blocks);
std::vector<int32_t> ilu0_diag_pos(const HostCSR &R);
HostFactors ilu0_split(const HostCSR &R);
HostFactors ilu0_split(const HostCSR &R, const double *val);  // R's pattern, these values
// Level sets of a row-by-row factorisation: row i waits for every row j < i
// with A[i][j] != 0 and, when `both`, with A[j][i] != 0.  rows[lptr[l] ..
// lptr[l + 1]) lists the rows of level l in increasing order; returns the
// number of levels.
int32_t factor_levels(const HostCSR &A, bool both, std::vector<int32_t> &lptr, std::vector<int32_t> &rows);
// fp64 -> fp16 bits, correctly rounded (round to nearest even)
uint16_t d2h_bits(double d);

// ------------------------------------------------------ device helpers ----
std::shared_ptr<SellShared> upload_sell_shared(Context *ctx, const HostSell &hs, bool internal);
DevPtr upload_sell_values(Context *ctx, const HostSell &hs, int dtype);

}  // namespace f3r
