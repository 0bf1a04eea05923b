// Runtime objects of libf3r_amd: contexts (one HIP stream each), device
// vectors, CSR / SELL-C-sigma matrices, the BLAS-1 kernels with single-pass
// grid reductions, and the SDAINV / ILU(0) preconditioner objects.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "internal.hpp"
#include "sell_kernel.hpp"

namespace f3r {

// ------------------------------------------------------------- errors ----
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

void DevBuf::alloc(size_t b) {
  bytes = b;
  if (b == 0) {
    p = nullptr;
    return;
  }
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    throw Error{F3R_ERR_ALLOC, "hipMalloc(" + std::to_string(b) + "): " + hipGetErrorString(e)};
  }
}
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

// ------------------------------------------------------------ kernels ----
template <class T> __global__ void k_fill(T *__restrict__ p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = v;
}

template <class S, class D> __global__ void k_cast(const S *__restrict__ s, D *__restrict__ d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    d[i] = cvt<D>(to_acc(s[i]));
}

template <class T> __global__ void k_axpy(int64_t n, double alpha, const T *__restrict__ x, T *__restrict__ y) {
  using A = typename Acc<T>::type;
  const A a = (A)alpha;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = cvt<T>(fma_<A>(a, to_acc(x[i]), to_acc(y[i])));
}

template <class T> __global__ void k_scal(int64_t n, double alpha, T *__restrict__ x) {
  using A = typename Acc<T>::type;
  const A a = (A)alpha;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    x[i] = cvt<T>(to_acc(x[i]) * a);
}

template <class A, class T1, class T2>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const T1 *__restrict__ a,
                                                const T2 *__restrict__ b, A *partial,
                                                unsigned *counter, double *out) {
  __shared__ A lds[kWavesPerBlock];
  A v[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    v[0] = fma_<A>((A)to_acc(a[i]), (A)to_acc(b[i]), v[0]);
  if (grid_sum<A, 1>(v, partial, counter, lds) && threadIdx.x == 0) *out = (double)v[0];
}

// CSR SpMV (CSR<T>::operate / compute_residual, csr.hpp:32-50): y = A x, or
// y = b - A x with b.  TPR consecutive lanes take a row: lane q sums entries
// q, q + TPR, ... in order, and the TPR sums are added by a butterfly over
// the lanes, so every lane's loads of val and idx are coalesced with its
// neighbours'.  TPR = 1 sums each row in CSR order (bit-identical to the SELL
// kernels and the oracle); TPR > 1 reassociates the sum.  Sums in fp64 for
// fp64 values, else fp32.
template <class V, class X, class Y, class ACC, int TPR>
__global__ __launch_bounds__(kBlock) void k_csr(int32_t n, const int32_t *__restrict__ rptr,
                                                const int32_t *__restrict__ idx, const V *__restrict__ val,
                                                const X *__restrict__ x, Y *__restrict__ y,
                                                const Y *__restrict__ b) {
  constexpr int kRowsPerBlock = kBlock / TPR;
  const int q = threadIdx.x % TPR;
  for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / TPR; row < n;
       row += (int64_t)gridDim.x * kRowsPerBlock) {
    ACC acc = 0;
    const int32_t k1 = rptr[row + 1];
    for (int32_t k = rptr[row] + q; k < k1; k += TPR)
      acc = fma_<ACC>((ACC)to_acc(val[k]), (ACC)to_acc(x[idx[k]]), acc);
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (q == 0) y[row] = b ? cvt<Y>((ACC)to_acc(b[row]) - acc) : cvt<Y>(acc);
  }
}

// ILU(0) triangular solves by levels: each launch solves one level of rows
// in parallel, every row in the order of the reference's block-Jacobi solve
// (trsv/bj.hpp:58-98): L forward with the unit diagonal stored last, U
// backward from the largest column, times the inverted diagonal.
template <class S>
__global__ void k_trsv_l(int32_t cnt, const int32_t *__restrict__ rows, const int32_t *__restrict__ rptr,
                         const int32_t *__restrict__ idx, const S *__restrict__ val,
                         const S *__restrict__ in, S *__restrict__ out) {
  using A = typename Acc<S>::type;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  A s = to_acc(in[i]);
  for (int k = rptr[i]; k < rptr[i + 1] - 1; ++k) s = fma_<A>(-to_acc(val[k]), to_acc(out[idx[k]]), s);
  out[i] = cvt<S>(s);
}
template <class S>
__global__ void k_trsv_u(int32_t cnt, const int32_t *__restrict__ rows, const int32_t *__restrict__ rptr,
                         const int32_t *__restrict__ idx, const S *__restrict__ val,
                         const S *__restrict__ dinv, const S *__restrict__ in, S *__restrict__ out) {
  using A = typename Acc<S>::type;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int r = rows[t];
  A s = to_acc(in[r]);
  for (int k = rptr[r + 1] - 1; k > rptr[r]; --k) s = fma_<A>(-to_acc(val[k]), to_acc(out[idx[k]]), s);
  out[r] = cvt<S>(s * to_acc(dinv[r]));
}

// The same solves for a block-Jacobi split: one workgroup per diagonal block
// walks that block's levels with a barrier between them (the blocks never
// read each other's rows), so a solve is one launch.  brows lists the rows
// block by block, level by level; block b's levels start at
// blptr[bstart[b] .. bstart[b + 1] - 1] (the last entry ends the block).
template <class S>
__global__ __launch_bounds__(1024) void k_bj_trsv_l(const int32_t *__restrict__ bstart, const int32_t *__restrict__ blptr,
                                                    const int32_t *__restrict__ brows, const int32_t *__restrict__ rptr,
                                                    const int32_t *__restrict__ idx, const S *__restrict__ val,
                                                    const S *__restrict__ in, S *__restrict__ out) {
  using A = typename Acc<S>::type;
  const int32_t q0 = bstart[blockIdx.x], q1 = bstart[blockIdx.x + 1];
  for (int32_t q = q0; q + 1 < q1; ++q) {
    for (int32_t t = blptr[q] + threadIdx.x; t < blptr[q + 1]; t += blockDim.x) {
      const int i = brows[t];
      A s = to_acc(in[i]);
      for (int k = rptr[i]; k < rptr[i + 1] - 1; ++k) s = fma_<A>(-to_acc(val[k]), to_acc(out[idx[k]]), s);
      out[i] = cvt<S>(s);
    }
    __syncthreads();
  }
}
template <class S>
__global__ __launch_bounds__(1024) void k_bj_trsv_u(const int32_t *__restrict__ bstart, const int32_t *__restrict__ blptr,
                                                    const int32_t *__restrict__ brows, const int32_t *__restrict__ rptr,
                                                    const int32_t *__restrict__ idx, const S *__restrict__ val,
                                                    const S *__restrict__ dinv, const S *__restrict__ in,
                                                    S *__restrict__ out) {
  using A = typename Acc<S>::type;
  const int32_t q0 = bstart[blockIdx.x], q1 = bstart[blockIdx.x + 1];
  for (int32_t q = q0; q + 1 < q1; ++q) {
    for (int32_t t = blptr[q] + threadIdx.x; t < blptr[q + 1]; t += blockDim.x) {
      const int r = brows[t];
      A s = to_acc(in[r]);
      for (int k = rptr[r + 1] - 1; k > rptr[r]; --k) s = fma_<A>(-to_acc(val[k]), to_acc(out[idx[k]]), s);
      out[r] = cvt<S>(s * to_acc(dinv[r]));
    }
    __syncthreads();
  }
}

// The same two solves fused in one launch with the block's vector in LDS:
// xs = in (the block's rows), the L levels in place, then the U levels in
// place, then out = xs.  Each row's entries are stored level by level as
// structure-of-arrays records (local column, value, in the order the solves
// above visit them; the U record also carries 1/u_ii), so the next level's
// records are loaded before the barrier that ends the current one and only
// LDS reads remain on the path from level to level.  Same arithmetic as
// k_trsv_l / k_trsv_u.
template <class S> struct BjRec {
  const int32_t *start, *lptr;  // per block: first level pointer; level -> slot range
  const int32_t *row, *cnt, *col;  // per slot: local row, entries, columns (R x slots)
  const S *val, *dinv;             // values (R x slots), 1/u_ii (U only)
  int32_t slots;
};
template <class S, int R> struct BjSlot {
  int32_t row, cnt;
  int32_t col[R];
  S val[R];
  S dinv;
};
// Loads the record of this thread's slot in level q (row -1: no slot, or q
// past the block's last level).  Every call issues the same loads (indices
// clamped into range; entries past the row's count are zero padding), so the
// compiler can count them and wait for exactly the oldest ones.
typedef const __attribute__((address_space(4))) int32_t *bj_const_i32;  // uniform read-only: scalar loads

template <class S, int R>
__device__ __forceinline__ void bj_load(const BjRec<S> &r, int32_t q, int32_t q1, bool upper, BjSlot<S, R> &sl) {
  const bool lvl = q + 1 < q1;
  const int32_t qq = __builtin_amdgcn_readfirstlane(lvl ? q : q1 - 2);
  // the level pointers by scalar loads: the vector-load counter then holds
  // only the records, and waiting for the oldest does not wait for the rest
  const bj_const_i32 lp = (bj_const_i32)r.lptr;
  const int32_t t0 = lp[qq] + (int32_t)threadIdx.x, te = lp[qq + 1];
  const bool on = lvl && t0 < te;
  const int32_t t = on ? t0 : te - 1;
  sl.row = on ? r.row[t] : -1;
  sl.cnt = r.cnt[t];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    sl.col[k] = r.col[(int64_t)k * r.slots + t];
    sl.val[k] = r.val[(int64_t)k * r.slots + t];
  }
  sl.dinv = upper ? r.dinv[t] : S(0);
}
// Barrier for LDS traffic only: this wave's LDS operations complete, then
// s_barrier.  __syncthreads() would also wait for every outstanding global
// load (its release fence), i.e. for the prefetched records below.  The
// memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The records of the next D levels stay in flight: level q + D is requested
// as level q finishes, so its latency overlaps D levels of work.  The loop
// is unrolled by D so the ring is indexed statically (registers).
template <class S, int R>
__device__ __forceinline__ void bj_levels(const BjRec<S> &r, bool upper, S *xs) {
  using A = typename Acc<S>::type;
  constexpr int D = R <= 4 ? 4 : R <= 8 ? 3 : 2;
  const bj_const_i32 st = (bj_const_i32)r.start;
  const int32_t q0 = st[blockIdx.x], q1 = st[blockIdx.x + 1];
  BjSlot<S, R> ring[D];
#pragma unroll
  for (int d = 0; d < D; ++d) bj_load<S, R>(r, q0 + d, q1, upper, ring[d]);
  // straight-line steps; the at most D - 1 steps past the last level have
  // no slots and only meet the barrier (the same count for every thread)
  for (int32_t q = q0; q + 1 < q1; q += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      BjSlot<S, R> &cur = ring[d];
      const int32_t row = cur.row;
      A s = to_acc(xs[row >= 0 ? row : 0]);
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (k < cur.cnt) s = fma_<A>(-to_acc(cur.val[k]), to_acc(xs[cur.col[k]]), s);
      const S res = upper ? cvt<S>(s * to_acc(cur.dinv)) : cvt<S>(s);
      bj_load<S, R>(r, q + d + D, q1, upper, cur);
      if (row >= 0) xs[row] = res;
      lds_barrier();
    }
  }
}

template <class S, int R>
__global__ __launch_bounds__(1024) void k_bj_lds(const int32_t *__restrict__ b0, const int32_t *__restrict__ bn,
                                                 BjRec<S> L, BjRec<S> U, const S *__restrict__ in,
                                                 S *__restrict__ out) {
  extern __shared__ unsigned char bj_smem[];
  S *xs = reinterpret_cast<S *>(bj_smem);
  const int32_t r0 = b0[blockIdx.x], nr = bn[blockIdx.x];
  for (int32_t i = threadIdx.x; i < nr; i += blockDim.x) xs[i] = in[r0 + i];
  __syncthreads();
  bj_levels<S, R>(L, false, xs);
  bj_levels<S, R>(U, true, xs);
  for (int32_t i = threadIdx.x; i < nr; i += blockDim.x) out[r0 + i] = xs[i];
}

static int grid_for(int64_t n, int cu) {
  int64_t g = (n + kBlock - 1) / kBlock;
  int64_t cap = (int64_t)cu * 16;
  if (g > cap) g = cap;
  return g < 1 ? 1 : (int)g;
}

// ------------------------------------------- allocation-free dot scratch ---
// f3r_vec_dot_noalloc keeps the reduction's arrival counters and per-block
// partial sums in this file-scope array instead of the context's buffers:
// [counters of one reduction slot | partials of up to kNaMaxBlocks blocks and
// their groups].  Each GPU the library runs on holds its own copy until the
// process ends.  Host code never passes `scratch` itself (a host-side
// stand-in); it uses the device address from hipGetSymbolAddress, resolved
// once per device when a context is created there.
constexpr int kNaMaxBlocks = 1024;
constexpr size_t kNaCounterBytes = sizeof(unsigned) * kCounterSlot;
constexpr size_t kNaPartialBytes = sizeof(double) * (kNaMaxBlocks + kNaMaxBlocks / kGroup + 1);
constexpr size_t kNoAllocScratchBytes = kNaCounterBytes + (kNaPartialBytes + 255) / 256 * 256;
static_assert(kNaMaxBlocks <= kMaxReduceGrid, "the scratch's counters cover every group");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// the device address of `scratch` on `device` (current), looked up once
static void *noalloc_scratch(int device) {
  F3R_CHECK(device >= 0 && device < kMaxDevices, F3R_ERR_RANGE, "more devices than the scratch table holds");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    F3R_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    F3R_CHECK(ptr != nullptr, F3R_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

__global__ void k_zero_u32(unsigned *p, int n) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = 0u;
}

// ------------------------------------------------------------- context ---
std::shared_ptr<Context> ctx_create(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw Error{F3R_ERR_NODEVICE, "no HIP device"};
  F3R_CHECK(device >= 0 && device < count, F3R_ERR_ARG, "device index out of range");
  F3R_HIP(hipSetDevice(device));
  auto c = std::make_shared<Context>();
  c->device = device;
  hipDeviceProp_t prop;
  F3R_HIP(hipGetDeviceProperties(&prop, device));
  c->num_cu = prop.multiProcessorCount;
  F3R_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->red.partial = dev_alloc(sizeof(double) * Reducer::kPartialDoubles);
  c->red.counters = dev_alloc(sizeof(unsigned) * Reducer::kSlots * kCounterSlot);
  c->red.scalars = dev_alloc(sizeof(double) * 64);
  F3R_HIP(hipMemsetAsync(c->red.counters->p, 0, sizeof(unsigned) * Reducer::kSlots * kCounterSlot, c->stream));
  F3R_HIP(hipHostMalloc((void **)&c->pinned, 4096, hipHostMallocDefault));
  c->na_scratch = noalloc_scratch(device);
  F3R_HIP(hipStreamSynchronize(c->stream));
  return c;
}

Context::~Context() {
  if (stream) (void)hipStreamSynchronize(stream);
  red = Reducer();
  if (pinned) (void)hipHostFree(pinned);
  if (stream) (void)hipStreamDestroy(stream);
}

double read_scalar(Context *c, const double *dev) {
  F3R_HIP(hipMemcpyAsync(c->pinned, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  F3R_HIP(hipStreamSynchronize(c->stream));
  return c->pinned[0];
}

// ------------------------------------------------------------- vectors ---
void vec_fill(Vec *v, double val) {
  int g = grid_for(v->n, v->ctx->num_cu);
  hipStream_t st = v->ctx->stream;
  switch (v->dtype) {
    case F3R_F64: hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kBlock), 0, st, (double *)v->ptr(), v->n, val); break;
    case F3R_F32: hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kBlock), 0, st, (float *)v->ptr(), v->n, (float)val); break;
    case F3R_F16: hipLaunchKernelGGL(k_fill<half_t>, dim3(g), dim3(kBlock), 0, st, (half_t *)v->ptr(), v->n, (half_t)(float)val); break;
    case F3R_BF16: {
      const uint16_t bits = d2bf_bits(val);
      hipLaunchKernelGGL(k_fill<uint16_t>, dim3(g), dim3(kBlock), 0, st, (uint16_t *)v->ptr(), v->n, bits);
      break;
    }
    case F3R_I32: hipLaunchKernelGGL(k_fill<int32_t>, dim3(g), dim3(kBlock), 0, st, (int32_t *)v->ptr(), v->n, (int32_t)val); break;
  }
  F3R_HIP(hipGetLastError());
}

template <class S> static void cast_from(const S *s, Vec *d, hipStream_t st, int g) {
  switch (d->dtype) {
    case F3R_F64: hipLaunchKernelGGL((k_cast<S, double>), dim3(g), dim3(kBlock), 0, st, s, (double *)d->ptr(), d->n); break;
    case F3R_F32: hipLaunchKernelGGL((k_cast<S, float>), dim3(g), dim3(kBlock), 0, st, s, (float *)d->ptr(), d->n); break;
    case F3R_F16: hipLaunchKernelGGL((k_cast<S, half_t>), dim3(g), dim3(kBlock), 0, st, s, (half_t *)d->ptr(), d->n); break;
    case F3R_BF16: hipLaunchKernelGGL((k_cast<S, bf16_t>), dim3(g), dim3(kBlock), 0, st, s, (bf16_t *)d->ptr(), d->n); break;
  }
}

void vec_copy(Vec *s, Vec *d) {
  F3R_CHECK(s->n == d->n, F3R_ERR_ARG, "copy: length mismatch");
  F3R_CHECK(s->dtype != F3R_I32 && d->dtype != F3R_I32, F3R_ERR_UNSUPPORTED, "copy: int32 vectors");
  hipStream_t st = d->ctx->stream;
  if (s->dtype == d->dtype) {
    F3R_HIP(hipMemcpyAsync(d->ptr(), s->ptr(), s->n * dtype_size(s->dtype), hipMemcpyDeviceToDevice, st));
    return;
  }
  int g = grid_for(s->n, d->ctx->num_cu);
  switch (s->dtype) {
    case F3R_F64: cast_from((const double *)s->ptr(), d, st, g); break;
    case F3R_F32: cast_from((const float *)s->ptr(), d, st, g); break;
    case F3R_F16: cast_from((const half_t *)s->ptr(), d, st, g); break;
    case F3R_BF16: cast_from((const bf16_t *)s->ptr(), d, st, g); break;
  }
  F3R_HIP(hipGetLastError());
}

void vec_axpy(double a, Vec *x, Vec *y) {
  F3R_CHECK(x->n == y->n && x->dtype == y->dtype, F3R_ERR_ARG, "axpy: x and y differ in length or type");
  int g = grid_for(x->n, x->ctx->num_cu);
  hipStream_t st = x->ctx->stream;
  switch (x->dtype) {
    case F3R_F64: hipLaunchKernelGGL(k_axpy<double>, dim3(g), dim3(kBlock), 0, st, x->n, a, (const double *)x->ptr(), (double *)y->ptr()); break;
    case F3R_F32: hipLaunchKernelGGL(k_axpy<float>, dim3(g), dim3(kBlock), 0, st, x->n, a, (const float *)x->ptr(), (float *)y->ptr()); break;
    case F3R_F16: hipLaunchKernelGGL(k_axpy<half_t>, dim3(g), dim3(kBlock), 0, st, x->n, a, (const half_t *)x->ptr(), (half_t *)y->ptr()); break;
    case F3R_BF16: hipLaunchKernelGGL(k_axpy<bf16_t>, dim3(g), dim3(kBlock), 0, st, x->n, a, (const bf16_t *)x->ptr(), (bf16_t *)y->ptr()); break;
    default: throw Error{F3R_ERR_UNSUPPORTED, "axpy: int32"};
  }
  F3R_HIP(hipGetLastError());
}

void vec_scale(Vec *x, double a) {
  int g = grid_for(x->n, x->ctx->num_cu);
  hipStream_t st = x->ctx->stream;
  switch (x->dtype) {
    case F3R_F64: hipLaunchKernelGGL(k_scal<double>, dim3(g), dim3(kBlock), 0, st, x->n, a, (double *)x->ptr()); break;
    case F3R_F32: hipLaunchKernelGGL(k_scal<float>, dim3(g), dim3(kBlock), 0, st, x->n, a, (float *)x->ptr()); break;
    case F3R_F16: hipLaunchKernelGGL(k_scal<half_t>, dim3(g), dim3(kBlock), 0, st, x->n, a, (half_t *)x->ptr()); break;
    case F3R_BF16: hipLaunchKernelGGL(k_scal<bf16_t>, dim3(g), dim3(kBlock), 0, st, x->n, a, (bf16_t *)x->ptr()); break;
    default: throw Error{F3R_ERR_UNSUPPORTED, "scale: int32"};
  }
  F3R_HIP(hipGetLastError());
}

template <class A, class T1, class T2>
static void dot_launch(Context *c, int64_t n, const T1 *a, const T2 *b, double *out) {
  int g = reduce_grid(n, c->num_cu);
  hipLaunchKernelGGL((k_dot<A, T1, T2>), dim3(g), dim3(kBlock), 0, c->stream, n, a, b,
                     c->red.partial->as<A>(), c->red.counters->as<unsigned>(), out);
}

double dot_raw(Context *c, int da, const void *pa, int db, const void *pb, int64_t n) {
  double *slot = c->red.scalars->as<double>();
  F3R_CHECK(da != F3R_I32 && db != F3R_I32, F3R_ERR_UNSUPPORTED, "dot: int32");
  // accumulate in the wider type; fp16 data in fp32
#define F3R_DOT(TA, TB)                                                         \
  if (std::is_same<TA, double>::value || std::is_same<TB, double>::value)       \
    dot_launch<double>(c, n, (const TA *)pa, (const TB *)pb, slot);             \
  else                                                                          \
    dot_launch<float>(c, n, (const TA *)pa, (const TB *)pb, slot);
  auto pick_b = [&](auto tagA) {
    using TA = decltype(tagA);
    switch (db) {
      case F3R_F64: { F3R_DOT(TA, double) break; }
      case F3R_F32: { F3R_DOT(TA, float) break; }
      case F3R_F16: { F3R_DOT(TA, half_t) break; }
      case F3R_BF16: { F3R_DOT(TA, bf16_t) break; }
    }
  };
  switch (da) {
    case F3R_F64: pick_b(double()); break;
    case F3R_F32: pick_b(float()); break;
    case F3R_F16: pick_b(half_t()); break;
    case F3R_BF16: pick_b(bf16_t()); break;
  }
#undef F3R_DOT
  F3R_HIP(hipGetLastError());
  return read_scalar(c, slot);
}

// dot of two device arrays into *out (device), with the scratch array as
// the reduction's counters and partials: the same kernel, grid and summation
// order as dot_raw, so the same bits.  Launches two kernels and allocates
// nothing.
static void dot_noalloc(Context *c, int da, const void *pa, int db, const void *pb, int64_t n, double *out) {
  F3R_CHECK(da != F3R_I32 && db != F3R_I32, F3R_ERR_UNSUPPORTED, "dot_noalloc: int32");
  F3R_CHECK(c->na_scratch, F3R_ERR_STATE, "dot_noalloc: the context has no scratch address");
  const int g = reduce_grid(n, c->num_cu);
  F3R_CHECK(g <= kNaMaxBlocks, F3R_ERR_RANGE, "dot_noalloc: the reduction needs more blocks than the scratch holds");
  unsigned char *base = static_cast<unsigned char *>(c->na_scratch);
  unsigned *cnt = reinterpret_cast<unsigned *>(base);
  void *part = base + kNaCounterBytes;
  // the counters this reduction uses start at zero (a fill kernel, not a
  // memset node, so the call can be captured in a graph)
  const int ng = (g + kGroup - 1) / kGroup;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kBlock), 0, c->stream, cnt, (1 + ng) * kCounterStride);
#define F3R_DOTNA(TA, TB)                                                                               \
  if (std::is_same<TA, double>::value || std::is_same<TB, double>::value)                             \
    hipLaunchKernelGGL((k_dot<double, TA, TB>), dim3(g), dim3(kBlock), 0, c->stream, n, (const TA *)pa, \
                       (const TB *)pb, (double *)part, cnt, out);                                     \
  else                                                                                                 \
    hipLaunchKernelGGL((k_dot<float, TA, TB>), dim3(g), dim3(kBlock), 0, c->stream, n, (const TA *)pa,  \
                       (const TB *)pb, (float *)part, cnt, out);
  auto pick_b = [&](auto tagA) {
    using TA = decltype(tagA);
    switch (db) {
      case F3R_F64: { F3R_DOTNA(TA, double) break; }
      case F3R_F32: { F3R_DOTNA(TA, float) break; }
      case F3R_F16: { F3R_DOTNA(TA, half_t) break; }
      case F3R_BF16: { F3R_DOTNA(TA, bf16_t) break; }
    }
  };
  switch (da) {
    case F3R_F64: pick_b(double()); break;
    case F3R_F32: pick_b(float()); break;
    case F3R_F16: pick_b(half_t()); break;
    case F3R_BF16: pick_b(bf16_t()); break;
  }
#undef F3R_DOTNA
  F3R_HIP(hipGetLastError());
}

double vec_dot(Vec *a, Vec *b) {
  F3R_CHECK(a->n == b->n, F3R_ERR_ARG, "dot: length mismatch");
  return dot_raw(a->ctx.get(), a->dtype, a->ptr(), b->dtype, b->ptr(), a->n);
}

// ------------------------------------------------------------ matrices ---
SellArgs sell_args(const SellShared &sh, bool with_perm) {
  SellArgs a;
  a.n = sh.n;
  int l = 0;
  while ((1 << l) < sh.C) ++l;
  a.log2C = l;
  a.cptr = sh.cptr->as<int32_t>();
  a.cwidth = sh.cwidth->as<int32_t>();
  a.idx = sh.idx->as<int32_t>();
  a.perm = (with_perm && sh.perm) ? sh.perm->as<int32_t>() : nullptr;
  a.iptr = sh.iptr ? sh.iptr->as<int32_t>() : nullptr;
  a.coff = sh.coff ? sh.coff->as<int32_t>() : nullptr;
  a.off_bias = sh.off_bias;
  a.lanes = sh.lanes;
  return a;
}

std::shared_ptr<SellShared> upload_sell_shared(Context *ctx, const HostSell &hs, bool internal) {
  auto sh = std::make_shared<SellShared>();
  sh->n = hs.n;
  sh->C = hs.C;
  sh->K = hs.K;
  sh->nchunks = hs.nchunks;
  sh->nslots = (int64_t)hs.val.size();
  sh->offset_chunks = hs.offset_chunks;
  sh->offset_slots = hs.offset_slots;
  sh->explicit_nnz = hs.explicit_nnz;
  sh->off_bias = hs.off_bias;
  sh->internal = internal;
  hipStream_t st = ctx->stream;
  sh->cptr = dev_alloc(sizeof(int32_t) * hs.cptr.size());
  sh->cwidth = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, hs.cwidth.size()));
  sh->idx = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, hs.idx.size()));
  F3R_HIP(hipMemcpyAsync(sh->cptr->p, hs.cptr.data(), sizeof(int32_t) * hs.cptr.size(), hipMemcpyHostToDevice, st));
  if (!hs.cwidth.empty())
    F3R_HIP(hipMemcpyAsync(sh->cwidth->p, hs.cwidth.data(), sizeof(int32_t) * hs.cwidth.size(), hipMemcpyHostToDevice, st));
  if (!hs.idx.empty())
    F3R_HIP(hipMemcpyAsync(sh->idx->p, hs.idx.data(), sizeof(int32_t) * hs.idx.size(), hipMemcpyHostToDevice, st));
  if (!hs.iptr.empty()) {
    sh->iptr = dev_alloc(sizeof(int32_t) * hs.iptr.size());
    sh->coff = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, hs.coff.size()));
    F3R_HIP(hipMemcpyAsync(sh->iptr->p, hs.iptr.data(), sizeof(int32_t) * hs.iptr.size(), hipMemcpyHostToDevice, st));
    if (!hs.coff.empty())
      F3R_HIP(hipMemcpyAsync(sh->coff->p, hs.coff.data(), sizeof(int32_t) * hs.coff.size(), hipMemcpyHostToDevice, st));
  }
  bool ident = true;
  for (int32_t r = 0; r < hs.n && ident; ++r) ident = hs.perm[r] == r;
  if (!ident) {
    sh->perm = dev_alloc(sizeof(int32_t) * hs.n);
    F3R_HIP(hipMemcpyAsync(sh->perm->p, hs.perm.data(), sizeof(int32_t) * hs.n, hipMemcpyHostToDevice, st));
  }
  sh->h_perm = hs.perm;
  F3R_HIP(hipStreamSynchronize(st));
  return sh;
}

DevPtr upload_values(Context *ctx, const std::vector<double> &v, int dtype) {
  DevPtr d = dev_alloc(dtype_size(dtype) * std::max<size_t>(1, v.size()));
  if (v.empty()) return d;
  hipStream_t st = ctx->stream;
  if (dtype == F3R_F64) {
    F3R_HIP(hipMemcpyAsync(d->p, v.data(), 8 * v.size(), hipMemcpyHostToDevice, st));
    F3R_HIP(hipStreamSynchronize(st));
  } else if (dtype == F3R_F32) {
    std::vector<float> f(v.size());
    for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i];
    F3R_HIP(hipMemcpyAsync(d->p, f.data(), 4 * f.size(), hipMemcpyHostToDevice, st));
    F3R_HIP(hipStreamSynchronize(st));
  } else if (dtype == F3R_F16 || dtype == F3R_BF16) {
    std::vector<uint16_t> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = dtype == F3R_F16 ? d2h_bits(v[i]) : d2bf_bits(v[i]);
    F3R_HIP(hipMemcpyAsync(d->p, h.data(), 2 * h.size(), hipMemcpyHostToDevice, st));
    F3R_HIP(hipStreamSynchronize(st));
  } else {
    throw Error{F3R_ERR_UNSUPPORTED, "matrix values must be fp64, fp32, fp16 or bf16"};
  }
  return d;
}

DevPtr upload_sell_values(Context *ctx, const HostSell &hs, int dtype) {
  return upload_values(ctx, hs.val, dtype);
}

// LDS of a block's x window in the LDS-staged SpMV: 32 KB, so that four
// blocks of 256 rows still fit a CU's 160 KB
constexpr size_t kLdsWindowBytes = 32 * 1024;

// per block of kBlock sorted rows, the window of columns its slots read
static std::vector<int2> block_windows(const HostSell &hs) {
  const int64_t nb = ((int64_t)hs.n + kBlock - 1) / kBlock;
  std::vector<int2> w(std::max<int64_t>(nb, 1), int2{0, 0});
  const int C = hs.C, K = hs.K;
  for (int32_t s = 0; s < hs.nchunks; ++s) {
    const int32_t r0 = s * C, tc = std::min(C, hs.n - r0);
    const int32_t groups = chunk_groups(hs.cwidth[s]);
    for (int32_t j = 0; j < tc; ++j) {
      const int64_t b = (r0 + j) / kBlock;
      int lo = w[b].y > 0 ? w[b].x : INT32_MAX, hi = w[b].y > 0 ? w[b].x + w[b].y : INT32_MIN;
      for (int32_t t = 0; t < groups * K; ++t) {
        const int32_t c = hs.idx[chunk_slot_pos(hs.cptr[s], hs.cwidth[s], tc, j, t, K)];
        lo = std::min(lo, c);
        hi = std::max(hi, c + 1);
      }
      if (hi > lo) w[b] = int2{lo, hi - lo};
    }
  }
  return w;
}

template <class V, class X, class Y>
static void sell_apply_t(const Mat *A, const void *x, void *y, const void *b, hipStream_t st) {
  using ACC = typename std::conditional<std::is_same<V, double>::value, double, float>::type;
  SellArgs a = sell_args(*A->sh, true);
  if (A->sh->bwin) {  // LDS-staged x gathers
    const int2 *w = A->sh->bwin->as<int2>();
    // LDS sized to the widest window that fits, so narrow bands keep the occupancy
    const int cap = A->sh->lds_fit[sizeof(X) == 2 ? 0 : sizeof(X) == 4 ? 1 : 2];
    if (b)
      launch_sell_lds<V, X, ACC>(A->sh->K, a, (const V *)A->s_val->p, (const X *)x, w, cap,
                                 EpiResid<Y, Y, ACC>{(Y *)y, (const Y *)b, a.perm}, st);
    else
      launch_sell_lds<V, X, ACC>(A->sh->K, a, (const V *)A->s_val->p, (const X *)x, w, cap,
                                 EpiStore<Y, ACC>{(Y *)y, a.perm}, st);
    return;
  }
  if (b)  // K threads per row when the layout has them (a.lanes)
    launch_sell<V, X, ACC>(A->sh->K, a, (const V *)A->s_val->p, (const X *)x,
                           EpiResid<Y, Y, ACC>{(Y *)y, (const Y *)b, a.perm}, st);
  else
    launch_sell<V, X, ACC>(A->sh->K, a, (const V *)A->s_val->p, (const X *)x,
                           EpiStore<Y, ACC>{(Y *)y, a.perm}, st);
}

// lanes per row of the CSR SpMV: the caller's choice, else the largest power
// of two not above half the mean row length (1 to 64)
static int csr_tpr(const Mat &A) {
  if (A.tpr > 0) return A.tpr;
  const double mean = A.host->n > 0 ? (double)A.host->nnz() / A.host->n : 1.0;
  int t = 1;
  while (t < 64 && 2.0 * t <= 0.5 * mean) t *= 2;
  return t;
}

void mat_apply(const Mat *A, Vec *x, Vec *y, Vec *b) {
  F3R_CHECK(x->n == (A->format == 0 ? A->host->ncols : A->sh->n) && y->n == x->n, F3R_ERR_ARG,
            "spmv: vector length does not match the matrix");
  F3R_CHECK(!b || (b->n == y->n && b->dtype == y->dtype), F3R_ERR_ARG, "residual: b must match r");
  hipStream_t st = A->ctx->stream;
  const void *bp = b ? b->ptr() : nullptr;
  if (A->format == 0) {
    const int32_t n = A->host->n;
    const int tpr = csr_tpr(*A);
    const void *val = A->dtype == F3R_F64 ? A->d_val->p : A->c_val->p;
    auto go = [&](auto vt, auto xt, auto yt) {
      using V = decltype(vt);
      using X = decltype(xt);
      using Y = decltype(yt);
      using ACC = typename std::conditional<std::is_same<V, double>::value, double, float>::type;
      const int64_t rows_per_block = kBlock / tpr;
      const int g = (int)std::max<int64_t>(1, std::min<int64_t>((n + rows_per_block - 1) / rows_per_block,
                                                                 (int64_t)A->ctx->num_cu * 64));
#define F3R_CSR_LAUNCH(T)                                                                                 \
  hipLaunchKernelGGL((k_csr<V, X, Y, ACC, T>), dim3(g), dim3(kBlock), 0, st, n, A->d_rptr->as<int32_t>(), \
                     A->d_idx->as<int32_t>(), (const V *)val, (const X *)x->ptr(), (Y *)y->ptr(), (const Y *)bp)
      switch (tpr) {
        case 1: F3R_CSR_LAUNCH(1); break;
        case 2: F3R_CSR_LAUNCH(2); break;
        case 4: F3R_CSR_LAUNCH(4); break;
        case 8: F3R_CSR_LAUNCH(8); break;
        case 16: F3R_CSR_LAUNCH(16); break;
        case 32: F3R_CSR_LAUNCH(32); break;
        default: F3R_CSR_LAUNCH(64); break;
      }
#undef F3R_CSR_LAUNCH
    };
    const int v = A->dtype, xd = x->dtype, yd = y->dtype;
    if (v == F3R_F64 && xd == F3R_F64 && yd == F3R_F64) go(double(), double(), double());
    else if (v == F3R_F64 && xd == F3R_F32 && yd == F3R_F32) go(double(), float(), float());
    else if (v == F3R_F64 && xd == F3R_F32 && yd == F3R_F64) go(double(), float(), double());
    else if (v == F3R_F64 && xd == F3R_F64 && yd == F3R_F32) go(double(), double(), float());
    else if (v == F3R_F32 && xd == F3R_F32 && yd == F3R_F32) go(float(), float(), float());
    else if (v == F3R_F32 && xd == F3R_F32 && yd == F3R_F64) go(float(), float(), double());
    else if (v == F3R_F16 && xd == F3R_F32 && yd == F3R_F32) go(half_t(), float(), float());
    else if (v == F3R_F16 && xd == F3R_F16 && yd == F3R_F16) go(half_t(), half_t(), half_t());
    else if (v == F3R_F16 && xd == F3R_F16 && yd == F3R_F32) go(half_t(), half_t(), float());
    else if (v == F3R_BF16 && xd == F3R_BF16 && yd == F3R_BF16) go(bf16_t(), bf16_t(), bf16_t());
    else if (v == F3R_BF16 && xd == F3R_F32 && yd == F3R_F32) go(bf16_t(), float(), float());
    else throw Error{F3R_ERR_UNSUPPORTED, "CSR spmv: unsupported (value, x, y) type combination"};
  } else {
    F3R_CHECK(!A->sh->internal, F3R_ERR_STATE, "spmv: internal-order matrix");
    const int v = A->dtype, xd = x->dtype, yd = y->dtype;
    if (v == F3R_F64 && xd == F3R_F64 && yd == F3R_F64) sell_apply_t<double, double, double>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_F64 && xd == F3R_F32 && yd == F3R_F32) sell_apply_t<double, float, float>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_F32 && xd == F3R_F32 && yd == F3R_F32) sell_apply_t<float, float, float>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_F32 && xd == F3R_F32 && yd == F3R_F64) sell_apply_t<float, float, double>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_F16 && xd == F3R_F32 && yd == F3R_F32) sell_apply_t<half_t, float, float>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_F16 && xd == F3R_F16 && yd == F3R_F16) sell_apply_t<half_t, half_t, half_t>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_F16 && xd == F3R_F16 && yd == F3R_F32) sell_apply_t<half_t, half_t, float>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_BF16 && xd == F3R_BF16 && yd == F3R_BF16) sell_apply_t<bf16_t, bf16_t, bf16_t>(A, x->ptr(), y->ptr(), bp, st);
    else if (v == F3R_BF16 && xd == F3R_F32 && yd == F3R_F32) sell_apply_t<bf16_t, float, float>(A, x->ptr(), y->ptr(), bp, st);
    else throw Error{F3R_ERR_UNSUPPORTED, "SELL spmv: unsupported (value, x, y) type combination"};
  }
  F3R_HIP(hipGetLastError());
}

// ------------------------------------------------------ preconditioners ---
static std::vector<int32_t> levels_lower(const HostCSR &L, std::vector<int32_t> &lptr) {
  const int32_t n = L.n;
  std::vector<int32_t> lev(n, 0);
  int32_t nl = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t m = -1;
    for (int32_t k = L.rptr[i]; k < L.rptr[i + 1] - 1; ++k) m = std::max(m, lev[L.idx[k]]);
    lev[i] = m + 1;
    nl = std::max(nl, lev[i] + 1);
  }
  lptr.assign(nl + 1, 0);
  for (int32_t i = 0; i < n; ++i) lptr[lev[i] + 1]++;
  for (int32_t l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
  std::vector<int32_t> rows(n), fill(lptr.begin(), lptr.end() - 1);
  for (int32_t i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
  return rows;
}
// Level-ordered rows regrouped by diagonal block of `bsz` rows (dependencies
// stay inside a block): brows block by block, each block level by level;
// blptr[bstart[b] + l] starts level l of block b, blptr[bstart[b + 1] - 1]
// ends the block.
static void group_by_block(const std::vector<int32_t> &rows, const std::vector<int32_t> &lptr, int32_t n,
                           int32_t bsz, std::vector<int32_t> &brows, std::vector<int32_t> &blptr,
                           std::vector<int32_t> &bstart) {
  const int32_t nb = (n + bsz - 1) / bsz;
  std::vector<int32_t> lev(n);
  for (size_t l = 0; l + 1 < lptr.size(); ++l)
    for (int32_t k = lptr[l]; k < lptr[l + 1]; ++k) lev[rows[k]] = (int32_t)l;
  // stable by block: rows keep their level order inside a block
  std::vector<int32_t> cnt(nb + 1, 0);
  for (int32_t i = 0; i < n; ++i) cnt[i / bsz + 1]++;
  for (int32_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
  brows.resize(n);
  std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
  for (int32_t r : rows) brows[fill[r / bsz]++] = r;
  blptr.clear();
  bstart.assign(nb + 1, 0);
  for (int32_t b = 0; b < nb; ++b) {
    bstart[b] = (int32_t)blptr.size();
    for (int32_t t = cnt[b]; t < cnt[b + 1]; ++t)
      if (t == cnt[b] || lev[brows[t]] != lev[brows[t - 1]]) blptr.push_back(t);
    blptr.push_back(cnt[b + 1]);
  }
  bstart[nb] = (int32_t)blptr.size();
}

// SoA records of a triangular factor in the grouped order `brows` (see
// k_bj_lds): entries in the visiting order of the solve, columns local to
// the row's block.  False when a row has more than `rmax` entries.
static bool bj_records(const HostCSR &M, bool upper, const std::vector<int32_t> &brows, int32_t bsz, int rmax,
                       std::vector<int32_t> &row, std::vector<int32_t> &cnt, std::vector<int32_t> &col,
                       std::vector<double> &val) {
  const int64_t ns = (int64_t)brows.size();
  row.assign(ns, 0);
  cnt.assign(ns, 0);
  col.assign(ns * rmax, 0);
  val.assign(ns * rmax, 0.0);
  for (int64_t t = 0; t < ns; ++t) {
    const int32_t i = brows[t], base = i / bsz * bsz;
    row[t] = i - base;
    int c = 0;
    auto put = [&](int32_t k) {
      if (c < rmax) {
        col[c * ns + t] = M.idx[k] - base;
        val[c * ns + t] = M.val[k];
      }
      ++c;
    };
    if (upper)
      for (int32_t k = M.rptr[i + 1] - 1; k > M.rptr[i]; --k) put(k);
    else
      for (int32_t k = M.rptr[i]; k < M.rptr[i + 1] - 1; ++k) put(k);
    if (c > rmax) return false;
    cnt[t] = c;
  }
  return true;
}

static std::vector<int32_t> levels_upper(const HostCSR &U, std::vector<int32_t> &lptr) {
  const int32_t n = U.n;
  std::vector<int32_t> lev(n, 0);
  int32_t nl = 0;
  for (int32_t r = n - 1; r >= 0; --r) {
    int32_t m = -1;
    for (int32_t k = U.rptr[r] + 1; k < U.rptr[r + 1]; ++k) m = std::max(m, lev[U.idx[k]]);
    lev[r] = m + 1;
    nl = std::max(nl, lev[r] + 1);
  }
  lptr.assign(nl + 1, 0);
  for (int32_t i = 0; i < n; ++i) lptr[lev[i] + 1]++;
  for (int32_t l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
  std::vector<int32_t> rows(n), fill(lptr.begin(), lptr.end() - 1);
  for (int32_t i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
  return rows;
}

static DevPtr upload_i32(Context *c, const std::vector<int32_t> &v) {
  DevPtr d = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, v.size()));
  if (!v.empty()) {
    F3R_HIP(hipMemcpyAsync(d->p, v.data(), 4 * v.size(), hipMemcpyHostToDevice, c->stream));
    F3R_HIP(hipStreamSynchronize(c->stream));
  }
  return d;
}

static bool is_diag(const HostCSR &M) {
  for (int32_t i = 0; i < M.n; ++i)
    if (M.rptr[i + 1] - M.rptr[i] != 1 || M.idx[M.rptr[i]] != i) return false;
  return true;
}

// Permuted copy P M P^T of a host CSR (row r of the result = row perm[r]),
// columns renumbered and sorted.
HostCSR permute_csr(const HostCSR &M, const std::vector<int32_t> &perm) {
  const int32_t n = M.n;
  std::vector<int32_t> inv(n);
  for (int32_t r = 0; r < n; ++r) inv[perm[r]] = r;
  HostCSR R;
  R.n = R.ncols = n;
  R.rptr.assign(n + 1, 0);
  R.idx.resize(M.nnz());
  R.val.resize(M.nnz());
  std::vector<std::pair<int32_t, double>> row;
  for (int32_t r = 0; r < n; ++r) {
    int32_t o = perm[r];
    row.clear();
    for (int32_t k = M.rptr[o]; k < M.rptr[o + 1]; ++k) row.push_back({inv[M.idx[k]], M.val[k]});
    std::sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first < b.first; });
    R.rptr[r + 1] = R.rptr[r] + (int32_t)row.size();
    for (size_t q = 0; q < row.size(); ++q) {
      R.idx[R.rptr[r] + q] = row[q].first;
      R.val[R.rptr[r] + q] = row[q].second;
    }
  }
  return R;
}

void prec_build_device(Prec *P, const std::vector<int32_t> *perm, int32_t C, int32_t K, bool offsets) {
  Context *c = P->ctx.get();
  const int32_t n = P->host.f[0].n;
  P->n = n;
  P->tmp = dev_alloc(dtype_size(P->dtype) * std::max(1, n));
  if (P->kind == 0) {
    const HostCSR &Z = P->host.f[0], &W = P->host.f[1];
    P->diag = is_diag(Z) && is_diag(W);
    if (P->diag) {
      std::vector<double> wd(n), zd(n);
      // under a symmetric permutation a diagonal stays diagonal
      for (int32_t r = 0; r < n; ++r) {
        int32_t o = perm ? (*perm)[r] : r;
        wd[r] = W.val[W.rptr[o]];
        zd[r] = Z.val[Z.rptr[o]];
      }
      P->dw = upload_values(c, wd, P->dtype);
      P->dz = upload_values(c, zd, P->dtype);
      if (P->dtype == F3R_F16) {
        std::vector<uint16_t> wz(2 * (size_t)n);
        for (int32_t r = 0; r < n; ++r) {
          wz[2 * (size_t)r] = d2h_bits(wd[r]);
          wz[2 * (size_t)r + 1] = d2h_bits(zd[r]);
        }
        P->dwz = dev_alloc(sizeof(uint16_t) * std::max<size_t>(2, wz.size()));
        if (n) F3R_HIP(hipMemcpyAsync(P->dwz->p, wz.data(), sizeof(uint16_t) * wz.size(), hipMemcpyHostToDevice, c->stream));
        F3R_HIP(hipStreamSynchronize(c->stream));  // wz is released at scope end
      }
    }
    for (int w = 0; w < 2; ++w) {
      const HostCSR &M = w == 0 ? W : Z;
      HostSell hs = build_sell(M, C, 1, K, perm, perm != nullptr, offsets);
      auto m = std::make_shared<Mat>();
      m->ctx = P->ctx;
      m->format = 1;
      m->dtype = P->dtype;
      m->sh = upload_sell_shared(c, hs, perm != nullptr);
      m->s_val = upload_sell_values(c, hs, P->dtype);
      (w == 0 ? P->W : P->Z) = m;
    }
  } else {
    HostCSR L = perm ? permute_csr(P->host.f[0], *perm) : P->host.f[0];
    HostCSR U = perm ? permute_csr(P->host.f[1], *perm) : P->host.f[1];
    // a permutation would break the triangular structure: only the identity
    F3R_CHECK(!perm, F3R_ERR_UNSUPPORTED, "ILU(0) with a sorted (sigma > 1) layout");
    const std::vector<int32_t> lrows = levels_lower(L, P->l_lptr), urows = levels_upper(U, P->u_lptr);
    P->l_rows = upload_i32(c, lrows);
    P->u_rows = upload_i32(c, urows);
    // many diagonal blocks: one workgroup per block, one launch per solve
    const int32_t blocks = (int32_t)P->tol;
    P->bj_blocks = blocks >= kBjBlockKernel ? blocks : 0;
    if (P->bj_blocks) {
      const int32_t bsz = (n + blocks - 1) / blocks;
      std::vector<int32_t> br, bl, bs;
      group_by_block(lrows, P->l_lptr, n, bsz, br, bl, bs);
      P->lb_rows = upload_i32(c, br);
      P->lb_ptr = upload_i32(c, bl);
      P->lb_start = upload_i32(c, bs);
      group_by_block(urows, P->u_lptr, n, bsz, br, bl, bs);
      P->ub_rows = upload_i32(c, br);
      P->ub_ptr = upload_i32(c, bl);
      P->ub_start = upload_i32(c, bs);
      P->bj_groups = (int32_t)bs.size() - 1;
      // the fused LDS kernel: the block's vector in LDS, at most 1024 rows
      // per level of a block and 16 entries per row
      int32_t maxrows = 0, maxw = 0, maxe = 0;
      for (int32_t b = 0; b < P->bj_groups; ++b) maxrows = std::max(maxrows, std::min(n, (b + 1) * bsz) - b * bsz);
      std::vector<int32_t> lbr, lbl, lbs, ubr, ubl, ubs;
      group_by_block(lrows, P->l_lptr, n, bsz, lbr, lbl, lbs);
      group_by_block(urows, P->u_lptr, n, bsz, ubr, ubl, ubs);
      for (const std::vector<int32_t> *v : {&lbl, &ubl})
        for (size_t q = 0; q + 1 < v->size(); ++q) maxw = std::max(maxw, (*v)[q + 1] - (*v)[q]);
      for (int32_t i = 0; i < n; ++i) {
        maxe = std::max(maxe, L.rptr[i + 1] - L.rptr[i] - 1);
        maxe = std::max(maxe, U.rptr[i + 1] - U.rptr[i] - 1);
      }
      const size_t lds = (size_t)maxrows * dtype_size(P->dtype);
      if (maxw <= 1024 && maxe <= 16 && lds <= kBjLdsBytes) {
        const int R = maxe <= 4 ? 4 : maxe <= 8 ? 8 : 16;
        std::vector<int32_t> row, cnt, col;
        std::vector<double> val;
        bj_records(L, false, lbr, bsz, R, row, cnt, col, val);
        P->lr_row = upload_i32(c, row);
        P->lr_cnt = upload_i32(c, cnt);
        P->lr_col = upload_i32(c, col);
        P->lr_val = upload_values(c, val, P->dtype);
        bj_records(U, true, ubr, bsz, R, row, cnt, col, val);
        P->ur_row = upload_i32(c, row);
        P->ur_cnt = upload_i32(c, cnt);
        P->ur_col = upload_i32(c, col);
        P->ur_val = upload_values(c, val, P->dtype);
        std::vector<double> dv(n);
        for (int32_t t = 0; t < n; ++t) dv[t] = 1. / U.val[U.rptr[ubr[t]]];
        P->ur_dinv = upload_values(c, dv, P->dtype);
        std::vector<int32_t> b0(P->bj_groups), bn(P->bj_groups);
        for (int32_t b = 0; b < P->bj_groups; ++b) {
          b0[b] = b * bsz;
          bn[b] = std::min(n, (b + 1) * bsz) - b * bsz;
        }
        P->bj_b0 = upload_i32(c, b0);
        P->bj_bn = upload_i32(c, bn);
        P->bj_R = R;
        P->bj_lds = lds;
        // one thread per row of the widest level: idle waves would still
        // load records and meet every barrier
        P->bj_threads = std::max(64, (maxw + 63) / 64 * 64);
      }
    }
    P->l_rptr = upload_i32(c, L.rptr);
    P->l_idx = upload_i32(c, L.idx);
    P->l_val = upload_values(c, L.val, P->dtype);
    P->u_rptr = upload_i32(c, U.rptr);
    P->u_idx = upload_i32(c, U.idx);
    P->u_val = upload_values(c, U.val, P->dtype);
    std::vector<double> inv(n);
    for (int32_t r = 0; r < n; ++r) inv[r] = 1. / U.val[U.rptr[r]];
    P->u_dinv = upload_values(c, inv, P->dtype);
  }
}

template <class S> static void ilu_apply_t(Prec *P, const S *in, S *out, hipStream_t st) {
  S *tmp = P->tmp->as<S>();
  if (P->bj_R) {
    const int32_t ns = P->n;
    BjRec<S> L{P->lb_start->as<int32_t>(), P->lb_ptr->as<int32_t>(), P->lr_row->as<int32_t>(),
               P->lr_cnt->as<int32_t>(), P->lr_col->as<int32_t>(), P->lr_val->as<S>(), nullptr, ns};
    BjRec<S> U{P->ub_start->as<int32_t>(), P->ub_ptr->as<int32_t>(), P->ur_row->as<int32_t>(),
               P->ur_cnt->as<int32_t>(), P->ur_col->as<int32_t>(), P->ur_val->as<S>(), P->ur_dinv->as<S>(), ns};
    const int32_t *b0 = P->bj_b0->as<int32_t>(), *bn = P->bj_bn->as<int32_t>();
    const unsigned lds = (unsigned)P->bj_lds;
    switch (P->bj_R) {
      case 4: hipLaunchKernelGGL((k_bj_lds<S, 4>), dim3(P->bj_groups), dim3(P->bj_threads), lds, st, b0, bn, L, U, in, out); break;
      case 8: hipLaunchKernelGGL((k_bj_lds<S, 8>), dim3(P->bj_groups), dim3(P->bj_threads), lds, st, b0, bn, L, U, in, out); break;
      default: hipLaunchKernelGGL((k_bj_lds<S, 16>), dim3(P->bj_groups), dim3(P->bj_threads), lds, st, b0, bn, L, U, in, out); break;
    }
    return;
  }
  if (P->bj_blocks) {
    hipLaunchKernelGGL(k_bj_trsv_l<S>, dim3(P->bj_groups), dim3(1024), 0, st, (const int32_t *)P->lb_start->as<int32_t>(),
                       (const int32_t *)P->lb_ptr->as<int32_t>(), (const int32_t *)P->lb_rows->as<int32_t>(),
                       (const int32_t *)P->l_rptr->as<int32_t>(), (const int32_t *)P->l_idx->as<int32_t>(),
                       (const S *)P->l_val->as<S>(), in, tmp);
    hipLaunchKernelGGL(k_bj_trsv_u<S>, dim3(P->bj_groups), dim3(1024), 0, st, (const int32_t *)P->ub_start->as<int32_t>(),
                       (const int32_t *)P->ub_ptr->as<int32_t>(), (const int32_t *)P->ub_rows->as<int32_t>(),
                       (const int32_t *)P->u_rptr->as<int32_t>(), (const int32_t *)P->u_idx->as<int32_t>(),
                       (const S *)P->u_val->as<S>(), (const S *)P->u_dinv->as<S>(), (const S *)tmp, out);
    return;
  }
  for (size_t l = 0; l + 1 < P->l_lptr.size(); ++l) {
    int32_t cnt = P->l_lptr[l + 1] - P->l_lptr[l];
    hipLaunchKernelGGL(k_trsv_l<S>, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cnt,
                       P->l_rows->as<int32_t>() + P->l_lptr[l], P->l_rptr->as<int32_t>(),
                       P->l_idx->as<int32_t>(), P->l_val->as<S>(), in, tmp);
  }
  for (size_t l = 0; l + 1 < P->u_lptr.size(); ++l) {
    int32_t cnt = P->u_lptr[l + 1] - P->u_lptr[l];
    hipLaunchKernelGGL(k_trsv_u<S>, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cnt,
                       P->u_rows->as<int32_t>() + P->u_lptr[l], P->u_rptr->as<int32_t>(),
                       P->u_idx->as<int32_t>(), P->u_val->as<S>(), P->u_dinv->as<S>(), tmp, out);
  }
}

template <class S> static void sdainv_apply_t(Prec *P, const S *in, S *out, hipStream_t st) {
  using ACC = typename Acc<S>::type;
  S *tmp = P->tmp->as<S>();
  const bool wp = !P->W->sh->internal;
  SellArgs aw = sell_args(*P->W->sh, wp), az = sell_args(*P->Z->sh, wp);
  launch_sell<S, S, ACC>(P->W->sh->K, aw, P->W->s_val->as<S>(), in, EpiStore<S, ACC>{tmp, aw.perm}, st);
  launch_sell<S, S, ACC>(P->Z->sh->K, az, P->Z->s_val->as<S>(), tmp, EpiStore<S, ACC>{out, az.perm}, st);
}

void prec_apply_raw(Prec *P, const void *in, void *out, hipStream_t st) {
  switch (P->dtype) {
    case F3R_F64:
      P->kind == 0 ? sdainv_apply_t<double>(P, (const double *)in, (double *)out, st)
                   : ilu_apply_t<double>(P, (const double *)in, (double *)out, st);
      break;
    case F3R_F32:
      P->kind == 0 ? sdainv_apply_t<float>(P, (const float *)in, (float *)out, st)
                   : ilu_apply_t<float>(P, (const float *)in, (float *)out, st);
      break;
    case F3R_F16:
      P->kind == 0 ? sdainv_apply_t<half_t>(P, (const half_t *)in, (half_t *)out, st)
                   : ilu_apply_t<half_t>(P, (const half_t *)in, (half_t *)out, st);
      break;
  }
  F3R_HIP(hipGetLastError());
}

HostCSR host_csr_from(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rptr,
                      const int32_t *col, const double *val) {
  F3R_CHECK(nrows >= 0 && ncols >= 0 && nnz >= 0 && rptr && (nnz == 0 || (col && val)),
            F3R_ERR_ARG, "CSR: null array or negative size");
  F3R_CHECK(nnz < INT32_MAX, F3R_ERR_RANGE, "CSR: nnz does not fit int32 indices");
  F3R_CHECK(rptr[0] == 0 && rptr[nrows] == nnz, F3R_ERR_ARG, "CSR: rptr[0] != 0 or rptr[n] != nnz");
  HostCSR h;
  h.n = nrows;
  h.ncols = ncols;
  h.rptr.assign(rptr, rptr + nrows + 1);
  h.idx.assign(col, col + nnz);
  h.val.assign(val, val + nnz);
  for (int32_t i = 0; i < nrows; ++i) {
    F3R_CHECK(h.rptr[i] <= h.rptr[i + 1] && h.rptr[i + 1] <= nnz, F3R_ERR_ARG, "CSR: rptr is not monotone");
    for (int32_t k = h.rptr[i]; k < h.rptr[i + 1]; ++k) {
      F3R_CHECK(h.idx[k] >= 0 && h.idx[k] < ncols, F3R_ERR_ARG, "CSR: column index out of range");
      F3R_CHECK(k == h.rptr[i] || h.idx[k - 1] < h.idx[k], F3R_ERR_ARG,
                "CSR: columns of a row must be sorted and unique");
    }
  }
  return h;
}

}  // namespace f3r

// ================================================================ C ABI ===
using namespace f3r;


struct f3r_ctx_s { std::shared_ptr<Context> c; };
struct f3r_vec_s { Vec v; };
struct f3r_mat_s { std::shared_ptr<Mat> m; };
struct f3r_prec_s { std::shared_ptr<Prec> p; };

namespace f3r {
std::shared_ptr<Context> ctx_of(f3r_ctx c) { return c ? c->c : nullptr; }
Vec *vec_of(f3r_vec v) { return v ? &v->v : nullptr; }
std::shared_ptr<Mat> mat_of(f3r_mat m) { return m ? m->m : nullptr; }
std::shared_ptr<Prec> prec_of(f3r_prec p) { return p ? p->p : nullptr; }
f3r_prec wrap_prec(std::shared_ptr<Prec> p) { return new f3r_prec_s{std::move(p)}; }
f3r_vec wrap_vec(const Vec &v) { return new f3r_vec_s{v}; }
void release_vec(f3r_vec v) { delete v; }
}  // namespace f3r

extern "C" {

const char *f3r_strerror(int s) {
  switch (s) {
    case F3R_OK: return "ok";
    case F3R_ERR_ARG: return "invalid argument";
    case F3R_ERR_HIP: return "HIP runtime error";
    case F3R_ERR_ALLOC: return "allocation failed";
    case F3R_ERR_UNSUPPORTED: return "unsupported type or format combination";
    case F3R_ERR_NODEVICE: return "no usable GPU";
    case F3R_ERR_STATE: return "invalid object state";
    case F3R_ERR_RANGE: return "size beyond the format's index range";
  }
  return "unknown status";
}
const char *f3r_last_error(void) { return g_last_error.c_str(); }

int f3r_device_count(int *count) {
  F3R_API_BEGIN
  F3R_CHECK(count, F3R_ERR_ARG, "null count");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  F3R_API_END
}

int f3r_ctx_create(int device, f3r_ctx *ctx) {
  F3R_API_BEGIN
  F3R_CHECK(ctx, F3R_ERR_ARG, "null ctx pointer");
  *ctx = nullptr;
  *ctx = new f3r_ctx_s{ctx_create(device)};
  F3R_API_END
}
int f3r_ctx_destroy(f3r_ctx ctx) {
  F3R_API_BEGIN
  if (!ctx) return F3R_OK;
  (void)hipStreamSynchronize(ctx->c->stream);
  delete ctx;  // objects created on the context keep it alive
  F3R_API_END
}
int f3r_ctx_sync(f3r_ctx ctx) {
  F3R_API_BEGIN
  F3R_CHECK(ctx, F3R_ERR_ARG, "null ctx");
  F3R_HIP(hipStreamSynchronize(ctx->c->stream));
  F3R_API_END
}

int f3r_vec_create(f3r_ctx ctx, int dtype, int64_t n, f3r_vec *v) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && v && n >= 0 && dtype_size(dtype), F3R_ERR_ARG, "vec_create: bad arguments");
  *v = nullptr;
  auto *h = new f3r_vec_s();
  h->v.ctx = ctx->c;
  h->v.dtype = dtype;
  h->v.n = n;
  try {
    h->v.buf = dev_alloc(dtype_size(dtype) * std::max<int64_t>(n, 1));
  } catch (...) {
    delete h;
    throw;
  }
  *v = h;
  F3R_API_END
}
int f3r_vec_view(f3r_vec base, int64_t offset, int64_t n, f3r_vec *v) {
  F3R_API_BEGIN
  F3R_CHECK(base && v && offset >= 0 && n >= 0 && offset + n <= base->v.n, F3R_ERR_ARG,
            "vec_view: range outside the base vector");
  auto *h = new f3r_vec_s();
  h->v = base->v;
  h->v.offset = base->v.offset + offset;
  h->v.n = n;
  *v = h;
  F3R_API_END
}
int f3r_vec_destroy(f3r_vec v) {
  F3R_API_BEGIN
  if (v) {
    (void)hipStreamSynchronize(v->v.ctx->stream);
    delete v;
  }
  F3R_API_END
}
int f3r_vec_upload(f3r_vec v, const void *host) {
  F3R_API_BEGIN
  F3R_CHECK(v && (host || v->v.n == 0), F3R_ERR_ARG, "vec_upload: null");
  F3R_HIP(hipMemcpyAsync(v->v.ptr(), host, v->v.n * dtype_size(v->v.dtype), hipMemcpyHostToDevice, v->v.ctx->stream));
  F3R_HIP(hipStreamSynchronize(v->v.ctx->stream));
  F3R_API_END
}
int f3r_vec_download(f3r_vec v, void *host) {
  F3R_API_BEGIN
  F3R_CHECK(v && (host || v->v.n == 0), F3R_ERR_ARG, "vec_download: null");
  F3R_HIP(hipMemcpyAsync(host, v->v.ptr(), v->v.n * dtype_size(v->v.dtype), hipMemcpyDeviceToHost, v->v.ctx->stream));
  F3R_HIP(hipStreamSynchronize(v->v.ctx->stream));
  F3R_API_END
}
int f3r_vec_info(f3r_vec v, int *dtype, int64_t *n, void **ptr) {
  F3R_API_BEGIN
  F3R_CHECK(v, F3R_ERR_ARG, "null vec");
  if (dtype) *dtype = v->v.dtype;
  if (n) *n = v->v.n;
  if (ptr) *ptr = v->v.ptr();
  F3R_API_END
}
int f3r_vec_fill(f3r_vec v, double value) {
  F3R_API_BEGIN
  F3R_CHECK(v, F3R_ERR_ARG, "null vec");
  vec_fill(&v->v, value);
  F3R_API_END
}
int f3r_vec_copy(f3r_vec s, f3r_vec d) {
  F3R_API_BEGIN
  F3R_CHECK(s && d, F3R_ERR_ARG, "null vec");
  vec_copy(&s->v, &d->v);
  F3R_API_END
}
int f3r_vec_axpy(double a, f3r_vec x, f3r_vec y) {
  F3R_API_BEGIN
  F3R_CHECK(x && y, F3R_ERR_ARG, "null vec");
  vec_axpy(a, &x->v, &y->v);
  F3R_API_END
}
int f3r_vec_scale(f3r_vec x, double a) {
  F3R_API_BEGIN
  F3R_CHECK(x, F3R_ERR_ARG, "null vec");
  vec_scale(&x->v, a);
  F3R_API_END
}
int f3r_vec_dot(f3r_vec a, f3r_vec b, double *out) {
  F3R_API_BEGIN
  F3R_CHECK(a && b && out, F3R_ERR_ARG, "null argument");
  *out = vec_dot(&a->v, &b->v);
  F3R_API_END
}
int f3r_vec_dot_noalloc(f3r_vec a, f3r_vec b, f3r_vec out) {
  F3R_API_BEGIN
  F3R_CHECK(a && b && out, F3R_ERR_ARG, "dot_noalloc: null argument");
  Vec *va = &a->v, *vb = &b->v, *vo = &out->v;
  F3R_CHECK(va->n == vb->n, F3R_ERR_ARG, "dot_noalloc: length mismatch");
  F3R_CHECK(vo->dtype == F3R_F64 && vo->n >= 1, F3R_ERR_ARG, "dot_noalloc: out must be an fp64 vector of at least one element");
  F3R_CHECK(va->ctx.get() == vb->ctx.get() && va->ctx.get() == vo->ctx.get(), F3R_ERR_ARG,
            "dot_noalloc: the vectors of one context");
  dot_noalloc(va->ctx.get(), va->dtype, va->ptr(), vb->dtype, vb->ptr(), va->n, static_cast<double *>(vo->ptr()));
  F3R_API_END
}
int f3r_vec_nrm2(f3r_vec a, double *out) {
  F3R_API_BEGIN
  F3R_CHECK(a && out, F3R_ERR_ARG, "null argument");
  *out = std::sqrt(vec_dot(&a->v, &a->v));
  F3R_API_END
}

int f3r_mat_create_csr(f3r_ctx ctx, int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rptr,
                       const int32_t *col, const double *val, f3r_mat *A) {
  F3R_API_BEGIN
  F3R_CHECK(A, F3R_ERR_ARG, "null argument");
  *A = nullptr;
  auto m = std::make_shared<Mat>();
  m->ctx = ctx ? ctx->c : nullptr;
  m->format = 0;
  m->dtype = F3R_F64;
  m->host = std::make_shared<HostCSR>(host_csr_from(nrows, ncols, nnz, rptr, col, val));
  if (ctx) {
    m->d_rptr = upload_i32(ctx->c.get(), m->host->rptr);
    m->d_idx = upload_i32(ctx->c.get(), m->host->idx);
    m->d_val = upload_values(ctx->c.get(), m->host->val, F3R_F64);
  }
  *A = new f3r_mat_s{m};
  F3R_API_END
}

int f3r_mat_to_sell(f3r_ctx ctx, f3r_mat csr, int dtype, int32_t C, int32_t sigma, int32_t K, f3r_mat *out) {
  return f3r_mat_to_sell_ex(ctx, csr, dtype, C, sigma, K, 0, out);
}

int f3r_mat_to_sell_ex(f3r_ctx ctx, f3r_mat csr, int dtype, int32_t C, int32_t sigma, int32_t K, int32_t flags,
                       f3r_mat *out) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && csr && out && csr->m->format == 0, F3R_ERR_ARG, "to_sell: needs a context and a CSR matrix");
  F3R_CHECK(C >= 16 && (C & (C - 1)) == 0, F3R_ERR_ARG, "to_sell: C must be a power of two >= 16");
  F3R_CHECK(K == 1 || K == 2 || K == 4 || K == 8, F3R_ERR_ARG, "to_sell: K must be 1, 2, 4 or 8");
  F3R_CHECK(sigma >= 1, F3R_ERR_ARG, "to_sell: sigma >= 1");
  F3R_CHECK(csr->m->host->n == csr->m->host->ncols, F3R_ERR_ARG, "to_sell: square matrices only");
  *out = nullptr;
  F3R_CHECK((flags & ~15) == 0, F3R_ERR_ARG, "to_sell: unknown flags");
  const bool lds = (flags & F3R_SELL_LDS) != 0;
  const bool lanes = (flags & F3R_SELL_TPR) != 0;
  F3R_CHECK(!lanes || ((K == 2 || K == 4 || K == 8) && (flags & (F3R_SELL_OFFSETS | F3R_SELL_LDS)) == 0),
            F3R_ERR_ARG, "to_sell: K lanes per row need K in {2, 4, 8} and the int32-index layout without LDS staging");
  F3R_CHECK(!lds || (flags & (F3R_SELL_OFFSETS | F3R_SELL_DEVICE)) == 0, F3R_ERR_ARG,
            "to_sell: LDS-staged gathers need the int32-index layout built on the host");
  if ((flags & 2) && csr->m->d_rptr) {  // build on the device
    std::shared_ptr<SellShared> sh;
    DevPtr v64;
    if (build_sell_device(ctx->c.get(), *csr->m->host, csr->m->d_rptr->as<int32_t>(), csr->m->d_idx->as<int32_t>(),
                          csr->m->d_val->as<double>(), C, sigma, K, (flags & 1) != 0, false, sh, v64)) {
      auto m = std::make_shared<Mat>();
      m->ctx = ctx->c;
      m->format = 1;
      m->dtype = dtype;
      m->host = csr->m->host;
      sh->internal = false;
      sh->nnz = csr->m->host->nnz();
      sh->lanes = lanes ? K : 1;
      m->sh = sh;
      m->s_val = cast_values_device(ctx->c.get(), v64, sh->nslots, dtype);
      m->sell_dev_val64 = v64;
      F3R_HIP(hipStreamSynchronize(ctx->c->stream));
      *out = new f3r_mat_s{m};
      return F3R_OK;
    }
  }
  HostSell hs = build_sell(*csr->m->host, C, sigma, K, nullptr, false, (flags & 1) != 0);
  auto m = std::make_shared<Mat>();
  m->ctx = ctx->c;
  m->format = 1;
  m->dtype = dtype;
  m->host = csr->m->host;
  m->sh = upload_sell_shared(ctx->c.get(), hs, false);
  m->sh->nnz = csr->m->host->nnz();
  m->sh->lanes = lanes ? K : 1;
  m->s_val = upload_sell_values(ctx->c.get(), hs, dtype);
  if (lds) {
    const std::vector<int2> w = block_windows(hs);
    const int32_t cap = (int32_t)(kLdsWindowBytes / dtype_size(dtype));  // x of the matrix's precision
    m->sh->lds_cap = cap;
    for (const int2 &q : w) m->sh->lds_blocks += q.y <= cap ? 1 : 0;
    for (int e = 0; e < 3; ++e) {
      const int32_t lim = (int32_t)(kLdsWindowBytes >> (e + 1));  // x of 2, 4, 8 bytes
      for (const int2 &q : w)
        if (q.y <= lim) m->sh->lds_fit[e] = std::max(m->sh->lds_fit[e], q.y);
    }
    m->sh->bwin = dev_alloc(sizeof(int2) * w.size());
    F3R_HIP(hipMemcpyAsync(m->sh->bwin->p, w.data(), sizeof(int2) * w.size(), hipMemcpyHostToDevice, ctx->c->stream));
    F3R_HIP(hipStreamSynchronize(ctx->c->stream));
  }
  m->sell_host_val = std::make_shared<std::vector<double>>(std::move(hs.val));
  *out = new f3r_mat_s{m};
  F3R_API_END
}

int f3r_sell_layout(int32_t n, const int32_t *rptr, const int32_t *col, const double *val, int32_t C,
                    int32_t sigma, int32_t K, int32_t *cptr, int32_t *cwidth, int32_t *sidx, double *sval,
                    int32_t *perm) {
  F3R_API_BEGIN
  F3R_CHECK(n >= 0 && rptr && cptr && cwidth && perm && C >= 1 && sigma >= 1 && K >= 1, F3R_ERR_ARG,
            "sell_layout: bad argument");
  HostCSR h = host_csr_from(n, n, rptr[n], rptr, col, val);
  HostSell hs = build_sell(h, C, sigma, K, nullptr, false);
  std::memcpy(cptr, hs.cptr.data(), sizeof(int32_t) * hs.cptr.size());
  std::memcpy(cwidth, hs.cwidth.data(), sizeof(int32_t) * hs.cwidth.size());
  std::memcpy(perm, hs.perm.data(), sizeof(int32_t) * hs.perm.size());
  if (!hs.idx.empty()) {
    F3R_CHECK(sidx && sval, F3R_ERR_ARG, "sell_layout: null output");
    std::memcpy(sidx, hs.idx.data(), sizeof(int32_t) * hs.idx.size());
    std::memcpy(sval, hs.val.data(), sizeof(double) * hs.val.size());
  }
  F3R_API_END
}

int f3r_sell_offset_layout(int32_t n, const int32_t *rptr, const int32_t *col, const double *val, int32_t C,
                           int32_t K, int64_t *nslots, int64_t *nidx, int64_t *ncoff, int32_t *cptr, int32_t *cwidth,
                           int32_t *iptr, int32_t *idx, int32_t *coff, double *sval) {
  F3R_API_BEGIN
  F3R_CHECK(n >= 0 && rptr && nslots && nidx && ncoff && C >= 64 && (C & (C - 1)) == 0 && K >= 1, F3R_ERR_ARG,
            "sell_offset_layout: bad argument (C a power of two >= 64)");
  HostCSR h = host_csr_from(n, n, rptr[n], rptr, col, val);
  HostSell hs = build_sell(h, C, 1, K, nullptr, false, true);
  *nslots = (int64_t)hs.val.size();
  *nidx = (int64_t)hs.idx.size();
  *ncoff = (int64_t)hs.coff.size();
  if (cptr) {
    F3R_CHECK(cwidth && iptr && (hs.idx.empty() || idx) && (hs.coff.empty() || coff) && (hs.val.empty() || sval),
              F3R_ERR_ARG, "sell_offset_layout: null output");
    std::memcpy(cptr, hs.cptr.data(), sizeof(int32_t) * hs.cptr.size());
    std::memcpy(cwidth, hs.cwidth.data(), sizeof(int32_t) * hs.cwidth.size());
    if (hs.iptr.empty())
      for (int32_t s = 0; s < hs.nchunks; ++s) iptr[s] = hs.cptr[s];
    else
      std::memcpy(iptr, hs.iptr.data(), sizeof(int32_t) * hs.iptr.size());
    if (!hs.idx.empty()) std::memcpy(idx, hs.idx.data(), sizeof(int32_t) * hs.idx.size());
    if (!hs.coff.empty()) std::memcpy(coff, hs.coff.data(), sizeof(int32_t) * hs.coff.size());
    if (!hs.val.empty()) std::memcpy(sval, hs.val.data(), sizeof(double) * hs.val.size());
  }
  F3R_API_END
}

int f3r_mat_cast(f3r_mat sell, int dtype, f3r_mat *out) {
  F3R_API_BEGIN
  F3R_CHECK(sell && out, F3R_ERR_ARG, "mat_cast: null argument");
  F3R_CHECK(dtype == F3R_F64 || dtype == F3R_F32 || dtype == F3R_F16 || dtype == F3R_BF16, F3R_ERR_ARG,
            "mat_cast: fp64, fp32, fp16 or bf16");
  if (sell->m->format == 0) {  // CSR<T>(CSR<double>): the index arrays shared, the values in T
    F3R_CHECK(sell->m->ctx && sell->m->d_rptr, F3R_ERR_ARG, "mat_cast: a CSR matrix on a context");
    auto m = std::make_shared<Mat>(*sell->m);
    m->dtype = dtype;
    m->c_val = dtype == F3R_F64 ? nullptr : upload_values(m->ctx.get(), m->host->val, dtype);
    F3R_HIP(hipStreamSynchronize(m->ctx->stream));
    *out = new f3r_mat_s{m};
    return F3R_OK;
  }
  F3R_CHECK(sell->m->format == 1 && (sell->m->sell_host_val || sell->m->sell_dev_val64), F3R_ERR_ARG,
            "mat_cast: needs a SELL matrix or a CSR matrix");
  auto m = std::make_shared<Mat>(*sell->m);
  m->dtype = dtype;
  if (m->sell_host_val)
    m->s_val = upload_values(m->ctx.get(), *m->sell_host_val, dtype);
  else
    m->s_val = cast_values_device(m->ctx.get(), m->sell_dev_val64, m->sh->nslots, dtype);
  *out = new f3r_mat_s{m};
  F3R_API_END
}

int f3r_mat_destroy(f3r_mat A) {
  F3R_API_BEGIN
  if (A) {
    if (A->m->ctx) (void)hipStreamSynchronize(A->m->ctx->stream);
    delete A;
  }
  F3R_API_END
}

int f3r_mat_info(f3r_mat A, int32_t *nrows, int64_t *nnz, int64_t *nslots, int *dtype, int *format) {
  F3R_API_BEGIN
  F3R_CHECK(A, F3R_ERR_ARG, "null matrix");
  const Mat &m = *A->m;
  if (nrows) *nrows = m.host ? m.host->n : m.sh->n;
  if (nnz) *nnz = m.host ? m.host->nnz() : m.sh->nnz;
  if (nslots) *nslots = m.format == 1 ? m.sh->nslots : (m.host ? m.host->nnz() : 0);
  if (dtype) *dtype = m.dtype;
  if (format) *format = m.format;
  F3R_API_END
}

int f3r_mat_set_threads_per_row(f3r_mat A, int32_t tpr) {
  F3R_API_BEGIN
  F3R_CHECK(A, F3R_ERR_ARG, "null matrix");
  F3R_CHECK(A->m->format == 0, F3R_ERR_ARG, "set_threads_per_row: a CSR matrix");
  F3R_CHECK(tpr == 0 || (tpr >= 1 && tpr <= 64 && (tpr & (tpr - 1)) == 0), F3R_ERR_ARG,
            "set_threads_per_row: 0 (automatic) or a power of two in [1, 64]");
  A->m->tpr = tpr;
  F3R_API_END
}

int f3r_mat_lds_blocks(f3r_mat A, int32_t *blocks, int32_t *staged) {
  F3R_API_BEGIN
  F3R_CHECK(A && blocks && staged, F3R_ERR_ARG, "null argument");
  F3R_CHECK(A->m->format == 1, F3R_ERR_ARG, "mat_lds_blocks: needs a SELL matrix");
  *blocks = (int32_t)((A->m->sh->n + kBlock - 1) / kBlock);
  *staged = A->m->sh->bwin ? A->m->sh->lds_blocks : 0;
  F3R_API_END
}

int f3r_mat_chunks(f3r_mat A, int32_t *chunks, int32_t *offset_chunks) {
  F3R_API_BEGIN
  F3R_CHECK(A && chunks && offset_chunks, F3R_ERR_ARG, "null argument");
  F3R_CHECK(A->m->format == 1, F3R_ERR_ARG, "mat_chunks: needs a SELL matrix");
  *chunks = A->m->sh->nchunks;
  *offset_chunks = A->m->sh->offset_chunks;
  F3R_API_END
}

int f3r_mat_setup_where(f3r_mat A, int *where) {
  F3R_API_BEGIN
  F3R_CHECK(A && where, F3R_ERR_ARG, "null argument");
  F3R_CHECK(A->m->format == 1, F3R_ERR_ARG, "mat_setup_where: needs a SELL matrix");
  *where = A->m->sell_dev_val64 ? 1 : 0;
  F3R_API_END
}

int f3r_spmv(f3r_mat A, f3r_vec x, f3r_vec y) {
  F3R_API_BEGIN
  F3R_CHECK(A && x && y, F3R_ERR_ARG, "null argument");
  F3R_CHECK(A->m->ctx, F3R_ERR_STATE, "spmv: host-only matrix (created without a context)");
  mat_apply(A->m.get(), &x->v, &y->v, nullptr);
  F3R_API_END
}
int f3r_residual(f3r_mat A, f3r_vec b, f3r_vec x, f3r_vec r) {
  F3R_API_BEGIN
  F3R_CHECK(A && b && x && r, F3R_ERR_ARG, "null argument");
  F3R_CHECK(A->m->ctx, F3R_ERR_STATE, "residual: host-only matrix (created without a context)");
  mat_apply(A->m.get(), &x->v, &r->v, &b->v);
  F3R_API_END
}
int f3r_spmv_time(f3r_mat A, f3r_vec x, f3r_vec y, int32_t reps, double *ms) {
  F3R_API_BEGIN
  F3R_CHECK(A && x && y && ms && reps > 0, F3R_ERR_ARG, "bad argument");
  hipStream_t st = A->m->ctx->stream;
  hipEvent_t e0, e1;
  F3R_HIP(hipEventCreate(&e0));
  F3R_HIP(hipEventCreate(&e1));
  mat_apply(A->m.get(), &x->v, &y->v, nullptr);  // warm the caches of the index arrays
  F3R_HIP(hipEventRecord(e0, st));
  for (int i = 0; i < reps; ++i) mat_apply(A->m.get(), &x->v, &y->v, nullptr);
  F3R_HIP(hipEventRecord(e1, st));
  F3R_HIP(hipEventSynchronize(e1));
  float t = 0;
  F3R_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  F3R_API_END
}

static int prec_create(f3r_ctx ctx, f3r_mat csr, int kind, double p1, double acc, int dtype, int flags,
                       f3r_prec *P) {
  F3R_API_BEGIN
  F3R_CHECK(csr && P && csr->m->host, F3R_ERR_ARG, "precond: needs a CSR matrix");
  F3R_CHECK(dtype == F3R_F64 || dtype == F3R_F32 || dtype == F3R_F16, F3R_ERR_ARG, "precond: dtype");
  *P = nullptr;
  auto p = std::make_shared<Prec>();
  p->ctx = ctx ? ctx->c : nullptr;
  p->kind = kind;
  p->dtype = dtype;
  p->tol = p1;
  p->acc = acc;
  const bool dev = (flags & F3R_SETUP_DEVICE) != 0;
  F3R_CHECK(!dev || (p->ctx && csr->m->d_rptr), F3R_ERR_ARG, "precond: a device set-up needs a context");
  bool built = false;
  if (dev) {
    const Mat &m = *csr->m;
    built = kind == 0 ? sdainv_device(p->ctx.get(), *m.host, m.d_rptr->as<int32_t>(), m.d_idx->as<int32_t>(),
                                      m.d_val->as<double>(), p1, acc, p->host)
                      : ilu0_device(p->ctx.get(), *m.host, m.d_rptr->as<int32_t>(), m.d_idx->as<int32_t>(),
                                    m.d_val->as<double>(), (int32_t)p1, acc, p->host);
    p->where = built ? 1 : 2;
  }
  if (!built) {
    if (kind == 0)
      p->host = sdainv_host(*csr->m->host, p1, acc);
    else
      p->host = ilu0_host(*csr->m->host, (int32_t)p1, acc);
  }
  p->n = csr->m->host->n;
  if (p->ctx) prec_build_device(p.get(), nullptr, 64, 1);
  *P = new f3r_prec_s{p};
  F3R_API_END
}

int f3r_sdainv_create(f3r_ctx ctx, f3r_mat csr, double tol, double acc, int dtype, f3r_prec *P) {
  return prec_create(ctx, csr, 0, tol, acc, dtype, 0, P);
}
int f3r_ilu0_create(f3r_ctx ctx, f3r_mat csr, int32_t blocks, double acc, int dtype, f3r_prec *P) {
  return prec_create(ctx, csr, 1, (double)blocks, acc, dtype, 0, P);
}
int f3r_sdainv_create_ex(f3r_ctx ctx, f3r_mat csr, double tol, double acc, int dtype, int flags, f3r_prec *P) {
  return prec_create(ctx, csr, 0, tol, acc, dtype, flags, P);
}
int f3r_ilu0_create_ex(f3r_ctx ctx, f3r_mat csr, int32_t blocks, double acc, int dtype, int flags, f3r_prec *P) {
  return prec_create(ctx, csr, 1, (double)blocks, acc, dtype, flags, P);
}
int f3r_factor_levels(f3r_mat csr, int both, int32_t *level, int32_t *nlevels) {
  F3R_API_BEGIN
  F3R_CHECK(csr && csr->m->host && level && nlevels, F3R_ERR_ARG, "factor_levels: needs a CSR matrix");
  std::vector<int32_t> lptr, rows;
  *nlevels = factor_levels(*csr->m->host, both != 0, lptr, rows);
  for (size_t l = 0; l + 1 < lptr.size(); ++l)
    for (int32_t k = lptr[l]; k < lptr[l + 1]; ++k) level[rows[k]] = (int32_t)l;
  F3R_API_END
}
int f3r_prec_setup_where(f3r_prec P, int *where) {
  F3R_API_BEGIN
  F3R_CHECK(P && where, F3R_ERR_ARG, "null argument");
  *where = P->p->where;
  F3R_API_END
}
int f3r_prec_apply(f3r_prec P, f3r_vec in, f3r_vec out) {
  F3R_API_BEGIN
  F3R_CHECK(P && in && out, F3R_ERR_ARG, "null argument");
  Prec *p = P->p.get();
  F3R_CHECK(p->ctx, F3R_ERR_STATE, "prec_apply: host-only preconditioner (created without a context)");
  F3R_CHECK(in->v.dtype == p->dtype && out->v.dtype == p->dtype && in->v.n == p->n && out->v.n == p->n,
            F3R_ERR_ARG, "prec_apply: vectors must have the preconditioner's type and size");
  prec_apply_raw(p, in->v.ptr(), out->v.ptr(), p->ctx->stream);
  F3R_API_END
}
int f3r_prec_factor_nnz(f3r_prec P, int which, int64_t *nnz) {
  F3R_API_BEGIN
  F3R_CHECK(P && nnz && (which == 0 || which == 1), F3R_ERR_ARG, "bad argument");
  *nnz = P->p->host.f[which].nnz();
  F3R_API_END
}
int f3r_prec_factor_get(f3r_prec P, int which, int32_t *rptr, int32_t *col, double *val) {
  F3R_API_BEGIN
  F3R_CHECK(P && rptr && (which == 0 || which == 1), F3R_ERR_ARG, "bad argument");
  const HostCSR &F = P->p->host.f[which];
  std::memcpy(rptr, F.rptr.data(), sizeof(int32_t) * (F.n + 1));
  if (F.nnz()) {
    F3R_CHECK(col && val, F3R_ERR_ARG, "null output");
    std::memcpy(col, F.idx.data(), sizeof(int32_t) * F.nnz());
    std::memcpy(val, F.val.data(), sizeof(double) * F.nnz());
  }
  F3R_API_END
}
int f3r_prec_destroy(f3r_prec P) {
  F3R_API_BEGIN
  if (P) {
    if (P->p->ctx) (void)hipStreamSynchronize(P->p->ctx->stream);
    delete P;
  }
  F3R_API_END
}

}  // extern "C"
