// The F3R nested solve on one GPU:
//   Restarted<double>(FGMRES<double>(A, L2, {m1, true}), {restart})
//   L2 = FGMRES<T2>(A2, L3, {m2, false})     fp32 basis (fp64 in DOUBLE)
//   L3 = FGMRES<T2>(A3, L4, {m3, false})     matrix values fp16 in HALF
//   L4 = adaptive Richardson(A4, M, m4, c)   fp16 storage in HALF
// as composed in work/src/f3r-gpu.cpp:210-266.
//
// Everything but the outer convergence test stays on the device: Hessenberg
// columns, Givens rotations, residual estimates and the adaptive Richardson
// weights are written by the last block of the reduction kernels that
// produce them, so no level below the outer one synchronises with the host.
// The body of level 2 (m2 level-3 solves and Arnoldi steps) is captured once
// per pattern of probe calls in a HIP graph and replayed.
//
// Vectors live in the internal row order of the SELL layout (rows sorted
// inside sigma windows, the matrix renumbered to P A P^T); b and x are
// permuted on entry and exit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "common.hpp"
#include "internal.hpp"
#include "sell_kernel.hpp"

namespace f3r {

// true when p is aligned to 4 elements (the vectorised kernels need it)
template <class T> inline bool aligned4(const T *p) { return ((uintptr_t)p % (4 * sizeof(T))) == 0; }


// ------------------------------------------------------- device state ---
// Adaptive Richardson weights of the innermost level (work/src/f3r-gpu.cpp:
// 224-227, 276-281): sum_k starts at 1, ome_k = sum_k / den, den = number of
// probes + 1.  Arithmetic in fp32 for the HALF / FLOAT variants.
struct RichDev {
  double sum[16];
  double ome[16];
  double alpha[16];
  int nprobe;
};

template <class T, int NV> struct Cols {
  const T *p[NV];
};

// Apply the previous rotations to column j of H, generate the rotation of
// step j and rotate e (Blas1::full_grot / final_grot, gmres.hpp:14-38).
template <class T>
__device__ void givens_step(T *H, T *cs, T *sn, T *e, int j, bool e0_one) {
  for (int i = 0; i < j; ++i) {
    T t = H[i], h1 = H[i + 1];
    H[i] = fma_<T>(cs[i], t, sn[i] * h1);
    H[i + 1] = fma_<T>(-sn[i], t, cs[i] * h1);
  }
  T t = (T)sqrt(fma_<T>(H[j], H[j], H[j + 1] * H[j + 1]));
  cs[j] = H[j] / t;
  sn[j] = H[j + 1] / t;
  H[j] = t;
  H[j + 1] = 0;
  T ej = (j == 0 && e0_one) ? (T)1 : e[j];
  e[j + 1] = -sn[j] * ej;
  e[j] = cs[j] * ej;
}

// back substitution of the j x j upper triangle of H (Blas1::trsv, gmres.hpp:41-53)
template <class T>
__device__ void hess_solve(const T *H, int ldh, const T *e, T *y, int j) {
  for (int i = j - 1; i >= 0; --i) {
    T t = e[i];
    for (int q = j - 1; q > i; --q) t = fma_<T>(-H[q * ldh + i], y[q], t);
    y[i] = t / H[i * ldh + i];
  }
}

// ------------------------------------------------- Arnoldi epilogues ----
// w = A z stored into V_{j+1}, and the dots <w, V_k>, k <= j, reduced in the
// same pass; the last block writes H(j, 0..j).
template <class T, int NV> struct EpiDots {
  struct State {
    T v[NV];
  };
  struct Pre {
    T v[NV];
  };
  T *__restrict__ w;
  Cols<T, NV> cols;
  int ncols;
  T *Hcol;
  T *partial;
  unsigned *counter;
  // x = z * s with s a power of two (fp16 storage of z): the sum is scaled
  // back by xs = 1/s, exactly, so w is bit-identical to the sum over z
  T xs = 1;
  __device__ Pre pre(int r) const {
    Pre p;
#pragma unroll
    for (int k = 0; k < NV; ++k) p.v[k] = k < ncols ? cols.p[k][r] : (T)0;
    return p;
  }
  __device__ void row(int r, T acc, const Pre &p, State &s) const {
    acc *= xs;
    w[r] = acc;
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k < ncols) s.v[k] = fma_<T>(acc, p.v[k], s.v[k]);
  }
  // Block sums only, into partial[k * gridDim.x + block]; k_dots_finish adds
  // them in the next launch.  A grid reduction finished in this kernel made
  // every block wait for its stores (the w rows included) and its arrival
  // atomic before retiring, which cost more than the SpMV's tail: 53 vs
  // 34 us per launch at 128^3.
  __device__ void finish(State &s) const {
    __shared__ T lds[NV * kWavesPerBlock];
    block_sum<T, NV>(s.v, lds);
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < NV; ++k) partial[(int64_t)k * gridDim.x + blockIdx.x] = s.v[k];
  }
};

// Hcol[k] = sum over the G block partials of dot k (block k of the launch,
// one per dot), in block order per thread, then a fixed tree over the 1024
// threads.  One block per dot keeps the dots' sums on different CUs: a single
// block summing 8 dots of 8192 partials took 17 us.
template <class T>
__global__ __launch_bounds__(1024) void k_dots_finish(const T *__restrict__ partial, int G, T *Hcol) {
  __shared__ T lds[16];
  // every load of a thread is issued before the first add (one memory
  // latency, not G / 1024 of them); G <= kMaxReduceGrid < 16 * 1024
  constexpr int kPer = 16;
  const T *__restrict__ p = partial + (int64_t)blockIdx.x * G;
  T t[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int b = threadIdx.x + q * 1024;
    t[q] = b < G ? p[b] : (T)0;
  }
  T v = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) v += t[q];
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T u = lds[0];
    for (int w = 1; w < 16; ++w) u += lds[w];
    Hcol[blockIdx.x] = u;
  }
}

// r = b - A x into V_0 and ||r||: e_0 = ||r||, inv = 1/||r|| (FGMRES::solve +
// the normalised start of FGMRES::kernel, gmres.hpp:84-102)
template <class T> struct EpiResidNorm {
  struct State {
    T v[1];
  };
  using Pre = NoPre;
  T *__restrict__ v0;
  const T *__restrict__ b;
  T *e;
  T *inv;
  double *status;
  T *partial;
  unsigned *counter;
  __device__ Pre pre(int) const { return {}; }
  __device__ void row(int r, T acc, const Pre &, State &s) const {
    T rv = b[r] - acc;
    v0[r] = rv;
    s.v[0] = fma_<T>(rv, rv, s.v[0]);
  }
  __device__ void finish(State &s) const {
    __shared__ T lds[kWavesPerBlock];
    if (grid_sum<T, 1>(s.v, partial, counter, lds) && threadIdx.x == 0) {
      T h = (T)sqrt(s.v[0]);
      e[0] = h;
      *inv = (T)1 / h;
      *status = fabs((double)h);
    }
  }
};

// V_{j+1} -= V H(j)[0..j] (MULMV order: V_0 h_0, then fused adds), ||V_{j+1}||,
// then the last block: H(j, j+1), 1/h, Givens, |e_{j+1}| for the host.
// VEC: every pointer is 4-element aligned; each lane then takes 4
// consecutive rows per step with 16-byte (fp32) loads and stores.
// store = 0: the updated vector is not written back (the last basis vector
// of an inner level, which nothing reads), only its norm is taken
template <class T, int NV, bool VEC = false>
__global__ __launch_bounds__(kBlock) void k_cgs_update(int64_t n, T *__restrict__ w, Cols<T, NV> cols,
                                                       int ncols, T *Hcol, T *cs, T *sn, T *e, int j,
                                                       int e0_one, T *inv, double *status, T *partial,
                                                       unsigned *counter, int store) {
  __shared__ T lds[kWavesPerBlock];
  T h[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) h[k] = k < ncols ? Hcol[k] : (T)0;
  T v[1] = {0};
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x, gstride = (int64_t)gridDim.x * kBlock;
  int64_t i0 = 0;
  if constexpr (VEC) {
    const int64_t n4 = n / 4;
    for (int64_t q = gid; q < n4; q += gstride) {
      V4<T> vk[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k)
        if (k < ncols) vk[k] = ld4(cols.p[k] + 4 * q);
      V4<T> wv = ld4(w + 4 * q);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T r = vk[0].v[c] * h[0];
#pragma unroll
        for (int k = 1; k < NV; ++k)
          if (k < ncols) r = fma_<T>(vk[k].v[c], h[k], r);
        wv.v[c] = wv.v[c] - r;
      }
      if (store) st4(w + 4 * q, wv);
#pragma unroll
      for (int c = 0; c < 4; ++c) v[0] = fma_<T>(wv.v[c], wv.v[c], v[0]);
    }
    i0 = 4 * n4;
  }
  for (int64_t i = i0 + gid; i < n; i += gstride) {
    T r = cols.p[0][i] * h[0];
#pragma unroll
    for (int k = 1; k < NV; ++k)
      if (k < ncols) r = fma_<T>(cols.p[k][i], h[k], r);
    T wn = w[i] - r;
    if (store) w[i] = wn;
    v[0] = fma_<T>(wn, wn, v[0]);
  }
  // the grid's sum of ||w||^2 in the last block (one wave of blocks: the
  // arrivals cost less than a one-block finishing launch, 0.12 ms per solve
  // at 128^3)
  if (grid_sum<T, 1>(v, partial, counter, lds) && threadIdx.x == 0) {
    T nv = (T)sqrt(v[0]);
    Hcol[j + 1] = nv;
    *inv = (T)1 / nv;
    givens_step<T>(Hcol, cs, sn, e, j, e0_one != 0);
    if (status) *status = fabs((double)e[j + 1]);
  }
}

// ||w|| from the G block sums (fixed order), then H(j, j+1), 1/h, Givens and
// |e_{j+1}| for the host
template <class T>
__global__ __launch_bounds__(1024) void k_cgs_finish(const T *__restrict__ partial, int G, T *Hcol, T *cs, T *sn,
                                                     T *e, int j, int e0_one, T *inv, double *status) {
  __shared__ T lds[16];
  constexpr int kPer = 16;  // G <= kMaxReduceGrid < 16 * 1024: all loads in flight at once
  T t[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) t[q] = threadIdx.x + q * 1024 < G ? partial[threadIdx.x + q * 1024] : (T)0;
  T v = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) v += t[q];
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = lds[0];
    for (int w = 1; w < 16; ++w) t += lds[w];
    T nv = (T)sqrt(t);
    Hcol[j + 1] = nv;
    *inv = (T)1 / nv;
    givens_step<T>(Hcol, cs, sn, e, j, e0_one != 0);
    if (status) *status = fabs((double)e[j + 1]);
  }
}

// the same for the outer level: columns V_0..V_j contiguous (ld = n), j <= 100
template <class T>
__global__ __launch_bounds__(kBlock) void k_cgs_update_strided(int64_t n, T *__restrict__ w,
                                                               const T *__restrict__ V, int ncols,
                                                               T *Hcol, T *cs, T *sn, T *e, int j, T *inv,
                                                               double *status, T *partial,
                                                               unsigned *counter) {
  __shared__ T lds[kWavesPerBlock];
  __shared__ T h[128];
  for (int k = threadIdx.x; k < ncols; k += kBlock) h[k] = Hcol[k];
  __syncthreads();
  T v[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    T r = V[i] * h[0];
    for (int k = 1; k < ncols; ++k) r = fma_<T>(V[(int64_t)k * n + i], h[k], r);
    T wn = w[i] - r;
    w[i] = wn;
    v[0] = fma_<T>(wn, wn, v[0]);
  }
  if (grid_sum<T, 1>(v, partial, counter, lds) && threadIdx.x == 0) {
    T nv = (T)sqrt(v[0]);
    Hcol[j + 1] = nv;
    *inv = (T)1 / nv;
    givens_step<T>(Hcol, cs, sn, e, j, false);
    *status = fabs((double)e[j + 1]);
  }
}

// <w, V_k> for k < ncols, V contiguous; blockIdx.y selects NV columns (NV
// sums reduced per block: 4 or 8 for the first columns, 16 beyond).
// VEC: every column 4-element aligned (n % 4 == 0); each lane then takes 4
// consecutive rows per step with 32-byte (fp64) loads
template <class T, bool VEC = false, int NV = 16>
__global__ __launch_bounds__(kBlock) void k_mdot_strided(int64_t n, const T *__restrict__ w,
                                                         const T *__restrict__ V, int ncols, T *Hcol,
                                                         T *partial, unsigned *counters) {
  __shared__ T lds[NV * kWavesPerBlock];
  const int k0 = blockIdx.y * NV;
  T v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0;
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x, gstride = (int64_t)gridDim.x * kBlock;
  if constexpr (VEC) {
    for (int64_t q = gid; q < n / 4; q += gstride) {
      const V4<T> wq = ld4(w + 4 * q);
#pragma unroll
      for (int k = 0; k < NV; ++k)
        if (k0 + k < ncols) {
          const V4<T> vq = ld4(V + (int64_t)(k0 + k) * n + 4 * q);
#pragma unroll
          for (int c = 0; c < 4; ++c) v[k] = fma_<T>(wq.v[c], vq.v[c], v[k]);
        }
    }
  } else {
    for (int64_t i = gid; i < n; i += gstride) {
      const T wi = w[i];
#pragma unroll
      for (int k = 0; k < NV; ++k)
        if (k0 + k < ncols) v[k] = fma_<T>(wi, V[(int64_t)(k0 + k) * n + i], v[k]);
    }
  }
  const int ng = (gridDim.x + kGroup - 1) / kGroup;
  if (grid_sum<T, NV>(v, partial + (int64_t)blockIdx.y * NV * (gridDim.x + ng), counters + (int64_t)blockIdx.y * kCounterSlot, lds) &&
      threadIdx.x == 0)
    for (int k = 0; k < NV && k0 + k < ncols; ++k) Hcol[k0 + k] = v[k];
}

template <class T>
__global__ void k_vscale(int64_t n, T *__restrict__ v, const T *__restrict__ inv) {
  const T a = *inv;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    v[i] *= a;
}

// out = Z y (with_init) of an inner level; every block solves the small
// triangular system itself (j <= NV).  Columns of type ZT hold z * s (fp16
// storage, s a power of two): z = column * zs is exact.
template <class T, class TO, int NV, bool VEC = false, class ZT = T>
__global__ __launch_bounds__(kBlock) void k_solution(int64_t n, Cols<ZT, NV> Z, int j, const T *H, int ldh,
                                                     const T *e, TO *__restrict__ out, T zs) {
  __shared__ T y[NV];
  if (threadIdx.x == 0) hess_solve<T>(H, ldh, e, y, j);
  __syncthreads();
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x, gstride = (int64_t)gridDim.x * kBlock;
  int64_t i0 = 0;
  if constexpr (VEC) {
    const int64_t n4 = n / 4;
    for (int64_t q = gid; q < n4; q += gstride) {
      V4<ZT> zk[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k)
        if (k < j) zk[k] = ld4(Z.p[k] + 4 * q);
      V4<TO> o;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T r = ((T)to_acc(zk[0].v[c]) * zs) * y[0];
#pragma unroll
        for (int k = 1; k < NV; ++k)
          if (k < j) r = fma_<T>((T)to_acc(zk[k].v[c]) * zs, y[k], r);
        o.v[c] = cvt<TO>(r);
      }
      st4(out + 4 * q, o);
    }
    i0 = 4 * n4;
  }
  for (int64_t i = i0 + gid; i < n; i += gstride) {
    T r = ((T)to_acc(Z.p[0][i]) * zs) * y[0];
#pragma unroll
    for (int k = 1; k < NV; ++k)
      if (k < j) r = fma_<T>((T)to_acc(Z.p[k][i]) * zs, y[k], r);
    out[i] = cvt<TO>(r);
  }
}

template <class T>
__global__ void k_hess_solve(const T *H, int ldh, const T *e, T *y, int j) {
  if (threadIdx.x == 0 && blockIdx.x == 0) hess_solve<T>(H, ldh, e, y, j);
}

// x += Z y for the outer level (Z contiguous)
template <class T>
__global__ void k_solution_strided(int64_t n, const T *__restrict__ Z, int j, const T *__restrict__ y,
                                   T *__restrict__ x) {
  __shared__ T ys[128];
  for (int k = threadIdx.x; k < j; k += kBlock) ys[k] = y[k];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    T r = Z[i] * ys[0];
    for (int k = 1; k < j; ++k) r = fma_<T>(Z[(int64_t)k * n + i], ys[k], r);
    x[i] = x[i] + r;
  }
}

template <class S, class D>
__global__ void k_cast2(int64_t n, const S *__restrict__ s, D *__restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    d[i] = cvt<D>(to_acc(s[i]));
}

__global__ void k_gather(int64_t n, const double *__restrict__ s, const int32_t *__restrict__ perm,
                         double *__restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    d[i] = s[perm[i]];
}
__global__ void k_scatter(int64_t n, const double *__restrict__ s, const int32_t *__restrict__ perm,
                          double *__restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    d[perm[i]] = s[i];
}

// ---------------------------------------------- Richardson (fp16 level) ---
// the interleaved diagonal factors of a row: (w, z) in one 32-bit word
__device__ __forceinline__ half_t dwz_w(uint32_t v) { return __builtin_bit_cast(half_t, (uint16_t)(v & 0xffffu)); }
__device__ __forceinline__ half_t dwz_z(uint32_t v) { return __builtin_bit_cast(half_t, (uint16_t)(v >> 16)); }
// Entry of one call (work/src/f3r-gpu.cpp:231-243 in this build's
// arithmetic): hin = fp16(in * s); with diagonal SDAINV factors also
// hout = fp16(z * fp16(w * hin)) and, outside probe calls, hout *= ome_0.
// `inv` (optional): `in` is a basis vector not yet normalised; it is scaled
// by *inv in place first (the normalisation of the Arnoldi step that made it).
// VEC: 4 consecutive rows per lane (every array 4-element aligned).
__device__ __forceinline__ void l4_entry_row(float &v, const float iv, bool has_inv, float s, half_t &h, half_t &o,
                                             half_t dwi, half_t dzi, bool diag, bool scale_ome, float om) {
  if (has_inv) v *= iv;
  h = cvt<half_t>(v * s);
  if (diag) {
    const half_t t = cvt<half_t>(__builtin_fmaf((float)dwi, (float)h, 0.f));
    o = cvt<half_t>(__builtin_fmaf((float)dzi, (float)t, 0.f));
    if (scale_ome) o = cvt<half_t>((float)o * om);
  }
}
template <bool VEC>
__global__ void k_l4_entry(int64_t n, float *__restrict__ in, const float *__restrict__ inv, float s,
                           half_t *__restrict__ hin, half_t *__restrict__ hout, const uint32_t *__restrict__ dwz,
                           const RichDev *st, int scale_ome) {
  const float om = scale_ome ? (float)st->ome[0] : 1.f;
  const float iv = inv ? *inv : 1.f;
  const bool diag = dwz != nullptr;
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x, gstride = (int64_t)gridDim.x * kBlock;
  int64_t i0 = 0;
  if constexpr (VEC) {
    const int64_t n4 = n / 4;
    for (int64_t q = gid; q < n4; q += gstride) {
      V4<float> v = ld4(in + 4 * q);
      V4<uint32_t> wz{};
      V4<half_t> h, o{};
      if (diag) wz = ld4(dwz + 4 * q);
#pragma unroll
      for (int c = 0; c < 4; ++c)
        l4_entry_row(v.v[c], iv, inv != nullptr, s, h.v[c], o.v[c], dwz_w(wz.v[c]), dwz_z(wz.v[c]), diag,
                     scale_ome != 0, om);
      if (inv) st4(in + 4 * q, v);
      st4(hin + 4 * q, h);
      if (diag) st4(hout + 4 * q, o);
    }
    i0 = 4 * n4;
  }
  for (int64_t i = i0 + gid; i < n; i += gstride) {
    float v = in[i];
    half_t h, o = (half_t)0;
    const uint32_t wz = diag ? dwz[i] : 0u;
    l4_entry_row(v, iv, inv != nullptr, s, h, o, dwz_w(wz), dwz_z(wz), diag, scale_ome != 0, om);
    if (inv) in[i] = v;
    hin[i] = h;
    if (diag) hout[i] = o;
  }
}

// hout *= ome_0 (or alpha_0 in probe calls)
__global__ void k_l4_scale(int64_t n, half_t *__restrict__ hout, const double *w) {
  const float a = (float)*w;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    hout[i] = cvt<half_t>((float)hout[i] * a);
}

// hout' = fp16(fma(w_k, tmp2, hout)); last step: out = fp32(hout') / s
__global__ void k_l4_update(int64_t n, const half_t *__restrict__ tmp2, const half_t *__restrict__ hout,
                            const double *w, half_t *__restrict__ hnext, float *__restrict__ out,
                            float inv_s) {
  const float a = (float)*w;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const half_t h = cvt<half_t>(__builtin_fmaf(a, (float)tmp2[i], (float)hout[i]));
    if (out)
      out[i] = (float)h * inv_s;
    else
      hnext[i] = h;
  }
}

// Richardson step k fused into the fp16 residual SpMV, diagonal factors:
// hr = fp16(hin - A16 hout), tmp2 = fp16(z * fp16(w * hr)), and outside
// probe calls hout' = fp16(fma(ome_k, tmp2, hout)) (last step: out = hout'/s).
struct EpiL4Diag {
  using State = NoState;
  struct Pre {
    half_t hin, hout;
    uint32_t wz;
  };
  const half_t *__restrict__ hin;
  const half_t *__restrict__ hout;
  const uint32_t *__restrict__ dwz;  // (w, z) diagonals interleaved
  const double *w;  // ome_k
  half_t *__restrict__ hr_out;    // probe calls: keep hr and tmp2
  half_t *__restrict__ tmp2_out;
  half_t *__restrict__ hnext;
  float *__restrict__ out;
  float inv_s;
  __device__ Pre pre(int r) const { return {hin[r], hr_out ? (half_t)0 : hout[r], dwz[r]}; }
  __device__ void row(int r, float acc, const Pre &p, State &) const {
    const half_t hr = cvt<half_t>((float)p.hin - acc);
    const half_t t = cvt<half_t>(__builtin_fmaf((float)dwz_w(p.wz), (float)hr, 0.f));
    const half_t t2 = cvt<half_t>(__builtin_fmaf((float)dwz_z(p.wz), (float)t, 0.f));
    if (hr_out) {
      hr_out[r] = hr;
      tmp2_out[r] = t2;
      return;
    }
    const half_t h = cvt<half_t>(__builtin_fmaf((float)*w, (float)t2, (float)p.hout));
    if (out)
      out[r] = (float)h * inv_s;
    else
      hnext[r] = h;
  }
  __device__ void finish(State &) const {}
};

// probe: alpha_k = <t, r> / <t, t> (fp32 sums of fp16 data); the last block
// updates sum_k, ome_k (work/src/f3r-gpu.cpp:234-240, 248-254)
__global__ __launch_bounds__(kBlock) void k_probe(int64_t n, const half_t *__restrict__ t,
                                                  const half_t *__restrict__ r, RichDev *st, int k,
                                                  float *partial, unsigned *counter) {
  __shared__ float lds[2 * kWavesPerBlock];
  float v[2] = {0.f, 0.f};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float ti = (float)t[i];
    v[0] = __builtin_fmaf(ti, ti, v[0]);
    v[1] = __builtin_fmaf(ti, (float)r[i], v[1]);
  }
  if (grid_sum<float, 2>(v, partial, counter, lds) && threadIdx.x == 0) {
    const float al = v[1] / v[0];
    if (k == 0) st->nprobe += 1;
    const float den = (float)(st->nprobe + 1);
    const float s = (float)st->sum[k] + al;
    st->sum[k] = s;
    st->ome[k] = s / den;
    st->alpha[k] = al;
  }
}

// weights 1 (adaptive start), or the static weight w > 0 for every step
__global__ void k_rich_reset(RichDev *st, double w) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int k = 0; k < 16; ++k) {
      st->sum[k] = 1.0;
      st->ome[k] = w > 0 ? w : 1.0;
      st->alpha[k] = 0.0;
    }
    st->nprobe = 0;
  }
}

// ---------------------------- Richardson of the FLOAT / DOUBLE builds ---
template <class T>
__global__ void k_scale_t(int64_t n, T *__restrict__ out, const double *w) {
  const T a = (T)*w;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    out[i] *= a;
}
template <class T>
__global__ void k_update_t(int64_t n, const T *__restrict__ t2, T *__restrict__ out, const double *w) {
  const T a = (T)*w;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    out[i] = fma_<T>(a, t2[i], out[i]);
}
// alpha_k = <t, r> / <t, t> summed in T; the last block updates sum_k, ome_k
template <class T>
__global__ __launch_bounds__(kBlock) void k_probe_t(int64_t n, const T *__restrict__ t, const T *__restrict__ r,
                                                    RichDev *st, int k, T *partial, unsigned *counter) {
  __shared__ T lds[2 * kWavesPerBlock];
  T v[2] = {0, 0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T ti = t[i];
    v[0] = fma_<T>(ti, ti, v[0]);
    v[1] = fma_<T>(ti, r[i], v[1]);
  }
  if (grid_sum<T, 2>(v, partial, counter, lds) && threadIdx.x == 0) {
    const T al = v[1] / v[0];
    if (k == 0) st->nprobe += 1;
    const T den = (T)(st->nprobe + 1);
    const T s = (T)st->sum[k] + al;
    st->sum[k] = s;
    st->ome[k] = s / den;
    st->alpha[k] = al;
  }
}

// ------------------------------------------ generic levels (depth study) ---
// FGMRES levels of the nesting-depth ablations (work/src/depth.cpp:122-167):
// basis storage S (fp32 or fp16), fp32 arithmetic.  An fp16 level keeps its
// vectors scaled by a power of two s (stored = s * value) so unit vectors
// stay clear of fp16 subnormals; dots and norms take 1/s^2, the output 1/s.

// out = SO(in * factor) (conversion into a level's storage)
template <class SI, class SO>
__global__ void k_g_entry(int64_t n, const SI *__restrict__ in, SO *__restrict__ out, double factor) {
  using L = decltype(to_acc(SI{}));
  const L f = (L)factor;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    out[i] = cvt<SO>(to_acc(in[i]) * f);
}

// block sums of <w, V_k> for the 16 columns of blockIdx.y (k < ncols) into
// partial[k * gridDim.x + block]
template <class S>
__global__ __launch_bounds__(kBlock) void k_g_mdot(int64_t n, const S *__restrict__ w, const S *__restrict__ V,
                                                   int64_t ld, int ncols, float *__restrict__ partial) {
  constexpr int NV = 16;
  __shared__ float lds[NV * kWavesPerBlock];
  const int k0 = blockIdx.y * NV;
  float v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float wi = to_acc(w[i]);
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k0 + k < ncols) v[k] = fma_<float>(wi, to_acc(V[(int64_t)(k0 + k) * ld + i]), v[k]);
  }
  block_sum<float, NV>(v, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV && k0 + k < ncols; ++k) partial[(int64_t)(k0 + k) * gridDim.x + blockIdx.x] = v[k];
}

// Hcol[k] = scale * the G block sums of column k (one block per column)
__global__ __launch_bounds__(kBlock) void k_g_dots_finish(const float *__restrict__ partial, int G, float *Hcol,
                                                          float scale) {
  __shared__ float lds[kWavesPerBlock];
  const int k = blockIdx.x;
  float v[1] = {0.f};
  for (int b = threadIdx.x; b < G; b += kBlock) v[0] += partial[(int64_t)k * G + b];
  block_sum<float, 1>(v, lds);
  if (threadIdx.x == 0) Hcol[k] = v[0] * scale;
}

// w = S(w - V h) over ncols <= 128 columns; block sums of scale * w^2 of the
// stored w into partial[block] (k_cgs_finish adds them)
template <class S>
__global__ __launch_bounds__(kBlock) void k_g_update(int64_t n, S *__restrict__ w, const S *__restrict__ V, int64_t ld,
                                                     int ncols, const float *Hcol, float scale,
                                                     float *__restrict__ partial) {
  __shared__ float h[128];
  __shared__ float lds[kWavesPerBlock];
  for (int k = threadIdx.x; k < ncols; k += kBlock) h[k] = Hcol[k];
  __syncthreads();
  float v[1] = {0.f};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    float r = to_acc(V[i]) * h[0];
    for (int k = 1; k < ncols; ++k) r = fma_<float>(to_acc(V[(int64_t)k * ld + i]), h[k], r);
    const S ws = cvt<S>(to_acc(w[i]) - r);
    w[i] = ws;
    const float wf = to_acc(ws);
    v[0] = fma_<float>(wf * scale, wf, v[0]);
  }
  block_sum<float, 1>(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

template <class S> __global__ void k_g_vscale(int64_t n, S *__restrict__ v, const float *inv) {
  const float a = *inv;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    v[i] = cvt<S>(to_acc(v[i]) * a);
}

// out = TO(factor * Z y), y from the level's m x m triangle (every block
// solves it; m <= 128)
template <class S, class TO>
__global__ __launch_bounds__(kBlock) void k_g_combine(int64_t n, const S *__restrict__ Z, int64_t ld, int m,
                                                      const float *H, int ldh, const float *e, TO *__restrict__ out,
                                                      float factor) {
  __shared__ float y[128];
  if (threadIdx.x == 0) hess_solve<float>(H, ldh, e, y, m);
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    float r = to_acc(Z[i]) * y[0];
    for (int k = 1; k < m; ++k) r = fma_<float>(to_acc(Z[(int64_t)k * ld + i]), y[k], r);
    out[i] = cvt<TO>(r * factor);
  }
}

// ------------------------------------------------------------ the solver ---
template <class T> struct InnerLevel {
  int m = 0;
  DevPtr Vbuf, Zbuf, H, cs, sn, e, inv;
  std::vector<T *> V, Z;  // column pointers; V[0] aliases the level's input
  // level 3 of the HALF variant: Z_j as the fp16 Richardson output h
  // (Z_j = h / s exactly, s the power-of-two range scale), half the bytes
  DevPtr Zhbuf;
  std::vector<half_t *> Zh;
  int ldh() const { return m + 1; }
  T *Hp() const { return H->as<T>(); }
};

struct Solver {
  std::shared_ptr<Context> ctx;
  f3r_params p{};
  int32_t n = 0;
  int64_t nnz = 0;
  std::shared_ptr<SellShared> sh;    // layout of A64 / A32
  std::shared_ptr<SellShared> sh16;  // layout of A16 (sh itself when K agrees)
  DevPtr A64, A32, A16;
  std::shared_ptr<Prec> P;
  float range_s = 1.f;
  // outer level (fp64)
  DevPtr V1, Z1, H1, c1, s1, e1, y1, inv1, status, b_int, x_int;
  // inner levels: fp32 basis (HALF, FLOAT) or fp64 basis (DOUBLE)
  InnerLevel<float> L2f, L3f;
  InnerLevel<double> L2d, L3d;
  // the nesting-depth ablations: FGMRES levels from the outermost inner one
  // down, then the preconditioner applied directly
  struct GLevel {
    int m = 0;
    int store = F3R_F32;  // basis storage: F3R_F32 or F3R_F16 (fp32 arithmetic)
    int mat = F3R_F32;    // the level's matrix values: F3R_F32 or F3R_F16
    float s = 1.f;        // storage scale (stored = s * value), fp16 levels
    int64_t ld = 0;
    DevPtr V, Z, H, cs, sn, e, inv;
    template <class S> S *Vc(int k) const { return V->as<S>() + (int64_t)k * ld; }
    template <class S> S *Zc(int k) const { return Z->as<S>() + (int64_t)k * ld; }
  };
  std::vector<GLevel> glev;
  // innermost operator under the generic levels: 0 M applied directly (the
  // depth ablations), 1 the Richardson of the F3R nest (m4, c, static_weight
  // of p), 2 the caller's operator (f3r_op)
  int inner = 0;
  f3r_apply_fn op_fn = nullptr;
  void *op_user = nullptr;
  int op_dtype = F3R_F32;
  DevPtr op_in, op_out;
  // stopping rule of the outer level: null = |e| / ||b|| < eps
  f3r_cond_fn cond_fn = nullptr;
  void *cond_user = nullptr;
  // Richardson work vectors (fp16 in HALF, else the basis type)
  DevPtr hin, hout, hout2, hr, tmp2, tt, rich;
  int64_t cnt = 0, calls = 0;
  std::vector<double> hist;
  std::map<uint64_t, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_src;
  // 1/||w|| of a basis vector left unnormalised by the last Arnoldi step;
  // the next innermost entry kernel (its first reader) applies it
  const float *pending_inv = nullptr;
  bool layouts_on_device = false;  // the matrix layouts were built on the device
  int red_grid = 1;   // capped grid of the outer level's strided reductions
  int dots_grid = 1;  // blocks of the streaming kernels with a grid reduction
  int spmv_grid = 0;  // blocks of the SpMV kernels with a fused grid reduction (0 = one row per thread)
  // capture of level-4 inputs (fp32, internal row order): every cap_every-th
  // call until cap_max vectors are held; level 2 runs eagerly meanwhile
  DevPtr cap;
  int32_t cap_max = 0, cap_every = 1, cap_n = 0;
  bool capturing() const { return cap_n < cap_max; }

  ~Solver() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    for (auto &g : graph_src) (void)hipGraphDestroy(g);
  }

  hipStream_t st() const { return ctx->stream; }
  int grid() const {
    int64_t g = (n + kBlock - 1) / kBlock;
    return (int)std::min<int64_t>(g, (int64_t)ctx->num_cu * 16);
  }
  SellArgs args() const { return sell_args(*sh, false); }
  // the layout of the matrix with value type V (the fp16 copy may differ in K)
  template <class V> const SellShared &shv() const {
    if constexpr (std::is_same<V, half_t>::value) return *sh16; else return *sh;
  }
  template <class V> SellArgs argsv() const { return sell_args(shv<V>(), false); }
  unsigned *counter(int i) const { return ctx->red.counters->as<unsigned>() + (int64_t)i * kCounterSlot; }
  template <class T> T *partial() const { return ctx->red.partial->as<T>(); }
  template <class T> InnerLevel<T> &L2() {
    if constexpr (std::is_same<T, float>::value) return L2f; else return L2d;
  }
  template <class T> InnerLevel<T> &L3() {
    if constexpr (std::is_same<T, float>::value) return L3f; else return L3d;
  }
  int m2() const { return p.variant == 2 ? L2d.m : L2f.m; }
  int m3() const { return p.variant == 2 ? L3d.m : L3f.m; }

  // ---------------------------------------------------- inner Arnoldi ---
  template <class T, class V, int NV>
  void spmv_dots(InnerLevel<T> &L, const V *val, int j) {
    Cols<T, NV> c{};
    for (int k = 0; k <= j; ++k) c.p[k] = L.V[k];
    EpiDots<T, NV> epi{L.V[j + 1], c, j + 1, L.Hp() + (int64_t)j * L.ldh(), partial<T>(), counter(0)};
    if (!L.Zh.empty()) {
      epi.xs = (T)(1.0 / range_s);
      launch_sell<V, half_t, T>(shv<V>().K, argsv<V>(), val, L.Zh[j], epi, st(), spmv_grid);
    } else {
      launch_sell<V, T, T>(shv<V>().K, argsv<V>(), val, L.Z[j], epi, st(), spmv_grid);
    }
    // block partials the SpMV + dots kernel wrote per dot: its grid
    hipLaunchKernelGGL(k_dots_finish<T>, dim3(j + 1), dim3(1024), 0, st(), (const T *)partial<T>(),
                       sell_blocks(argsv<V>(), spmv_grid), L.Hp() + (int64_t)j * L.ldh());
  }
  template <class T, int NV> void cgs_update(InnerLevel<T> &L, int j) {
    Cols<T, NV> c{};
    for (int k = 0; k <= j; ++k) c.p[k] = L.V[k];
    bool vec = aligned4(L.V[j + 1]);
    for (int k = 0; k <= j; ++k) vec = vec && aligned4(L.V[k]);
    auto kern = vec ? k_cgs_update<T, NV, true> : k_cgs_update<T, NV, false>;
    hipLaunchKernelGGL(kern, dim3(dots_grid), dim3(kBlock), 0, st(), (int64_t)n,
                       L.V[j + 1], c, j + 1, L.Hp() + (int64_t)j * L.ldh(), L.cs->template as<T>(),
                       L.sn->template as<T>(), L.e->template as<T>(), j, 1, L.inv->template as<T>(),
                       (double *)nullptr, partial<T>(), counter(1), j + 1 < L.m ? 1 : 0);
  }
  template <class T, class V> void arnoldi(InnerLevel<T> &L, const V *val, int j) {
    if (j + 1 <= 4) {
      spmv_dots<T, V, 4>(L, val, j);
      cgs_update<T, 4>(L, j);
    } else if (j + 1 <= 8) {
      spmv_dots<T, V, 8>(L, val, j);
      cgs_update<T, 8>(L, j);
    } else {
      spmv_dots<T, V, 16>(L, val, j);
      cgs_update<T, 16>(L, j);
    }
    // the last basis vector of an inner level is never read (the level's
    // output is Z y), so it is left unnormalised.  In the HALF variant the
    // first reader of V[j+1] is the next innermost entry kernel (level 3's
    // next step, or level 3's start when this is level 2), which normalises
    // it in passing.
    if (j + 1 < L.m) {
      if constexpr (std::is_same<T, float>::value) {
        if (p.variant == 0) {
          pending_inv = L.inv->template as<float>();
          return;
        }
      }
      hipLaunchKernelGGL(k_vscale<T>, dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n, L.V[j + 1],
                         (const T *)L.inv->template as<T>());
    }
  }
  template <class T, class TO> void solution(InnerLevel<T> &L, TO *out) {
    const int j = L.m;
    auto go = [&](auto nvtag) {
      constexpr int NV = decltype(nvtag)::value;
      if (!L.Zh.empty()) {
        Cols<half_t, NV> c{};
        for (int k = 0; k < j; ++k) c.p[k] = L.Zh[k];
        bool vec = aligned4(out);
        for (int k = 0; k < j; ++k) vec = vec && aligned4(L.Zh[k]);
        auto kern = vec ? k_solution<T, TO, NV, true, half_t> : k_solution<T, TO, NV, false, half_t>;
        hipLaunchKernelGGL(kern, dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n, c, j, (const T *)L.Hp(),
                           L.ldh(), (const T *)L.e->template as<T>(), out, (T)(1.0 / range_s));
        return;
      }
      Cols<T, NV> c{};
      for (int k = 0; k < j; ++k) c.p[k] = L.Z[k];
      bool vec = aligned4(out);
      for (int k = 0; k < j; ++k) vec = vec && aligned4(L.Z[k]);
      auto kern = vec ? k_solution<T, TO, NV, true> : k_solution<T, TO, NV, false>;
      hipLaunchKernelGGL(kern, dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n, c, j,
                         (const T *)L.Hp(), L.ldh(), (const T *)L.e->template as<T>(), out, (T)1);
    };
    if (j <= 4) go(std::integral_constant<int, 4>());
    else if (j <= 8) go(std::integral_constant<int, 8>());
    else go(std::integral_constant<int, 16>());
  }

  // ------------------------------------------------- level 4 (fp16) ---
  // out: fp32 result h / s; or, with hz, the fp16 result h itself
  void level4_half(const float *in, float *out, half_t *hz = nullptr) {
    ++cnt;
    ++calls;
    const bool probe = p.static_weight <= 0 && (cnt % p.c) == 0;
    const int64_t nn = n;
    const dim3 g(grid()), b(kBlock);
    RichDev *rs = rich->as<RichDev>();
    half_t *hi = hin->as<half_t>(), *ho = hout->as<half_t>(), *ho2 = hout2->as<half_t>();
    const float inv_s = 1.f / range_s;
    const bool diag = P->kind == 0 && P->diag;
    const uint32_t *dwz = diag ? P->dwz->as<uint32_t>() : nullptr;
    const half_t *a16 = A16->as<half_t>();
    SellArgs a = argsv<half_t>();
    const int K16 = sh16->K;
    // entry (normalising `in` first when the Arnoldi step that made it did not)
    const float *pinv = pending_inv;
    pending_inv = nullptr;
    const bool vec = aligned4(in) && aligned4(hi) && aligned4(ho) && (!diag || aligned4(dwz));
    if (vec)
      hipLaunchKernelGGL(k_l4_entry<true>, g, b, 0, st(), nn, const_cast<float *>(in), pinv, range_s, hi, ho, dwz,
                         (const RichDev *)rs, (diag && !probe) ? 1 : 0);
    else
      hipLaunchKernelGGL(k_l4_entry<false>, g, b, 0, st(), nn, const_cast<float *>(in), pinv, range_s, hi, ho, dwz,
                         (const RichDev *)rs, (diag && !probe) ? 1 : 0);
    if (!diag) {
      prec_apply_raw(P.get(), hi, ho, st());
      if (!probe) hipLaunchKernelGGL(k_l4_scale, g, b, 0, st(), nn, ho, (const double *)&rs->ome[0]);
    }
    if (probe) {
      half_t *t = tt->as<half_t>();
      launch_sell<half_t, half_t, float>(K16, a, a16, ho, EpiStore<half_t, float>{t, nullptr}, st());
      hipLaunchKernelGGL(k_probe, dim3(dots_grid), b, 0, st(), nn, (const half_t *)t, (const half_t *)hi, rs,
                         0, partial<float>(), counter(2));
      hipLaunchKernelGGL(k_l4_scale, g, b, 0, st(), nn, ho, (const double *)&rs->alpha[0]);
    }
    for (int k = 1; k < p.m4; ++k) {
      const bool last = k == p.m4 - 1;
      half_t *hr_ = hr->as<half_t>(), *t2 = tmp2->as<half_t>();
      half_t *hn = last && hz ? hz : ho2;     // the step's fp16 result
      float *of = last && !hz ? out : nullptr;  // or, last step, its fp32 h / s
      if (diag) {
        EpiL4Diag epi{hi, ho, dwz, &rs->ome[k], probe ? hr_ : nullptr, probe ? t2 : nullptr,
                      hn, of, inv_s};
        launch_sell<half_t, half_t, float>(K16, a, a16, ho, epi, st());
      } else {
        launch_sell<half_t, half_t, float>(K16, a, a16, ho, EpiResid<half_t, half_t, float>{hr_, hi, nullptr},
                                           st());
        prec_apply_raw(P.get(), hr_, t2, st());
      }
      if (probe) {
        half_t *t = tt->as<half_t>();
        launch_sell<half_t, half_t, float>(K16, a, a16, t2, EpiStore<half_t, float>{t, nullptr}, st());
        hipLaunchKernelGGL(k_probe, dim3(dots_grid), b, 0, st(), nn, (const half_t *)t, (const half_t *)hr_, rs,
                           k, partial<float>(), counter(2));
        hipLaunchKernelGGL(k_l4_update, g, b, 0, st(), nn, (const half_t *)t2, (const half_t *)ho,
                           (const double *)&rs->alpha[k], hn, of, inv_s);
      } else if (!diag) {
        hipLaunchKernelGGL(k_l4_update, g, b, 0, st(), nn, (const half_t *)t2, (const half_t *)ho,
                           (const double *)&rs->ome[k], hn, of, inv_s);
      }
      if (!last) std::swap(ho, ho2);
    }
    if (p.m4 == 1) {  // no Richardson step: out = fp32(hout) / s, or hz = hout
      if (hz)
        hipLaunchKernelGGL((k_cast2<half_t, half_t>), g, b, 0, st(), nn, (const half_t *)ho, hz);
      else
        hipLaunchKernelGGL((k_g_entry<half_t, float>), g, b, 0, st(), nn, (const half_t *)ho, out, (double)inv_s);
    }
  }

  // ------------------------------------ level 4 of the FLOAT / DOUBLE builds ---
  // Vectors and weights in T, preconditioner in T, SpMVs with the fp64 A
  // summed in fp64 (work/src/f3r-gpu.cpp:123-152, 174-203); `out` is updated
  // in place.
  template <class T> void level4_t(const T *in, T *out) {
    ++cnt;
    ++calls;
    const bool probe = p.static_weight <= 0 && (cnt % p.c) == 0;
    const int64_t nn = n;
    const dim3 g(grid()), b(kBlock);
    RichDev *rs = rich->as<RichDev>();
    const double *a64 = A64->as<double>();
    SellArgs a = args();
    T *t = tt->as<T>(), *hr_ = hr->as<T>(), *t2 = tmp2->as<T>();
    prec_apply_raw(P.get(), in, out, st());
    if (probe) {
      launch_sell<double, T, double>(sh->K, a, a64, out, EpiStore<T, double>{t, nullptr}, st());
      hipLaunchKernelGGL(k_probe_t<T>, dim3(dots_grid), b, 0, st(), nn, (const T *)t, in, rs, 0, partial<T>(),
                         counter(2));
      hipLaunchKernelGGL(k_scale_t<T>, g, b, 0, st(), nn, out, (const double *)&rs->alpha[0]);
    } else {
      hipLaunchKernelGGL(k_scale_t<T>, g, b, 0, st(), nn, out, (const double *)&rs->ome[0]);
    }
    for (int k = 1; k < p.m4; ++k) {
      launch_sell<double, T, double>(sh->K, a, a64, out, EpiResid<T, T, double>{hr_, in, nullptr}, st());
      prec_apply_raw(P.get(), hr_, t2, st());
      if (probe) {
        launch_sell<double, T, double>(sh->K, a, a64, t2, EpiStore<T, double>{t, nullptr}, st());
        hipLaunchKernelGGL(k_probe_t<T>, dim3(dots_grid), b, 0, st(), nn, (const T *)t, (const T *)hr_, rs, k,
                           partial<T>(), counter(2));
        hipLaunchKernelGGL(k_update_t<T>, g, b, 0, st(), nn, (const T *)t2, out, (const double *)&rs->alpha[k]);
      } else {
        hipLaunchKernelGGL(k_update_t<T>, g, b, 0, st(), nn, (const T *)t2, out, (const double *)&rs->ome[k]);
      }
    }
  }

  // hz (HALF variant, level 3): the fp16 result instead of `out`
  template <class T> void level4(const T *in, T *out, half_t *hz = nullptr) {
    if (capturing() && calls % cap_every == 0) {
      float *dst = cap->as<float>() + (int64_t)cap_n * n;
      if constexpr (std::is_same<T, float>::value)
        F3R_HIP(hipMemcpyAsync(dst, in, sizeof(float) * n, hipMemcpyDeviceToDevice, st()));
      else
        hipLaunchKernelGGL((k_cast2<T, float>), dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n, in, dst);
      ++cap_n;
    }
    if constexpr (std::is_same<T, float>::value) {
      if (p.variant == 0) {
        level4_half(in, out, hz);
        return;
      }
    }
    level4_t<T>(in, out);
  }

  // ----------------------------------------------- level 3 and level 2 ---
  template <class T> void level3(const T *in, T *out) {
    InnerLevel<T> &L = L3<T>();
    L.V[0] = const_cast<T *>(in);
    for (int j = 0; j < L.m; ++j) {
      level4<T>(L.V[j], L.Z[j], L.Zh.empty() ? nullptr : L.Zh[j]);
      if constexpr (std::is_same<T, double>::value) {
        arnoldi<T, double>(L, A64->as<double>(), j);
      } else {
        if (p.variant == 0)
          arnoldi<T, half_t>(L, A16->as<half_t>(), j);
        else
          arnoldi<T, float>(L, A32->as<float>(), j);
      }
    }
    solution<T, T>(L, out);
  }

  template <class T> void level2_body() {
    InnerLevel<T> &L = L2<T>();
    for (int j = 0; j < L.m; ++j) {
      level3<T>(L.V[j], L.Z[j]);
      if constexpr (std::is_same<T, double>::value)
        arnoldi<T, double>(L, A64->as<double>(), j);
      else
        arnoldi<T, float>(L, A32->as<float>(), j);
    }
  }

  // probe positions of the next level-2 body as a bit mask (bit q: the q-th
  // level-4 call probes); UINT64_MAX when not representable
  uint64_t next_pattern() const {
    const int64_t calls_per = (int64_t)m2() * m3();
    if (calls_per > 63) return UINT64_MAX;
    uint64_t mask = 0;
    for (int64_t q = 0; q < calls_per && p.static_weight <= 0; ++q)
      if (((cnt + q + 1) % p.c) == 0) mask |= 1ULL << q;
    return mask | (1ULL << 63);
  }

  template <class T> void level2(const double *in, double *out) {
    InnerLevel<T> &L = L2<T>();
    hipLaunchKernelGGL((k_cast2<double, T>), dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n, in, L.V[0]);
    uint64_t key = p.use_graph && !capturing() ? next_pattern() : UINT64_MAX;
    if (key == UINT64_MAX || graphs.size() > 32) {
      level2_body<T>();
    } else {
      auto it = graphs.find(key);
      if (it == graphs.end()) {
        const int64_t cnt0 = cnt, calls0 = calls;
        hipGraph_t gr;
        F3R_HIP(hipStreamBeginCapture(st(), hipStreamCaptureModeThreadLocal));
        try {
          level2_body<T>();
        } catch (...) {
          (void)hipStreamEndCapture(st(), &gr);
          throw;
        }
        F3R_HIP(hipStreamEndCapture(st(), &gr));
        hipGraphExec_t ex;
        F3R_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        graph_src.push_back(gr);
        it = graphs.emplace(key, ex).first;
        cnt = cnt0;  // capture only recorded the calls
        calls = calls0;
      }
      F3R_HIP(hipGraphLaunch(it->second, st()));
      cnt += (int64_t)m2() * m3();
      calls += (int64_t)m2() * m3();
    }
    solution<T, double>(L, out);
  }

  // ------------------------------------ generic levels (depth ablations) ---
  // M applied directly to a vector of the caller's storage SI, scaled by s_in
  template <class SI> void mdirect(const SI *in, SI *out, float s_in) {
    ++cnt;
    ++calls;
    if constexpr (std::is_same<SI, half_t>::value) {
      prec_apply_raw(P.get(), in, out, st());  // M is linear: the scale carries through
    } else {
      half_t *hi = hin->as<half_t>(), *ho = hout->as<half_t>();
      hipLaunchKernelGGL((k_g_entry<SI, half_t>), dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n, in, hi,
                         (double)range_s / s_in);
      prec_apply_raw(P.get(), hi, ho, st());
      hipLaunchKernelGGL((k_g_entry<half_t, SI>), dim3(grid()), dim3(kBlock), 0, st(), (int64_t)n,
                         (const half_t *)ho, out, (double)s_in / range_s);
    }
  }
  // the caller's operator on copies of the level's vectors in the operator's
  // dtype: fp16 operands carry the range factor range_s (as M's operands do
  // in mdirect), fp32 / fp64 operands none
  template <class SI> void op_apply(const SI *in, SI *out, float s_in) {
    ++cnt;
    ++calls;
    const int64_t nn = n;
    const dim3 g(grid()), b(kBlock);
    const double s_op = op_dtype == F3R_F16 ? (double)range_s : 1.0;
    auto to_op = [&](auto tag) {
      using T = decltype(tag);
      hipLaunchKernelGGL((k_g_entry<SI, T>), g, b, 0, st(), nn, in, op_in->as<T>(), s_op / s_in);
    };
    auto from_op = [&](auto tag) {
      using T = decltype(tag);
      hipLaunchKernelGGL((k_g_entry<T, SI>), g, b, 0, st(), nn, (const T *)op_out->as<T>(), out, s_in / s_op);
    };
    if (op_dtype == F3R_F16) to_op(half_t()); else if (op_dtype == F3R_F32) to_op(float()); else to_op(double());
    Vec vi{ctx, op_dtype, n, op_in, 0}, vo{ctx, op_dtype, n, op_out, 0};
    f3r_vec hi = wrap_vec(vi), ho = wrap_vec(vo);
    const int rc = op_fn(op_user, hi, ho);
    release_vec(hi);
    release_vec(ho);
    F3R_CHECK(rc == F3R_OK, rc, "the caller's operator returned an error");
    if (op_dtype == F3R_F16) from_op(half_t()); else if (op_dtype == F3R_F32) from_op(float()); else from_op(double());
  }
  template <class SI> void innermost(const SI *in, SI *out, float s_in) {
    if (inner == 2) {
      op_apply<SI>(in, out, s_in);
    } else if (inner == 1) {
      // the fp16 Richardson of the F3R nest on an fp32 level's vectors
      if constexpr (std::is_same<SI, float>::value) level4<float>(in, out, nullptr);
      else throw Error{F3R_ERR_UNSUPPORTED, "the Richardson level needs an fp32-basis level above it"};
    } else {
      mdirect<SI>(in, out, s_in);
    }
  }
  // generic level l applied to `in` (storage SI, scaled by s_in), out likewise
  template <class SI> void glevel(int l, const SI *in, SI *out, float s_in) {
    if (l == (int)glev.size()) {
      innermost<SI>(in, out, s_in);
      return;
    }
    if (glev[l].store == F3R_F16)
      glevel_t<SI, half_t>(l, in, out, s_in);
    else
      glevel_t<SI, float>(l, in, out, s_in);
  }
  // FGMRES(m) with normalize = false (solver/gmres2.hpp:84-145): V_0 = in,
  // e_0 = 1; CGS with the dots, the update and the norm as separate passes
  template <class SI, class S> void glevel_t(int l, const SI *in, SI *out, float s_in) {
    GLevel &L = glev[l];
    const int64_t nn = n;
    const dim3 g(grid()), b(kBlock);
    const float sc = 1.f / (L.s * L.s);
    const int G = dots_grid;
    float *part = partial<float>();
    hipLaunchKernelGGL((k_g_entry<SI, S>), g, b, 0, st(), nn, in, L.Vc<S>(0), (double)L.s / s_in);
    for (int j = 0; j < L.m; ++j) {
      glevel<S>(l + 1, L.Vc<S>(j), L.Zc<S>(j), L.s);
      if (L.mat == F3R_F16)
        launch_sell<half_t, S, float>(sh16->K, argsv<half_t>(), A16->as<half_t>(), L.Zc<S>(j),
                                      EpiStore<S, float>{L.Vc<S>(j + 1), nullptr}, st());
      else
        launch_sell<float, S, float>(sh->K, args(), A32->as<float>(), L.Zc<S>(j),
                                     EpiStore<S, float>{L.Vc<S>(j + 1), nullptr}, st());
      const int ncols = j + 1;
      float *Hcol = L.H->as<float>() + (int64_t)j * (L.m + 1);
      hipLaunchKernelGGL(k_g_mdot<S>, dim3(G, (ncols + 15) / 16), b, 0, st(), nn, (const S *)L.Vc<S>(j + 1),
                         (const S *)L.Vc<S>(0), L.ld, ncols, part);
      hipLaunchKernelGGL(k_g_dots_finish, dim3(ncols), b, 0, st(), (const float *)part, G, Hcol, sc);
      hipLaunchKernelGGL(k_g_update<S>, dim3(G), b, 0, st(), nn, L.Vc<S>(j + 1), (const S *)L.Vc<S>(0), L.ld,
                         ncols, (const float *)Hcol, sc, part);
      hipLaunchKernelGGL(k_cgs_finish<float>, dim3(1), dim3(1024), 0, st(), (const float *)part, G, Hcol,
                         L.cs->as<float>(), L.sn->as<float>(), L.e->as<float>(), j, 1, L.inv->as<float>(),
                         (double *)nullptr);
      hipLaunchKernelGGL(k_g_vscale<S>, g, b, 0, st(), nn, L.Vc<S>(j + 1), (const float *)L.inv->as<float>());
    }
    hipLaunchKernelGGL((k_g_combine<S, SI>), g, b, 0, st(), nn, (const S *)L.Zc<S>(0), L.ld, L.m,
                       (const float *)L.H->as<float>(), L.m + 1, (const float *)L.e->as<float>(), out,
                       s_in / L.s);
  }

  void level2_any(const double *in, double *out) {
    if (!glev.empty()) {
      glevel<double>(0, in, out, 1.f);
      return;
    }
    if (p.variant == 2)
      level2<double>(in, out);
    else
      level2<float>(in, out);
  }

  // -------------------------------------------------------- outer level ---
  struct Res {
    bool solved;
    int iters;
    double nrm;
  };

  Res outer_cycle(double nrm_b, bool record) {
    const int m1 = p.m1;
    const int64_t nn = n;
    double *V = V1->as<double>(), *Z = Z1->as<double>(), *H = H1->as<double>();
    double *stat = status->as<double>();
    SellArgs a = args();
    EpiResidNorm<double> er{V, b_int->as<double>(), e1->as<double>(), inv1->as<double>(), stat,
                            partial<double>(), counter(3)};
    launch_sell<double, double, double>(sh->K, a, A64->as<double>(), x_int->as<double>(), er, st(), spmv_grid);
    hipLaunchKernelGGL(k_vscale<double>, dim3(grid()), dim3(kBlock), 0, st(), nn, V, (const double *)inv1->as<double>());
    double e = read_scalar(ctx.get(), stat);
    // |e| / ||b||; a zero right-hand side is met only by a zero residual (its
    // solution from x = 0 is x = 0, with no step)
    auto rel = [&](double v) { return nrm_b > 0 ? v / nrm_b : (v == 0 ? 0.0 : HUGE_VAL); };
    if (record) hist.push_back(rel(e));
    // the caller's cond(i, |e|) of Solver::solve (sparse/models.hpp:191-267),
    // called as FGMRES::kernel calls it (i = 0 first, then the step count)
    auto stop = [&](int i) { return cond_fn ? cond_fn(cond_user, i, e) != 0 : rel(e) < p.eps; };
    if (stop(0)) return {true, 0, e};
    int j = 0;
    bool flag = false;
    for (; j < m1; ++j) {
      level2_any(V + (int64_t)j * nn, Z + (int64_t)j * nn);
      launch_sell<double, double, double>(sh->K, a, A64->as<double>(), Z + (int64_t)j * nn,
                                          EpiStore<double, double>{V + (int64_t)(j + 1) * nn, nullptr}, st());
      const int ncols = j + 1;
      // one grid row of NV columns up to 8 columns (counter slot 8), then
      // rows of 16 (slots 8.. 8 + (m1 + 15) / 16 - 1)
      const bool v4 = nn % 4 == 0 && aligned4(V);
      auto mdot = ncols <= 4   ? (v4 ? k_mdot_strided<double, true, 4> : k_mdot_strided<double, false, 4>)
                  : ncols <= 8 ? (v4 ? k_mdot_strided<double, true, 8> : k_mdot_strided<double, false, 8>)
                               : (v4 ? k_mdot_strided<double, true, 16> : k_mdot_strided<double, false, 16>);
      dim3 gm(red_grid, ncols <= 8 ? 1 : (ncols + 15) / 16);
      hipLaunchKernelGGL(mdot, gm, dim3(kBlock), 0, st(), nn, (const double *)(V + (int64_t)(j + 1) * nn),
                         (const double *)V, ncols, H + (int64_t)j * (m1 + 1), partial<double>(), counter(8));
      hipLaunchKernelGGL(k_cgs_update_strided<double>, dim3(red_grid), dim3(kBlock), 0, st(), nn,
                         V + (int64_t)(j + 1) * nn, (const double *)V, ncols, H + (int64_t)j * (m1 + 1),
                         c1->as<double>(), s1->as<double>(), e1->as<double>(), j, inv1->as<double>(), stat,
                         partial<double>(), counter(4));
      hipLaunchKernelGGL(k_vscale<double>, dim3(grid()), dim3(kBlock), 0, st(), nn, V + (int64_t)(j + 1) * nn,
                         (const double *)inv1->as<double>());
      e = read_scalar(ctx.get(), stat);
      if (record) hist.push_back(rel(e));
      if (stop(j + 1)) {
        ++j;
        flag = true;
        break;
      }
    }
    if (j > 0) {
      hipLaunchKernelGGL(k_hess_solve<double>, dim3(1), dim3(64), 0, st(), (const double *)H, m1 + 1,
                         (const double *)e1->as<double>(), y1->as<double>(), j);
      hipLaunchKernelGGL(k_solution_strided<double>, dim3(grid()), dim3(kBlock), 0, st(), nn, (const double *)Z,
                         j, (const double *)y1->as<double>(), x_int->as<double>());
    }
    F3R_HIP(hipGetLastError());
    return flag ? Res{true, j, e} : Res{false, m1, e};
  }
};

// ----------------------------------------------------------- set-up ------
static int ilog2_floor(double v) { return v < 1 ? 0 : (int)std::floor(std::log2(v)); }

// a composed nest (f3r_nest_create): generic FGMRES levels from the
// outermost inner one down, and the innermost operator
struct NestSpec {
  std::vector<f3r_level> levels;
  int inner = 0;  // Solver::inner
  f3r_apply_fn fn = nullptr;
  void *user = nullptr;
  int op_dtype = F3R_F32;
};

Solver *solver_create(std::shared_ptr<Context> ctx, const HostCSR &A, const Prec *Pin, const f3r_params &p,
                      const Mat *dev = nullptr, const NestSpec *nest = nullptr) {
  F3R_CHECK(Pin || (nest && nest->inner == 2), F3R_ERR_ARG, "solver: a preconditioner is needed");
  F3R_CHECK(p.variant >= 0 && p.variant <= 2, F3R_ERR_ARG, "solver: variant must be 0 (HALF), 1 (FLOAT) or 2 (DOUBLE)");
  F3R_CHECK(p.m1 >= 1 && p.m1 <= 120, F3R_ERR_ARG, "solver: m1 must be in [1, 120]");
  F3R_CHECK(p.m2 >= 1 && p.m2 <= 16 && p.m3 >= 1 && p.m3 <= 16, F3R_ERR_ARG, "solver: m2, m3 in [1, 16]");
  F3R_CHECK(p.m4 >= 1 && p.m4 <= 16 && p.c >= 1 && p.restart >= 1, F3R_ERR_ARG, "solver: m4 in [1, 16], c, restart >= 1");
  F3R_CHECK(A.n == A.ncols && A.n > 0, F3R_ERR_ARG, "solver: square non-empty matrix");
  F3R_CHECK(p.depth >= 0 && p.depth <= 5, F3R_ERR_ARG, "solver: depth must be 0 (F3R) or 1..5 (F2 F2H F3 F3H F4)");
  F3R_CHECK(p.depth == 0 || p.variant == 0, F3R_ERR_ARG, "solver: the depth ablations run in the HALF build (variant 0)");
  F3R_CHECK(p.static_weight >= 0, F3R_ERR_ARG, "solver: static_weight must be >= 0 (0: adaptive weights)");
  auto S = std::make_unique<Solver>();
  S->ctx = ctx;
  S->p = p;
  S->n = A.n;
  S->nnz = A.nnz();
  const int C = p.C > 0 ? p.C : 64;
  const int sigma = p.sigma > 0 ? p.sigma : 1;
  // K = 0: four slots per lane for the fp64 / fp32 matrices (16- or 32-byte
  // loads); for the fp16 matrix eight when rows are long (27-point: 4 slot
  // groups per row instead of 7, 27 vs 30 us at 128^3), else four.  The two
  // layouts share the row order; their slot padding differs.
  const int K = p.K > 0 ? p.K : 4;
  const int K16 = p.K > 0 ? p.K : (A.n > 0 && A.nnz() >= (int64_t)16 * A.n ? 8 : 4);
  F3R_CHECK(C >= 16 && (C & (C - 1)) == 0, F3R_ERR_ARG, "solver: C must be a power of two >= 16");
  F3R_CHECK(K == 1 || K == 2 || K == 4 || K == 8, F3R_ERR_ARG, "solver: K must be 1, 2, 4 or 8");
  F3R_CHECK(!(p.flags & 32) || (p.index == 0 && K >= 2 && K16 >= 2), F3R_ERR_ARG,
            "solver: K threads per row (flags bit 5) need the int32-index layout (index 0) and K >= 2");
  // flags bit 2: build the layouts on the device from the CSR's device copy
  // (the same arrays as the host build)
  const bool on_dev = (p.flags & 4) && dev && dev->d_rptr && dev->ctx == ctx;
  HostSell hs;
  bool built = false;
  if (on_dev) {
    std::shared_ptr<SellShared> sh, sh16;
    DevPtr v64, v64_16;
    built = build_sell_device(ctx.get(), A, dev->d_rptr->as<int32_t>(), dev->d_idx->as<int32_t>(),
                              dev->d_val->as<double>(), C, sigma, K, p.index != 0, true, sh, v64);
    if (built && p.variant == 0 && K16 != K)
      built = build_sell_device(ctx.get(), A, dev->d_rptr->as<int32_t>(), dev->d_idx->as<int32_t>(),
                                dev->d_val->as<double>(), C, sigma, K16, p.index != 0, true, sh16, v64_16);
    if (built) {
      S->sh = sh;
      S->sh->nnz = A.nnz();
      S->sh16 = S->sh;
      S->A64 = cast_values_device(ctx.get(), v64, sh->nslots, F3R_F64);
      if (p.variant != 2) S->A32 = cast_values_device(ctx.get(), v64, sh->nslots, F3R_F32);
      if (p.variant == 0) {
        if (K16 == K) {
          S->A16 = cast_values_device(ctx.get(), v64, sh->nslots, F3R_F16);
        } else {
          S->sh16 = sh16;
          S->sh16->nnz = A.nnz();
          S->A16 = cast_values_device(ctx.get(), v64_16, sh16->nslots, F3R_F16);
        }
      }
      F3R_HIP(hipStreamSynchronize(ctx->stream));
      hs.perm = sh->h_perm;
      S->layouts_on_device = true;
    }
  }
  if (!built) {
    hs = build_sell(A, C, sigma, K, nullptr, true, p.index != 0);
    S->sh = upload_sell_shared(ctx.get(), hs, true);
    S->sh->nnz = A.nnz();
    S->sh->sigma = sigma;
    S->sh16 = S->sh;
    // the level matrices: fp64 always (outer level; Richardson of FLOAT and
    // DOUBLE), fp32 for levels 2-3 of HALF / FLOAT, fp16 for HALF
    S->A64 = upload_sell_values(ctx.get(), hs, F3R_F64);
    if (p.variant != 2) S->A32 = upload_sell_values(ctx.get(), hs, F3R_F32);
    if (p.variant == 0) {
      if (K16 == K) {
        S->A16 = upload_sell_values(ctx.get(), hs, F3R_F16);
      } else {
        HostSell h16 = build_sell(A, C, sigma, K16, nullptr, true, p.index != 0);
        S->sh16 = upload_sell_shared(ctx.get(), h16, true);
        S->sh16->nnz = A.nnz();
        S->sh16->sigma = sigma;
        S->A16 = upload_sell_values(ctx.get(), h16, F3R_F16);
      }
    }
  }
  // flags bit 5: every SpMV with K threads per row (the sums reassociated)
  if (p.flags & 32) {
    S->sh->lanes = S->sh->K;
    S->sh16->lanes = S->sh16->K;
  }
  bool ident = true;
  for (int32_t r = 0; r < A.n && ident; ++r) ident = hs.perm[r] == r;
  F3R_CHECK(!nest || nest->inner != 2 || ident, F3R_ERR_UNSUPPORTED,
            "nest: a caller's operator needs the matrix's own row order (sigma 1)");
  // innermost preconditioner in the internal order and the innermost precision
  if (Pin) {
    const int pdt = p.variant == 0 ? F3R_F16 : p.variant == 1 ? F3R_F32 : F3R_F64;
    auto P = std::make_shared<Prec>();
    P->ctx = ctx;
    P->kind = Pin->kind;
    P->dtype = pdt;
    P->tol = Pin->tol;
    P->acc = Pin->acc;
    P->host = Pin->host;
    F3R_CHECK(P->host.f[0].n == A.n, F3R_ERR_ARG, "solver: preconditioner size differs from A");
    prec_build_device(P.get(), ident ? nullptr : &hs.perm, C, pdt == F3R_F16 ? K16 : K, p.index != 0);
    S->P = P;
  }
  S->range_s = p.variant != 0 ? 1.f
               : p.range_scale > 0 ? (float)p.range_scale
                                   : std::ldexp(1.f, ilog2_floor(std::sqrt((double)A.n) / 4.0));
  const int64_t n = A.n;
  S->red_grid = reduce_grid(n, ctx->num_cu);
  // fused SpMV + reduction kernels: eight blocks per CU (each arrival at a
  // grid reduction costs an atomic round trip, so fewer, longer blocks)
  S->dots_grid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, std::min(kMaxReduceGrid, 8 * ctx->num_cu));
  if (p.grid > 0) S->dots_grid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, std::min((int)p.grid, kMaxReduceGrid));
  // the SpMV kernels with a fused reduction: one row per thread (the gathers
  // need the waves; the arrivals of 8x more blocks cost less than that), up
  // to the reduction's block limit; flags bit 1 caps them like the others
  S->spmv_grid = (p.flags & 2) ? S->dots_grid
                               : (int)std::min<int64_t>((n + kBlock - 1) / kBlock, kMaxReduceGrid);
  auto dd = [](int64_t cnt) { return dev_alloc(sizeof(double) * std::max<int64_t>(cnt, 1)); };
  S->V1 = dd(n * (p.m1 + 1));
  S->Z1 = dd(n * p.m1);
  S->H1 = dd((int64_t)(p.m1 + 1) * p.m1);
  S->c1 = dd(p.m1);
  S->s1 = dd(p.m1);
  S->e1 = dd(p.m1 + 1);
  S->y1 = dd(p.m1);
  S->inv1 = dd(1);
  S->status = dd(4);
  S->b_int = dd(n);
  S->x_int = dd(n);
  auto mk = [&](auto &L, int m) {
    using T = typename std::remove_reference<decltype(*L.V.data())>::type;
    using E = typename std::remove_pointer<T>::type;
    L.m = m;
    // columns 64-element aligned, so the streaming kernels can use 16-byte accesses
    const int64_t ld = (n + 63) / 64 * 64;
    L.Vbuf = dev_alloc(sizeof(E) * ld * (m + 1));
    L.Zbuf = dev_alloc(sizeof(E) * ld * m);
    L.H = dev_alloc(sizeof(E) * (m + 1) * m);
    L.cs = dev_alloc(sizeof(E) * m);
    L.sn = dev_alloc(sizeof(E) * m);
    L.e = dev_alloc(sizeof(E) * (m + 1));
    L.inv = dev_alloc(sizeof(E) * 4);
    L.V.resize(m + 1);
    L.Z.resize(m);
    for (int k = 0; k <= m; ++k) L.V[k] = L.Vbuf->template as<E>() + (int64_t)k * ld;
    for (int k = 0; k < m; ++k) L.Z[k] = L.Zbuf->template as<E>() + (int64_t)k * ld;
  };
  if (p.depth != 0 || nest) {
    // the nests of work/src/depth.cpp:122-167, or the caller's levels:
    // {m, basis storage, matrix values}
    struct Spec {
      int m, store, mat;
    };
    std::vector<Spec> spec;
    if (nest) {
      for (const f3r_level &q : nest->levels) spec.push_back({q.m, q.basis, q.matrix});
      S->inner = nest->inner;
      S->op_fn = nest->fn;
      S->op_user = nest->user;
      S->op_dtype = nest->op_dtype;
      if (nest->inner == 2) {
        S->op_in = dev_alloc(dtype_size(nest->op_dtype) * std::max<int64_t>(1, n));
        S->op_out = dev_alloc(dtype_size(nest->op_dtype) * std::max<int64_t>(1, n));
      }
    } else switch (p.depth) {
      case 1: spec = {{64, F3R_F32, F3R_F32}}; break;                                           // F2
      case 2: spec = {{64, F3R_F16, F3R_F16}}; break;                                           // F2H
      case 3: spec = {{8, F3R_F32, F3R_F32}, {8, F3R_F32, F3R_F16}}; break;                     // F3
      case 4: spec = {{8, F3R_F32, F3R_F32}, {8, F3R_F16, F3R_F16}}; break;                     // F3H
      default: spec = {{8, F3R_F32, F3R_F32}, {4, F3R_F32, F3R_F16}, {2, F3R_F16, F3R_F16}};  // F4
    }
    const int64_t ld = (n + 63) / 64 * 64;
    for (const Spec &q : spec) {
      Solver::GLevel L;
      L.m = q.m;
      L.store = q.store;
      L.mat = q.mat;
      L.s = q.store == F3R_F16 ? S->range_s : 1.f;
      L.ld = ld;
      const size_t es = q.store == F3R_F16 ? sizeof(half_t) : sizeof(float);
      L.V = dev_alloc(es * ld * (q.m + 1));
      L.Z = dev_alloc(es * ld * q.m);
      L.H = dev_alloc(sizeof(float) * (q.m + 1) * q.m);
      L.cs = dev_alloc(sizeof(float) * q.m);
      L.sn = dev_alloc(sizeof(float) * q.m);
      L.e = dev_alloc(sizeof(float) * (q.m + 1));
      L.inv = dev_alloc(sizeof(float) * 4);
      S->glev.push_back(L);
    }
  } else if (p.variant == 2) {
    mk(S->L2d, p.m2);
    mk(S->L3d, p.m3);
  } else {
    mk(S->L2f, p.m2);
    mk(S->L3f, p.m3);
    int e2 = 0;
    if (p.variant == 0 && std::frexp(S->range_s, &e2) == 0.5f) {  // exact only for a power-of-two scale
      const int64_t ld = (n + 63) / 64 * 64;
      S->L3f.Zhbuf = dev_alloc(sizeof(half_t) * ld * p.m3);
      S->L3f.Zh.resize(p.m3);
      for (int k = 0; k < p.m3; ++k) S->L3f.Zh[k] = S->L3f.Zhbuf->as<half_t>() + (int64_t)k * ld;
    }
  }
  const size_t wsz = p.variant == 0 ? sizeof(half_t) : p.variant == 1 ? sizeof(float) : sizeof(double);
  auto dh = [&](void) { return dev_alloc(wsz * n); };
  S->hin = dh();
  S->hout = dh();
  S->hout2 = dh();
  S->hr = dh();
  S->tmp2 = dh();
  S->tt = dh();
  S->rich = dev_alloc(sizeof(RichDev));
  hipLaunchKernelGGL(k_rich_reset, dim3(1), dim3(64), 0, ctx->stream, S->rich->as<RichDev>(), p.static_weight);
  F3R_HIP(hipStreamSynchronize(ctx->stream));
  return S.release();
}

void solver_solve(Solver *S, Vec *b, Vec *x, f3r_result *res) {
  F3R_CHECK(b->dtype == F3R_F64 && x->dtype == F3R_F64 && b->n == S->n && x->n == S->n, F3R_ERR_ARG,
            "solve: b and x must be fp64 vectors of the matrix size");
  Context *c = S->ctx.get();
  const int64_t n = S->n;
  const dim3 g(S->grid()), bl(kBlock);
  const bool permuted = (bool)S->sh->perm;
  // nrm(b) before the timer, as the reference driver does (src/f3r-gpu.cpp:82)
  const double nrm_b = std::sqrt(dot_raw(c, F3R_F64, b->ptr(), F3R_F64, b->ptr(), n));
  if (permuted) {
    hipLaunchKernelGGL(k_gather, g, bl, 0, c->stream, n, (const double *)b->ptr(), S->sh->perm->as<int32_t>(),
                       S->b_int->as<double>());
    hipLaunchKernelGGL(k_gather, g, bl, 0, c->stream, n, (const double *)x->ptr(), S->sh->perm->as<int32_t>(),
                       S->x_int->as<double>());
  } else {
    F3R_HIP(hipMemcpyAsync(S->b_int->p, b->ptr(), 8 * n, hipMemcpyDeviceToDevice, c->stream));
    F3R_HIP(hipMemcpyAsync(S->x_int->p, x->ptr(), 8 * n, hipMemcpyDeviceToDevice, c->stream));
  }
  // reset the adaptive weights (src/f3r-gpu.cpp:277-281)
  S->cnt = S->p.c;
  S->calls = 0;
  hipLaunchKernelGGL(k_rich_reset, dim3(1), dim3(64), 0, c->stream, S->rich->as<RichDev>(), S->p.static_weight);
  F3R_HIP(hipStreamSynchronize(c->stream));
  S->hist.clear();
  const bool record = (S->p.flags & 1) != 0;
  auto t0 = std::chrono::steady_clock::now();
  Solver::Res tot{false, 0, 0.0};
  int cycles = 0;
  for (int cyc = 0; cyc < S->p.restart; ++cyc) {
    Solver::Res r = S->outer_cycle(nrm_b, record);
    ++cycles;
    tot.solved = tot.solved || r.solved;
    tot.iters += r.iters;
    tot.nrm = r.nrm;
    if (r.solved) break;
  }
  F3R_HIP(hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  if (permuted)
    hipLaunchKernelGGL(k_scatter, g, bl, 0, c->stream, n, (const double *)S->x_int->as<double>(),
                       S->sh->perm->as<int32_t>(), (double *)x->ptr());
  else
    F3R_HIP(hipMemcpyAsync(x->ptr(), S->x_int->p, 8 * n, hipMemcpyDeviceToDevice, c->stream));
  F3R_HIP(hipStreamSynchronize(c->stream));
  res->is_solved = tot.solved ? 1 : 0;
  res->res_iter = tot.iters;
  res->res_nrm2 = tot.nrm;
  res->nrm_b = nrm_b;
  res->solve_s = std::chrono::duration<double>(t1 - t0).count();
  res->precond_calls = S->calls;
  res->outer_cycles = cycles;
}

}  // namespace f3r

// ================================================================ C ABI ===
using namespace f3r;

struct f3r_solver_s {
  Solver *s;
};
struct f3r_op_s {
  std::shared_ptr<Context> ctx;
  int32_t n;
  int dtype;
  f3r_apply_fn fn;
  void *user;
};


extern "C" {

int f3r_params_default(f3r_params *p) {
  if (!p) return F3R_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->m1 = 100;
  p->m2 = 8;
  p->m3 = 4;
  p->m4 = 2;
  p->c = 64;
  p->restart = 20;
  p->eps = 1e-8;
  p->variant = 0;
  p->range_scale = 0;
  p->C = 0;
  p->sigma = 0;
  p->K = 0;
  p->use_graph = 1;
  p->flags = 1;
  p->grid = 0;
  p->index = 1;
  return F3R_OK;
}

int f3r_solver_create(f3r_ctx ctx, f3r_mat A, f3r_prec P, const f3r_params *params, f3r_solver *S) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && A && P && params && S, F3R_ERR_ARG, "solver_create: null argument");
  auto m = mat_of(A);
  F3R_CHECK(m->host, F3R_ERR_ARG, "solver_create: A must carry its CSR");
  *S = nullptr;
  Solver *s = solver_create(ctx_of(ctx), *m->host, prec_of(P).get(), *params, m.get());
  *S = new f3r_solver_s{s};
  F3R_API_END
}

int f3r_solve(f3r_solver S, f3r_vec b, f3r_vec x, f3r_result *res) {
  F3R_API_BEGIN
  F3R_CHECK(S && b && x && res, F3R_ERR_ARG, "solve: null argument");
  std::memset(res, 0, sizeof(*res));
  solver_solve(S->s, vec_of(b), vec_of(x), res);
  F3R_API_END
}

int f3r_solve_cond(f3r_solver S, f3r_vec b, f3r_vec x, f3r_cond_fn cond, void *user, f3r_result *res) {
  F3R_API_BEGIN
  F3R_CHECK(S && b && x && res && cond, F3R_ERR_ARG, "solve_cond: null argument");
  std::memset(res, 0, sizeof(*res));
  S->s->cond_fn = cond;
  S->s->cond_user = user;
  try {
    solver_solve(S->s, vec_of(b), vec_of(x), res);
  } catch (...) {
    S->s->cond_fn = nullptr;
    S->s->cond_user = nullptr;
    throw;
  }
  S->s->cond_fn = nullptr;
  S->s->cond_user = nullptr;
  F3R_API_END
}

int f3r_op_create(f3r_ctx ctx, int32_t n, int dtype, f3r_apply_fn fn, void *user, f3r_op *op) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && fn && op && n > 0, F3R_ERR_ARG, "op_create: bad argument");
  F3R_CHECK(dtype == F3R_F64 || dtype == F3R_F32 || dtype == F3R_F16, F3R_ERR_ARG, "op_create: fp64, fp32 or fp16 vectors");
  *op = new f3r_op_s{ctx_of(ctx), n, dtype, fn, user};
  F3R_API_END
}

int f3r_op_destroy(f3r_op op) {
  delete op;
  return F3R_OK;
}

int f3r_nest_create(f3r_ctx ctx, f3r_mat A, const f3r_level *levels, int32_t nlevels, const f3r_inner *inner,
                    const f3r_params *params, f3r_solver *S) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && A && levels && inner && params && S, F3R_ERR_ARG, "nest_create: null argument");
  *S = nullptr;
  auto m = mat_of(A);
  F3R_CHECK(m->host, F3R_ERR_ARG, "nest_create: A must carry its CSR");
  F3R_CHECK(nlevels >= 1 && nlevels <= 4, F3R_ERR_ARG, "nest_create: 1 to 4 FGMRES levels");
  for (int32_t l = 0; l < nlevels; ++l) {
    const f3r_level &q = levels[l];
    F3R_CHECK(q.m >= 1 && q.m <= 64, F3R_ERR_ARG, "nest_create: a level's m in [1, 64]");
    F3R_CHECK((q.basis == F3R_F32 || q.basis == F3R_F16) && (q.matrix == F3R_F32 || q.matrix == F3R_F16),
              F3R_ERR_ARG, "nest_create: a level's basis and matrix are F3R_F32 or F3R_F16");
  }
  F3R_CHECK(inner->kind >= F3R_INNER_PREC && inner->kind <= F3R_INNER_OP, F3R_ERR_ARG, "nest_create: inner kind");
  f3r_params p = *params;
  p.variant = 0;
  p.depth = 0;
  p.static_weight = 0;
  const Prec *P = nullptr;
  if (inner->kind != F3R_INNER_OP) {
    F3R_CHECK(inner->P, F3R_ERR_ARG, "nest_create: the innermost level needs a preconditioner");
    P = prec_of(inner->P).get();
  }
  if (inner->kind == F3R_INNER_RICHARDSON) {
    F3R_CHECK(inner->m4 >= 1 && inner->m4 <= 16 && inner->c >= 1 && inner->static_weight >= 0, F3R_ERR_ARG,
              "nest_create: Richardson m4 in [1, 16], c >= 1, static_weight >= 0");
    F3R_CHECK(levels[nlevels - 1].basis == F3R_F32, F3R_ERR_ARG,
              "nest_create: the level above the Richardson needs an fp32 basis");
    p.m4 = inner->m4;
    p.c = inner->c;
    p.static_weight = inner->static_weight;
    // the F3R nest itself: f3r_solver_create's fused solver
    if (nlevels == 2 && levels[0].basis == F3R_F32 && levels[0].matrix == F3R_F32 && levels[1].basis == F3R_F32 &&
        levels[1].matrix == F3R_F16 && levels[0].m <= 16 && levels[1].m <= 16) {
      p.m2 = levels[0].m;
      p.m3 = levels[1].m;
      *S = new f3r_solver_s{solver_create(ctx_of(ctx), *m->host, P, p, m.get())};
      return F3R_OK;
    }
  }
  NestSpec nest;
  nest.levels.assign(levels, levels + nlevels);
  nest.inner = inner->kind;
  if (inner->kind == F3R_INNER_OP) {
    F3R_CHECK(inner->op && inner->op->n == m->host->n, F3R_ERR_ARG, "nest_create: an operator of the matrix's size");
    F3R_CHECK(inner->op->ctx.get() == ctx_of(ctx).get(), F3R_ERR_ARG, "nest_create: the operator's context");
    nest.fn = inner->op->fn;
    nest.user = inner->op->user;
    nest.op_dtype = inner->op->dtype;
  }
  p.m2 = p.m3 = 1;  // not used by the generic levels
  *S = new f3r_solver_s{solver_create(ctx_of(ctx), *m->host, P, p, m.get(), &nest)};
  F3R_API_END
}

int f3r_solver_setup_where(f3r_solver S, int *where) {
  F3R_API_BEGIN
  F3R_CHECK(S && where, F3R_ERR_ARG, "setup_where: null argument");
  *where = S->s->layouts_on_device ? 1 : 0;
  F3R_API_END
}

int f3r_solver_history(f3r_solver S, double *hist, int32_t cap, int32_t *len) {
  F3R_API_BEGIN
  F3R_CHECK(S && len, F3R_ERR_ARG, "history: null argument");
  const auto &h = S->s->hist;
  *len = (int32_t)h.size();
  if (hist)
    for (int32_t i = 0; i < cap && i < (int32_t)h.size(); ++i) hist[i] = h[i];
  F3R_API_END
}

int f3r_solver_capture(f3r_solver S, int32_t max_vectors, int32_t every) {
  F3R_API_BEGIN
  F3R_CHECK(S && max_vectors >= 0 && every >= 1, F3R_ERR_ARG, "capture: bad argument");
  Solver *s = S->s;
  F3R_HIP(hipStreamSynchronize(s->st()));
  s->cap.reset();
  s->cap_n = 0;
  s->cap_max = 0;
  s->cap_every = every;
  if (max_vectors > 0) {
    s->cap = dev_alloc(sizeof(float) * s->n * (int64_t)max_vectors);
    s->cap_max = max_vectors;
  }
  F3R_API_END
}

int f3r_solver_captured(f3r_solver S, int32_t *count, float *out) {
  F3R_API_BEGIN
  F3R_CHECK(S && count, F3R_ERR_ARG, "captured: null argument");
  Solver *s = S->s;
  F3R_HIP(hipStreamSynchronize(s->st()));
  *count = s->cap_n;
  if (out && s->cap_n > 0) {
    const int64_t n = s->n;
    std::vector<float> h((size_t)n * s->cap_n);
    F3R_HIP(hipMemcpy(h.data(), s->cap->p, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    const std::vector<int32_t> &perm = s->sh->h_perm;
    for (int32_t k = 0; k < s->cap_n; ++k) {
      const float *src = h.data() + (int64_t)k * n;
      float *dst = out + (int64_t)k * n;
      if (s->sh->perm)
        for (int64_t i = 0; i < n; ++i) dst[perm[i]] = src[i];
      else
        std::memcpy(dst, src, sizeof(float) * n);
    }
  }
  F3R_API_END
}

int f3r_solver_level_apply(f3r_solver S, int32_t level, f3r_vec in, f3r_vec out) {
  F3R_API_BEGIN
  F3R_CHECK(S && in && out && level >= 2 && level <= 4, F3R_ERR_ARG, "level_apply: bad argument");
  Solver *s = S->s;
  Vec *vi = vec_of(in), *vo = vec_of(out);
  F3R_CHECK(vi->dtype == F3R_F64 && vo->dtype == F3R_F64 && vi->n == s->n && vo->n == s->n, F3R_ERR_ARG,
            "level_apply: fp64 vectors of the matrix size");
  F3R_CHECK(!s->sh->perm, F3R_ERR_UNSUPPORTED, "level_apply: sorted layouts");
  F3R_CHECK(s->glev.empty() || level == 2, F3R_ERR_UNSUPPORTED,
            "level_apply: a depth ablation applies its whole nest as level 2 only");
  const int64_t n = s->n;
  const dim3 g(s->grid()), b(kBlock);
  s->cnt = s->p.c;
  hipLaunchKernelGGL(k_rich_reset, dim3(1), dim3(64), 0, s->st(), s->rich->as<RichDev>(), s->p.static_weight);
  if (level == 2) {
    const int32_t keep = s->p.use_graph;
    s->p.use_graph = 0;
    s->level2_any((const double *)vi->ptr(), (double *)vo->ptr());
    s->p.use_graph = keep;
  } else {
    auto go = [&](auto tag) {
      using T = decltype(tag);
      InnerLevel<T> &L = s->template L2<T>();
      T *a = L.V[0], *r = L.Z[0];
      hipLaunchKernelGGL((k_cast2<double, T>), g, b, 0, s->st(), n, (const double *)vi->ptr(), a);
      if (level == 3)
        s->template level3<T>(a, r);
      else
        s->template level4<T>(a, r);
      hipLaunchKernelGGL((k_cast2<T, double>), g, b, 0, s->st(), n, (const T *)r, (double *)vo->ptr());
    };
    if (s->p.variant == 2) go(double()); else go(float());
  }
  F3R_HIP(hipGetLastError());
  F3R_HIP(hipStreamSynchronize(s->st()));
  F3R_API_END
}

int f3r_solver_time_kernel(f3r_solver S, int32_t which, int32_t reps, double *ms, double *bytes) {
  F3R_API_BEGIN
  const bool cold = (which & F3R_TIME_COLD) != 0;
  which &= ~F3R_TIME_COLD;
  F3R_CHECK(S && ms && reps > 0 && which >= 0 && which <= 9, F3R_ERR_ARG, "time_kernel: bad argument");
  F3R_CHECK(which < 4 || (S->s->p.variant == 0 && S->s->L3f.m <= 4), F3R_ERR_UNSUPPORTED,
            "time_kernel: the Arnoldi-step kernels need the HALF variant with m3 <= 4");
  Solver *s = S->s;
  F3R_CHECK(which < 4 || s->p.depth == 0, F3R_ERR_UNSUPPORTED, "time_kernel: the Arnoldi-step kernels of the F3R nest");
  SellArgs a = s->args(), a16 = s->argsv<half_t>();
  F3R_CHECK(s->p.variant == 0 || which == 3, F3R_ERR_UNSUPPORTED, "time_kernel: fp16/fp32 kernels need the HALF variant");
  half_t *x16 = s->hout->as<half_t>(), *y16 = s->tt->as<half_t>();
  float *x32 = s->L2f.V.empty() ? nullptr : s->L2f.V[0], *y32 = s->L2f.V.empty() ? nullptr : s->L2f.V[1];
  double *x64 = s->x_int->as<double>(), *y64 = s->V1->as<double>();
  double by = 0;
  // launches per launch() call: the Arnoldi-step mixes run steps 0 .. m3 - 1
  const int per = (which == 6 || which == 7) ? s->L3f.m : 1;
  const bool diag = s->P->kind == 0 && s->P->diag;
  F3R_CHECK(which < 8 || diag, F3R_ERR_UNSUPPORTED, "time_kernel: the fused Richardson kernels need diagonal SDAINV");
  DevPtr one;  // 1/||w|| = 1 for the entry kernel's in-place normalisation
  if (which == 9) {
    one = dev_alloc(sizeof(float));
    const float f1 = 1.f;
    F3R_HIP(hipMemcpyAsync(one->p, &f1, sizeof(float), hipMemcpyHostToDevice, s->st()));
  }
  auto launch = [&]() {
    switch (which) {
      case 0:  // fp16 values, fp16 x and y (the metric's fp16 SpMV)
        launch_sell<half_t, half_t, float>(s->sh16->K, a16, s->A16->as<half_t>(), x16,
                                           EpiStore<half_t, float>{y16, nullptr}, s->st());
        by = s->sh16->algo_bytes(2, 2, 2);
        break;
      case 1:  // fp16 values, fp32 x and y (level 3)
        launch_sell<half_t, float, float>(s->sh16->K, a16, s->A16->as<half_t>(), x32,
                                          EpiStore<float, float>{y32, nullptr}, s->st());
        by = s->sh16->algo_bytes(2, 4, 4);
        break;
      case 2:  // fp32 (level 2)
        launch_sell<float, float, float>(s->sh->K, a, s->A32->as<float>(), x32,
                                         EpiStore<float, float>{y32, nullptr}, s->st());
        by = s->sh->algo_bytes(4, 4, 4);
        break;
      case 3:  // fp64 (outer level)
        launch_sell<double, double, double>(s->sh->K, a, s->A64->as<double>(), x64,
                                            EpiStore<double, double>{y64, nullptr}, s->st());
        by = s->sh->algo_bytes(8, 8, 8);
        break;
      case 4:  // level-3 Arnoldi step j = m3 - 1: fp16 SpMV of Z_j + CGS dots with V_0..V_j, and their finish
        s->template spmv_dots<float, half_t, 4>(s->L3f, s->A16->as<half_t>(), s->L3f.m - 1);
        by = s->sh16->algo_bytes(2, s->L3f.Zh.empty() ? 4 : 2, 4) + 4.0 * s->n * s->L3f.m;
        break;
      case 5:  // level-3 CGS update + norm + Givens of that step (the last: w is not stored)
        s->template cgs_update<float, 4>(s->L3f, s->L3f.m - 1);
        by = 4.0 * s->n * (s->L3f.m + 1);
        break;
      case 6: {  // the solve's mix of level-3 SpMV + CGS dots launches, steps j = 0 .. m3 - 1 (no finish)
        const int m = s->L3f.m;
        by = 0;
        for (int j = 0; j < m; ++j) {
          Cols<float, 4> c{};
          for (int k = 0; k <= j; ++k) c.p[k] = s->L3f.V[k];
          EpiDots<float, 4> epi{s->L3f.V[j + 1], c, j + 1, s->L3f.Hp() + (int64_t)j * s->L3f.ldh(),
                                s->template partial<float>(), s->counter(0)};
          if (!s->L3f.Zh.empty()) {
            epi.xs = 1.f / s->range_s;
            launch_sell<half_t, half_t, float>(s->sh16->K, a16, s->A16->as<half_t>(), s->L3f.Zh[j], epi, s->st(),
                                               s->spmv_grid);
          } else {
            launch_sell<half_t, float, float>(s->sh16->K, a16, s->A16->as<half_t>(), s->L3f.Z[j], epi, s->st(),
                                              s->spmv_grid);
          }
          // values, Z_j gathered, w written, V_0 .. V_j read
          by += s->sh16->algo_bytes(2, s->L3f.Zh.empty() ? 4 : 2, 4) + 4.0 * s->n * (j + 1);
        }
        by /= m;
        break;
      }
      case 7: {  // the solve's mix of CGS update + norm + Givens launches, steps j = 0 .. m3 - 1
        const int m = s->L3f.m;
        by = 0;
        for (int j = 0; j < m; ++j) {
          s->template cgs_update<float, 4>(s->L3f, j);
          // V_0 .. V_j read, w read, and written except in the last step
          by += 4.0 * s->n * (j + (j + 1 < m ? 3 : 2));
        }
        by /= m;
        break;
      }
      case 8: {  // the fused Richardson step (residual SpMV + diagonal SDAINV + weighted update), fp16 out
        RichDev *rs = s->rich->as<RichDev>();
        EpiL4Diag epi{s->hin->as<half_t>(), s->hout->as<half_t>(), s->P->dwz->as<uint32_t>(), &rs->ome[1],
                      nullptr, nullptr, s->hout2->as<half_t>(), nullptr, 1.f / s->range_s};
        launch_sell<half_t, half_t, float>(s->sh16->K, a16, s->A16->as<half_t>(), s->hout->as<half_t>(), epi,
                                           s->st());
        // values, h gathered, hin / h / (w, z) read per row, h' written
        by = s->sh16->algo_bytes(2, 2, 2) + (double)s->n * (2 + 2 + 4);
        break;
      }
      case 9: {  // the innermost entry: normalise the basis vector, fp16 input, first weighted M application
        const int64_t nn = s->n;
        float *in = s->L3f.V[s->L3f.m];
        hipLaunchKernelGGL(k_l4_entry<true>, dim3(s->grid()), dim3(kBlock), 0, s->st(), nn, in,
                           (const float *)one->as<float>(), s->range_s, s->hin->as<half_t>(),
                           s->hout2->as<half_t>(), (const uint32_t *)s->P->dwz->as<uint32_t>(),
                           (const RichDev *)s->rich->as<RichDev>(), 1);
        by = (double)s->n * (4 + 4 + 4 + 2 + 2);  // in read and written, (w, z) read, hin and hout written
        break;
      }
    }
  };
  hipEvent_t e0, e1;
  F3R_HIP(hipEventCreate(&e0));
  F3R_HIP(hipEventCreate(&e1));
  if (cold) {
    // 256 MiB read + 256 MiB written before every launch: twice the
    // last-level cache, so the launch finds none of its operands cached
    const int64_t nf = (int64_t)1 << 26;
    DevPtr fa = dev_alloc(sizeof(float) * nf), fb = dev_alloc(sizeof(float) * nf);
    double tot = 0;
    for (int i = 0; i < reps; ++i) {
      hipLaunchKernelGGL((k_cast2<float, float>), dim3(4096), dim3(kBlock), 0, s->st(), nf,
                         (const float *)fa->as<float>(), fb->as<float>());
      F3R_HIP(hipEventRecord(e0, s->st()));
      launch();
      F3R_HIP(hipEventRecord(e1, s->st()));
      F3R_HIP(hipEventSynchronize(e1));
      float t = 0;
      F3R_HIP(hipEventElapsedTime(&t, e0, e1));
      tot += t;
    }
    *ms = tot / ((double)reps * per);
  } else {
    launch();
    F3R_HIP(hipEventRecord(e0, s->st()));
    for (int i = 0; i < reps; ++i) launch();
    F3R_HIP(hipEventRecord(e1, s->st()));
    F3R_HIP(hipEventSynchronize(e1));
    float t = 0;
    F3R_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = (double)t / ((double)reps * per);
  }
  if (bytes) *bytes = by;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  F3R_API_END
}

int f3r_solver_destroy(f3r_solver S) {
  F3R_API_BEGIN
  if (S) {
    (void)hipStreamSynchronize(S->s->st());
    delete S->s;
    delete S;
  }
  F3R_API_END
}

}  // extern "C"
