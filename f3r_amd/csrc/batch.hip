// Batched F3R: many independent systems (and right-hand sides) solved in
// lockstep, every launch covering all of them.
//
// Each segment (one system with one right-hand side) occupies a contiguous
// row range of one block-diagonal matrix, padded to a whole number of
// 2048-row tiles, so a workgroup never straddles two segments.  The SpMVs are
// the single solver's SELL kernels on the concatenated matrices (fp64, fp32,
// fp16 values; the SDAINV factors likewise).  Every reduction is per
// segment: each block writes its partial sums, and one block per segment
// adds its blocks' partials in a fixed order and does that segment's
// Hessenberg / Givens / Richardson-weight work.  Each outer step of all
// segments is replayed from one HIP graph, after which the host reads one
// array of residual norms for all segments.  A segment that has
// converged stops counting steps and keeps its solution; its vectors may run
// on with the others, and the breakdown guards below (a zero norm, a zero
// rotation, a zero pivot) keep them finite.
//
// The nest is the HALF build's: fp64 FGMRES(m1) with restarts > fp32
// FGMRES(m2) on fp32 values > fp32-basis FGMRES(m3) on fp16 values >
// adaptive Richardson(m4, c) in fp16 with SDAINV (work/src/f3r-gpu.cpp:
// 210-266), each segment with its own probes' weights.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <map>
#include <numeric>

#include "common.hpp"
#include "device.hpp"
#include "internal.hpp"
#include "sell_kernel.hpp"

namespace f3r {
namespace {

// ---------------------------------------------------- device helpers ---
template <class T> __device__ void givens_guarded(T *H, T *cs, T *sn, T *e, int j, bool e0_one) {
  for (int i = 0; i < j; ++i) {
    T t = H[i], h1 = H[i + 1];
    H[i] = fma_<T>(cs[i], t, sn[i] * h1);
    H[i + 1] = fma_<T>(-sn[i], t, cs[i] * h1);
  }
  T t = (T)sqrt(fma_<T>(H[j], H[j], H[j + 1] * H[j + 1]));
  if (t > (T)0) {
    cs[j] = H[j] / t;
    sn[j] = H[j + 1] / t;
  } else {  // a zero column (a segment that has run dry): identity rotation
    cs[j] = 1;
    sn[j] = 0;
  }
  H[j] = t;
  H[j + 1] = 0;
  T ej = (j == 0 && e0_one) ? (T)1 : e[j];
  e[j + 1] = -sn[j] * ej;
  e[j] = cs[j] * ej;
}

template <class T> __device__ void hess_solve_guarded(const T *H, int ldh, const T *e, T *y, int j) {
  for (int i = j - 1; i >= 0; --i) {
    T t = e[i];
    for (int q = j - 1; q > i; --q) t = fma_<T>(-H[q * ldh + i], y[q], t);
    const T d = H[i * ldh + i];
    y[i] = d != (T)0 ? t / d : (T)0;
  }
}

// Per-segment Richardson state (work/src/f3r-gpu.cpp:224-260)
struct RichSeg {
  double sum[16];
  double ome[16];
  double alpha[16];
  int nprobe;
};

__global__ void k_b_rich_reset(int nseg, RichSeg *st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  for (int k = 0; k < 16; ++k) {
    st[s].sum[k] = 1.0;
    st[s].ome[k] = 1.0;
    st[s].alpha[k] = 0.0;
  }
  st[s].nprobe = 0;
}

// ----------------------------------------------------------- kernels ---
// Block b covers the tile of rows [b * kTile, (b + 1) * kTile), all in
// segment bseg[b]; thread t takes rows b * kTile + t + q * kBlock, q < kRows
// (coalesced).  Segments are padded to whole tiles, so the grid is exactly
// the row count / kTile.
constexpr int kRows = 8;
constexpr int64_t kTile = (int64_t)kBlock * kRows;

__device__ __forceinline__ int64_t tile_row(int q) { return (int64_t)blockIdx.x * kTile + threadIdx.x + (int64_t)q * kBlock; }
// the same tile as kRows / 4 groups of 4 consecutive rows per thread (one
// 8- or 16-byte access per vector and group; segments and vectors start on
// tile boundaries, so every group is aligned)
constexpr int kQuads = kRows / 4;
__device__ __forceinline__ int64_t tile_quad(int q) {
  return (int64_t)blockIdx.x * kTile + ((int64_t)threadIdx.x + (int64_t)q * kBlock) * 4;
}

// block partial sums of <w, V_k> for the NV columns of blockIdx.y (k < ncols):
// partial[k * gridDim.x + block]
// Basis columns: column 0 at V0, column k >= 1 at V + k * ld, its values
// scaled by csc[k * nsegs + seg] when csc is given (a column left
// unnormalised in memory; the fp32 product is what a normalised copy holds)
template <class S> __device__ __forceinline__ const S *bcol(const S *V0, const S *V, int64_t ld, int k) {
  return k == 0 ? V0 : V + (int64_t)k * ld;
}
template <class A> __device__ __forceinline__ A bscale(const A *csc, int nsegs, int k, int seg) {
  return csc && k > 0 ? csc[(int64_t)k * nsegs + seg] : (A)1;
}

template <class S, class A, int NV>
__global__ __launch_bounds__(kBlock) void k_b_mdot(const S *__restrict__ w, const S *__restrict__ V0,
                                                   const S *__restrict__ V, int64_t ld, int ncols,
                                                   const A *__restrict__ csc, int nsegs,
                                                   const int32_t *__restrict__ bseg, A *__restrict__ partial) {
  __shared__ A lds[NV * kWavesPerBlock];
  const int k0 = blockIdx.y * NV;
  const int seg = bseg[blockIdx.x];
  A sk[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) sk[k] = k0 + k < ncols ? bscale(csc, nsegs, k0 + k, seg) : (A)1;
  A v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = (A)0;
  for (int q = 0; q < kQuads; ++q) {
    const int64_t i = tile_quad(q);
    const V4<S> wq = ld4(w + i);
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k0 + k < ncols) {
        const V4<S> vq = ld4(bcol(V0, V, ld, k0 + k) + i);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[k] = fma_<A>((A)to_acc(wq.v[c]), (A)to_acc(vq.v[c]) * sk[k], v[k]);
      }
  }
  block_sum<A, NV>(v, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV && k0 + k < ncols; ++k) partial[(int64_t)(k0 + k) * gridDim.x + blockIdx.x] = v[k];
}

// H[seg * hs + k] = the partials of segment seg's blocks, added in block
// order; grid (segments, columns)
template <class A>
__global__ __launch_bounds__(kBlock) void k_b_dots_finish(const A *__restrict__ partial, int G,
                                                          const int32_t *__restrict__ sb0,
                                                          const int32_t *__restrict__ snb, A *H, int64_t hs) {
  __shared__ A lds[kWavesPerBlock];
  const int seg = blockIdx.x, k = blockIdx.y;
  const int32_t b0 = sb0[seg], nb = snb[seg];
  A v[1] = {(A)0};
  for (int32_t b = threadIdx.x; b < nb; b += kBlock) v[0] += partial[(int64_t)k * G + b0 + b];
  block_sum<A, 1>(v, lds);
  if (threadIdx.x == 0) H[(int64_t)seg * hs + k] = v[0];
}

// w = S(w - V h) with h of the block's segment (MULMV order); block partial
// of ||w||^2 of the stored w
template <class S, class A>
__global__ __launch_bounds__(kBlock) void k_b_update(S *__restrict__ w, const S *__restrict__ V0,
                                                     const S *__restrict__ V, int64_t ld, int ncols,
                                                     const A *__restrict__ csc, int nsegs, const A *H, int64_t hs,
                                                     const int32_t *__restrict__ bseg, A *__restrict__ partial) {
  __shared__ A h[128], sk[128];
  __shared__ A lds[kWavesPerBlock];
  const int seg = bseg[blockIdx.x];
  for (int k = threadIdx.x; k < ncols; k += kBlock) {
    h[k] = H[(int64_t)seg * hs + k];
    sk[k] = bscale(csc, nsegs, k, seg);
  }
  __syncthreads();
  A v[1] = {(A)0};
  for (int q = 0; q < kQuads; ++q) {
    const int64_t i = tile_quad(q);
    const V4<S> v0 = ld4(V0 + i);
    V4<S> wq = ld4(w + i);
    A r[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = ((A)to_acc(v0.v[c]) * sk[0]) * h[0];
    for (int k = 1; k < ncols; ++k) {
      const V4<S> vk = ld4(V + (int64_t)k * ld + i);
#pragma unroll
      for (int c = 0; c < 4; ++c) r[c] = fma_<A>((A)to_acc(vk.v[c]) * sk[k], h[k], r[c]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      wq.v[c] = cvt<S>((A)to_acc(wq.v[c]) - r[c]);
      const A wf = (A)to_acc(wq.v[c]);
      v[0] = fma_<A>(wf, wf, v[0]);
    }
    st4(w + i, wq);
  }
  block_sum<A, 1>(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// block partial of ||v||^2 (the outer level's residual norm)
template <class A>
__global__ __launch_bounds__(kBlock) void k_b_norm2(const A *__restrict__ v, A *__restrict__ partial) {
  __shared__ A lds[kWavesPerBlock];
  A s[1] = {(A)0};
  for (int q = 0; q < kRows; ++q) {
    const A x = v[tile_row(q)];
    s[0] = fma_<A>(x, x, s[0]);
  }
  block_sum<A, 1>(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// One block per segment: ||w|| from its blocks' partials, then H(j+1, j),
// 1/||w|| (0 for a zero vector), the guarded Givens step and |e_{j+1}|.
// H, e, cs, sn point at segment 0's column j / arrays; strides per segment.
template <class A>
__global__ __launch_bounds__(kBlock) void k_b_norm_finish(const A *__restrict__ partial, const int32_t *__restrict__ sb0,
                                                          const int32_t *__restrict__ snb, A *H, int64_t hs, A *cs,
                                                          A *sn, int64_t ms, A *e, int64_t es, int j, int e0_one,
                                                          A *inv, double *status) {
  __shared__ A lds[kWavesPerBlock];
  const int seg = blockIdx.x;
  const int32_t b0 = sb0[seg], nb = snb[seg];
  A v[1] = {(A)0};
  for (int32_t b = threadIdx.x; b < nb; b += kBlock) v[0] += partial[b0 + b];
  block_sum<A, 1>(v, lds);
  if (threadIdx.x == 0) {
    A *Hc = H + (int64_t)seg * hs;
    const A nv = (A)sqrt(v[0]);
    Hc[j + 1] = nv;
    inv[seg] = nv > (A)0 ? (A)1 / nv : (A)0;
    if (cs) {
      A *ee = e + (int64_t)seg * es;
      givens_guarded<A>(Hc, cs + (int64_t)seg * ms, sn + (int64_t)seg * ms, ee, j, e0_one != 0);
      if (status) status[seg] = fabs((double)ee[j + 1]);
    } else if (status) {
      status[seg] = (double)nv;  // the residual norm itself (start of a cycle)
    }
  }
}

// v *= inv[segment of the tile]
template <class S, class A>
__global__ __launch_bounds__(kBlock) void k_b_vscale(S *__restrict__ v, const A *__restrict__ inv,
                                                     const int32_t *__restrict__ bseg) {
  const A a = inv[bseg[blockIdx.x]];
  for (int q = 0; q < kQuads; ++q) {
    const int64_t i = tile_quad(q);
    V4<S> x = ld4(v + i);
#pragma unroll
    for (int c = 0; c < 4; ++c) x.v[c] = cvt<S>((A)to_acc(x.v[c]) * a);
    st4(v + i, x);
  }
}

// out = TO(Z y) (accumulate: out += Z y) with y solved from the block's
// segment's j x j triangle; j = jseg[segment] or jall (jseg null)
template <class S, class A, class TO>
__global__ __launch_bounds__(kBlock) void k_b_combine(const S *__restrict__ Z, int64_t ld, const A *H, int64_t hs,
                                                      int ldh, const A *e, int64_t es, const int32_t *jseg, int jall,
                                                      const int32_t *__restrict__ bseg, TO *__restrict__ out,
                                                      int accumulate, const float *__restrict__ zsc) {
  __shared__ A y[128];
  const int seg = bseg[blockIdx.x];
  const int j = jseg ? jseg[seg] : jall;
  if (j == 0) {
    if (!accumulate)
      for (int q = 0; q < kRows; ++q) out[tile_row(q)] = cvt<TO>((A)0);
    return;
  }
  if (threadIdx.x == 0) hess_solve_guarded<A>(H + (int64_t)seg * hs, ldh, e + (int64_t)seg * es, y, j);
  __syncthreads();
  // zsc: Z holds z * s (fp16 storage, s a power of two): z = Z * zsc exactly
  const A zs = zsc ? (A)zsc[seg] : (A)1;
  for (int q = 0; q < kQuads; ++q) {
    const int64_t i = tile_quad(q);
    const V4<S> z0 = ld4(Z + i);
    A r[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = ((A)to_acc(z0.v[c]) * zs) * y[0];
    for (int k = 1; k < j; ++k) {
      const V4<S> zk = ld4(Z + (int64_t)k * ld + i);
#pragma unroll
      for (int c = 0; c < 4; ++c) r[c] = fma_<A>((A)to_acc(zk.v[c]) * zs, y[k], r[c]);
    }
    V4<TO> o;
    if (accumulate) {
      o = ld4(out + i);
#pragma unroll
      for (int c = 0; c < 4; ++c) o.v[c] = cvt<TO>((A)to_acc(o.v[c]) + r[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) o.v[c] = cvt<TO>(r[c]);
    }
    st4(out + i, o);
  }
}

template <class SI, class SO>
__global__ __launch_bounds__(kBlock) void k_b_cast(const SI *__restrict__ in, SO *__restrict__ out) {
  for (int q = 0; q < kRows; ++q) {
    const int64_t i = tile_row(q);
    out[i] = cvt<SO>(to_acc(in[i]));
  }
}

// out = SO(in * f[segment]) (conversions into and out of the fp16 level)
template <class SI, class SO>
__global__ __launch_bounds__(kBlock) void k_b_entry(const SI *__restrict__ in, SO *__restrict__ out,
                                                    const float *__restrict__ f, const int32_t *__restrict__ bseg) {
  const float a = f[bseg[blockIdx.x]];
  for (int q = 0; q < kRows; ++q) {
    const int64_t i = tile_row(q);
    out[i] = cvt<SO>(to_acc(in[i]) * a);
  }
}

// block partials of <t, t> and <t, r> (Richardson probe): partial[q * G + b]
__global__ __launch_bounds__(kBlock) void k_b_probe_dots(const half_t *__restrict__ t, const half_t *__restrict__ r,
                                                         float *__restrict__ partial) {
  __shared__ float lds[2 * kWavesPerBlock];
  float v[2] = {0.f, 0.f};
  for (int q = 0; q < kRows; ++q) {
    const int64_t i = tile_row(q);
    const float ti = to_acc(t[i]);
    v[0] = fma_<float>(ti, ti, v[0]);
    v[1] = fma_<float>(ti, to_acc(r[i]), v[1]);
  }
  block_sum<float, 2>(v, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = v[0];
    partial[(int64_t)gridDim.x + blockIdx.x] = v[1];
  }
}

// one block per segment: alpha_k = <t, r> / <t, t>, then the running mean of
// the weights (probe count den = probes + 1: the initial weight 1 counts)
__global__ __launch_bounds__(kBlock) void k_b_probe_finish(const float *__restrict__ partial, int G,
                                                           const int32_t *__restrict__ sb0,
                                                           const int32_t *__restrict__ snb, RichSeg *st, int k) {
  __shared__ float lds[2 * kWavesPerBlock];
  const int seg = blockIdx.x;
  const int32_t b0 = sb0[seg], nb = snb[seg];
  float v[2] = {0.f, 0.f};
  for (int32_t b = threadIdx.x; b < nb; b += kBlock) {
    v[0] += partial[b0 + b];
    v[1] += partial[(int64_t)G + b0 + b];
  }
  block_sum<float, 2>(v, lds);
  if (threadIdx.x == 0) {
    RichSeg &s = st[seg];
    const float al = v[0] > 0.f ? v[1] / v[0] : 0.f;
    if (k == 0) s.nprobe += 1;
    const float den = (float)(s.nprobe + 1);
    const float sm = (float)s.sum[k] + al;
    s.sum[k] = sm;
    s.ome[k] = sm / den;
    s.alpha[k] = al;
  }
}

// ho = fp16(ho * w) (step 0) or ho2 = fp16(w * t2 + ho) (step k), w the
// segment's weight (alpha on a probe call, else the running mean)
__global__ __launch_bounds__(kBlock) void k_b_l4_scale(half_t *__restrict__ ho, const RichSeg *__restrict__ st,
                                                       int probe, const int32_t *__restrict__ bseg) {
  const RichSeg &s = st[bseg[blockIdx.x]];
  const float w = (float)(probe ? s.alpha[0] : s.ome[0]);
  for (int q = 0; q < kRows; ++q) {
    const int64_t i = tile_row(q);
    ho[i] = cvt<half_t>(to_acc(ho[i]) * w);
  }
}
__global__ __launch_bounds__(kBlock) void k_b_l4_update(const half_t *__restrict__ t2, const half_t *__restrict__ ho,
                                                        half_t *__restrict__ ho2, const RichSeg *__restrict__ st, int k,
                                                        int probe, const int32_t *__restrict__ bseg) {
  const RichSeg &s = st[bseg[blockIdx.x]];
  const float w = (float)(probe ? s.alpha[k] : s.ome[k]);
  for (int q = 0; q < kRows; ++q) {
    const int64_t i = tile_row(q);
    ho2[i] = cvt<half_t>(fma_<float>(w, to_acc(t2[i]), to_acc(ho[i])));
  }
}

// SpMV of a block-diagonal matrix stored once per system, applied to the NR
// copies (right-hand sides) of each system's vectors: copy c of the rows of
// chunk s lies at vector row = matrix row + shift[s] + c * stride[s].  One
// lane takes a matrix row and keeps NR sums, so each value and column index
// is read once for all copies; every sum runs in slot order, as in k_sell.
// Epilogue: y = acc (b null) or y = b - acc, at the vector rows.
// Offset chunks gather x through a buffer resource over x's nx elements: the
// per-lane byte offset is the row's, the chunk offset and the copy's stride
// ride in the scalar offset (no per-lane address arithmetic per gather).
template <class V, class X, class ACC, int K, int NR, class Y>
__global__ __launch_bounds__(kBlock) void k_sell_mr(SellArgs a, const V *__restrict__ val, const X *__restrict__ x,
                                                    int64_t nx, const int32_t *__restrict__ cshift,
                                                    const int32_t *__restrict__ cstride, Y *__restrict__ y,
                                                    const Y *__restrict__ b, const int32_t *__restrict__ cseg,
                                                    const float *__restrict__ ysc) {
  typedef const __attribute__((address_space(4))) int32_t *const_i32;
  for (int r = xcd_block() * kBlock + threadIdx.x; r < a.n; r += gridDim.x * kBlock) {
    const int s = r >> a.log2C;
    const int j = r - (s << a.log2C);
    const int C = 1 << a.log2C;
    const int rem = a.n - (s << a.log2C);
    const int tc = rem < C ? rem : C;
    const int su = __builtin_amdgcn_readfirstlane(s);  // C >= 64: one chunk per wavefront
    const int cw = a.cwidth[su];
    const int groups = chunk_groups(cw);
    const int32_t shift = cshift[su], stride = cstride[su];
    int64_t pos = (int64_t)a.cptr[su] + (int64_t)j * K;
    const int64_t step = (int64_t)tc * K;
    ACC acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = 0;
    const int64_t vr = (int64_t)r + shift;
    const int32_t start = a.iptr ? a.iptr[su] : 0;
    if (a.iptr && start < 0) {
      const const_i32 off = (const_i32)(a.coff + (-start - 1));
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void *)((uintptr_t)x - (uintptr_t)a.off_bias * sizeof(X)), (short)0,
          (int)((nx + a.off_bias) * (int64_t)sizeof(X)), kBufferWord3);
      const int vo = (int)(vr * (int64_t)sizeof(X));
      const int bias = a.off_bias, cst = stride * (int)sizeof(X);
#pragma unroll 2
      for (int g = 0; g < groups; ++g, pos += step) {
        const Pack<V, K> pv = load_pack<V, K>(val + pos);
        X xv[K][NR];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int so = (off[g * K + k] + bias) * (int)sizeof(X);
#pragma unroll
          for (int c = 0; c < NR; ++c) xv[k][c] = buffer_gather<X>(rs, vo, so + c * cst);
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xv[k][c]), acc[c]);
      }
      if constexpr (K >= 2) {
        if (cw & kHalfTail) {  // the half group (K / 2 slots per row, rows consecutive)
          constexpr int H = K / 2;
          const Pack<V, H> pv = load_pack<V, H>(val + ((int64_t)a.cptr[su] + (int64_t)groups * step + (int64_t)j * H));
          X xv[H][NR];
#pragma unroll
          for (int k = 0; k < H; ++k) {
            const int so = (off[groups * K + k] + bias) * (int)sizeof(X);
#pragma unroll
            for (int c = 0; c < NR; ++c) xv[k][c] = buffer_gather<X>(rs, vo, so + c * cst);
          }
#pragma unroll
          for (int k = 0; k < H; ++k)
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[c] = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xv[k][c]), acc[c]);
        }
      }
    } else {
      const int32_t *__restrict__ ip = a.idx + (a.iptr ? (int64_t)start : (int64_t)a.cptr[su]) + (int64_t)j * K;
      for (int g = 0; g < groups; ++g, pos += step, ip += step) {
        const Pack<V, K> pv = load_pack<V, K>(val + pos);
        const Pack<int32_t, K> pi = load_pack<int32_t, K>(ip);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int64_t col = (int64_t)pi.v[k] + shift;
#pragma unroll
          for (int c = 0; c < NR; ++c)
            acc[c] = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(x[col + (int64_t)c * stride]), acc[c]);
        }
      }
    }
    // ysc: x held z * s (fp16 storage, s a power of two per segment), so
    // y = sum * (1/s) is the sum over z exactly
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const int64_t o = vr + (int64_t)c * stride;
      const ACC sum = ysc ? acc[c] * (ACC)ysc[cseg[su] + c] : acc[c];
      y[o] = b ? cvt<Y>((ACC)to_acc(b[o]) - sum) : cvt<Y>(sum);
    }
  }
}

// The multi-copy SpMV with K threads per row (int32-index layout): lane k of
// a matrix row sums slot k of each of the row's groups for all NR copies,
// then the K lanes add their sums by shuffles (the sums are reassociated).
template <class V, class X, class ACC, int K, int NR, class Y>
__global__ __launch_bounds__(kBlock) void k_sell_mr_lanes(SellArgs a, const V *__restrict__ val,
                                                          const X *__restrict__ x, const int32_t *__restrict__ cshift,
                                                          const int32_t *__restrict__ cstride, Y *__restrict__ y,
                                                          const Y *__restrict__ b, const int32_t *__restrict__ cseg,
                                                          const float *__restrict__ ysc) {
  static_assert(K >= 2 && kWave % K == 0, "K lanes of a row lie in one wavefront");
  constexpr int kRowsPerBlock = kBlock / K;
  const int k = threadIdx.x % K;
  for (int r0 = xcd_block() * kRowsPerBlock; r0 < a.n; r0 += gridDim.x * kRowsPerBlock) {
    const int r = r0 + threadIdx.x / K;
    ACC acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = 0;
    int s = 0;
    int64_t vr = 0, stride = 0;
    if (r < a.n) {
      s = r >> a.log2C;
      const int j = r - (s << a.log2C);
      const int C = 1 << a.log2C;
      const int rem = a.n - (s << a.log2C);
      const int tc = rem < C ? rem : C;
      const int groups = chunk_groups(a.cwidth[s]);
      const int64_t shift = cshift[s];
      stride = cstride[s];
      vr = (int64_t)r + shift;
      int64_t pos = (int64_t)a.cptr[s] + (int64_t)j * K + k;
      const int64_t step = (int64_t)tc * K;
      for (int g = 0; g < groups; ++g, pos += step) {
        const ACC v = (ACC)to_acc(val[pos]);
        const int64_t col = (int64_t)a.idx[pos] + shift;
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = fma_<ACC>(v, (ACC)to_acc(x[col + (int64_t)c * stride]), acc[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < NR; ++c)
#pragma unroll
      for (int off = K / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, kWave);
    if (r < a.n && k == 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const int64_t o = vr + (int64_t)c * stride;
        const ACC sum = ysc ? acc[c] * (ACC)ysc[cseg[s] + c] : acc[c];
        y[o] = b ? cvt<Y>((ACC)to_acc(b[o]) - sum) : cvt<Y>(sum);
      }
    }
  }
}

// An SDAINV factor split by row kind: a row whose only entry is its diagonal
// is an elementwise product (dval), any other row is a CSR row summed in
// column order with fp32 accumulation, as the SELL SpMV sums it; dval holds
// an fp16 NaN for those rows.  M = Z W^T x runs as two launches:
//   k_b_mw: t = W^T x for every row (tile blocks: diagonal rows; row blocks:
//           the CSR rows), and y = Z t for the rows where Z is diagonal;
//   k_b_mz: y = Z t for the CSR rows of Z.  They read t of other rows: t of
//           a diagonal row of W is recomputed from x, t of a CSR row of W
//           is read back (k_b_mw stores t for those rows only).
// Each row's products are those of applying W^T, then Z, one at a time.
struct Factor16 {
  DevPtr dval, rows, rseg, rptr, col, val;  // rseg: segment of each CSR row
  int32_t nr = 0;
};
// dval is indexed by vector row + dsh[segment]: the right-hand sides of one
// system share one copy of its diagonals when they share its factors

__device__ __forceinline__ bool is_nan16(half_t v) {
  return (__builtin_bit_cast(uint16_t, v) & 0x7fffu) > 0x7c00u;
}

// input x_i of M: fp16 x, or the level-4 entry fp16(in * scale[seg]) computed
// here (and stored into hin by the tile blocks)
struct MIn {
  const half_t *x;
  const float *in;
  const float *sc;
  half_t *hin;
  const float *inv;  // `in` is an unnormalised basis column: in * inv[seg] (null: 1)
  __device__ __forceinline__ float iv(int seg) const { return inv ? inv[seg] : 1.f; }
  __device__ __forceinline__ half_t get(int64_t i, int seg) const {
    return in ? cvt<half_t>((to_acc(in[i]) * iv(seg)) * sc[seg]) : x[i];
  }
};

// what becomes of y_i = (M x)_i: 0 stored; 1 scaled by the segment's ome_0
// and stored; 2 the Richardson update fp16(fma(ome_k, y_i, ho_i)) stored;
// 3 that update converted out of the fp16 level, out_i = fp32(.) * iscale
struct MOut {
  int mode;
  half_t *y;
  const RichSeg *rs;
  int k;
  const half_t *ho;
  float *out;
  const float *iscale;
  // the segment's weight (modes 1-3)
  __device__ __forceinline__ float weight(int seg) const {
    return mode == 0 ? 0.f : (float)rs[seg].ome[mode == 1 ? 0 : k];
  }
  // hov = ho[i] (modes 2, 3), w = weight(seg)
  __device__ __forceinline__ void put(int64_t i, int seg, half_t v, half_t hov, float w) const {
    if (mode == 0) {
      y[i] = v;
    } else if (mode == 1) {
      y[i] = cvt<half_t>(to_acc(v) * w);
    } else {
      const half_t h = cvt<half_t>(fma_<float>(w, to_acc(v), to_acc(hov)));
      if (mode == 2)
        y[i] = h;
      else
        out[i] = cvt<float>(to_acc(h) * iscale[seg]);
    }
  }
  __device__ __forceinline__ void put(int64_t i, int seg, half_t v) const {
    put(i, seg, v, mode >= 2 ? ho[i] : (half_t)0, weight(seg));
  }
  // four consecutive rows (i aligned to 4): the same values as put() per row
  __device__ __forceinline__ void put4(int64_t i, int seg, const V4<half_t> &v, const V4<half_t> &hov,
                                       float w) const {
    if (mode <= 2) {
      V4<half_t> r;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        r.v[c] = mode == 0   ? v.v[c]
                 : mode == 1 ? cvt<half_t>(to_acc(v.v[c]) * w)
                             : cvt<half_t>(fma_<float>(w, to_acc(v.v[c]), to_acc(hov.v[c])));
      st4(y + i, r);
    } else {
      const float is = iscale[seg];
      V4<float> r;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        r.v[c] = cvt<float>(to_acc(cvt<half_t>(fma_<float>(w, to_acc(v.v[c]), to_acc(hov.v[c])))) * is);
      st4(out + i, r);
    }
  }
};

struct FacArgs {
  const half_t *__restrict__ d;
  const int32_t *__restrict__ rows;
  const int32_t *__restrict__ rseg;
  const int32_t *__restrict__ rptr;
  const int32_t *__restrict__ col;
  const half_t *__restrict__ val;
  int32_t nr;
};

__global__ __launch_bounds__(kBlock) void k_b_mw(MIn xin, FacArgs W, const half_t *__restrict__ zd,
                                                 const int64_t *__restrict__ dsh, half_t *__restrict__ t, int G,
                                                 const int32_t *__restrict__ bseg, MOut o) {
  if ((int)blockIdx.x < G) {
    // every load of the thread's rows first, then the stores (a branch
    // between a row's loads would keep one row's loads in flight at a time)
    const int seg = bseg[blockIdx.x];
    const float w = o.weight(seg);
    const float sc = xin.in ? xin.sc[seg] : 0.f, iv = xin.iv(seg);
    const int64_t sh = dsh[seg];
    V4<half_t> xi[kQuads], dw[kQuads], dz[kQuads], hov[kQuads];
#pragma unroll
    for (int q = 0; q < kQuads; ++q) {
      const int64_t i = tile_quad(q);
      if (xin.in) {
        const V4<float> f = ld4(xin.in + i);
#pragma unroll
        for (int c = 0; c < 4; ++c) xi[q].v[c] = cvt<half_t>((f.v[c] * iv) * sc);
      } else {
        xi[q] = ld4(xin.x + i);
      }
      dw[q] = ld4(W.d + i + sh);
      dz[q] = ld4(zd + i + sh);
      if (o.mode >= 2) hov[q] = ld4(o.ho + i);
    }
#pragma unroll
    for (int q = 0; q < kQuads; ++q) {
      const int64_t i = tile_quad(q);
      if (xin.in) st4(xin.hin + i, xi[q]);
      bool plain = true;  // all four rows diagonal in W and Z
      V4<half_t> yv;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        plain = plain && !is_nan16(dw[q].v[c]) && !is_nan16(dz[q].v[c]);
        const half_t ti = cvt<half_t>(to_acc(dw[q].v[c]) * to_acc(xi[q].v[c]));
        yv.v[c] = cvt<half_t>(to_acc(dz[q].v[c]) * to_acc(ti));
      }
      if (plain) {
        o.put4(i, seg, yv, o.mode >= 2 ? hov[q] : yv, w);
        continue;
      }
      // a CSR row of W is the row blocks' (and a CSR row of Z k_b_mz's)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (!is_nan16(dw[q].v[c]) && !is_nan16(dz[q].v[c])) o.put(i + c, seg, yv.v[c], hov[q].v[c], w);
    }
    return;
  }
  const int32_t r = (int32_t)(blockIdx.x - G) * kBlock + threadIdx.x;
  if (r >= W.nr) return;
  const int64_t i = W.rows[r];
  const int seg = W.rseg[r];
  float acc = 0.f;
  for (int32_t k = W.rptr[r]; k < W.rptr[r + 1]; ++k)
    acc = fma_<float>(to_acc(W.val[k]), to_acc(xin.get(W.col[k], seg)), acc);
  const half_t ti = cvt<half_t>(acc);
  t[i] = ti;
  const half_t dz = zd[i + dsh[seg]];
  if (!is_nan16(dz)) o.put(i, seg, cvt<half_t>(to_acc(dz) * to_acc(ti)));
}

__global__ __launch_bounds__(kBlock) void k_b_mz(FacArgs Z, MIn xin, const half_t *__restrict__ wd,
                                                 const int64_t *__restrict__ dsh, const half_t *__restrict__ t,
                                                 MOut o) {
  const int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= Z.nr) return;
  const int seg = Z.rseg[r];
  const int64_t sh = dsh[seg];
  float acc = 0.f;
  for (int32_t k = Z.rptr[r]; k < Z.rptr[r + 1]; ++k) {
    const int32_t c = Z.col[k];
    const half_t dw = wd[c + sh];
    const half_t tc = is_nan16(dw) ? t[c] : cvt<half_t>(to_acc(dw) * to_acc(xin.get(c, seg)));
    acc = fma_<float>(to_acc(Z.val[k]), to_acc(tc), acc);
  }
  o.put(Z.rows[r], seg, cvt<half_t>(acc));
}

// ----------------------------------------------------------- solver ---
template <class T> struct BLevel {
  int m = 0;
  DevPtr V, Z, H, cs, sn, e, inv;  // V, Z column-major (ld = N); per segment H, cs, sn, e, inv
  DevPtr csc;  // per column and segment 1/||w|| of columns left unnormalised (level 3)
  T *Vc(int64_t N, int k) const { return V->as<T>() + (int64_t)k * N; }
  T *Zc(int64_t N, int k) const { return Z->as<T>() + (int64_t)k * N; }
  int64_t hs() const { return (int64_t)(m + 1) * m; }
};

HostCSR concat(const std::vector<const HostCSR *> &parts, const std::vector<int64_t> &off, int64_t N) {
  HostCSR R;
  int64_t nnz = 0;
  for (const HostCSR *p : parts) nnz += p->nnz();
  F3R_CHECK(N < INT32_MAX && nnz < INT32_MAX, F3R_ERR_RANGE, "batch: more than 2^31 rows or nonzeros in all");
  R.n = R.ncols = (int32_t)N;
  R.rptr.assign(N + 1, 0);
  R.idx.resize(nnz);
  R.val.resize(nnz);
  int64_t q = 0;
  for (size_t s = 0; s < parts.size(); ++s) {
    const HostCSR &A = *parts[s];
    const int32_t o = (int32_t)off[s];
    for (int32_t i = 0; i < A.n; ++i) {
      for (int32_t k = A.rptr[i]; k < A.rptr[i + 1]; ++k, ++q) {
        R.idx[q] = A.idx[k] + o;
        R.val[q] = A.val[k];
      }
      R.rptr[o + i + 1] = (int32_t)q;
    }
    for (int64_t r = o + A.n; r < off[s + 1]; ++r) R.rptr[r + 1] = (int32_t)q;  // padding rows: empty
  }
  return R;
}

}  // namespace

struct BatchSolver {
  std::shared_ptr<Context> ctx;
  f3r_params p{};
  int nseg = 0;
  int64_t N = 0;
  int G = 0;  // blocks of kBlock rows
  std::vector<int64_t> off;
  std::vector<int32_t> nrow;
  DevPtr bseg, sb0, snb, scale, iscale;
  std::shared_ptr<SellShared> sh, sh16;
  DevPtr A64, A32, A16;
  // NR > 1: each system's matrix is stored once for its NR right-hand sides
  // (chunk -> vector-row shift and copy stride, per layout)
  int NR = 1;
  // flags bit 4: the SpMVs with K threads per row (int32-index layouts)
  bool lanes = false;
  DevPtr csh, cst, csh16, cst16;
  Factor16 W, Z;
  DevPtr dsh;  // per segment: vector row -> row of the factor diagonals
  DevPtr cseg;  // NR > 1: per chunk, the segment of its first copy
  // level 3's Z as the fp16 Richardson output h (Z = h / s exactly, s the
  // segment's power-of-two range scale): NR > 1 and every s a power of two
  bool zh = false;
  DevPtr Zh3;
  half_t *zh3(int k) const { return Zh3->as<half_t>() + (int64_t)k * N; }
  DevPtr partial, status, jseg, rich;
  DevPtr b_int, x_int, r_int;
  // outer level (fp64)
  BLevel<double> L1;
  BLevel<float> L2, L3;
  DevPtr hin, hout, hout2, hr, tmp, tt, t2;
  int64_t cnt = 0, calls = 0;
  std::map<uint64_t, hipGraphExec_t> graphs;  // level-2 body per probe pattern
  // a whole outer step per (probe pattern, step index j)
  std::map<std::pair<uint64_t, int>, hipGraphExec_t> step_graphs;
  static constexpr size_t kMaxStepGraphs = 64;
  std::vector<hipGraph_t> graph_src;

  BatchSolver() = default;
  BatchSolver(const BatchSolver &) = delete;
  BatchSolver &operator=(const BatchSolver &) = delete;
  ~BatchSolver() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    for (auto &g : step_graphs) (void)hipGraphExecDestroy(g.second);
    for (auto &g : graph_src) (void)hipGraphDestroy(g);
  }

  hipStream_t st() const { return ctx->stream; }
  int32_t *bs() const { return bseg->as<int32_t>(); }

  // y = A x (b null) or y = b - A x over all segments, A the fp64/fp32 copy
  // (layout sh) or the fp16 one (sh16)
  template <class V, class X, class Y, class ACC>
  // ysc (NR > 1 only): per-segment factor on the sums (x stored scaled)
  void spmv(const V *val, bool half_layout, const X *x, Y *y, const Y *b, const float *ysc = nullptr) {
    const SellShared &L = half_layout ? *sh16 : *sh;
    SellArgs a = sell_args(L, false);
    if (NR == 1) {
      F3R_CHECK(ysc == nullptr, F3R_ERR_STATE, "batch: scaled SpMV needs shared matrices");
      a.lanes = lanes ? L.K : 1;  // K threads per row
      if (b)
        launch_sell<V, X, ACC>(L.K, a, val, x, EpiResid<Y, Y, ACC>{y, b, nullptr}, st());
      else
        launch_sell<V, X, ACC>(L.K, a, val, x, EpiStore<Y, ACC>{y, nullptr}, st());
      return;
    }
    const int32_t *shf = (half_layout ? csh16 : csh)->as<int32_t>(), *str = (half_layout ? cst16 : cst)->as<int32_t>();
    const dim3 g((a.n + kBlock - 1) / kBlock);
    auto go = [&](auto kt, auto nt) {
      constexpr int K = decltype(kt)::value, R = decltype(nt)::value;
      if constexpr (K >= 2) {
        if (lanes) {  // K threads per matrix row
          const dim3 gl((unsigned)(((int64_t)a.n * K + kBlock - 1) / kBlock));
          hipLaunchKernelGGL((k_sell_mr_lanes<V, X, ACC, K, R, Y>), gl, dim3(kBlock), 0, st(), a, val, x, shf, str,
                             y, b, (const int32_t *)cseg->as<int32_t>(), ysc);
          return;
        }
      }
      hipLaunchKernelGGL((k_sell_mr<V, X, ACC, K, R, Y>), g, dim3(kBlock), 0, st(), a, val, x, (int64_t)N, shf, str, y,
                         b, (const int32_t *)cseg->as<int32_t>(), ysc);
    };
    auto byk = [&](auto nt) {
      switch (L.K) {
        case 1: go(std::integral_constant<int, 1>(), nt); break;
        case 2: go(std::integral_constant<int, 2>(), nt); break;
        case 4: go(std::integral_constant<int, 4>(), nt); break;
        default: go(std::integral_constant<int, 8>(), nt); break;
      }
    };
    if (NR == 2)
      byk(std::integral_constant<int, 2>());
    else
      byk(std::integral_constant<int, 4>());
  }

  template <class T> void alloc_level(BLevel<T> &L, int m) {
    L.m = m;
    L.V = dev_alloc(sizeof(T) * N * (m + 1));
    L.Z = dev_alloc(sizeof(T) * N * m);
    L.H = dev_alloc(sizeof(T) * nseg * L.hs());
    L.cs = dev_alloc(sizeof(T) * nseg * m);
    L.sn = dev_alloc(sizeof(T) * nseg * m);
    L.e = dev_alloc(sizeof(T) * nseg * (m + 1));
    L.inv = dev_alloc(sizeof(T) * nseg);
    L.csc = dev_alloc(sizeof(T) * nseg * (m + 1));
  }

  // one Arnoldi step j of level L (w = V_{j+1} already holds A z_j).  V0:
  // column 0 (the level's input, not copied).  lazy: V_{j+1} stays
  // unnormalised, its 1/||w|| per segment goes to csc column j + 1 and every
  // reader applies it (no scaling pass)
  template <class T> void arnoldi(BLevel<T> &L, int j, int e0_one, double *stat, const T *V0, bool lazy) {
    const int ncols = j + 1;
    T *H = L.H->template as<T>();
    T *part = partial->as<T>();
    const T *csc = lazy ? L.csc->template as<T>() : nullptr;
    const T *Vb = L.Vc(N, 0);
    if (ncols <= 4)
      hipLaunchKernelGGL((k_b_mdot<T, T, 4>), dim3(G, 1), dim3(kBlock), 0, st(), (const T *)L.Vc(N, j + 1), V0, Vb,
                         N, ncols, csc, nseg, (const int32_t *)bs(), part);
    else if (ncols <= 8)
      hipLaunchKernelGGL((k_b_mdot<T, T, 8>), dim3(G, 1), dim3(kBlock), 0, st(), (const T *)L.Vc(N, j + 1), V0, Vb,
                         N, ncols, csc, nseg, (const int32_t *)bs(), part);
    else
      hipLaunchKernelGGL((k_b_mdot<T, T, 16>), dim3(G, (ncols + 15) / 16), dim3(kBlock), 0, st(),
                         (const T *)L.Vc(N, j + 1), V0, Vb, N, ncols, csc, nseg, (const int32_t *)bs(), part);
    hipLaunchKernelGGL(k_b_dots_finish<T>, dim3(nseg, ncols), dim3(kBlock), 0, st(), (const T *)part, G,
                       (const int32_t *)sb0->as<int32_t>(), (const int32_t *)snb->as<int32_t>(),
                       H + (int64_t)j * (L.m + 1), L.hs());
    hipLaunchKernelGGL((k_b_update<T, T>), dim3(G), dim3(kBlock), 0, st(), L.Vc(N, j + 1), V0, Vb, N, ncols, csc,
                       nseg, (const T *)(H + (int64_t)j * (L.m + 1)), L.hs(), (const int32_t *)bs(), part);
    hipLaunchKernelGGL(k_b_norm_finish<T>, dim3(nseg), dim3(kBlock), 0, st(), (const T *)part,
                       (const int32_t *)sb0->as<int32_t>(), (const int32_t *)snb->as<int32_t>(),
                       H + (int64_t)j * (L.m + 1), L.hs(), L.cs->template as<T>(), L.sn->template as<T>(),
                       (int64_t)L.m, L.e->template as<T>(), (int64_t)(L.m + 1), j, e0_one,
                       lazy ? L.csc->template as<T>() + (int64_t)(j + 1) * nseg : L.inv->template as<T>(), stat);
    if (lazy) return;
    hipLaunchKernelGGL((k_b_vscale<T, T>), dim3(G), dim3(kBlock), 0, st(), L.Vc(N, j + 1),
                       (const T *)L.inv->template as<T>(), (const int32_t *)bs());
  }

  static FacArgs fac(const Factor16 &F) {
    return {F.dval->as<half_t>(), F.rows->as<int32_t>(), F.rseg->as<int32_t>(), F.rptr->as<int32_t>(),
            F.col->as<int32_t>(), F.val->as<half_t>(), F.nr};
  }
  // y = M x with M = Z W^T (ainv.hpp:110-114): W^T first, then Z; x given
  // by `xin`, y consumed by `o`
  void precond(const MIn &xin, const MOut &o) {
    half_t *t = tmp->as<half_t>();
    const int gw = (W.nr + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_b_mw, dim3(G + gw), dim3(kBlock), 0, st(), xin, fac(W), (const half_t *)Z.dval->as<half_t>(),
                       (const int64_t *)dsh->as<int64_t>(), t, G, (const int32_t *)bs(), o);
    if (Z.nr > 0) {
      // k_b_mz reads x as k_b_mw does (in the entry call: fp16(in * s) from
      // the fp32 input)
      hipLaunchKernelGGL(k_b_mz, dim3((Z.nr + kBlock - 1) / kBlock), dim3(kBlock), 0, st(), fac(Z), xin,
                         (const half_t *)W.dval->as<half_t>(), (const int64_t *)dsh->as<int64_t>(), (const half_t *)t,
                         o);
    }
  }
  void precond(const half_t *in, half_t *out) {
    precond(MIn{in, nullptr, nullptr, nullptr, nullptr}, MOut{0, out, nullptr, 0, nullptr, nullptr, nullptr});
  }

  void probe(const half_t *x, const half_t *r, int k) {
    half_t *t = tt->as<half_t>();
    spmv<half_t, half_t, half_t, float>(A16->as<half_t>(), true, x, t, nullptr);
    float *part = partial->as<float>();
    hipLaunchKernelGGL(k_b_probe_dots, dim3(G), dim3(kBlock), 0, st(), (const half_t *)t, r, part);
    hipLaunchKernelGGL(k_b_probe_finish, dim3(nseg), dim3(kBlock), 0, st(), (const float *)part, G,
                       (const int32_t *)sb0->as<int32_t>(), (const int32_t *)snb->as<int32_t>(),
                       rich->as<RichSeg>(), k);
  }

  // adaptive Richardson in fp16 (work/src/f3r-gpu.cpp:229-260), per segment
  // in_inv: per-segment 1/||w|| of an unnormalised input column (null: 1)
  // hz: the fp16 result h itself instead of out = h / s
  void level4(const float *in, const float *in_inv, float *out, half_t *hz = nullptr) {
    ++cnt;
    ++calls;
    const bool pr = (cnt % p.c) == 0;
    half_t *hi = hin->as<half_t>(), *ho = hout->as<half_t>(), *ho2 = hout2->as<half_t>();
    half_t *hr_ = hr->as<half_t>(), *tw = t2->as<half_t>();
    RichSeg *rs = rich->as<RichSeg>();
    // entry hin = fp16(in * s) fused into the first application of M; off
    // probe calls the weighting by ome_0 is fused too
    const float *isc = iscale->as<float>();
    precond(MIn{nullptr, in, scale->as<float>(), hi, in_inv},
            MOut{pr ? 0 : 1, ho, rs, 0, nullptr, nullptr, nullptr});
    if (pr) {
      probe(ho, hi, 0);
      hipLaunchKernelGGL(k_b_l4_scale, dim3(G), dim3(kBlock), 0, st(), ho, (const RichSeg *)rs, 1,
                         (const int32_t *)bs());
    }
    bool done = false;  // the last step already wrote the result
    for (int k = 1; k < p.m4; ++k) {
      const bool last = k == p.m4 - 1;
      half_t *hn = last && hz ? hz : ho2;  // the step's fp16 result
      spmv<half_t, half_t, half_t, float>(A16->as<half_t>(), true, (const half_t *)ho, hr_, (const half_t *)hi);
      if (pr) {
        precond(hr_, tw);
        probe(tw, hr_, k);
        hipLaunchKernelGGL(k_b_l4_update, dim3(G), dim3(kBlock), 0, st(), (const half_t *)tw, (const half_t *)ho,
                           hn, (const RichSeg *)rs, k, 1, (const int32_t *)bs());
        done = last && hz;
      } else {
        // the weighted update (and on the last step the way out of the
        // fp16 level, unless hz takes h itself) fused into M
        precond(MIn{hr_, nullptr, nullptr, nullptr, nullptr},
                MOut{last && !hz ? 3 : 2, hn, rs, k, (const half_t *)ho, out, isc});
        done = last;
      }
      std::swap(ho, ho2);
    }
    if (done) return;
    if (hz)
      hipLaunchKernelGGL((k_b_cast<half_t, half_t>), dim3(G), dim3(kBlock), 0, st(), (const half_t *)ho, hz);
    else
      hipLaunchKernelGGL((k_b_entry<half_t, float>), dim3(G), dim3(kBlock), 0, st(), (const half_t *)ho, out, isc,
                         (const int32_t *)bs());
  }

  // FGMRES(m3) on the fp16 values, fp32 basis, normalize = false
  // (V_0 is the level-2 column itself; the other columns stay unnormalised)
  void level3(const float *in, float *out) {
    BLevel<float> &L = L3;
    const float *isc = iscale->as<float>();
    for (int j = 0; j < L.m; ++j) {
      const float *vin = j == 0 ? in : L.Vc(N, j);
      const float *vinv = j == 0 ? nullptr : L.csc->as<float>() + (int64_t)j * nseg;
      if (zh) {
        level4(vin, vinv, nullptr, zh3(j));
        spmv<half_t, half_t, float, float>(A16->as<half_t>(), true, (const half_t *)zh3(j), L.Vc(N, j + 1), nullptr,
                                           isc);
      } else {
        level4(vin, vinv, L.Zc(N, j));
        spmv<half_t, float, float, float>(A16->as<half_t>(), true, (const float *)L.Zc(N, j), L.Vc(N, j + 1),
                                          nullptr);
      }
      arnoldi(L, j, 1, nullptr, in, true);
    }
    if (zh)
      hipLaunchKernelGGL((k_b_combine<half_t, float, float>), dim3(G), dim3(kBlock), 0, st(), (const half_t *)zh3(0),
                         N, (const float *)L.H->as<float>(), L.hs(), L.m + 1, (const float *)L.e->as<float>(),
                         (int64_t)(L.m + 1), (const int32_t *)nullptr, L.m, (const int32_t *)bs(), out, 0, isc);
    else
      hipLaunchKernelGGL((k_b_combine<float, float, float>), dim3(G), dim3(kBlock), 0, st(), (const float *)L.Zc(N, 0),
                         N, (const float *)L.H->as<float>(), L.hs(), L.m + 1, (const float *)L.e->as<float>(),
                         (int64_t)(L.m + 1), (const int32_t *)nullptr, L.m, (const int32_t *)bs(), out, 0,
                         (const float *)nullptr);
  }

  // the level-2 Arnoldi loop: every buffer it touches is the solver's own, so
  // the same launches serve every call with the same probe positions
  void level2_body() {
    BLevel<float> &L = L2;
    for (int j = 0; j < L.m; ++j) {
      level3(L.Vc(N, j), L.Zc(N, j));
      spmv<float, float, float, float>(A32->as<float>(), false, (const float *)L.Zc(N, j), L.Vc(N, j + 1), nullptr);
      arnoldi(L, j, 1, nullptr, (const float *)L.Vc(N, 0), false);
    }
  }

  // probe positions of the next level-2 body as a bit mask (bit q: the q-th
  // level-4 call probes); UINT64_MAX when not representable
  uint64_t next_pattern() const {
    const int64_t calls_per = (int64_t)p.m2 * p.m3;
    if (calls_per > 63) return UINT64_MAX;
    uint64_t mask = 0;
    for (int64_t q = 0; q < calls_per; ++q)
      if (((cnt + q + 1) % p.c) == 0) mask |= 1ULL << q;
    return mask | (1ULL << 63);
  }

  // FGMRES(m2) on the fp32 values, fp32 basis, normalize = false.  With
  // use_graph the body is captured once per probe pattern (at most 33) and
  // replayed: it holds no host synchronisation, memset or host copy, and the
  // Richardson weights and probe results stay in device memory between calls
  // body_graph false: the body is launched eagerly (inside a capture of the
  // whole outer step)
  void level2(const double *in, double *out, bool body_graph) {
    BLevel<float> &L = L2;
    hipLaunchKernelGGL((k_b_cast<double, float>), dim3(G), dim3(kBlock), 0, st(), in, L.Vc(N, 0));
    const uint64_t key = p.use_graph && body_graph ? next_pattern() : UINT64_MAX;
    if (key == UINT64_MAX || (graphs.size() > 32 && graphs.find(key) == graphs.end())) {
      level2_body();
    } else {
      auto it = graphs.find(key);
      if (it == graphs.end()) {
        const int64_t cnt0 = cnt, calls0 = calls;
        hipGraph_t gr;
        F3R_HIP(hipStreamBeginCapture(st(), hipStreamCaptureModeThreadLocal));
        try {
          level2_body();
        } catch (...) {
          (void)hipStreamEndCapture(st(), &gr);
          throw;
        }
        F3R_HIP(hipStreamEndCapture(st(), &gr));
        hipGraphExec_t ex;
        const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        if (e != hipSuccess) (void)hipGraphDestroy(gr);
        F3R_HIP(e);
        graph_src.push_back(gr);
        it = graphs.emplace(key, ex).first;
        cnt = cnt0;  // the capture only recorded the calls
        calls = calls0;
      }
      F3R_HIP(hipGraphLaunch(it->second, st()));
      cnt += (int64_t)p.m2 * p.m3;
      calls += (int64_t)p.m2 * p.m3;
    }
    hipLaunchKernelGGL((k_b_combine<float, float, double>), dim3(G), dim3(kBlock), 0, st(),
                       (const float *)L.Zc(N, 0), N, (const float *)L.H->as<float>(), L.hs(), L.m + 1,
                       (const float *)L.e->as<float>(), (int64_t)(L.m + 1), (const int32_t *)nullptr, L.m,
                       (const int32_t *)bs(), out, 0, (const float *)nullptr);
  }

  // outer Arnoldi step j of every segment: level 2 on V_j into Z_j, the fp64
  // SpMV into V_{j+1} and the outer Gram-Schmidt / Givens step, whose |e_{j+1}|
  // per segment lands in `status`.  With use_graph the whole step is
  // captured once per (probe pattern, j) and replayed as one graph launch
  // (beyond kMaxStepGraphs of them, only the level-2 bodies are replayed).
  void outer_step(int j) {
    BLevel<double> &L = L1;
    double *V = L.V->as<double>();
    auto body = [&](bool body_graph) {
      level2(V + (int64_t)j * N, L.Zc(N, j), body_graph);
      spmv<double, double, double, double>(A64->as<double>(), false, (const double *)L.Zc(N, j),
                                           V + (int64_t)(j + 1) * N, nullptr);
      arnoldi(L, j, 0, status->as<double>(), (const double *)L.Vc(N, 0), false);
    };
    const uint64_t pat = p.use_graph ? next_pattern() : UINT64_MAX;
    const auto key = std::make_pair(pat, j);
    auto it = step_graphs.find(key);
    if (pat == UINT64_MAX || (it == step_graphs.end() && step_graphs.size() >= kMaxStepGraphs)) {
      body(true);
      return;
    }
    if (it == step_graphs.end()) {
      const int64_t cnt0 = cnt, calls0 = calls;
      hipGraph_t gr;
      F3R_HIP(hipStreamBeginCapture(st(), hipStreamCaptureModeThreadLocal));
      try {
        body(false);
      } catch (...) {
        (void)hipStreamEndCapture(st(), &gr);
        throw;
      }
      F3R_HIP(hipStreamEndCapture(st(), &gr));
      hipGraphExec_t ex;
      const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      if (e != hipSuccess) (void)hipGraphDestroy(gr);
      F3R_HIP(e);
      graph_src.push_back(gr);
      it = step_graphs.emplace(key, ex).first;
      cnt = cnt0;  // the capture only recorded the calls
      calls = calls0;
    }
    F3R_HIP(hipGraphLaunch(it->second, st()));
    cnt += (int64_t)p.m2 * p.m3;
    calls += (int64_t)p.m2 * p.m3;
  }

  std::vector<double> read_status() {
    std::vector<double> h(nseg);
    F3R_HIP(hipMemcpyAsync(h.data(), status->p, sizeof(double) * nseg, hipMemcpyDeviceToHost, st()));
    F3R_HIP(hipStreamSynchronize(st()));
    return h;
  }

  // Restarted<double>(FGMRES<double>(A, L2, {m1, true}), {restart}) per segment
  // |e| / ||b|| against eps; a zero right-hand side is met only by a zero residual
  static double rel(double e, double nb) { return nb > 0 ? e / nb : (e == 0 ? 0.0 : HUGE_VAL); }

  void solve(const std::vector<double> &nrm_b, std::vector<f3r_result> &res) {
    BLevel<double> &L = L1;
    const int m1 = L.m;
    double *V = L.V->as<double>(), *Hd = L.H->as<double>();
    std::vector<char> solved(nseg, 0);
    std::vector<int32_t> iters(nseg, 0), cycles(nseg, 0);
    std::vector<double> last(nseg, 0.0);
    cnt = p.c;
    calls = 0;
    hipLaunchKernelGGL(k_b_rich_reset, dim3((nseg + 63) / 64), dim3(64), 0, st(), nseg, rich->as<RichSeg>());
    for (int cyc = 0; cyc < p.restart; ++cyc) {
      std::vector<char> active(nseg, 0);
      for (int s = 0; s < nseg; ++s) active[s] = !solved[s];
      // r = b - A x, e_0 = ||r||, V_0 = r / ||r||
      spmv<double, double, double, double>(A64->as<double>(), false, (const double *)x_int->as<double>(), V,
                                           (const double *)b_int->as<double>());
      hipLaunchKernelGGL(k_b_norm2<double>, dim3(G), dim3(kBlock), 0, st(), (const double *)V, partial->as<double>());
      // no rotation (j = -1): e_0 = ||r|| is written where H(j + 1) would go
      hipLaunchKernelGGL(k_b_norm_finish<double>, dim3(nseg), dim3(kBlock), 0, st(),
                         (const double *)partial->as<double>(), (const int32_t *)sb0->as<int32_t>(),
                         (const int32_t *)snb->as<int32_t>(), L.e->as<double>(), (int64_t)(m1 + 1),
                         (double *)nullptr, (double *)nullptr, (int64_t)0, (double *)nullptr, (int64_t)0, -1, 0,
                         L.inv->as<double>(), status->as<double>());
      hipLaunchKernelGGL((k_b_vscale<double, double>), dim3(G), dim3(kBlock), 0, st(), V,
                         (const double *)L.inv->as<double>(), (const int32_t *)bs());
      std::vector<double> e = read_status();
      std::vector<int32_t> jdone(nseg, 0);
      bool any = false;
      for (int s = 0; s < nseg; ++s) {
        if (!active[s]) continue;
        last[s] = e[s];
        ++cycles[s];
        if (rel(e[s], nrm_b[s]) < p.eps) {
          solved[s] = 1;
          active[s] = 0;
        } else {
          any = true;
        }
      }
      if (!any) break;
      int j = 0;
      for (; j < m1; ++j) {
        outer_step(j);
        std::vector<double> st_h = read_status();
        bool left = false;
        for (int s = 0; s < nseg; ++s) {
          if (!active[s]) continue;
          ++iters[s];
          jdone[s] = j + 1;
          last[s] = st_h[s];
          if (rel(st_h[s], nrm_b[s]) < p.eps) {
            solved[s] = 1;
            active[s] = 0;
          } else {
            left = true;
          }
        }
        if (!left) break;
      }
      // x += Z y(j_s) for every segment that stepped in this cycle
      F3R_HIP(hipMemcpyAsync(jseg->p, jdone.data(), sizeof(int32_t) * nseg, hipMemcpyHostToDevice, st()));
      hipLaunchKernelGGL((k_b_combine<double, double, double>), dim3(G), dim3(kBlock), 0, st(),
                         (const double *)L.Zc(N, 0), N, (const double *)Hd, L.hs(), m1 + 1,
                         (const double *)L.e->as<double>(), (int64_t)(m1 + 1), (const int32_t *)jseg->as<int32_t>(),
                         0, (const int32_t *)bs(), x_int->as<double>(), 1, (const float *)nullptr);
      F3R_HIP(hipStreamSynchronize(st()));  // jdone is host memory of this scope
    }
    F3R_HIP(hipGetLastError());
    res.resize(nseg);
    for (int s = 0; s < nseg; ++s) {
      res[s] = f3r_result{};
      res[s].is_solved = solved[s];
      res[s].res_iter = iters[s];
      res[s].res_nrm2 = last[s];
      res[s].nrm_b = nrm_b[s];
      res[s].precond_calls = calls;
      res[s].outer_cycles = cycles[s];
    }
  }
};

BatchSolver *batch_create(std::shared_ptr<Context> ctx, const std::vector<std::shared_ptr<Mat>> &A,
                          const std::vector<std::shared_ptr<Prec>> &P, const f3r_params &p) {
  F3R_CHECK(!A.empty() && A.size() == P.size(), F3R_ERR_ARG, "batch: one matrix and one preconditioner per segment");
  F3R_CHECK(p.m1 >= 1 && p.m1 <= 120 && p.m2 >= 1 && p.m2 <= 16 && p.m3 >= 1 && p.m3 <= 16 && p.m4 >= 1 &&
                p.m4 <= 16 && p.c >= 1 && p.restart >= 1,
            F3R_ERR_ARG, "batch: m1 in [1, 120], m2, m3, m4 in [1, 16], c, restart >= 1");
  // the lockstep nest is the adaptive-weight HALF F3R nest only
  F3R_CHECK(p.variant == 0 && p.static_weight == 0.0 && p.depth == 0, F3R_ERR_ARG,
            "batch: only the HALF F3R nest with adaptive weights (variant 0, static_weight 0, depth 0)");
  // k_sell_mr reads a chunk's header once per wavefront (C >= 64) and a tile
  // holds whole chunks (C <= kTile); K selects a built kernel
  F3R_CHECK(p.C == 0 || (p.C >= 64 && p.C <= kTile && (p.C & (p.C - 1)) == 0), F3R_ERR_ARG,
            "batch: C must be a power of two in [64, 2048] (0 = 64)");
  F3R_CHECK(p.K == 0 || p.K == 1 || p.K == 2 || p.K == 4 || p.K == 8, F3R_ERR_ARG,
            "batch: K must be 1, 2, 4 or 8 (0 = 4)");
  F3R_CHECK(!(p.flags & 16) || (p.index == 0 && p.K != 1), F3R_ERR_ARG,
            "batch: K threads per row (flags bit 4) need the int32-index layout (index 0) and K >= 2");
  auto B = std::make_unique<BatchSolver>();
  B->ctx = ctx;
  B->p = p;
  B->lanes = (p.flags & 16) != 0;
  B->nseg = (int)A.size();
  const int nseg = B->nseg;
  B->off.assign(nseg + 1, 0);
  B->nrow.resize(nseg);
  std::vector<const HostCSR *> pa(nseg), pw(nseg), pz(nseg);
  std::vector<float> sc(nseg), isc(nseg);
  for (int s = 0; s < nseg; ++s) {
    F3R_CHECK(A[s] && A[s]->host && A[s]->format == 0, F3R_ERR_ARG, "batch: CSR matrices");
    const HostCSR &h = *A[s]->host;
    F3R_CHECK(h.n > 0 && h.n == h.ncols, F3R_ERR_ARG, "batch: square non-empty matrices");
    F3R_CHECK(P[s] && P[s]->kind == 0 && P[s]->host.f[0].n == h.n, F3R_ERR_ARG,
              "batch: an SDAINV preconditioner of each segment's matrix");
    B->nrow[s] = h.n;
    B->off[s + 1] = B->off[s] + (h.n + kTile - 1) / kTile * kTile;
    pa[s] = &h;
    pw[s] = &P[s]->host.f[1];
    pz[s] = &P[s]->host.f[0];
    const double rs = p.range_scale > 0 ? p.range_scale
                                        : std::ldexp(1.0, (int)std::floor(std::log2(std::max(1.0, std::sqrt((double)h.n) / 4.0))));
    sc[s] = (float)rs;
    isc[s] = (float)(1.0 / rs);
  }
  const int64_t N = B->off[nseg];
  F3R_CHECK(N / kTile < INT32_MAX, F3R_ERR_RANGE, "batch: too many rows");
  B->N = N;
  B->G = (int)(N / kTile);
  Context *c = ctx.get();
  // block -> segment, segment -> first block and block count
  std::vector<int32_t> bseg(B->G), sb0(nseg), snb(nseg);
  for (int s = 0; s < nseg; ++s) {
    sb0[s] = (int32_t)(B->off[s] / kTile);
    snb[s] = (int32_t)((B->off[s + 1] - B->off[s]) / kTile);
    for (int32_t b = 0; b < snb[s]; ++b) bseg[sb0[s] + b] = s;
  }
  auto up32 = [&](const std::vector<int32_t> &v) {
    DevPtr d = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, v.size()));
    F3R_HIP(hipMemcpyAsync(d->p, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice, c->stream));
    return d;
  };
  auto upf = [&](const std::vector<float> &v) {
    DevPtr d = dev_alloc(sizeof(float) * std::max<size_t>(1, v.size()));
    F3R_HIP(hipMemcpyAsync(d->p, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice, c->stream));
    return d;
  };
  B->bseg = up32(bseg);
  B->sb0 = up32(sb0);
  B->snb = up32(snb);
  B->scale = upf(sc);
  B->iscale = upf(isc);
  F3R_HIP(hipStreamSynchronize(c->stream));
  // the block-diagonal matrices; one row order (sigma = 1) keeps segments contiguous
  const int C = p.C > 0 ? p.C : 64;
  const int K = p.K > 0 ? p.K : 4;
  const bool offsets = p.index != 0;
  // consecutive segments of one matrix (a system's right-hand sides): when
  // every system has the same count NR in {2, 4}, its matrix is laid out
  // once and the SpMVs serve the NR copies together
  {
    std::vector<int> runs;
    for (int s = 0; s < nseg;) {
      int e = s + 1;
      while (e < nseg && A[e].get() == A[s].get()) ++e;
      runs.push_back(e - s);
      s = e;
    }
    const int r0 = runs[0];
    bool same = r0 == 2 || r0 == 4;
    for (int r : runs) same = same && r == r0;
    B->NR = same ? r0 : 1;
  }
  const int NR = B->NR;
  std::vector<const HostCSR *> msys;  // one part per stored matrix
  std::vector<int64_t> moff(1, 0);     // matrix row offsets (whole tiles)
  for (int s = 0; s < nseg; s += NR) {
    msys.push_back(pa[s]);
    moff.push_back(moff.back() + (B->off[s + 1] - B->off[s]));
  }
  const int64_t NM = moff.back();
  // chunk (of C rows) -> shift from matrix row to vector row, copy stride
  auto chunk_tables = [&](DevPtr &dsh, DevPtr &dst) {
    const int64_t nch = (NM + C - 1) / C;
    std::vector<int32_t> sh_(std::max<int64_t>(1, nch)), st_(std::max<int64_t>(1, nch)), sg_(std::max<int64_t>(1, nch));
    for (size_t m = 0; m < msys.size(); ++m) {
      const int64_t seg0 = (int64_t)m * NR;
      for (int64_t ch = moff[m] / C; ch < moff[m + 1] / C; ++ch) {
        sh_[ch] = (int32_t)(B->off[seg0] - moff[m]);
        st_[ch] = (int32_t)(B->off[seg0 + 1] - B->off[seg0]);
        sg_[ch] = (int32_t)seg0;
      }
    }
    dsh = up32(sh_);
    dst = up32(st_);
    B->cseg = up32(sg_);
    F3R_HIP(hipStreamSynchronize(c->stream));  // the host tables are released here
  };
  if (NR > 1) {
    chunk_tables(B->csh, B->cst);
    B->csh16 = B->csh;  // both layouts have chunks of C rows in one order
    B->cst16 = B->cst;
  }
  {
    HostCSR big = NR > 1 ? concat(msys, moff, NM) : concat(pa, B->off, N);
    // fp16 copy: K = 8 once rows average 12 entries (a 27-point row then
    // takes 4 slot groups instead of 7; a 7-point row takes 1 either way)
    const int K16 = p.K > 0 ? p.K : (big.nnz() >= (int64_t)12 * big.n ? 8 : 4);
    HostSell hs = build_sell(big, C, 1, K, nullptr, false, offsets);
    B->sh = upload_sell_shared(c, hs, false);
    B->A64 = upload_sell_values(c, hs, F3R_F64);
    B->A32 = upload_sell_values(c, hs, F3R_F32);
    if (K16 == K) {
      B->sh16 = B->sh;
      B->A16 = upload_sell_values(c, hs, F3R_F16);
    } else {
      HostSell h16 = build_sell(big, C, 1, K16, nullptr, false, offsets);
      B->sh16 = upload_sell_shared(c, h16, false);
      B->A16 = upload_sell_values(c, h16, F3R_F16);
    }
    // k_sell_mr gathers x (fp64 at most) of offset chunks through a buffer
    // resource whose size and byte offsets are 32-bit
    const int64_t bias = std::max(B->sh->off_bias, B->sh16->off_bias);
    F3R_CHECK(NR == 1 || (N + bias) * (int64_t)sizeof(double) < ((int64_t)1 << 31), F3R_ERR_RANGE,
              "batch: the padded rows of all segments exceed the 2^28 a multi-copy SpMV can address");
  }
  // the SDAINV factors: diagonal-only rows as a vector, the rest as CSR rows
  // a system's right-hand sides share its diagonals when they share its
  // preconditioner
  bool share_d = NR > 1;
  for (int s = 0; s < nseg && share_d; ++s) share_d = P[s].get() == P[s - s % NR].get();
  std::vector<int64_t> dshift(nseg, 0);
  if (share_d)
    for (int s = 0; s < nseg; ++s) dshift[s] = moff[s / NR] - B->off[s];
  auto factor = [&](const std::vector<const HostCSR *> &parts) {
    HostCSR M = concat(parts, B->off, N);
    std::vector<double> d(N, 0.0);
    std::vector<int32_t> rows, rp(1, 0), cl;
    std::vector<double> vl;
    for (int32_t i = 0; i < M.n; ++i) {
      const int32_t b = M.rptr[i], e = M.rptr[i + 1];
      if (e - b == 1 && M.idx[b] == i) {
        d[i] = M.val[b];
      } else if (e > b) {
        d[i] = std::numeric_limits<double>::quiet_NaN();  // marks a CSR row
        rows.push_back(i);
        for (int32_t k = b; k < e; ++k) {
          cl.push_back(M.idx[k]);
          vl.push_back(M.val[k]);
        }
        rp.push_back((int32_t)cl.size());
      }
    }
    std::vector<int32_t> rseg(rows.size());
    for (size_t q = 0; q < rows.size(); ++q)
      rseg[q] = (int32_t)(std::upper_bound(B->off.begin(), B->off.end(), (int64_t)rows[q]) - B->off.begin() - 1);
    if (share_d) {  // segment s's diagonals at matrix rows: those of its first copy
      std::vector<double> ds(NM, 0.0);
      for (int s = 0; s < nseg; s += NR)
        std::copy(d.begin() + B->off[s], d.begin() + B->off[s + 1], ds.begin() + moff[s / NR]);
      d.swap(ds);
    }
    Factor16 F;
    F.dval = upload_values(c, d, F3R_F16);
    F.nr = (int32_t)rows.size();
    F.rows = up32(rows);
    F.rseg = up32(rseg);
    F.rptr = up32(rp);
    F.col = up32(cl);
    F.val = upload_values(c, vl, F3R_F16);
    F3R_HIP(hipStreamSynchronize(c->stream));  // the host arrays above are released here
    return F;
  };
  {
    DevPtr d = dev_alloc(sizeof(int64_t) * std::max(1, nseg));
    F3R_HIP(hipMemcpyAsync(d->p, dshift.data(), sizeof(int64_t) * nseg, hipMemcpyHostToDevice, c->stream));
    F3R_HIP(hipStreamSynchronize(c->stream));
    B->dsh = d;
  }
  B->W = factor(pw);
  B->Z = factor(pz);
  F3R_HIP(hipStreamSynchronize(c->stream));
  // work space
  B->alloc_level(B->L1, p.m1);
  B->alloc_level(B->L2, p.m2);
  B->alloc_level(B->L3, p.m3);
  {
    bool pow2 = true;
    for (float v : sc) {
      int e = 0;
      pow2 = pow2 && std::frexp(v, &e) == 0.5f;
    }
    B->zh = NR > 1 && pow2;
    if (B->zh) B->Zh3 = dev_alloc(sizeof(half_t) * N * p.m3);
  }
  auto hv = [&]() { return dev_alloc(sizeof(half_t) * N); };
  B->hin = hv();
  B->hout = hv();
  B->hout2 = hv();
  B->hr = hv();
  B->tmp = hv();
  B->tt = hv();
  B->t2 = hv();
  B->b_int = dev_alloc(sizeof(double) * N);
  B->x_int = dev_alloc(sizeof(double) * N);
  B->partial = dev_alloc(sizeof(double) * (int64_t)B->G * (p.m1 + 1 + 16));
  B->status = dev_alloc(sizeof(double) * nseg);
  B->jseg = dev_alloc(sizeof(int32_t) * nseg);
  B->rich = dev_alloc(sizeof(RichSeg) * nseg);
  F3R_HIP(hipStreamSynchronize(c->stream));
  return B.release();
}

}  // namespace f3r

// ================================================================ C ABI ===
using namespace f3r;

struct f3r_batch_s {
  BatchSolver *b;
};

extern "C" {

int f3r_batch_create(f3r_ctx ctx, int32_t nseg, const f3r_mat *A, const f3r_prec *P, const f3r_params *params,
                     f3r_batch *out) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && A && P && params && out && nseg > 0, F3R_ERR_ARG, "batch_create: null argument");
  *out = nullptr;
  auto c = ctx_of(ctx);
  std::vector<std::shared_ptr<Mat>> am(nseg);
  std::vector<std::shared_ptr<Prec>> pp(nseg);
  for (int s = 0; s < nseg; ++s) {
    F3R_CHECK(A[s] && P[s], F3R_ERR_ARG, "batch_create: null matrix or preconditioner");
    am[s] = mat_of(A[s]);
    pp[s] = prec_of(P[s]);
  }
  BatchSolver *b = batch_create(c, am, pp, *params);
  *out = new f3r_batch_s{b};
  F3R_API_END
}

int f3r_batch_solve(f3r_batch B, const f3r_vec *b, f3r_vec *x, f3r_result *res) {
  F3R_API_BEGIN
  F3R_CHECK(B && b && x && res, F3R_ERR_ARG, "batch_solve: null argument");
  BatchSolver *S = B->b;
  hipStream_t st = S->ctx->stream;
  const int nseg = S->nseg;
  std::vector<double> nrm_b(nseg);
  F3R_HIP(hipMemsetAsync(S->b_int->p, 0, sizeof(double) * S->N, st));
  F3R_HIP(hipMemsetAsync(S->x_int->p, 0, sizeof(double) * S->N, st));
  for (int s = 0; s < nseg; ++s) {
    Vec *bv = vec_of(b[s]), *xv = vec_of(x[s]);
    F3R_CHECK(bv && xv && bv->dtype == F3R_F64 && xv->dtype == F3R_F64 && bv->n == S->nrow[s] && xv->n == S->nrow[s],
              F3R_ERR_ARG, "batch_solve: b and x of each segment: fp64 vectors of its matrix size");
    F3R_HIP(hipMemcpyAsync(S->b_int->as<double>() + S->off[s], bv->ptr(), sizeof(double) * S->nrow[s],
                           hipMemcpyDeviceToDevice, st));
    F3R_HIP(hipMemcpyAsync(S->x_int->as<double>() + S->off[s], xv->ptr(), sizeof(double) * S->nrow[s],
                           hipMemcpyDeviceToDevice, st));
    nrm_b[s] = std::sqrt(dot_raw(S->ctx.get(), F3R_F64, bv->ptr(), F3R_F64, bv->ptr(), S->nrow[s]));
  }
  F3R_HIP(hipStreamSynchronize(st));
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<f3r_result> r;
  S->solve(nrm_b, r);
  F3R_HIP(hipStreamSynchronize(st));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int s = 0; s < nseg; ++s) {
    Vec *xv = vec_of(x[s]);
    F3R_HIP(hipMemcpyAsync(xv->ptr(), S->x_int->as<double>() + S->off[s], sizeof(double) * S->nrow[s],
                           hipMemcpyDeviceToDevice, st));
    r[s].solve_s = secs;
    res[s] = r[s];
  }
  F3R_HIP(hipStreamSynchronize(st));
  F3R_API_END
}

int f3r_batch_info(f3r_batch B, int64_t *rows, int32_t *segments) {
  F3R_API_BEGIN
  F3R_CHECK(B && rows && segments, F3R_ERR_ARG, "batch_info: null argument");
  *rows = B->b->N;
  *segments = B->b->nseg;
  F3R_API_END
}

int f3r_batch_destroy(f3r_batch B) {
  F3R_API_BEGIN
  if (B) {
    (void)hipStreamSynchronize(B->b->ctx->stream);
    delete B->b;
    delete B;
  }
  F3R_API_END
}

}  // extern "C"
