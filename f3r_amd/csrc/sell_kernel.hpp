// SELL-C-sigma SpMV on CDNA4: one row per lane, chunks of C rows, K slots of
// a row interleaved so that each lane issues one K-wide vector load of values
// and one of column indices per slot group; a wavefront of 64 consecutive
// rows therefore reads 64*K contiguous values and indices per step.  The row
// sum is a fused multiply-add chain in slot order (= the CSR order of the
// row, as SELL::operate sums it, sparse/matrix/sell.hpp:63-69).  What happens
// with the sum is the epilogue's business: plain store, residual, or the
// fused Krylov / Richardson updates of the solver.
#pragma once

#include "common.hpp"
#include "device.hpp"

// groups of K slots of an offset chunk issued back to back (memory-level
// parallelism of one lane)
#ifndef F3R_OFF_UNROLL
#define F3R_OFF_UNROLL 2
#endif
#define F3R_PRAGMA_(x) _Pragma(#x)
#define F3R_PRAGMA(x) F3R_PRAGMA_(x)
// x gathers of offset chunks through a buffer resource: the per-lane byte
// offset r * sizeof(X) is fixed for the row and the chunk's column offset
// rides in the scalar offset, so a gather is one instruction without
// per-lane address arithmetic
#ifndef F3R_X_BUFFER
#define F3R_X_BUFFER 1
#endif

namespace f3r {

struct SellArgs {
  int32_t n;        // rows
  int32_t log2C;    // chunk height C = 1 << log2C
  const int32_t *__restrict__ cptr;    // element offset of each chunk
  const int32_t *__restrict__ cwidth;  // slot groups of each chunk (| kHalfTail: and a half group)
  const int32_t *__restrict__ idx;     // column of each slot
  const int32_t *__restrict__ perm;    // sorted row -> original row (null = identity)
  // chunk-offset format (null = every chunk explicit): iptr[s] >= 0 is the
  // start of chunk s in idx; iptr[s] < 0: slot t of row r reads column
  // r + coff[-iptr[s] - 1 + t], and the chunk has no column indices.  Used
  // only with C >= 64, so that a wavefront's rows share one chunk.
  const int32_t *__restrict__ iptr;
  const int32_t *__restrict__ coff;
  int32_t off_bias;  // -(smallest column offset), >= 0: makes every scalar offset non-negative
  // threads per row: 1, or K (int32-index chunks only; launch_sell then runs
  // k_sell_lanes, whose sums are reassociated)
  int32_t lanes = 1;
};

// word 3 of a raw buffer resource on gfx9-family GPUs (no format, no swizzle)
constexpr int kBufferWord3 = 0x00020000;

template <class X>
__device__ __forceinline__ X buffer_gather(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  if constexpr (sizeof(X) == 2) {
    return __builtin_bit_cast(X, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(rs, voff, soff, 0));
  } else if constexpr (sizeof(X) == 4) {
    return __builtin_bit_cast(X, (unsigned)__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
  } else {
    return __builtin_bit_cast(X, (unsigned long long)__builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
  }
}

#ifndef F3R_FIXED_ROWS
#define F3R_FIXED_ROWS 1
#endif
// Row of an offset chunk of G full groups and, if HALF, a half group, with
// every value load and x gather issued before the first multiply-add (a
// runtime loop keeps one or two groups in flight).  Same sum, same order.
template <int G, bool HALF, class V, class X, class ACC, int K>
__device__ __forceinline__ ACC sell_row_fixed(const V *__restrict__ val, int64_t pos, int64_t step, int64_t hpos,
                                              __amdgpu_buffer_rsrc_t rs, int vo, int bias,
                                              const __attribute__((address_space(4))) int32_t *off) {
  constexpr int H = K / 2;
  Pack<V, K> pv[G];
  X xv[G][K];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    pv[g] = load_pack<V, K>(val + pos + g * step);
#pragma unroll
    for (int k = 0; k < K; ++k) xv[g][k] = buffer_gather<X>(rs, vo, (off[g * K + k] + bias) * (int)sizeof(X));
  }
  Pack<V, (HALF ? H : 1)> ph{};
  X xh[HALF ? H : 1];
  if constexpr (HALF) {
    ph = load_pack<V, H>(val + hpos);
#pragma unroll
    for (int k = 0; k < H; ++k) xh[k] = buffer_gather<X>(rs, vo, (off[G * K + k] + bias) * (int)sizeof(X));
  }
  ACC acc = 0;
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int k = 0; k < K; ++k) acc = fma_<ACC>((ACC)to_acc(pv[g].v[k]), (ACC)to_acc(xv[g][k]), acc);
  if constexpr (HALF) {
#pragma unroll
    for (int k = 0; k < H; ++k) acc = fma_<ACC>((ACC)to_acc(ph.v[k]), (ACC)to_acc(xh[k]), acc);
  }
  return acc;
}

// Sum of row r (in sorted order) of A*x with accumulator ACC.
template <class V, class X, class ACC, int K>
__device__ __forceinline__ ACC sell_row(const SellArgs &a, const V *__restrict__ val,
                                        const X *__restrict__ x, int r) {
  const int s = r >> a.log2C;
  const int j = r - (s << a.log2C);
  const int C = 1 << a.log2C;
  const int rem = a.n - (s << a.log2C);
  const int tc = rem < C ? rem : C;
  const int cw = a.cwidth[s];
  const int groups = chunk_groups(cw);
  int64_t pos = (int64_t)a.cptr[s] + (int64_t)j * K;
  const int64_t step = (int64_t)tc * K;
  ACC acc = 0;
  const int32_t *__restrict__ ip = a.idx + pos;
  if (a.iptr) {
    // the chunk index is uniform over the wavefront: scalar loads of the
    // chunk header and of the offset table
    const int su = __builtin_amdgcn_readfirstlane(s);
    const int32_t start = a.iptr[su];
    if (start < 0) {
      // the offset table is read-only for the whole kernel: constant
      // address space, so the uniform reads become scalar loads
      typedef const __attribute__((address_space(4))) int32_t *const_i32;
      const_i32 off = (const_i32)(a.coff + (-start - 1));
#if F3R_X_BUFFER
      // base x - off_bias, so that row + offset = (r) + (offset + off_bias) - off_bias
      // with a non-negative scalar part; out-of-range reads would return 0
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void *)((uintptr_t)x - (uintptr_t)a.off_bias * sizeof(X)), (short)0,
          (int)(((int64_t)a.n + a.off_bias) * (int64_t)sizeof(X)), kBufferWord3);
      const int vo = r * (int)sizeof(X);
      const int bias = a.off_bias;
#if F3R_FIXED_ROWS
      {
        // the widths of 27- and 7-point stencil chunks (uniform per wavefront)
        const int cwu = __builtin_amdgcn_readfirstlane(cw);
        const int64_t hpos = (int64_t)a.cptr[s] + (int64_t)groups * step + (int64_t)j * (K / 2);
        if constexpr (K == 8) {
          if (cwu == (3 | kHalfTail)) return sell_row_fixed<3, true, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
          if (cwu == 4) return sell_row_fixed<4, false, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
          if (cwu == 1) return sell_row_fixed<1, false, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
        } else if constexpr (K == 4) {
          if (cwu == 7) return sell_row_fixed<7, false, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
          if (cwu == 2) return sell_row_fixed<2, false, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
        } else if constexpr (K == 2) {
          if (cwu == (13 | kHalfTail)) return sell_row_fixed<13, true, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
          if (cwu == 14) return sell_row_fixed<14, false, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
          if (cwu == (3 | kHalfTail)) return sell_row_fixed<3, true, V, X, ACC, K>(val, pos, step, hpos, rs, vo, bias, off);
        }
      }
#endif
      F3R_PRAGMA(unroll F3R_OFF_UNROLL)
      for (int g = 0; g < groups; ++g, pos += step) {
        const Pack<V, K> pv = load_pack<V, K>(val + pos);
        X xv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) xv[k] = buffer_gather<X>(rs, vo, (off[g * K + k] + bias) * (int)sizeof(X));
#pragma unroll
        for (int k = 0; k < K; ++k) acc = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xv[k]), acc);
      }
      if constexpr (K >= 2) {
        if (cw & kHalfTail) {  // the half group: K / 2 slots per row, rows consecutive
          constexpr int H = K / 2;
          const Pack<V, H> pv = load_pack<V, H>(val + ((int64_t)a.cptr[s] + (int64_t)groups * step + (int64_t)j * H));
          X xv[H];
#pragma unroll
          for (int k = 0; k < H; ++k) xv[k] = buffer_gather<X>(rs, vo, (off[groups * K + k] + bias) * (int)sizeof(X));
#pragma unroll
          for (int k = 0; k < H; ++k) acc = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xv[k]), acc);
        }
      }
#else
      const X *__restrict__ xr = x + r;
      F3R_PRAGMA(unroll F3R_OFF_UNROLL)
      for (int g = 0; g < groups; ++g, pos += step) {
        const Pack<V, K> pv = load_pack<V, K>(val + pos);
        X xv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) xv[k] = xr[off[g * K + k]];
#pragma unroll
        for (int k = 0; k < K; ++k) acc = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xv[k]), acc);
      }
      if constexpr (K >= 2) {
        if (cw & kHalfTail) {
          constexpr int H = K / 2;
          const Pack<V, H> pv = load_pack<V, H>(val + ((int64_t)a.cptr[s] + (int64_t)groups * step + (int64_t)j * H));
#pragma unroll
          for (int k = 0; k < H; ++k)
            acc = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xr[off[groups * K + k]]), acc);
        }
      }
#endif
      return acc;
    }
    ip = a.idx + (int64_t)start + (int64_t)j * K;
  }
  for (int g = 0; g < groups; ++g, pos += step, ip += step) {
    const Pack<V, K> pv = load_pack<V, K>(val + pos);
    const Pack<int32_t, K> pi = load_pack<int32_t, K>(ip);
#pragma unroll
    for (int k = 0; k < K; ++k) acc = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(x[pi.v[k]]), acc);
  }
  return acc;
}

// An epilogue provides
//   struct State;                                   per-thread carry (may be empty)
//   struct Pre;                                     per-row operands loaded early
//   __device__ Pre pre(int r);                      issued before the row's SpMV,
//                                                   so its loads overlap the gathers
//   __device__ void row(int r, ACC acc, const Pre &p, State &s);  for every row r < n
//   __device__ void finish(State &s);               called once by every thread
// finish() may reduce over the block and the grid.
struct NoState {};
struct NoPre {};

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin by
// blockIdx.  Renumbering them so that XCD k takes the k-th contiguous eighth
// of the blocks keeps the rows one XCD works on, and so the x entries it
// gathers, together in that XCD's L2 instead of every XCD fetching all of x.
constexpr int kXcds = 8;
#ifndef F3R_XCD_MAP
#define F3R_XCD_MAP 1
#endif
__device__ __forceinline__ int xcd_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if (!F3R_XCD_MAP || g % kXcds) return b;
  return (b % kXcds) * (g / kXcds) + b / kXcds;
}

template <class V, class X, class ACC, int K, class Epi>
__global__ __launch_bounds__(kBlock) void k_sell(SellArgs a, const V *__restrict__ val,
                                                 const X *__restrict__ x, Epi epi) {
  typename Epi::State st{};
  for (int r = xcd_block() * kBlock + threadIdx.x; r < a.n; r += gridDim.x * kBlock) {
    const typename Epi::Pre p = epi.pre(r);
    epi.row(r, sell_row<V, X, ACC, K>(a, val, x, r), p, st);
  }
  epi.finish(st);
}

// SELL SpMV with K lanes per row (int32-index chunks only): lane k of row r
// sums slot k of each of the row's groups, K contiguous values and indices
// per group of the K lanes (coalesced, as the one-lane-per-row kernel reads
// them), then the K lanes add their sums by butterfly shuffles.  The sum is
// reassociated: not the CSR order.  For rows much longer than K slots.
template <class V, class X, class ACC, int K, class Epi>
__global__ __launch_bounds__(kBlock) void k_sell_lanes(SellArgs a, const V *__restrict__ val,
                                                       const X *__restrict__ x, Epi epi) {
  static_assert(K >= 2 && kWave % K == 0, "K lanes of a row lie in one wavefront");
  constexpr int kRowsPerBlock = kBlock / K;
  const int k = threadIdx.x % K;
  typename Epi::State st{};
  for (int r0 = xcd_block() * kRowsPerBlock; r0 < a.n; r0 += gridDim.x * kRowsPerBlock) {
    const int r = r0 + threadIdx.x / K;
    ACC acc = 0;
    if (r < a.n) {
      const int s = r >> a.log2C;
      const int j = r - (s << a.log2C);
      const int C = 1 << a.log2C;
      const int rem = a.n - (s << a.log2C);
      const int tc = rem < C ? rem : C;
      const int groups = chunk_groups(a.cwidth[s]);
      int64_t pos = (int64_t)a.cptr[s] + (int64_t)j * K + k;
      const int64_t step = (int64_t)tc * K;
      for (int g = 0; g < groups; ++g, pos += step)
        acc = fma_<ACC>((ACC)to_acc(val[pos]), (ACC)to_acc(x[a.idx[pos]]), acc);
    }
#pragma unroll
    for (int off = K / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (r < a.n && k == 0) {
      const typename Epi::Pre p = epi.pre(r);
      epi.row(r, acc, p, st);
    }
  }
  epi.finish(st);
}

// Row r of an int32-index SELL with x staged in LDS: xs[c - lo] holds x[c]
// for the block's column window [lo, lo + len).  The slots and the sum are
// those of sell_row (same bits); only the source of x differs.
template <class V, class X, class ACC, int K>
__device__ __forceinline__ ACC sell_row_lds(const SellArgs &a, const V *__restrict__ val, const X *xs, int lo, int r) {
  const int s = r >> a.log2C;
  const int j = r - (s << a.log2C);
  const int C = 1 << a.log2C;
  const int rem = a.n - (s << a.log2C);
  const int tc = rem < C ? rem : C;
  const int groups = chunk_groups(a.cwidth[s]);
  int64_t pos = (int64_t)a.cptr[s] + (int64_t)j * K;
  const int64_t step = (int64_t)tc * K;
  const int32_t *__restrict__ ip = a.idx + pos;
  ACC acc = 0;
  for (int g = 0; g < groups; ++g, pos += step, ip += step) {
    const Pack<V, K> pv = load_pack<V, K>(val + pos);
    const Pack<int32_t, K> pi = load_pack<int32_t, K>(ip);
#pragma unroll
    for (int k = 0; k < K; ++k) acc = fma_<ACC>((ACC)to_acc(pv.v[k]), (ACC)to_acc(xs[pi.v[k] - lo]), acc);
  }
  return acc;
}

// SELL SpMV with LDS-staged x gathers (int32-index chunks only): one row per
// thread; a block whose column window fits the LDS (bwin[b] = {lo, len},
// len <= cap) first copies x[lo, lo + len) into LDS with coalesced loads, then
// gathers from LDS; a wider block gathers from global memory as k_sell does.
// Worth it when the blocks' windows are narrow (banded matrices); the
// autotuner decides.
template <class V, class X, class ACC, int K, class Epi>
__global__ __launch_bounds__(kBlock) void k_sell_lds(SellArgs a, const V *__restrict__ val, const X *__restrict__ x,
                                                     const int2 *__restrict__ bwin, int cap, Epi epi) {
  extern __shared__ unsigned char sell_lds_raw[];
  X *xs = reinterpret_cast<X *>(sell_lds_raw);
  const int b = xcd_block();
  const int2 win = bwin[b];
  const bool staged = win.y <= cap;
  if (staged) {
    for (int i = threadIdx.x; i < win.y; i += kBlock) xs[i] = x[win.x + i];
    __syncthreads();
  }
  typename Epi::State st{};
  const int r = b * kBlock + threadIdx.x;
  if (r < a.n) {
    const typename Epi::Pre p = epi.pre(r);
    epi.row(r, staged ? sell_row_lds<V, X, ACC, K>(a, val, xs, win.x, r) : sell_row<V, X, ACC, K>(a, val, x, r), p,
            st);
  }
  epi.finish(st);
}

template <class V, class X, class ACC, class Epi>
inline void launch_sell_lds(int K, const SellArgs &a, const V *val, const X *x, const int2 *bwin, int cap, Epi epi,
                            hipStream_t st) {
  if (a.n == 0) return;
  const dim3 g((a.n + kBlock - 1) / kBlock);
  const size_t lds = (size_t)cap * sizeof(X);
  switch (K) {
    case 1: hipLaunchKernelGGL((k_sell_lds<V, X, ACC, 1, Epi>), g, dim3(kBlock), lds, st, a, val, x, bwin, cap, epi); break;
    case 2: hipLaunchKernelGGL((k_sell_lds<V, X, ACC, 2, Epi>), g, dim3(kBlock), lds, st, a, val, x, bwin, cap, epi); break;
    case 4: hipLaunchKernelGGL((k_sell_lds<V, X, ACC, 4, Epi>), g, dim3(kBlock), lds, st, a, val, x, bwin, cap, epi); break;
    case 8: hipLaunchKernelGGL((k_sell_lds<V, X, ACC, 8, Epi>), g, dim3(kBlock), lds, st, a, val, x, bwin, cap, epi); break;
    default: break;
  }
}

// blocks launch_sell runs for `grid` (0 = one row per thread, or per K
// threads with K threads per row; otherwise that block count when smaller)
inline int sell_blocks(const SellArgs &a, int grid) {
  const int64_t full = ((int64_t)a.n * (a.lanes > 1 ? a.lanes : 1) + kBlock - 1) / kBlock;
  return (int)(grid > 0 && grid < full ? grid : full);
}

// grid: 0 = one row per thread; otherwise the given (capped) block count
template <class V, class X, class ACC, class Epi>
inline void launch_sell(int K, const SellArgs &a, const V *val, const X *x, Epi epi,
                        hipStream_t st, int grid = 0) {
  if (a.n == 0) return;
  const dim3 g(sell_blocks(a, grid));
  if (a.lanes > 1) {
    switch (K) {
      case 2: hipLaunchKernelGGL((k_sell_lanes<V, X, ACC, 2, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
      case 4: hipLaunchKernelGGL((k_sell_lanes<V, X, ACC, 4, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
      case 8: hipLaunchKernelGGL((k_sell_lanes<V, X, ACC, 8, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
      default: break;
    }
    return;
  }
  switch (K) {
    case 1: hipLaunchKernelGGL((k_sell<V, X, ACC, 1, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
    case 2: hipLaunchKernelGGL((k_sell<V, X, ACC, 2, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
    case 4: hipLaunchKernelGGL((k_sell<V, X, ACC, 4, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
    case 8: hipLaunchKernelGGL((k_sell<V, X, ACC, 8, Epi>), g, dim3(kBlock), 0, st, a, val, x, epi); break;
    default: break;
  }
}

// ------------------------------------------------------------ epilogues ---
template <class Y, class ACC> struct EpiStore {
  using State = NoState;
  using Pre = NoPre;
  Y *__restrict__ y;
  const int32_t *__restrict__ perm;
  __device__ Pre pre(int) const { return {}; }
  __device__ void row(int r, ACC acc, const Pre &, State &) const {
    const int row = perm ? perm[r] : r;
    y[row] = cvt<Y>(acc);
  }
  __device__ void finish(State &) const {}
};

template <class Y, class B, class ACC> struct EpiResid {
  using State = NoState;
  struct Pre {
    B b;
  };
  Y *__restrict__ y;
  const B *__restrict__ b;
  const int32_t *__restrict__ perm;
  __device__ Pre pre(int r) const { return {b[perm ? perm[r] : r]}; }
  __device__ void row(int r, ACC acc, const Pre &p, State &) const {
    const int row = perm ? perm[r] : r;
    y[row] = cvt<Y>((ACC)to_acc(p.b) - acc);
  }
  __device__ void finish(State &) const {}
};

}  // namespace f3r
