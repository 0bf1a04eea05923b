// Device-side building blocks for CDNA4 (gfx950): 64-wide wavefront
// reductions, LDS block reductions, the single-pass grid reduction with a
// last-block finish, vector loads of SELL slot groups and precision casts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace f3r {

using half_t = _Float16;
using bf16_t = __bf16;

constexpr int kBlock = 256;          // threads per block of the streaming kernels
constexpr int kWave = 64;            // CDNA wavefront
constexpr int kWavesPerBlock = kBlock / kWave;

// ------------------------------------------------------------ precision ---
template <class T> struct Acc { using type = float; };   // fp16 / fp32 data: fp32 sums
template <> struct Acc<double> { using type = double; };

__device__ __forceinline__ float to_acc(half_t v) { return (float)v; }
__device__ __forceinline__ float to_acc(bf16_t v) { return (float)v; }
__device__ __forceinline__ float to_acc(float v) { return v; }
__device__ __forceinline__ double to_acc(double v) { return v; }

template <class T> __device__ __forceinline__ T fma_(T a, T b, T c);
template <> __device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Keeps v as a rounded fp32 value in a register: the compiler may otherwise
// fold the arithmetic that produced v and a following fp32 -> fp16 conversion
// into one mixed-precision FMA that rounds once, straight to fp16.  The
// arithmetic here is "fp32 result, then fp16 store" (two roundings), as the
// oracle computes it.
__device__ __forceinline__ float fp32_rounded(float v) {
  asm volatile("" : "+v"(v));
  return v;
}

// store conversions, round to nearest even
template <class D> __device__ __forceinline__ D cvt(float v);
template <class D> __device__ __forceinline__ D cvt(double v);
template <> __device__ __forceinline__ half_t cvt<half_t>(float v) { return (half_t)fp32_rounded(v); }
template <> __device__ __forceinline__ float cvt<float>(float v) { return v; }
template <> __device__ __forceinline__ double cvt<double>(float v) { return (double)v; }
template <> __device__ __forceinline__ float cvt<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ double cvt<double>(double v) { return v; }
// fp64 -> fp16 correctly rounded, without the double rounding of a detour
// through fp32 (integer arithmetic on the fp64 bit pattern)
__device__ __forceinline__ half_t d2h(double d) {
  uint64_t u = __double_as_longlong(d);
  uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
  uint64_t mag = u & 0x7fffffffffffffffULL;
  uint16_t bits;
  int e = (int)(mag >> 52) - 1023;
  if (mag >= 0x7ff0000000000000ULL) {
    bits = sign | 0x7c00u | (mag > 0x7ff0000000000000ULL ? 0x0200u : 0u);
  } else if ((mag >> 52) == 0 || e < -25) {
    bits = sign;
  } else if (e > 15) {
    bits = sign | 0x7c00u;
  } else {
    uint64_t m = (mag & 0xfffffffffffffULL) | (1ULL << 52);
    int drop = e >= -14 ? 42 : 42 + (-14 - e);
    uint64_t q = m >> drop;
    uint64_t rest = m & ((1ULL << drop) - 1), hw = 1ULL << (drop - 1);
    q += (rest > hw || (rest == hw && (q & 1))) ? 1 : 0;
    if (e >= -14) {
      uint32_t ex = (uint32_t)(e + 15);
      if (q >> 11) {
        q >>= 1;
        ++ex;
      }
      bits = ex >= 31 ? (uint16_t)(sign | 0x7c00u) : (uint16_t)(sign | (ex << 10) | (q & 0x3ffu));
    } else {
      bits = (uint16_t)(sign | q);
    }
  }
  return __builtin_bit_cast(half_t, bits);
}
template <> __device__ __forceinline__ half_t cvt<half_t>(double v) { return d2h(v); }
// fp64 -> bfloat16 bits, correctly rounded (round to nearest even), from the
// fp64 bit pattern: 8 significant bits for normals (exponent >= -126), fewer
// for subnormals (down to 2^-133)
__host__ __device__ inline uint16_t d2bf_bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, 8);
  const uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
  const uint64_t mag = u & 0x7fffffffffffffffULL;
  if (mag >= 0x7ff0000000000000ULL)  // inf / nan
    return sign | 0x7f80u | (mag > 0x7ff0000000000000ULL ? 0x0040u : 0u);
  if ((mag >> 52) == 0) return sign;  // fp64 subnormals: below half the smallest bf16
  const int e = (int)(mag >> 52) - 1023;
  if (e > 127) return sign | 0x7f80u;
  if (e < -134) return sign;
  const uint64_t m = (mag & 0xfffffffffffffULL) | (1ULL << 52);
  const int drop = e >= -126 ? 45 : 45 + (-126 - e);
  uint64_t q = m >> drop;
  const uint64_t rest = m & ((1ULL << drop) - 1), half = 1ULL << (drop - 1);
  q += (rest > half || (rest == half && (q & 1))) ? 1 : 0;
  if (e >= -126) {
    uint32_t ex = (uint32_t)(e + 127);
    if (q >> 8) {  // mantissa overflow into the next binade
      q >>= 1;
      ++ex;
    }
    if (ex >= 255) return sign | 0x7f80u;
    return (uint16_t)(sign | (ex << 7) | (q & 0x7fu));
  }
  return (uint16_t)(sign | q);  // q == 0x80 is the smallest normal
}
// bfloat16: fp32 -> bf16 round to nearest even from the fp32 bit pattern;
// fp64 -> bf16 correctly rounded from the fp64 bits (no double rounding)
template <> __device__ __forceinline__ bf16_t cvt<bf16_t>(float v) {
  uint32_t u = __builtin_bit_cast(uint32_t, fp32_rounded(v));
  uint16_t b;
  if ((u & 0x7fffffffu) > 0x7f800000u)
    b = (uint16_t)((u >> 16) | 0x0040u);  // nan stays nan
  else
    b = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  return __builtin_bit_cast(bf16_t, b);
}
template <> __device__ __forceinline__ bf16_t cvt<bf16_t>(double v) { return __builtin_bit_cast(bf16_t, d2bf_bits(v)); }

// -------------------------------------------------------------- loads ----
// K consecutive elements starting at an address aligned to K elements.
template <class T, int K> struct Pack { T v[K]; };
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Matrix values are streamed with non-temporal loads except for the element
// sizes in F3R_TEMPORAL_SIZES (bit 0: 2 bytes, bit 1: 4, bit 2: 8), read
// with ordinary loads.  fp16 values are reused: the solver's two fp16 SpMVs
// per Arnoldi step read the same 117 MB (128^3), which the 256 MB last-level
// cache then keeps between launches, while the fp32 copy (233 MB, read once
// per level-2 step) streams past without evicting it.  fp64 values stream
// faster with ordinary loads.  Measured at 128^3: 22.6 ms per solve with
// every load non-temporal, 20.8 with none, 20.3 with this choice.
#ifndef F3R_TEMPORAL_SIZES
#define F3R_TEMPORAL_SIZES 5
#endif
template <class T> constexpr bool temporal_values() {
  return (F3R_TEMPORAL_SIZES & (sizeof(T) == 2 ? 1 : sizeof(T) == 4 ? 2 : sizeof(T) == 8 ? 4 : 0)) != 0;
}
template <bool TEMPORAL, class W> __device__ __forceinline__ W vload(const W *p) {
  if constexpr (TEMPORAL) return *p; else return __builtin_nontemporal_load(p);
}
template <class T, int K>
__device__ __forceinline__ Pack<T, K> load_pack(const T *__restrict__ p) {
  Pack<T, K> r;
  if constexpr (K * sizeof(T) == 16) {
    u32x4 w = vload<temporal_values<T>()>(reinterpret_cast<const u32x4 *>(p));
    __builtin_memcpy(&r, &w, 16);
  } else if constexpr (K * sizeof(T) == 8) {
    u32x2 w = vload<temporal_values<T>()>(reinterpret_cast<const u32x2 *>(p));
    __builtin_memcpy(&r, &w, 8);
  } else if constexpr (K * sizeof(T) == 4) {
    uint32_t w = vload<temporal_values<T>()>(reinterpret_cast<const uint32_t *>(p));
    __builtin_memcpy(&r, &w, 4);
  } else if constexpr (K * sizeof(T) == 32) {
    const u32x4 *q = reinterpret_cast<const u32x4 *>(p);
    u32x4 w0 = vload<temporal_values<T>()>(q), w1 = vload<temporal_values<T>()>(q + 1);
    __builtin_memcpy(&r, &w0, 16);
    __builtin_memcpy(reinterpret_cast<char *>(&r) + 16, &w1, 16);
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) r.v[k] = vload<temporal_values<T>()>(p + k);
  }
  return r;
}

// four consecutive elements, loaded and stored as one 8- or 16-byte access
// (p aligned to 4 elements)
template <class T> struct V4 {
  T v[4];
};
template <class T> __device__ __forceinline__ V4<T> ld4(const T *p) {
  V4<T> r;
  __builtin_memcpy(&r, __builtin_assume_aligned(p, 4 * sizeof(T)), sizeof(r));
  return r;
}
template <class T> __device__ __forceinline__ void st4(T *p, const V4<T> &r) {
  __builtin_memcpy(__builtin_assume_aligned(p, 4 * sizeof(T)), &r, sizeof(r));
}

// --------------------------------------------------------- reductions ----
template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sums NV values per thread over the block; thread 0 receives the totals.
template <class T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T *lds /* NV * kWavesPerBlock */) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[i * kWavesPerBlock + wid] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      T s = lds[i * kWavesPerBlock];
#pragma unroll
      for (int w = 1; w < kWavesPerBlock; ++w) s += lds[i * kWavesPerBlock + w];
      v[i] = s;
    }
  }
  __syncthreads();
}

constexpr int kGroup = 64;            // blocks per first-level group of grid_sum
constexpr int kCounterStride = 64;    // 256 B between counters: one cache line each
constexpr int kCounterSlot = 256 * kCounterStride;  // one reduction slot: 1 + 255 groups

// Single-pass grid reduction of NV sums, two-level.  Every block writes its
// NV block sums to partial[i * G + block]; the last block to arrive in each
// group of kGroup blocks adds that group's partials (in block order) into
// partial[NV * G + i * ng + group]; the last group to finish adds the group
// sums (in group order) and returns true in thread 0, with the totals in its v.
// The order of every addition is fixed, so the result does not depend on
// scheduling.  counters[0] counts groups, counters[(1 + g) * kCounterStride]
// the blocks of group g;
// each is reset by the block that reads it last, so no memset is ever needed
// between launches (nor recorded in a graph).  Each counter has a cache line
// of its own: atomics on one line are serialised, and 64 groups sharing four
// lines made the arrivals, not the sums, the cost of the reduction.
//
// Publication across XCDs: the partial sums are agent-scope atomic stores,
// completed (vmcnt 0) before the counter increment, and the block that wins
// a counter reads them with agent-scope atomic loads, issued after the
// increment returned.  Only atomic accesses touch the partials, so neither
// side needs a cache-wide fence: an agent-scope release writes back the
// whole L2 of the XCD and an acquire invalidates it, and one of either per
// block cost more than the reduction itself.
template <class T>
__device__ __forceinline__ void publish(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T collect(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until this wave's stores are complete, and keep the compiler from
// moving memory operations across
__device__ __forceinline__ void stores_complete() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}
__device__ __forceinline__ unsigned arrive(unsigned *c) {
  return __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, int NV>
__device__ __forceinline__ bool grid_sum(T (&v)[NV], T *__restrict__ partial,
                                         unsigned *__restrict__ counters, T *lds) {
  // block totals into thread 0; after the one barrier only wave 0 stays, so
  // a block waiting for its arrival to come back holds one wave slot, not
  // four (the SpMV kernels with fused reductions keep their occupancy)
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[i * kWavesPerBlock + wid] = v[i];
  }
  __syncthreads();
  if (wid != 0) return false;
  const int G = gridDim.x;
  const int ng = (G + kGroup - 1) / kGroup;
  const int grp = blockIdx.x / kGroup;
  const int b0 = grp * kGroup;
  const int gsize = (G - b0) < kGroup ? (G - b0) : kGroup;
  int last = 0;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      T t = lds[i * kWavesPerBlock];
#pragma unroll
      for (int w = 1; w < kWavesPerBlock; ++w) t += lds[i * kWavesPerBlock + w];
      publish(&partial[(int64_t)i * G + blockIdx.x], t);
    }
    stores_complete();
    const unsigned prev = arrive(&counters[(1 + grp) * kCounterStride]);
    last = prev == (unsigned)gsize - 1;
  }
  if (!__builtin_amdgcn_readfirstlane(last)) return false;
  // last block of its group: add the group's partials in block order
  T acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    acc[i] = lane < gsize ? collect(&partial[(int64_t)i * G + b0 + lane]) : (T)0;
    acc[i] = wave_sum(acc[i]);
  }
  T *gpart = partial + (int64_t)NV * G;
  last = 0;
  if (lane == 0) {
    __hip_atomic_store(&counters[(1 + grp) * kCounterStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int i = 0; i < NV; ++i) publish(&gpart[(int64_t)i * ng + grp], acc[i]);
    stores_complete();
    const unsigned prev = arrive(&counters[0]);
    last = prev == (unsigned)ng - 1;
  }
  if (!__builtin_amdgcn_readfirstlane(last)) return false;
  // last group: add the group sums in group order
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    T t = 0;
    for (int g = lane; g < ng; g += kWave) t += collect(&gpart[(int64_t)i * ng + g]);
    acc[i] = wave_sum(t);
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = acc[i];
    __hip_atomic_store(&counters[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return lane == 0;
}

// largest grid grid_sum supports (one counter per group in a slot)
constexpr int kMaxReduceGrid = (kCounterSlot / kCounterStride - 1) * kGroup;

// Grid size of the reduction kernels: enough blocks to stream from HBM,
// capped so that the last block's pass over the partials stays short.
inline int reduce_grid(int64_t n, int num_cu) {
  int64_t want = (n + kBlock * 8 - 1) / (kBlock * 8);
  int64_t cap = (int64_t)num_cu * 4;
  if (cap > 1024) cap = 1024;
  if (want > cap) want = cap;
  return want < 1 ? 1 : (int)want;
}

}  // namespace f3r
