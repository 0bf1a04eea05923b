// Learned correction of SDAINV for the innermost fp16 level.
//
// A per-row MLP, 12 -> 64 -> 64 -> 1 with ReLU, reads local features of the
// scaled matrix and of the SDAINV factors and outputs y_i; the corrected
// preconditioner is M = Z diag(s) W^T with s_i = exp(clamp(y_i)), i.e.
// SDAINV's values rescaled column by column on the same sparsity pattern
// (ainv.hpp:57-64 builds Z = (Z^t D^-1)^T; this scales D^-1 per row).  With
// the last layer at zero, s = 1 and M is plain SDAINV.
//
// Training minimises L = sum_p ||(I - A M) z_p||^2 / ||z_p||^2 over probe
// vectors z_p that the solver feeds to its innermost level.  Forward and
// backward are hand-written kernels: the dense layers run on bf16 MFMA
// (v_mfma_f32_16x16x16_bf16, v_mfma_f32_16x16x32_bf16) with fp32
// accumulation, the weight gradients are MFMA products over rows, reduced
// over blocks in a fixed order; Adam keeps fp32 master weights.
#include <cmath>
#include <cstring>
#include <random>

#include "common.hpp"
#include "internal.hpp"
#include "sell_kernel.hpp"

namespace f3r {

constexpr int kNin = 16;   // 12 features + 4 zero columns (MFMA K = 16)
constexpr int kFeat = 12;
constexpr int kH = 64;
// parameter vector layout (fp32): W1 [64][16], b1 [64], W2 [64][64], b2 [64], w3 [64], b3
constexpr int kOffW1 = 0, kOffB1 = kH * kNin, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH, kNParams = kOffB3 + 1;
constexpr float kYClamp = 3.0f;

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

struct MLPDev {
  const uint16_t *w1;  // bf16 [64][16]
  const uint16_t *w2;  // bf16 [64][64]
  const float *p;      // fp32 master parameters (biases, w3, b3 used in fp32)
};

// Shared state of one block: weights and the per-wave activation tiles.
struct MLPShared {  // 16-byte aligned rows: the MFMA operands are read as 8- and 16-byte vectors
  alignas(16) uint16_t w1[kH * kNin];
  alignas(16) uint16_t w2[kH * kH];
  alignas(16) uint16_t w2t[kH * kH];   // transposed, for dh1 = dz2 W2
  alignas(16) uint16_t h1[4][16 * kH];  // per wave: 16 rows x 64 units (bf16)
  alignas(16) uint16_t dz[4][16 * kH];  // per wave: dz2 tile (backward)
  float b1[kH], b2[kH], w3[kH];
  float b3;
};

__device__ void load_mlp(const MLPDev &m, MLPShared &sm, bool transpose) {
  for (int t = threadIdx.x; t < kH * kNin; t += blockDim.x) sm.w1[t] = m.w1[t];
  for (int t = threadIdx.x; t < kH * kH; t += blockDim.x) {
    sm.w2[t] = m.w2[t];
    if (transpose) sm.w2t[(t % kH) * kH + t / kH] = m.w2[t];
  }
  for (int t = threadIdx.x; t < kH; t += blockDim.x) {
    sm.b1[t] = m.p[kOffB1 + t];
    sm.b2[t] = m.p[kOffB2 + t];
    sm.w3[t] = m.p[kOffW3 + t];
  }
  if (threadIdx.x == 0) sm.b3 = m.p[kOffB3];
  __syncthreads();
}

// Forward of the 16 rows [r0, r0+16) of one wave.  MFMA operand layouts
// (lane l, i = l % 16, g = l / 16):
//   16x16x16: A[i][4g..4g+3], B[4g..4g+3][i], D[4g..4g+3][i]
//   16x16x32: A[i][8g..8g+7], B[8g..8g+7][i], D[4g..4g+3][i]
// Returns the layer-2 pre-activations acc[nt][q] (row 4g+q, unit 16nt+i) and
// leaves relu(layer 1) in sm.h1[wave] as bf16.
__device__ __forceinline__ void mlp_rows(const MLPShared &sm, uint16_t *h1t, const uint16_t *__restrict__ feat,
                                         int64_t n, int64_t r0, f32x4 (&acc)[4], s16x4 &xa) {
  const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
  // layer 1: X [16 x 16] * W1^T [16 x 64]
  const int64_t r = r0 + i;
  xa = s16x4{0, 0, 0, 0};
  if (r < n) {
    const uint64_t w = *reinterpret_cast<const uint64_t *>(feat + r * kNin + 4 * g);
    __builtin_memcpy(&xa, &w, 8);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    s16x4 b;
    __builtin_memcpy(&b, &sm.w1[(nt * 16 + i) * kNin + 4 * g], 8);
    f32x4 c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xa, b, f32x4{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float h = fmaxf(c[q] + sm.b1[nt * 16 + i], 0.f);
      h1t[(4 * g + q) * kH + nt * 16 + i] = f2bf(h);
    }
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS writes done before the wave reads them
  __builtin_amdgcn_wave_barrier();
  // layer 2: H1 [16 x 64] * W2^T [64 x 64]
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0, 0, 0, 0};
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    bf16x8 a;
    __builtin_memcpy(&a, &h1t[i * kH + kb * 32 + 8 * g], 16);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      bf16x8 b;
      __builtin_memcpy(&b, &sm.w2[(nt * 16 + i) * kH + kb * 32 + 8 * g], 16);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[nt][q] += sm.b2[nt * 16 + i];
}

// y[r] = MLP(features[r]) for every row
__global__ __launch_bounds__(256) void k_mlp_forward(MLPDev m, const uint16_t *__restrict__ feat, int64_t n,
                                                     float *__restrict__ y) {
  __shared__ MLPShared sm;
  load_mlp(m, sm, false);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
  const int64_t ntiles = (n + 63) / 64;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r0 = t * 64 + wave * 16;
    f32x4 acc[4];
    s16x4 xa;
    mlp_rows(sm, sm.h1[wave], feat, n, r0, acc, xa);
    float part[4] = {0, 0, 0, 0};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int q = 0; q < 4; ++q) part[q] = __builtin_fmaf(fmaxf(acc[nt][q], 0.f), sm.w3[nt * 16 + i], part[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) part[q] += __shfl_xor(part[q], off, 64);
    }
    if (i == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t r = r0 + 4 * g + q;
        if (r < n) y[r] = part[q] + sm.b3;
      }
    }
    __syncthreads();
  }
}

// Backward: given gy[r] = dL/dy_r, every wave accumulates its share of the
// parameter gradient and writes it to part[(block * 4 + wave) * kNParams + .];
// k_grad_reduce then adds the slices in a fixed order.
__global__ __launch_bounds__(256) void k_mlp_backward(MLPDev m, const uint16_t *__restrict__ feat, int64_t n,
                                                      const float *__restrict__ gy, float *__restrict__ part) {
  __shared__ MLPShared sm;
  load_mlp(m, sm, true);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
  f32x4 gw2[4][4];  // [mt][it]: gW2[16mt + 4g + q][16it + i]
  f32x4 gw1[4];     // [mt]: gW1[16mt + 4g + q][i]
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gw1[a] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < 4; ++b) gw2[a][b] = f32x4{0, 0, 0, 0};
  }
  uint16_t *h1t = sm.h1[wave], *dzt = sm.dz[wave];
  const int64_t ntiles = (n + 63) / 64;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r0 = t * 64 + wave * 16;
    f32x4 acc[4];
    s16x4 xa;
    mlp_rows(sm, h1t, feat, n, r0, acc, xa);
    float grow[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = r0 + 4 * g + q;
      grow[q] = r < n ? gy[r] : 0.f;
    }
    if (i == 0) gb3 += grow[0] + grow[1] + grow[2] + grow[3];
    // dz2 = g * w3 * relu'(z2); gw3 += g * h2; gb2 += dz2
    f32x4 dz2[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const float w = sm.w3[nt * 16 + i];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float z2 = acc[nt][q];
        gw3[nt] = __builtin_fmaf(grow[q], fmaxf(z2, 0.f), gw3[nt]);
        const float d = z2 > 0.f ? grow[q] * w : 0.f;
        dz2[nt][q] = d;
        gb2[nt] += d;
        dzt[(4 * g + q) * kH + nt * 16 + i] = f2bf(d);
      }
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    // gW2[o][k] += sum_rows dz2[r][o] h1[r][k]: A = dz2^T (16 out x 16 rows), B = h1 (16 rows x 16 in)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      s16x4 a;
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = (short)f2bf(dz2[mt][q]);
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        s16x4 b;
#pragma unroll
        for (int q = 0; q < 4; ++q) b[q] = (short)h1t[(4 * g + q) * kH + it * 16 + i];
        gw2[mt][it] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, gw2[mt][it], 0, 0, 0);
      }
    }
    // dh1 = (dz2 W2) * relu'(z1): A = dz2 (16 rows x 64 out), B = W2 (64 out x 64 in)
    f32x4 dh1[4];
#pragma unroll
    for (int it = 0; it < 4; ++it) dh1[it] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      bf16x8 a;
      __builtin_memcpy(&a, &dzt[i * kH + kb * 32 + 8 * g], 16);
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        bf16x8 b;
        __builtin_memcpy(&b, &sm.w2t[(it * 16 + i) * kH + kb * 32 + 8 * g], 16);
        dh1[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, dh1[it], 0, 0, 0);
      }
    }
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float h = bf2f(h1t[(4 * g + q) * kH + it * 16 + i]);
        if (!(h > 0.f)) dh1[it][q] = 0.f;
        gb1[it] += dh1[it][q];
      }
    // gW1[o][k] += sum_rows dh1[r][o] x[r][k]: A = dh1^T (16 out x 16 rows), B = X (16 rows x 16 feat)
    s16x4 bx;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = r0 + 4 * g + q;
      bx[q] = r < n ? (short)feat[r * kNin + i] : (short)0;
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      s16x4 a;
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = (short)f2bf(dh1[mt][q]);
      gw1[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx, gw1[mt], 0, 0, 0);
    }
    __syncthreads();
  }
  // reduce the per-lane bias / w3 partials over the four lane groups g
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    gb1[nt] += __shfl_xor(gb1[nt], 16, 64);
    gb1[nt] += __shfl_xor(gb1[nt], 32, 64);
    gb2[nt] += __shfl_xor(gb2[nt], 16, 64);
    gb2[nt] += __shfl_xor(gb2[nt], 32, 64);
    gw3[nt] += __shfl_xor(gw3[nt], 16, 64);
    gw3[nt] += __shfl_xor(gw3[nt], 32, 64);
  }
  gb3 += __shfl_xor(gb3, 16, 64);
  gb3 += __shfl_xor(gb3, 32, 64);
  float *out = part + (int64_t)(blockIdx.x * 4 + wave) * kNParams;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int o = mt * 16 + 4 * g + q;
      out[kOffW1 + o * kNin + i] = gw1[mt][q];
#pragma unroll
      for (int it = 0; it < 4; ++it) out[kOffW2 + o * kH + it * 16 + i] = gw2[mt][it][q];
    }
  }
  if (g == 0) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      out[kOffB1 + nt * 16 + i] = gb1[nt];
      out[kOffB2 + nt * 16 + i] = gb2[nt];
      out[kOffW3 + nt * 16 + i] = gw3[nt];
    }
    if (i == 0) out[kOffB3] = gb3;
  }
}

// grad[k] += sum over slices s of part[s * kNParams + k], slices in order
__global__ void k_grad_reduce(const float *__restrict__ part, int nslices, float *__restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNParams) return;
  float s = 0.f;
  for (int j = 0; j < nslices; ++j) s += part[(int64_t)j * kNParams + k];
  grad[k] += s;
}

// Adam on the fp32 master parameters, then refresh the bf16 copies
__global__ void k_adam(float *__restrict__ p, float *__restrict__ mom, float *__restrict__ vel,
                       const float *__restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                       float c2, uint16_t *__restrict__ w1b, uint16_t *__restrict__ w2b) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNParams) return;
  const float gk = grad[k];
  const float mk = b1 * mom[k] + (1.f - b1) * gk;
  const float vk = b2 * vel[k] + (1.f - b2) * gk * gk;
  mom[k] = mk;
  vel[k] = vk;
  float pk = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  if (k >= kOffW1 && k < kOffB1 && (k % kNin) >= kFeat) pk = 0.f;  // padded feature columns
  p[k] = pk;
  if (k < kOffB1) w1b[k] = f2bf(pk);
  if (k >= kOffW2 && k < kOffB2) w2b[k - kOffW2] = f2bf(pk);
}

__global__ void k_refresh_bf16(const float *__restrict__ p, uint16_t *__restrict__ w1b, uint16_t *__restrict__ w2b) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kH * kNin) w1b[k] = f2bf(p[kOffW1 + k]);
  if (k < kH * kH) w2b[k] = f2bf(p[kOffW2 + k]);
}

// s = exp(clamp(y)); v = s * u
__global__ void k_scale_from_y(int64_t n, const float *__restrict__ y, float *__restrict__ s) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s[i] = expf(fminf(fmaxf(y[i], -kYClamp), kYClamp));
}
__global__ void k_mul(int64_t n, const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ c) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    c[i] = a[i] * b[i];
}
// r = z - q and ||r||^2, ||z||^2 (single-pass grid reduction)
__global__ __launch_bounds__(kBlock) void k_resid_norms(int64_t n, const float *__restrict__ z,
                                                        const float *__restrict__ q, float *__restrict__ r,
                                                        float *partial, unsigned *counters, float *out) {
  __shared__ float lds[2 * kWavesPerBlock];
  float v[2] = {0.f, 0.f};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float ri = z[i] - q[i];
    r[i] = ri;
    v[0] = __builtin_fmaf(ri, ri, v[0]);
    v[1] = __builtin_fmaf(z[i], z[i], v[1]);
  }
  if (grid_sum<float, 2>(v, partial, counters, lds) && threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
  }
}
// <z, q>, <q, q>, <z, z> (single-pass grid reduction)
__global__ __launch_bounds__(kBlock) void k_qdots(int64_t n, const float *__restrict__ z,
                                                  const float *__restrict__ q, float *partial, unsigned *counters,
                                                  float *out) {
  __shared__ float lds[3 * kWavesPerBlock];
  float v[3] = {0.f, 0.f, 0.f};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    v[0] = __builtin_fmaf(z[i], q[i], v[0]);
    v[1] = __builtin_fmaf(q[i], q[i], v[1]);
    v[2] = __builtin_fmaf(z[i], z[i], v[2]);
  }
  if (grid_sum<float, 3>(v, partial, counters, lds) && threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
  }
}
// r = <q,q> z - <z,q> q: minus the q-gradient of 1 - <z,q>^2 / (<z,z><q,q>),
// up to the factor 2 <z,q> / (<z,z> <q,q>^2)
__global__ void k_rscaled(int64_t n, const float *__restrict__ z, const float *__restrict__ q,
                          const float *__restrict__ d, float *__restrict__ r) {
  const float a = d[0], b = d[1];
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    r[i] = b * z[i] - a * q[i];
}
// gy += c * g2 * u * s, zero where the clamp is active; c = -2 wgt / ||z||^2
// (kind 0) or -2 wgt <z,q> / (<z,z> <q,q>^2) (kind 1)
__global__ void k_gy(int64_t n, const float *__restrict__ g2, const float *__restrict__ u,
                     const float *__restrict__ s, const float *__restrict__ y, const float *__restrict__ norms,
                     int kind, float wgt, float *__restrict__ gy) {
  const float c = kind == 0 ? -2.f * wgt / norms[1] : -2.f * wgt * norms[0] / (norms[2] * norms[1] * norms[1]);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const bool live = fabsf(y[i]) < kYClamp;
    gy[i] += live ? c * g2[i] * u[i] * s[i] : 0.f;
  }
}

}  // namespace f3r

// ================================================================= host ===
namespace f3r {

struct MLP {
  std::shared_ptr<Context> ctx;
  DevPtr p, mom, vel, w1b, w2b, part, grad;
  int grid = 256;
  int64_t step = 0;
  MLPDev dev() const { return MLPDev{w1b->as<uint16_t>(), w2b->as<uint16_t>(), p->as<float>()}; }
};

// One training system: A, A^T, Z (already holding D^-1), Z^T, W^T as fp32
// SELL matrices, the bf16 features and the probe vectors.
struct LSys {
  std::shared_ptr<Context> ctx;
  int32_t n = 0;
  struct M {
    std::shared_ptr<SellShared> sh;
    DevPtr val;
  } A, At, Z, Zt, Wt;
  DevPtr feat, y, s, u, v, pz, q, r, g1, g2, gy, norms, probes;
  int32_t nprobes = 0;
};

static LSys::M upload_f32(Context *c, const HostCSR &M) {
  HostSell hs = build_sell(M, 64, 1, 1, nullptr, false, true);
  LSys::M m;
  m.sh = upload_sell_shared(c, hs, false);
  m.val = upload_sell_values(c, hs, F3R_F32);
  return m;
}

static void spmv32(const LSys::M &m, const float *x, float *y, hipStream_t st) {
  SellArgs a = sell_args(*m.sh, false);
  launch_sell<float, float, float>(m.sh->K, a, m.val->as<float>(), x, EpiStore<float, float>{y, nullptr}, st);
}

static HostCSR transpose_csr(const HostCSR &A) {
  HostCSR T;
  T.n = T.ncols = A.n;
  T.rptr.assign(A.n + 1, 0);
  for (int32_t k : A.idx) T.rptr[k + 1]++;
  for (int32_t i = 0; i < A.n; ++i) T.rptr[i + 1] += T.rptr[i];
  T.idx.resize(A.nnz());
  T.val.resize(A.nnz());
  std::vector<int32_t> fill(T.rptr.begin(), T.rptr.end() - 1);
  for (int32_t i = 0; i < A.n; ++i)
    for (int32_t k = A.rptr[i]; k < A.rptr[i + 1]; ++k) {
      int32_t p = fill[A.idx[k]]++;
      T.idx[p] = i;
      T.val[p] = A.val[k];
    }
  return T;
}

// Twelve features of row i (the rest of the 16 columns zero):
//  0  log|a_ii| of the unscaled matrix (0 when not given)
//  1-3 the three entries of the scaled row nearest below the diagonal in
//      column order, farthest first, 0-padded in front when fewer
//  4-6 the three nearest above it, nearest first, 0-padded behind
//      (an interior 7-point row: -z, -y, -x and +x, +y, +z)
//  7  sum |a_ij|, j != i     8  sum a_ij (row sum)
//  9  Z_ii = 1/d_i (SDAINV pivot)   10 sum_{j!=i} |W^T_ij|   11 sum_{j!=i} |Z_ij|
static void row_features(const HostCSR &A, const HostFactors &F, const double *diag, float *out) {
  const HostCSR &Z = F.f[0], &W = F.f[1];
  for (int32_t i = 0; i < A.n; ++i) {
    float *f = out + (int64_t)i * kNin;
    for (int k = 0; k < kNin; ++k) f[k] = 0.f;
    f[0] = diag ? (float)std::log(std::max(std::fabs(diag[i]), 1e-300)) : 0.f;
    int32_t dpos = -1;
    double absoff = 0, rowsum = 0;
    for (int32_t k = A.rptr[i]; k < A.rptr[i + 1]; ++k) {
      rowsum += A.val[k];
      if (A.idx[k] == i) dpos = k;
      else absoff += std::fabs(A.val[k]);
    }
    int32_t lo = A.rptr[i], hi = A.rptr[i + 1];
    int32_t split = dpos >= 0 ? dpos : lo;
    if (dpos < 0)
      while (split < hi && A.idx[split] < i) ++split;
    for (int t = 0; t < 3; ++t) {
      const int32_t kl = split - 3 + t;  // below: the three nearest, farthest first
      if (kl >= lo && A.idx[kl] < i) f[1 + t] = (float)A.val[kl];
      const int32_t ku = (dpos >= 0 ? dpos + 1 : split) + t;
      if (ku < hi && A.idx[ku] > i) f[4 + t] = (float)A.val[ku];
    }
    f[7] = (float)absoff;
    f[8] = (float)rowsum;
    double zii = 0, zfill = 0, wfill = 0;
    for (int32_t k = Z.rptr[i]; k < Z.rptr[i + 1]; ++k) {
      if (Z.idx[k] == i) zii = Z.val[k];
      else zfill += std::fabs(Z.val[k]);
    }
    for (int32_t k = W.rptr[i]; k < W.rptr[i + 1]; ++k)
      if (W.idx[k] != i) wfill += std::fabs(W.val[k]);
    f[9] = (float)zii;
    f[10] = (float)wfill;
    f[11] = (float)zfill;
  }
}

static uint16_t f2bf_host(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

MLP *mlp_create(std::shared_ptr<Context> ctx, uint64_t seed) {
  auto m = std::make_unique<MLP>();
  m->ctx = ctx;
  std::vector<float> p(kNParams, 0.f);
  std::mt19937_64 eng(seed);
  // He-uniform hidden layers; the last layer starts at zero, so the corrected
  // preconditioner starts as plain SDAINV
  std::uniform_real_distribution<float> u1(-std::sqrt(6.f / kFeat), std::sqrt(6.f / kFeat));
  std::uniform_real_distribution<float> u2(-std::sqrt(6.f / kH), std::sqrt(6.f / kH));
  for (int o = 0; o < kH; ++o)
    for (int k = 0; k < kFeat; ++k) p[kOffW1 + o * kNin + k] = u1(eng);
  for (int k = 0; k < kH * kH; ++k) p[kOffW2 + k] = u2(eng);
  for (int k = 0; k < kH; ++k) {
    p[kOffB1 + k] = 0.01f;
    p[kOffB2 + k] = 0.01f;
  }
  m->p = dev_alloc(4 * kNParams);
  m->mom = dev_alloc(4 * kNParams);
  m->vel = dev_alloc(4 * kNParams);
  m->grad = dev_alloc(4 * kNParams);
  m->w1b = dev_alloc(2 * kH * kNin);
  m->w2b = dev_alloc(2 * kH * kH);
  m->part = dev_alloc((size_t)4 * kNParams * m->grid * 4);
  hipStream_t st = ctx->stream;
  F3R_HIP(hipMemcpyAsync(m->p->p, p.data(), 4 * kNParams, hipMemcpyHostToDevice, st));
  F3R_HIP(hipMemsetAsync(m->mom->p, 0, 4 * kNParams, st));
  F3R_HIP(hipMemsetAsync(m->vel->p, 0, 4 * kNParams, st));
  F3R_HIP(hipMemsetAsync(m->grad->p, 0, 4 * kNParams, st));
  hipLaunchKernelGGL(k_refresh_bf16, dim3(16), dim3(256), 0, st, (const float *)m->p->as<float>(),
                     m->w1b->as<uint16_t>(), m->w2b->as<uint16_t>());
  F3R_HIP(hipStreamSynchronize(st));
  return m.release();
}

static int fwd_grid(Context *c, int64_t n) {
  int64_t t = (n + 63) / 64;
  return (int)std::max<int64_t>(1, std::min<int64_t>(t, (int64_t)c->num_cu * 4));
}

void mlp_forward_dev(MLP *m, const uint16_t *feat, int64_t n, float *y, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_forward, dim3(fwd_grid(m->ctx.get(), n)), dim3(256), 0, st, m->dev(), feat, n, y);
}

void mlp_backward_dev(MLP *m, const uint16_t *feat, int64_t n, const float *gy, hipStream_t st) {
  const int grid = std::min<int>(m->grid, (int)std::max<int64_t>(1, (n + 63) / 64));
  hipLaunchKernelGGL(k_mlp_backward, dim3(grid), dim3(256), 0, st, m->dev(), feat, n, gy, m->part->as<float>());
  hipLaunchKernelGGL(k_grad_reduce, dim3((kNParams + 255) / 256), dim3(256), 0, st,
                     (const float *)m->part->as<float>(), grid * 4, m->grad->as<float>());
}

LSys *lsys_create(std::shared_ptr<Context> ctx, const HostCSR &A, const HostFactors &F, const float *feat) {
  auto L = std::make_unique<LSys>();
  L->ctx = ctx;
  L->n = A.n;
  Context *c = ctx.get();
  L->A = upload_f32(c, A);
  L->At = upload_f32(c, transpose_csr(A));
  L->Z = upload_f32(c, F.f[0]);
  L->Zt = upload_f32(c, transpose_csr(F.f[0]));
  L->Wt = upload_f32(c, F.f[1]);
  const int64_t n = A.n;
  std::vector<uint16_t> fb((size_t)n * kNin);
  for (size_t k = 0; k < fb.size(); ++k) fb[k] = f2bf_host(feat[k]);
  L->feat = dev_alloc(2 * fb.size());
  F3R_HIP(hipMemcpyAsync(L->feat->p, fb.data(), 2 * fb.size(), hipMemcpyHostToDevice, c->stream));
  for (DevPtr *v : {&L->y, &L->s, &L->u, &L->v, &L->pz, &L->q, &L->r, &L->g1, &L->g2, &L->gy}) *v = dev_alloc(4 * n);
  F3R_HIP(hipStreamSynchronize(c->stream));
  return L.release();
}

// Mean over the probes z of, with q = A Z diag(s) W^T z,
//   kind 0: ||z - q||^2 / ||z||^2
//   kind 1: min_w ||z - w q||^2 / ||z||^2 = 1 - <z,q>^2 / (<z,z> <q,q>)
// (kind 1 is blind to a global scale of M, which the adaptive Richardson
// weights absorb anyway).  With `grad` the gradient of the mean is added to
// the MLP's gradient buffer.
double lsys_loss(LSys *L, MLP *m, int kind, bool grad) {
  Context *c = L->ctx.get();
  hipStream_t st = c->stream;
  const int64_t n = L->n;
  const dim3 g((unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, (int64_t)c->num_cu * 16)), b(kBlock);
  const uint16_t *feat = L->feat->as<uint16_t>();
  float *y = L->y->as<float>(), *s = L->s->as<float>(), *u = L->u->as<float>(), *v = L->v->as<float>();
  float *pz = L->pz->as<float>(), *q = L->q->as<float>(), *r = L->r->as<float>();
  float *g1 = L->g1->as<float>(), *g2 = L->g2->as<float>(), *gy = L->gy->as<float>();
  float *norms = L->norms->as<float>();
  mlp_forward_dev(m, feat, n, y, st);
  hipLaunchKernelGGL(k_scale_from_y, g, b, 0, st, n, (const float *)y, s);
  if (grad) F3R_HIP(hipMemsetAsync(gy, 0, 4 * n, st));
  const int rgrid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, kMaxReduceGrid);
  for (int k = 0; k < L->nprobes; ++k) {
    const float *z = L->probes->as<float>() + (int64_t)k * n;
    float *nk = norms + 4 * k;
    spmv32(L->Wt, z, u, st);
    hipLaunchKernelGGL(k_mul, g, b, 0, st, n, (const float *)s, (const float *)u, v);
    spmv32(L->Z, v, pz, st);
    spmv32(L->A, pz, q, st);
    unsigned *cnt = c->red.counters->as<unsigned>() + 20 * kCounterSlot;
    if (kind == 0) {
      hipLaunchKernelGGL(k_resid_norms, dim3(rgrid), b, 0, st, n, z, (const float *)q, r,
                         c->red.partial->as<float>(), cnt, nk);
    } else {
      hipLaunchKernelGGL(k_qdots, dim3(rgrid), b, 0, st, n, z, (const float *)q, c->red.partial->as<float>(), cnt,
                         nk);
      if (grad) hipLaunchKernelGGL(k_rscaled, g, b, 0, st, n, z, (const float *)q, (const float *)nk, r);
    }
    if (grad) {
      spmv32(L->At, r, g1, st);
      spmv32(L->Zt, g1, g2, st);
      hipLaunchKernelGGL(k_gy, g, b, 0, st, n, (const float *)g2, (const float *)u, (const float *)s,
                         (const float *)y, (const float *)nk, kind, 1.f / (float)L->nprobes, gy);
    }
  }
  if (grad && L->nprobes > 0) mlp_backward_dev(m, feat, n, gy, st);
  F3R_HIP(hipGetLastError());
  std::vector<float> h(4 * (size_t)L->nprobes);
  F3R_HIP(hipMemcpyAsync(h.data(), norms, 4 * h.size(), hipMemcpyDeviceToHost, st));
  F3R_HIP(hipStreamSynchronize(st));
  double total = 0;
  for (int k = 0; k < L->nprobes; ++k) {
    const float *e = &h[4 * (size_t)k];
    total += kind == 0 ? (double)e[0] / (double)e[1]
                       : 1.0 - (double)e[0] * (double)e[0] / ((double)e[2] * (double)e[1]);
  }
  return L->nprobes > 0 ? total / L->nprobes : 0.0;
}

}  // namespace f3r

// ================================================================ C ABI ===
using namespace f3r;

struct f3r_mlp_s {
  std::unique_ptr<MLP> m;
};
struct f3r_lsys_s {
  std::unique_ptr<LSys> L;
};

namespace {

void sync(Context *c) { F3R_HIP(hipStreamSynchronize(c->stream)); }

// host fp32 feature rows (n x 16) -> device bf16
DevPtr upload_features(Context *c, int64_t n, const float *feat) {
  std::vector<uint16_t> fb((size_t)n * kNin);
  for (size_t k = 0; k < fb.size(); ++k) fb[k] = f2bf_host(feat[k]);
  DevPtr d = dev_alloc(2 * std::max<size_t>(fb.size(), 1));
  F3R_HIP(hipMemcpyAsync(d->p, fb.data(), 2 * fb.size(), hipMemcpyHostToDevice, c->stream));
  sync(c);
  return d;
}

void check_features(const float *feat, int64_t n) {
  for (int64_t i = 0; i < n * kNin; ++i)
    F3R_CHECK(std::isfinite(feat[i]), F3R_ERR_ARG, "learn: features must be finite");
}

}  // namespace

extern "C" {

int f3r_mlp_nparams(int64_t *n) {
  F3R_API_BEGIN
  F3R_CHECK(n, F3R_ERR_ARG, "null argument");
  *n = kNParams;
  F3R_API_END
}

int f3r_mlp_create(f3r_ctx ctx, uint64_t seed, f3r_mlp *out) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && out, F3R_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<MLP> m(mlp_create(ctx_of(ctx), seed));
  *out = new f3r_mlp_s{std::move(m)};
  F3R_API_END
}

int f3r_mlp_destroy(f3r_mlp m) {
  F3R_API_BEGIN
  if (m) {
    (void)hipStreamSynchronize(m->m->ctx->stream);
    delete m;
  }
  F3R_API_END
}

int f3r_mlp_get(f3r_mlp m, float *params) {
  F3R_API_BEGIN
  F3R_CHECK(m && params, F3R_ERR_ARG, "null argument");
  Context *c = m->m->ctx.get();
  F3R_HIP(hipMemcpyAsync(params, m->m->p->p, 4 * kNParams, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  F3R_API_END
}

int f3r_mlp_set(f3r_mlp m, const float *params) {
  F3R_API_BEGIN
  F3R_CHECK(m && params, F3R_ERR_ARG, "null argument");
  for (int k = 0; k < kNParams; ++k) F3R_CHECK(std::isfinite(params[k]), F3R_ERR_ARG, "mlp_set: non-finite");
  MLP *M = m->m.get();
  Context *c = M->ctx.get();
  std::vector<float> p(params, params + kNParams);
  for (int o = 0; o < kH; ++o)
    for (int k = kFeat; k < kNin; ++k) p[kOffW1 + o * kNin + k] = 0.f;
  F3R_HIP(hipMemcpyAsync(M->p->p, p.data(), 4 * kNParams, hipMemcpyHostToDevice, c->stream));
  F3R_HIP(hipMemsetAsync(M->mom->p, 0, 4 * kNParams, c->stream));
  F3R_HIP(hipMemsetAsync(M->vel->p, 0, 4 * kNParams, c->stream));
  M->step = 0;
  hipLaunchKernelGGL(k_refresh_bf16, dim3(16), dim3(256), 0, c->stream, (const float *)M->p->as<float>(),
                     M->w1b->as<uint16_t>(), M->w2b->as<uint16_t>());
  F3R_HIP(hipGetLastError());
  sync(c);
  F3R_API_END
}

int f3r_mlp_forward(f3r_mlp m, int64_t n, const float *feat, float *y) {
  F3R_API_BEGIN
  F3R_CHECK(m && feat && y && n > 0, F3R_ERR_ARG, "bad argument");
  check_features(feat, n);
  Context *c = m->m->ctx.get();
  DevPtr f = upload_features(c, n, feat);
  DevPtr dy = dev_alloc(4 * n);
  mlp_forward_dev(m->m.get(), f->as<uint16_t>(), n, dy->as<float>(), c->stream);
  F3R_HIP(hipGetLastError());
  F3R_HIP(hipMemcpyAsync(y, dy->p, 4 * n, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  F3R_API_END
}

int f3r_mlp_grad_zero(f3r_mlp m) {
  F3R_API_BEGIN
  F3R_CHECK(m, F3R_ERR_ARG, "null argument");
  Context *c = m->m->ctx.get();
  F3R_HIP(hipMemsetAsync(m->m->grad->p, 0, 4 * kNParams, c->stream));
  sync(c);
  F3R_API_END
}

int f3r_mlp_backward(f3r_mlp m, int64_t n, const float *feat, const float *gy) {
  F3R_API_BEGIN
  F3R_CHECK(m && feat && gy && n > 0, F3R_ERR_ARG, "bad argument");
  check_features(feat, n);
  Context *c = m->m->ctx.get();
  DevPtr f = upload_features(c, n, feat);
  DevPtr dg = dev_alloc(4 * n);
  F3R_HIP(hipMemcpyAsync(dg->p, gy, 4 * n, hipMemcpyHostToDevice, c->stream));
  mlp_backward_dev(m->m.get(), f->as<uint16_t>(), n, dg->as<float>(), c->stream);
  F3R_HIP(hipGetLastError());
  sync(c);
  F3R_API_END
}

int f3r_mlp_grad_get(f3r_mlp m, float *grad) {
  F3R_API_BEGIN
  F3R_CHECK(m && grad, F3R_ERR_ARG, "null argument");
  Context *c = m->m->ctx.get();
  F3R_HIP(hipMemcpyAsync(grad, m->m->grad->p, 4 * kNParams, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  F3R_API_END
}

int f3r_mlp_adam(f3r_mlp m, float lr, float beta1, float beta2, float eps) {
  F3R_API_BEGIN
  F3R_CHECK(m && lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, F3R_ERR_ARG,
            "mlp_adam: bad hyper-parameters");
  MLP *M = m->m.get();
  Context *c = M->ctx.get();
  M->step += 1;
  const float c1 = 1.f - std::pow(beta1, (float)M->step), c2 = 1.f - std::pow(beta2, (float)M->step);
  hipLaunchKernelGGL(k_adam, dim3((kNParams + 255) / 256), dim3(256), 0, c->stream, M->p->as<float>(),
                     M->mom->as<float>(), M->vel->as<float>(), (const float *)M->grad->as<float>(), lr, beta1, beta2,
                     eps, c1, c2, M->w1b->as<uint16_t>(), M->w2b->as<uint16_t>());
  F3R_HIP(hipGetLastError());
  sync(c);
  F3R_API_END
}

int f3r_learn_features(f3r_mat A, f3r_prec P, const double *diag, float *feat) {
  F3R_API_BEGIN
  F3R_CHECK(A && P && feat, F3R_ERR_ARG, "null argument");
  auto a = mat_of(A);
  auto p = prec_of(P);
  F3R_CHECK(a->host, F3R_ERR_ARG, "learn_features: needs the CSR matrix");
  F3R_CHECK(p->kind == 0, F3R_ERR_ARG, "learn_features: needs an SDAINV preconditioner");
  F3R_CHECK(p->host.f[0].n == a->host->n, F3R_ERR_ARG, "learn_features: sizes differ");
  row_features(*a->host, p->host, diag, feat);
  F3R_API_END
}

int f3r_lsys_create(f3r_ctx ctx, f3r_mat A, f3r_prec P, const float *feat, f3r_lsys *out) {
  F3R_API_BEGIN
  F3R_CHECK(ctx && A && P && feat && out, F3R_ERR_ARG, "null argument");
  *out = nullptr;
  auto a = mat_of(A);
  auto p = prec_of(P);
  F3R_CHECK(a->host && p->kind == 0 && p->host.f[0].n == a->host->n, F3R_ERR_ARG,
            "lsys_create: needs the CSR matrix and its SDAINV preconditioner");
  check_features(feat, a->host->n);
  std::unique_ptr<LSys> L(lsys_create(ctx_of(ctx), *a->host, p->host, feat));
  *out = new f3r_lsys_s{std::move(L)};
  F3R_API_END
}

int f3r_lsys_destroy(f3r_lsys L) {
  F3R_API_BEGIN
  if (L) {
    (void)hipStreamSynchronize(L->L->ctx->stream);
    delete L;
  }
  F3R_API_END
}

int f3r_lsys_set_probes(f3r_lsys L, int32_t k, const float *probes) {
  F3R_API_BEGIN
  F3R_CHECK(L && k >= 0 && (k == 0 || probes), F3R_ERR_ARG, "bad argument");
  LSys *S = L->L.get();
  Context *c = S->ctx.get();
  const int64_t n = S->n;
  for (int32_t j = 0; j < k; ++j) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) {
      F3R_CHECK(std::isfinite(probes[j * n + i]), F3R_ERR_ARG, "lsys_set_probes: non-finite");
      s += (double)probes[j * n + i] * probes[j * n + i];
    }
    F3R_CHECK(s > 0, F3R_ERR_ARG, "lsys_set_probes: zero probe vector");
  }
  S->probes = dev_alloc(4 * std::max<int64_t>(n * k, 1));
  if (k) F3R_HIP(hipMemcpyAsync(S->probes->p, probes, 4 * n * k, hipMemcpyHostToDevice, c->stream));
  S->norms = dev_alloc(16 * std::max<int32_t>(k, 1));
  S->nprobes = k;
  sync(c);
  F3R_API_END
}

int f3r_lsys_loss(f3r_lsys L, f3r_mlp m, int32_t kind, int32_t grad, double *loss) {
  F3R_API_BEGIN
  F3R_CHECK(L && m && loss, F3R_ERR_ARG, "null argument");
  F3R_CHECK(kind == 0 || kind == 1, F3R_ERR_ARG, "lsys_loss: kind must be 0 or 1");
  F3R_CHECK(L->L->ctx == m->m->ctx, F3R_ERR_ARG, "lsys_loss: different contexts");
  F3R_CHECK(L->L->nprobes > 0, F3R_ERR_STATE, "lsys_loss: no probe vectors");
  *loss = lsys_loss(L->L.get(), m->m.get(), kind, grad != 0);
  F3R_API_END
}

int f3r_lsys_scales(f3r_lsys L, f3r_mlp m, int32_t normalize, float *s) {
  F3R_API_BEGIN
  F3R_CHECK(L && m && s, F3R_ERR_ARG, "null argument");
  F3R_CHECK(L->L->ctx == m->m->ctx, F3R_ERR_ARG, "lsys_scales: different contexts");
  LSys *S = L->L.get();
  Context *c = S->ctx.get();
  const int64_t n = S->n;
  const dim3 g((unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, (int64_t)c->num_cu * 16)), b(kBlock);
  mlp_forward_dev(m->m.get(), S->feat->as<uint16_t>(), n, S->y->as<float>(), c->stream);
  hipLaunchKernelGGL(k_scale_from_y, g, b, 0, c->stream, n, (const float *)S->y->as<float>(), S->s->as<float>());
  F3R_HIP(hipGetLastError());
  F3R_HIP(hipMemcpyAsync(s, S->s->p, 4 * n, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  if (normalize) {  // divide by the geometric mean: the loss of kind 1 cannot see a global scale
    double lm = 0;
    for (int64_t i = 0; i < n; ++i) lm += std::log((double)s[i]);
    const double g = std::exp(lm / (double)n);
    for (int64_t i = 0; i < n; ++i) s[i] = (float)((double)s[i] / g);
  }
  F3R_API_END
}

int f3r_sdainv_rescale(f3r_ctx ctx, f3r_prec P, const float *s, f3r_prec *out) {
  F3R_API_BEGIN
  F3R_CHECK(P && s && out, F3R_ERR_ARG, "null argument");
  *out = nullptr;
  auto p = prec_of(P);
  F3R_CHECK(p->kind == 0, F3R_ERR_ARG, "sdainv_rescale: needs an SDAINV preconditioner");
  const int32_t n = p->n;
  for (int32_t i = 0; i < n; ++i)
    F3R_CHECK(std::isfinite(s[i]) && s[i] > 0, F3R_ERR_ARG, "sdainv_rescale: scales must be positive");
  auto q = std::make_shared<Prec>();
  q->ctx = ctx ? ctx_of(ctx) : nullptr;
  q->kind = 0;
  q->dtype = p->dtype;
  q->tol = p->tol;
  q->acc = p->acc;
  q->n = n;
  q->host = p->host;
  HostCSR &Z = q->host.f[0];
  for (int32_t i = 0; i < Z.n; ++i)
    for (int32_t k = Z.rptr[i]; k < Z.rptr[i + 1]; ++k) Z.val[k] *= (double)s[Z.idx[k]];
  if (q->ctx) prec_build_device(q.get(), nullptr, 64, 1);
  *out = wrap_prec(q);
  F3R_API_END
}

}  // extern "C"
