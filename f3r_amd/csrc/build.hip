// Device build of the SELL-C-sigma layout from a CSR matrix already on the
// device, producing the same arrays as the host builder (build_sell): the
// row permutation, chunk pointers and widths, and for the chunk-offset
// format the per-chunk offset tables, the compacted int32 indices of the
// other chunks, and the fp64 values in slot order.
//
//   sigma > 1 (one block per window of sigma rows): the rows of the window
//          by descending length, ties in row order (bitonic sort in LDS of
//          (length, row) keys: the host's stable sort); with renumbering,
//          the inverse permutation
//   pass 1 (one block per chunk): the widest row; for the offset format the
//          sorted union of col - row over the chunk (bitonic sort in LDS) and
//          the eligibility test of the host builder
//   host:  prefix sums over the chunks (cptr, iptr, table offsets)
//   pass 2 (one block per chunk): the offset tables
//   pass 3 (one thread per row): values and indices into their slots
//
// A chunk whose entries do not fit the LDS sort buffer is stored explicitly,
// and a sigma window wider than kSigmaCap rows does not fit its sort; the
// host builder has no such limits, so the build refuses the device path
// when either happens and the caller builds on the host instead.
#include <algorithm>
#include <climits>
#include <numeric>

#include "common.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace f3r {

constexpr int kSortCap = 4096;   // entries of one chunk sorted in LDS (16 KB)
constexpr int kSigmaCap = 4096;  // rows of one sigma window sorted in LDS (32 KB)

struct ChunkInfo {
  int32_t width;   // slots of the widest row, or |union| for an offset chunk
  int32_t offset;  // 1: offset chunk
  int32_t ulen;    // |union|
  int32_t umin;    // smallest offset (offset chunks)
  int32_t toobig;  // 1: the chunk's entries exceed kSortCap (host must build)
};

__device__ __forceinline__ int block_scan_excl(int v, int *lds, int &total) {
  // exclusive prefix sum over the block (blockDim.x <= 1024), fixed order
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    const int a = t >= off ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += a;
    __syncthreads();
  }
  total = lds[blockDim.x - 1];
  const int incl = lds[t];
  __syncthreads();
  return incl - v;
}

// perm[w0 + i] over the window of rows [w0, w0 + sigma): its rows by
// descending length, ties in row order, as sigma_perm sorts them on the host
// (csr_to_sell_sigma, sell.hpp:314-327)
__global__ __launch_bounds__(1024) void k_sigma_perm(int32_t n, int32_t sigma, const int32_t *__restrict__ rptr,
                                                     int32_t *__restrict__ perm) {
  __shared__ unsigned long long keys[kSigmaCap];
  const int64_t w0 = (int64_t)blockIdx.x * sigma;
  const int len = (int)min<int64_t>(sigma, n - w0);
  int m = 1;
  while (m < len) m <<= 1;
  // key = (2^32 - 1 - length, position in the window): ascending keys are
  // descending lengths, ties by position; pad keys sort last
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    unsigned long long k = ~0ull;
    if (i < len) {
      const int64_t r = w0 + i;
      const unsigned rl = (unsigned)(rptr[r + 1] - rptr[r]);
      k = ((unsigned long long)(0xFFFFFFFFu - rl) << 32) | (unsigned)i;
    }
    keys[i] = k;
  }
  __syncthreads();
  for (int k = 2; k <= m; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const int l = i ^ jj;
        if (l > i) {
          const unsigned long long a = keys[i], b = keys[l];
          if ((a > b) == ((i & k) == 0)) {
            keys[i] = b;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < len; i += blockDim.x) perm[w0 + i] = (int32_t)(w0 + (int64_t)(keys[i] & 0xFFFFFFFFu));
}

__global__ void k_inverse_perm(int32_t n, const int32_t *__restrict__ perm, int32_t *__restrict__ inv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) inv[perm[r]] = r;
}

// perm (null: identity): row j of chunk s is row perm[s * C + j] of the CSR;
// inv (null: none): columns renumbered into the sorted order
__global__ __launch_bounds__(256) void k_chunk_scan(int32_t n, int32_t C, const int32_t *__restrict__ rptr,
                                                    const int32_t *__restrict__ col, const int32_t *__restrict__ perm,
                                                    const int32_t *__restrict__ inv, int offsets,
                                                    ChunkInfo *__restrict__ info, int32_t *__restrict__ ubuf) {
  __shared__ int keys[kSortCap];
  __shared__ int scan[256];
  __shared__ int s_mx, s_bad;
  const int s = blockIdx.x;
  const int64_t r0 = (int64_t)s * C;
  const int tc = (int)min<int64_t>(C, n - r0);
  if (threadIdx.x == 0) {
    s_mx = 0;
    s_bad = 0;
  }
  __syncthreads();
  // rows of this thread: j = threadIdx.x + k * 256 (C <= 256 in practice)
  int cnt = 0;
  for (int j = threadIdx.x; j < tc; j += blockDim.x) {
    const int r = (int)(r0 + j);
    const int o = perm ? perm[r] : r;
    const int len = rptr[o + 1] - rptr[o];
    atomicMax(&s_mx, len);
    cnt += len;
  }
  int total = 0;
  const int base = block_scan_excl(cnt, scan, total);
  const int mx = s_mx;
  ChunkInfo ci{mx, 0, 0, 0, 0};
  if (offsets && mx > 0 && total > kSortCap) ci.toobig = 1;
  if (offsets && mx > 0 && total <= kSortCap) {
    // (col - row) of every entry, and the column order of every row
    int p = base;
    for (int j = threadIdx.x; j < tc; j += blockDim.x) {
      const int r = (int)(r0 + j);
      const int o = perm ? perm[r] : r;
      int prev = INT_MIN;
      for (int k = rptr[o]; k < rptr[o + 1]; ++k) {
        const int c = inv ? inv[col[k]] : col[k];
        if (c <= prev) s_bad = 1;
        prev = c;
        keys[p++] = c - r;
      }
    }
    int m = 1;
    while (m < total) m <<= 1;
    for (int i = total + threadIdx.x; i < m; i += blockDim.x) keys[i] = INT_MAX;
    __syncthreads();
    // bitonic sort of keys[0, m)
    for (int k = 2; k <= m; k <<= 1)
      for (int jj = k >> 1; jj > 0; jj >>= 1) {
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
          const int l = i ^ jj;
          if (l > i) {
            const int a = keys[i], b = keys[l];
            const bool up = (i & k) == 0;
            if ((a > b) == up) {
              keys[i] = b;
              keys[l] = a;
            }
          }
        }
        __syncthreads();
      }
    // unique: first occurrences, compacted in order (one pass per thread range)
    const int per = (total + blockDim.x - 1) / blockDim.x;
    const int i0 = min(total, (int)threadIdx.x * per), i1 = min(total, i0 + per);
    int firsts = 0;
    for (int i = i0; i < i1; ++i) firsts += (i == 0 || keys[i] != keys[i - 1]);
    int u = 0;
    const int ub = block_scan_excl(firsts, scan, u);
    int q = ub;
    int32_t *out = ubuf + (int64_t)s * kSortCap;
    for (int i = i0; i < i1; ++i)
      if (i == 0 || keys[i] != keys[i - 1]) out[q++] = keys[i];
    const int umin = keys[0], umax = keys[total - 1];
    const bool ok = !s_bad && (int64_t)u * 4 <= (int64_t)mx * 5 && r0 + umin >= 0 && r0 + tc - 1 + umax < n;
    if (ok) {
      ci.width = u;
      ci.offset = 1;
      ci.ulen = u;
      ci.umin = umin;
    }
  }
  if (threadIdx.x == 0) info[s] = ci;
}

// same[s] = 1: offset chunks s - 1 and s have the same union (chunk s then
// shares the table of chunk s - 1, as in the host builder)
__global__ void k_union_same(int32_t nch, const ChunkInfo *__restrict__ info, const int32_t *__restrict__ ubuf,
                             int32_t *__restrict__ same) {
  __shared__ int diff;
  const int s = blockIdx.x;
  if (threadIdx.x == 0) diff = 0;
  __syncthreads();
  const ChunkInfo a = info[s];
  const bool cand = s > 0 && a.offset && info[s - 1].offset && info[s - 1].ulen == a.ulen;
  if (cand)
    for (int t = threadIdx.x; t < a.ulen; t += blockDim.x)
      if (ubuf[(int64_t)s * kSortCap + t] != ubuf[(int64_t)(s - 1) * kSortCap + t]) diff = 1;
  __syncthreads();
  if (threadIdx.x == 0) same[s] = cand && !diff ? 1 : 0;
}

// offset tables: coff[tpos[s] + t] = union[t] (0 past |union|, up to the
// chunk's slot width); tpos[s] < 0: the chunk shares an earlier table
__global__ void k_coff_fill(const ChunkInfo *__restrict__ info, const int32_t *__restrict__ ubuf,
                            const int32_t *__restrict__ tpos, const int32_t *__restrict__ cw, int K,
                            int32_t *__restrict__ coff) {
  const int s = blockIdx.x;
  const ChunkInfo ci = info[s];
  if (!ci.offset || tpos[s] < 0) return;
  const int w = chunk_slots(cw[s], K);
  for (int t = threadIdx.x; t < w; t += blockDim.x)
    coff[tpos[s] + t] = t < ci.ulen ? ubuf[(int64_t)s * kSortCap + t] : 0;
}

// values and indices of row r into their slots (val zeroed beforehand)
__global__ void k_sell_fill(int32_t n, int32_t C, int K, const int32_t *__restrict__ rptr,
                            const int32_t *__restrict__ col, const int32_t *__restrict__ perm,
                            const int32_t *__restrict__ inv, const double *__restrict__ a,
                            const ChunkInfo *__restrict__ info, const int32_t *__restrict__ ubuf,
                            const int32_t *__restrict__ cptr, const int32_t *__restrict__ cw,
                            const int32_t *__restrict__ iptr, int32_t *__restrict__ idx, double *__restrict__ val) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = r / C, j = r - s * C;
  const int tc = min(C, n - s * C);
  const ChunkInfo ci = info[s];
  const int64_t c0 = cptr[s];
  const int o = perm ? perm[r] : r;
  if (ci.offset) {
    const int32_t *u = ubuf + (int64_t)s * kSortCap;
    int t = 0;
    for (int k = rptr[o]; k < rptr[o + 1]; ++k) {
      const int d = (inv ? inv[col[k]] : col[k]) - r;
      while (u[t] < d) ++t;  // columns ascend: the slot only moves forward
      val[chunk_slot_pos(c0, cw[s], tc, j, t, K)] = a[k];
    }
    return;
  }
  const int width = cw[s] * K;
  const int len = rptr[o + 1] - rptr[o];
  const int64_t ib = (int64_t)iptr[s] - c0;
  int last = 0;
  for (int q = 0; q < width; ++q) {
    const int64_t pos = c0 + ((int64_t)(q / K) * tc + j) * K + (q % K);
    if (q < len) {
      const int c = col[rptr[o] + q];
      last = inv ? inv[c] : c;
      val[pos] = a[rptr[o] + q];
    } else {
      val[pos] = 0.0;
    }
    idx[pos + ib] = last;
  }
}

__global__ void k_zero_f64(int64_t n, double *p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

template <class D> __global__ void k_vals_cast(int64_t n, const double *__restrict__ s, D *__restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = cvt<D>(s[i]);
}

DevPtr cast_values_device(Context *c, const DevPtr &v64, int64_t n, int dtype) {
  if (dtype == F3R_F64) {
    DevPtr d = dev_alloc(sizeof(double) * std::max<int64_t>(1, n));
    if (n) F3R_HIP(hipMemcpyAsync(d->p, v64->p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    return d;
  }
  DevPtr d = dev_alloc(dtype_size(dtype) * std::max<int64_t>(1, n));
  if (dtype == F3R_F32)
    hipLaunchKernelGGL(k_vals_cast<float>, dim3(2048), dim3(256), 0, c->stream, n, (const double *)v64->as<double>(),
                       d->as<float>());
  else if (dtype == F3R_BF16)
    hipLaunchKernelGGL(k_vals_cast<bf16_t>, dim3(2048), dim3(256), 0, c->stream, n,
                       (const double *)v64->as<double>(), d->as<bf16_t>());
  else
    hipLaunchKernelGGL(k_vals_cast<half_t>, dim3(2048), dim3(256), 0, c->stream, n,
                       (const double *)v64->as<double>(), d->as<half_t>());
  F3R_HIP(hipGetLastError());
  return d;
}

static DevPtr upload_vec(Context *c, const std::vector<int32_t> &v) {
  DevPtr d = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, v.size()));
  if (!v.empty()) F3R_HIP(hipMemcpyAsync(d->p, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice, c->stream));
  return d;
}

// Builds the layout of the device CSR (d_rptr, d_col, d_val; host rptr for
// the row counts) on the device: rows sorted inside windows of sigma rows,
// columns renumbered into the sorted order when `renumber` (the solver's
// internal order).  Returns false, building nothing, when a chunk's entries
// exceed the sort buffer or sigma exceeds kSigmaCap.  Values come back as
// fp64 in slot order (sval); the caller converts them.
bool build_sell_device(Context *c, const HostCSR &A, const int32_t *d_rptr, const int32_t *d_col,
                       const double *d_val, int32_t C, int32_t sigma, int32_t K, bool offsets, bool renumber,
                       std::shared_ptr<SellShared> &sh, DevPtr &sval) {
  hipStream_t st = c->stream;
  const int32_t n = A.n;
  if (sigma > kSigmaCap) return false;
  offsets = offsets && C >= 64 && n < (1 << 27);
  const int32_t nch = (n + C - 1) / C;
  // the row permutation (sigma > 1), kept only when it is not the identity
  std::vector<int32_t> h_perm(n);
  DevPtr d_perm, d_inv;
  bool ident = true;
  if (sigma > 1 && n > 0) {
    d_perm = dev_alloc(sizeof(int32_t) * n);
    hipLaunchKernelGGL(k_sigma_perm, dim3((n + sigma - 1) / sigma), dim3(1024), 0, st, n, sigma, d_rptr,
                       d_perm->as<int32_t>());
    F3R_HIP(hipGetLastError());
    F3R_HIP(hipMemcpyAsync(h_perm.data(), d_perm->p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    F3R_HIP(hipStreamSynchronize(st));
    for (int32_t r = 0; r < n && ident; ++r) ident = h_perm[r] == r;
  }
  if (ident) {
    std::iota(h_perm.begin(), h_perm.end(), 0);
    d_perm.reset();
  } else if (renumber) {
    d_inv = dev_alloc(sizeof(int32_t) * n);
    hipLaunchKernelGGL(k_inverse_perm, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       (const int32_t *)d_perm->as<int32_t>(), d_inv->as<int32_t>());
    F3R_HIP(hipGetLastError());
  }
  const int32_t *dp = d_perm ? d_perm->as<int32_t>() : nullptr;
  const int32_t *di = d_inv ? d_inv->as<int32_t>() : nullptr;
  DevPtr dinfo = dev_alloc(sizeof(ChunkInfo) * std::max(1, nch));
  DevPtr ubuf = dev_alloc(offsets ? sizeof(int32_t) * (size_t)kSortCap * nch : 4);
  if (nch > 0)
    hipLaunchKernelGGL(k_chunk_scan, dim3(nch), dim3(256), 0, st, n, C, d_rptr, d_col, dp, di, offsets ? 1 : 0,
                       dinfo->as<ChunkInfo>(), ubuf->as<int32_t>());
  F3R_HIP(hipGetLastError());
  std::vector<ChunkInfo> info(nch);
  std::vector<int32_t> same(nch, 0);
  if (nch) F3R_HIP(hipMemcpyAsync(info.data(), dinfo->p, sizeof(ChunkInfo) * nch, hipMemcpyDeviceToHost, st));
  DevPtr dsame = dev_alloc(sizeof(int32_t) * std::max(1, nch));
  if (offsets && nch) {
    hipLaunchKernelGGL(k_union_same, dim3(nch), dim3(64), 0, st, nch, (const ChunkInfo *)dinfo->as<ChunkInfo>(),
                       (const int32_t *)ubuf->as<int32_t>(), dsame->as<int32_t>());
    F3R_HIP(hipGetLastError());
    F3R_HIP(hipMemcpyAsync(same.data(), dsame->p, sizeof(int32_t) * nch, hipMemcpyDeviceToHost, st));
  }
  F3R_HIP(hipStreamSynchronize(st));
  for (const ChunkInfo &ci : info)
    if (ci.toobig) return false;
  auto round_up = [](int32_t v, int32_t m) { return (v + m - 1) / m * m; };
  // nonzeros of chunk s (its rows in sorted order)
  auto chunk_nnz = [&](int32_t s) {
    int64_t t = 0;
    const int64_t r1 = std::min<int64_t>(n, (int64_t)(s + 1) * C);
    if (ident) return (int64_t)A.rptr[r1] - A.rptr[(int64_t)s * C];
    for (int64_t r = (int64_t)s * C; r < r1; ++r) t += A.rptr[h_perm[r] + 1] - A.rptr[h_perm[r]];
    return t;
  };
  std::vector<int32_t> cptr(nch + 1, 0), cw(nch), iptr(offsets ? nch : 0), tpos(nch, 0);
  int64_t tot = 0, itot = 0, ctot = 0, off_slots = 0, expl_nnz = 0;
  int32_t off_chunks = 0, off_bias = 0;
  for (int32_t s = 0; s < nch; ++s) {
    const int32_t tc = std::min<int32_t>(C, n - s * C);
    cw[s] = offsets && info[s].offset ? offset_chunk_width(info[s].width, K) : round_up(info[s].width, K) / K;
    tot += (int64_t)chunk_slots(cw[s], K) * tc;
    F3R_CHECK(tot < INT32_MAX, F3R_ERR_RANGE, "SELL: more than 2^31 slots");
    cptr[s + 1] = (int32_t)tot;
    if (offsets) {
      if (info[s].offset) {
        if (same[s]) {
          iptr[s] = iptr[s - 1];
          tpos[s] = -1;
        } else {
          iptr[s] = -(int32_t)ctot - 1;
          tpos[s] = (int32_t)ctot;
          ctot += chunk_slots(cw[s], K);
          off_bias = std::max(off_bias, -info[s].umin);
        }
        off_slots += (int64_t)chunk_slots(cw[s], K) * tc;
        ++off_chunks;
      } else {
        iptr[s] = (int32_t)itot;
        itot += cptr[s + 1] - cptr[s];
        expl_nnz += chunk_nnz(s);
      }
    } else {
      expl_nnz += chunk_nnz(s);
    }
  }
  if (!offsets) itot = tot;
  sh = std::make_shared<SellShared>();
  sh->n = n;
  sh->C = C;
  sh->sigma = sigma;
  sh->K = K;
  sh->nchunks = nch;
  sh->nslots = tot;
  sh->internal = renumber;
  sh->offset_chunks = off_chunks;
  sh->offset_slots = off_slots;
  sh->explicit_nnz = expl_nnz;
  sh->off_bias = off_bias;
  sh->cptr = upload_vec(c, cptr);
  sh->cwidth = upload_vec(c, cw);
  sh->idx = dev_alloc(sizeof(int32_t) * std::max<int64_t>(1, itot));
  sh->h_perm = h_perm;
  sh->perm = d_perm;
  DevPtr d_iptr;
  if (offsets) {
    sh->iptr = upload_vec(c, iptr);
    d_iptr = sh->iptr;
    sh->coff = dev_alloc(sizeof(int32_t) * std::max<int64_t>(1, ctot));
    DevPtr d_tpos = upload_vec(c, tpos);
    if (nch)
      hipLaunchKernelGGL(k_coff_fill, dim3(nch), dim3(64), 0, st, (const ChunkInfo *)dinfo->as<ChunkInfo>(),
                         (const int32_t *)ubuf->as<int32_t>(), (const int32_t *)d_tpos->as<int32_t>(),
                         (const int32_t *)sh->cwidth->as<int32_t>(), K, sh->coff->as<int32_t>());
    F3R_HIP(hipGetLastError());
    F3R_HIP(hipStreamSynchronize(st));  // d_tpos is released at scope end
  } else {
    d_iptr = sh->cptr;  // explicit everywhere: indices at the value offsets
  }
  sval = dev_alloc(sizeof(double) * std::max<int64_t>(1, tot));
  hipLaunchKernelGGL(k_zero_f64, dim3(1024), dim3(256), 0, st, std::max<int64_t>(1, tot), sval->as<double>());
  if (n > 0)
    hipLaunchKernelGGL(k_sell_fill, dim3((n + 255) / 256), dim3(256), 0, st, n, C, K, d_rptr, d_col, dp, di, d_val,
                       (const ChunkInfo *)dinfo->as<ChunkInfo>(), (const int32_t *)ubuf->as<int32_t>(),
                       (const int32_t *)sh->cptr->as<int32_t>(), (const int32_t *)sh->cwidth->as<int32_t>(),
                       (const int32_t *)d_iptr->as<int32_t>(), sh->idx->as<int32_t>(), sval->as<double>());
  F3R_HIP(hipGetLastError());
  F3R_HIP(hipStreamSynchronize(st));
  return true;
}

}  // namespace f3r
