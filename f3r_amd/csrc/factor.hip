// Device factorisation of the innermost preconditioners, bit-identical to
// the host restatements (sdainv_host, ilu0_host):
//
//   SDAINV (left_sdainv, ainv.hpp:268-329): rows are processed in level
//   sets (factor_levels with both triangles), one kernel launch per level,
//   one 64-lane wavefront per row.  The sparse rows z_i, w_i live in LDS as
//   sorted (column, value) lists; a sparse dot is a binary search per entry
//   of the CSR row plus an in-order sum of the matched products, a sparse
//   axpy is a binary search per entry of either list plus two prefix counts
//   that place every kept entry.  Every value is made by the same single
//   fp64 operation as on the host, in the same order, so the factors match
//   bit for bit.  Rows live in fixed-capacity slots (64, 256 or 1024
//   entries); a row that outgrows its slot sets a flag and the build is
//   retried with the next capacity, then handed back to the host.
//
//   ILU(0) (ilu.hpp:128-164): the block-diagonal part of A eliminated in
//   level sets of its lower triangle, one wavefront per row; each step
//   divides by the pivot of an earlier row and updates the later entries of
//   the row that row k's upper part matches.
//
// The transpose of A that SDAINV's z-side reads is built here too (column
// counts by atomics, host prefix sum, atomic fill, per-column sort).
#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>

#include "common.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace f3r {
namespace {

constexpr int kLanes = 64;
constexpr int kIluRowCap = 512;  // entries of one ILU(0) row held in LDS

// ------------------------------------------------------------- transpose --
__global__ void k_t_count(int32_t n, const int32_t *__restrict__ rptr, const int32_t *__restrict__ col,
                          int32_t *__restrict__ cnt) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t k = rptr[i]; k < rptr[i + 1]; ++k) atomicAdd(&cnt[col[k]], 1);
}

__global__ void k_t_fill(int32_t n, const int32_t *__restrict__ rptr, const int32_t *__restrict__ col,
                         const double *__restrict__ val, int32_t *__restrict__ fill, int32_t *__restrict__ tcol,
                         double *__restrict__ tval) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int32_t k = rptr[i]; k < rptr[i + 1]; ++k) {
    const int32_t p = atomicAdd(&fill[col[k]], 1);
    tcol[p] = i;
    tval[p] = val[k];
  }
}

// the atomic fill leaves each column's rows in arrival order: sort them, one
// wave per column (rank of each row among the column's rows; rows are
// distinct), lane 0 alone for a column of more than 64 entries
__global__ __launch_bounds__(256) void k_t_sort(int32_t n, const int32_t *__restrict__ trptr, int32_t *__restrict__ tcol,
                                                double *__restrict__ tval) {
  const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes;
  const int lane = threadIdx.x % kLanes;
  if (c >= n) return;
  const int32_t b = trptr[c], e = trptr[c + 1], len = e - b;
  if (len <= kLanes) {
    const int32_t key = lane < len ? tcol[b + lane] : INT32_MAX;
    const double v = lane < len ? tval[b + lane] : 0.0;
    int rank = 0;
    for (int q = 0; q < len; ++q) rank += __shfl(key, q, kLanes) < key;
    if (lane < len) {
      tcol[b + rank] = key;
      tval[b + rank] = v;
    }
  } else if (lane == 0) {
    for (int32_t p = b + 1; p < e; ++p) {
      const int32_t key = tcol[p];
      const double v = tval[p];
      int32_t q = p - 1;
      while (q >= b && tcol[q] > key) {
        tcol[q + 1] = tcol[q];
        tval[q + 1] = tval[q];
        --q;
      }
      tcol[q + 1] = key;
      tval[q + 1] = v;
    }
  }
}

// ------------------------------------------------------ wave primitives --
__device__ __forceinline__ int lower_bound_lds(const int32_t *c, int len, int32_t key) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int wave_scan_excl(int v, int &total) {
  const int lane = threadIdx.x;
  if (__builtin_amdgcn_readfirstlane(__ballot(v > 1) == 0)) {
    // every lane holds 0 or 1: a mask count
    const uint64_t m = __ballot(v != 0);
    total = __builtin_popcountll(m);
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
  }
  int x = v;
  for (int off = 1; off < kLanes; off <<= 1) {
    const int y = __shfl_up(x, off, kLanes);
    if (lane >= off) x += y;
  }
  total = __shfl(x, kLanes - 1, kLanes);
  return x - v;
}

// Sum over the columns shared by the LDS list (xc, xv)[0, lx) and the CSR
// row (col, val)[0, len), in column order, each term x * a (merge_dot).
__device__ double sparse_dot(const int32_t *xc, const double *xv, int lx, const int32_t *__restrict__ col,
                             const double *__restrict__ val, int len) {
  double s = 0.0;
  for (int base = 0; base < len; base += kLanes) {
    const int b = base + (int)threadIdx.x;
    double p = 0.0;
    bool m = false;
    if (b < len) {
      const int32_t cb = col[b];
      const int pos = lower_bound_lds(xc, lx, cb);
      if (pos < lx && xc[pos] == cb) {
        m = true;
        p = xv[pos] * val[b];
      }
    }
    uint64_t mask = __ballot(m);
    while (mask) {
      const int l = __builtin_ctzll(mask);
      mask &= mask - 1;
      s += readlane_f64(p, l);
    }
  }
  return s;
}

// The same with the CSR row held in registers: lane b has entry b (rlen <= 64)
__device__ double sparse_dot_reg(const int32_t *xc, const double *xv, int lx, int32_t cb, double vb, int rlen) {
  double s = 0.0;
  double p = 0.0;
  bool m = false;
  if ((int)threadIdx.x < rlen) {
    const int pos = lower_bound_lds(xc, lx, cb);
    if (pos < lx && xc[pos] == cb) {
      m = true;
      p = xv[pos] * vb;
    }
  }
  uint64_t mask = __ballot(m);
  while (mask) {
    const int l = __builtin_ctzll(mask);
    mask &= mask - 1;
    s += readlane_f64(p, l);
  }
  return s;
}

// out = y + alpha * x with the drop rule |v| >= tol && v != 0 (merge_axpy);
// all lists in LDS, sorted by column.  Lane l owns entries [l*E, l*E + E)
// of x and of y.  Returns the length of out, or -1 when it exceeds 64 E.
template <int E>
__device__ int sparse_axpy(double alpha, const int32_t *xc, const double *xv, int lx, const int32_t *yc,
                           const double *yv, int ly, double tol, int32_t *oc, double *ov, int32_t *px,
                           int32_t *py) {
  const int lane = threadIdx.x;
  int xlb[E], ylb[E];
  double xval[E];
  bool xk[E], yk[E];
  int nx = 0, ny = 0;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int a = lane * E + q;
    xk[q] = false;
    xlb[q] = 0;
    xval[q] = 0.0;
    if (a < lx) {
      const int32_t c = xc[a];
      const int lb = lower_bound_lds(yc, ly, c);
      double v;
      if (lb < ly && yc[lb] == c)
        v = yv[lb] + alpha * xv[a];
      else
        v = xv[a] * alpha;
      xk[q] = fabs(v) >= tol && v != 0.0;
      xlb[q] = lb;
      xval[q] = v;
      nx += xk[q];
    }
    const int b = lane * E + q;
    yk[q] = false;
    ylb[q] = 0;
    if (b < ly) {
      const int32_t c = yc[b];
      const int lb = lower_bound_lds(xc, lx, c);
      const bool matched = lb < lx && xc[lb] == c;
      const double v = yv[b];
      yk[q] = !matched && fabs(v) >= tol && v != 0.0;
      ylb[q] = lb;
      ny += yk[q];
    }
  }
  int tx, ty;
  int rx = wave_scan_excl(nx, tx);
  int ry = wave_scan_excl(ny, ty);
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = lane * E + q;
    if (e < lx) px[e] = rx;
    if (e < ly) py[e] = ry;
    rx += xk[q];
    ry += yk[q];
  }
  if (lane == 0) {
    px[lx] = tx;
    py[ly] = ty;
  }
  __syncthreads();
  const int total = tx + ty;
  if (total > kLanes * E) return -1;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = lane * E + q;
    if (xk[q]) {
      const int pos = px[e] + py[xlb[q]];
      oc[pos] = xc[e];
      ov[pos] = xval[q];
    }
    if (yk[q]) {
      const int pos = py[e] + px[ylb[q]];
      oc[pos] = yc[e];
      ov[pos] = yv[e];
    }
  }
  __syncthreads();
  return total;
}

template <int E>
__device__ __forceinline__ void store_row(int32_t i, int l, const int32_t *c, const double *v,
                                          int32_t *__restrict__ len, int32_t *__restrict__ pc,
                                          double *__restrict__ pv) {
  constexpr int CAP = kLanes * E;
  const int64_t base = (int64_t)i * CAP;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = q * kLanes + (int)threadIdx.x;
    if (e < l) {
      pc[base + e] = c[e];
      pv[base + e] = v[e];
    }
  }
  if (threadIdx.x == 0) len[i] = l;
}

struct SdainvArgs {
  const int32_t *rows;  // rows of this level
  int32_t cnt;          // rows in this level
  const int32_t *rptr, *col;
  const double *val;
  const int32_t *trptr, *tcol;  // A^T
  const double *tval;
  double tol, acc;
  int32_t *wlen, *wc, *zlen, *zc;
  double *wv, *zv, *d;
  int32_t *overflow;
};

// One side of one row of the biconjugation (ainv.hpp:280-325) per 64-lane
// block: blocks [0, cnt) build w_i against the z_j of the lower part of row
// i, then the pivot d_i; blocks [cnt, 2 cnt) build z_i against the w_j of
// the upper part of column i.  The two sides of a row read only rows of
// earlier levels, so they run side by side.  The slots of NB neighbours j
// (the dot operand, the axpy source and d_j) are fetched together, whole
// slots without waiting for their lengths, so a batch costs one round trip.
template <int E, int NB> __global__ __launch_bounds__(64) void k_sdainv_side(SdainvArgs a) {
  constexpr int CAP = kLanes * E;
  __shared__ int32_t nc[NB][2][CAP];  // neighbour slots: 0 dot operand, 1 axpy source
  __shared__ double nv[NB][2][CAP];
  __shared__ int32_t yc[2][CAP];      // the row being built (ping-pong)
  __shared__ double yv[2][CAP];
  __shared__ int32_t spx[CAP + 1], spy[CAP + 1];
  const int lane = threadIdx.x;
  const bool zside = (int32_t)blockIdx.x >= a.cnt;
  const int32_t i = a.rows[zside ? blockIdx.x - a.cnt : blockIdx.x];
  // w-side: row i of A, dot operands z_j, sources w_j; z-side: row i of A^T,
  // dot operands w_j, sources z_j
  const int32_t r0 = zside ? a.trptr[i] : a.rptr[i];
  const int32_t rlen = (zside ? a.trptr[i + 1] : a.rptr[i + 1]) - r0;
  const int32_t *rcol = (zside ? a.tcol : a.col) + r0;
  const double *rval = (zside ? a.tval : a.val) + r0;
  const int32_t *dlen = zside ? a.wlen : a.zlen, *dc = zside ? a.wc : a.zc;
  const double *dv = zside ? a.wv : a.zv;
  int32_t *slen = zside ? a.zlen : a.wlen, *sc = zside ? a.zc : a.wc;
  double *sv = zside ? a.zv : a.wv;
  // entries of the row before column i (sorted row)
  const bool short_row = rlen <= kLanes;
  int32_t cb = INT32_MAX;
  double vb = 0.0;
  if ((int)threadIdx.x < rlen) {
    cb = rcol[threadIdx.x];
    vb = rval[threadIdx.x];
  }
  int nl;
  if (short_row) {
    nl = __builtin_popcountll(__ballot(cb < i));
  } else {
    nl = 0;
    while (nl < rlen && rcol[nl] < i) ++nl;
  }

  int cur = 0, ly = 1;
  if (lane == 0) {
    yc[0][0] = i;
    yv[0][0] = 1.0;
  }
  for (int base = 0; base < nl; base += NB) {
    const int nb = min(NB, nl - base);
    int32_t lc[NB][2][E];
    double lv[NB][2][E];
    int ll[NB][2];
    double ld[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (q < nb) {
        const int32_t j = rcol[base + q];
        const int64_t off = (int64_t)j * CAP;
        ll[q][0] = dlen[j];
        ll[q][1] = slen[j];
        ld[q] = a.d[j];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          lc[q][0][e] = dc[off + e * kLanes + lane];
          lv[q][0][e] = dv[off + e * kLanes + lane];
          lc[q][1][e] = sc[off + e * kLanes + lane];
          lv[q][1][e] = sv[off + e * kLanes + lane];
        }
      }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (q < nb) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          nc[q][0][e * kLanes + lane] = lc[q][0][e];
          nv[q][0][e * kLanes + lane] = lv[q][0][e];
          nc[q][1][e * kLanes + lane] = lc[q][1][e];
          nv[q][1][e * kLanes + lane] = lv[q][1][e];
        }
      }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q)
      if (q < nb) {
        const int l0 = min(max(ll[q][0], 0), CAP), l1 = min(max(ll[q][1], 0), CAP);
        const double dot = short_row ? sparse_dot_reg(nc[q][0], nv[q][0], l0, cb, vb, rlen)
                                     : sparse_dot(nc[q][0], nv[q][0], l0, rcol, rval, rlen);
        const double alpha = -dot / ld[q];
        const int lo = sparse_axpy<E>(alpha, nc[q][1], nv[q][1], l1, yc[cur], yv[cur], ly, a.tol, yc[1 - cur],
                                      yv[1 - cur], spx, spy);
        if (lo < 0) {
          if (lane == 0) a.overflow[0] = 1;
          return;
        }
        ly = lo;
        cur = 1 - cur;
      }
  }
  __syncthreads();
  if (!zside) {
    // pivot with the weight acc on w_i's last entry (ainv.hpp:310-316)
    const int32_t t0 = a.trptr[i], clen = a.trptr[i + 1] - t0;
    double di = 0.0;
    if (ly > 0) {
      if (lane == 0) yv[cur][ly - 1] = yv[cur][ly - 1] * a.acc;
      __syncthreads();
      di = sparse_dot(yc[cur], yv[cur], ly, a.tcol + t0, a.tval + t0, clen);
      __syncthreads();
      if (lane == 0) yv[cur][ly - 1] = yv[cur][ly - 1] / a.acc;
      __syncthreads();
    }
    if (di == 0.0) di = 1.0;
    if (lane == 0) a.d[i] = di;
  }
  store_row<E>(i, ly, yc[cur], yv[cur], slen, sc, sv);
}

__global__ void k_sdainv_init(int32_t n, int32_t *wlen, int32_t *zlen, double *d) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  wlen[i] = 0;
  zlen[i] = 0;
  d[i] = 1.0;
}

// pools -> compact rows (z_i scaled by 1 / d_i, ainv.hpp:57-64); one wave per row
__global__ void k_sdainv_compact(int32_t n, int cap, const int32_t *__restrict__ len, const int32_t *__restrict__ pc,
                                 const double *__restrict__ pv, const double *__restrict__ d,
                                 const int64_t *__restrict__ ptr, int32_t *__restrict__ oc, double *__restrict__ ov) {
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes;
  const int lane = threadIdx.x % kLanes;
  if (w >= n) return;
  const int32_t i = (int32_t)w;
  const int l = len[i];
  const int64_t src = (int64_t)i * cap, dst = ptr[i];
  for (int e = lane; e < l; e += kLanes) {
    oc[dst + e] = pc[src + e];
    ov[dst + e] = d ? pv[src + e] / d[i] : pv[src + e];
  }
}

// ------------------------------------------------------------------ ILU(0) --
// One row of IKJ ILU(0) per 64-lane block (ilu.hpp:140-160).
__global__ __launch_bounds__(64) void k_ilu0_row(const int32_t *__restrict__ rows, const int32_t *__restrict__ rptr,
                                                 const int32_t *__restrict__ col, double *__restrict__ val,
                                                 const int32_t *__restrict__ dpos, double acc) {
  __shared__ int32_t rc[kIluRowCap], kc[kIluRowCap];
  __shared__ double rv[kIluRowCap], kv[kIluRowCap];
  const int lane = threadIdx.x;
  const int32_t i = rows[blockIdx.x];
  const int32_t p0 = rptr[i], len = rptr[i + 1] - p0;
  for (int e = lane; e < len; e += kLanes) {
    rc[e] = col[p0 + e];
    rv[e] = val[p0 + e];
  }
  __syncthreads();
  if (lane == 0) rv[dpos[i] - p0] = rv[dpos[i] - p0] * acc;
  __syncthreads();
  for (int q = 0; q < len; ++q) {
    const int32_t k = rc[q];
    if (k >= i) break;
    const int32_t r = dpos[k], ke = rptr[k + 1];
    const int ku = ke - r;  // row k from its diagonal on
    for (int e = lane; e < ku; e += kLanes) {
      kc[e] = col[r + e];
      kv[e] = val[r + e];
    }
    __syncthreads();
    const double lq = rv[q] / kv[0];
    for (int t = q + 1 + lane; t < len; t += kLanes) {
      const int pos = lower_bound_lds(kc, ku, rc[t]);
      if (pos < ku && kc[pos] == rc[t]) rv[t] = rv[t] - lq * kv[pos];
    }
    __syncthreads();
    if (lane == 0) rv[q] = lq;
    __syncthreads();
  }
  for (int e = lane; e < len; e += kLanes) val[p0 + e] = rv[e];
}

DevPtr upload_i32(Context *c, const int32_t *p, size_t n) {
  DevPtr d = dev_alloc(sizeof(int32_t) * std::max<size_t>(1, n));
  if (n) F3R_HIP(hipMemcpyAsync(d->p, p, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  return d;
}

DevPtr upload_f64(Context *c, const double *p, size_t n) {
  DevPtr d = dev_alloc(sizeof(double) * std::max<size_t>(1, n));
  if (n) F3R_HIP(hipMemcpyAsync(d->p, p, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  return d;
}

template <int E>
bool sdainv_pass(Context *c, const HostCSR &A, const std::vector<int32_t> &lptr, const DevPtr &d_rows,
                 SdainvArgs a, HostFactors &F) {
  constexpr int CAP = kLanes * E;
  hipStream_t st = c->stream;
  const int32_t n = A.n;
  DevPtr wlen = dev_alloc(sizeof(int32_t) * n), zlen = dev_alloc(sizeof(int32_t) * n);
  DevPtr d = dev_alloc(sizeof(double) * n);
  DevPtr wc = dev_alloc(sizeof(int32_t) * (size_t)n * CAP), zc = dev_alloc(sizeof(int32_t) * (size_t)n * CAP);
  DevPtr wv = dev_alloc(sizeof(double) * (size_t)n * CAP), zv = dev_alloc(sizeof(double) * (size_t)n * CAP);
  DevPtr flag = dev_alloc(sizeof(int32_t));
  F3R_HIP(hipMemsetAsync(flag->p, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_sdainv_init, dim3((n + 255) / 256), dim3(256), 0, st, n, wlen->as<int32_t>(),
                     zlen->as<int32_t>(), d->as<double>());
  a.wlen = wlen->as<int32_t>();
  a.zlen = zlen->as<int32_t>();
  a.wc = wc->as<int32_t>();
  a.zc = zc->as<int32_t>();
  a.wv = wv->as<double>();
  a.zv = zv->as<double>();
  a.d = d->as<double>();
  a.overflow = flag->as<int32_t>();
  constexpr int NB = E == 1 ? 8 : E == 4 ? 2 : 1;
  for (size_t l = 0; l + 1 < lptr.size(); ++l) {
    a.cnt = lptr[l + 1] - lptr[l];
    a.rows = d_rows->as<int32_t>() + lptr[l];
    hipLaunchKernelGGL((k_sdainv_side<E, NB>), dim3(2 * a.cnt), dim3(kLanes), 0, st, a);
  }
  F3R_HIP(hipGetLastError());
  int32_t over = 0;
  F3R_HIP(hipMemcpyAsync(&over, flag->p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  F3R_HIP(hipStreamSynchronize(st));
  if (over) return false;
  // row pointers of the compact factors
  std::vector<int32_t> hw(n), hz(n);
  F3R_HIP(hipMemcpyAsync(hw.data(), wlen->p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  F3R_HIP(hipMemcpyAsync(hz.data(), zlen->p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  F3R_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> pw(n + 1, 0), pz(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) {
    pw[i + 1] = pw[i] + hw[i];
    pz[i + 1] = pz[i] + hz[i];
  }
  F3R_CHECK(pw[n] < INT32_MAX && pz[n] < INT32_MAX, F3R_ERR_RANGE, "SDAINV: more than 2^31 factor entries");
  HostCSR Wt, Zr;
  for (int w = 0; w < 2; ++w) {
    const std::vector<int64_t> &p = w ? pz : pw;
    HostCSR &M = w ? Zr : Wt;
    const int64_t nnz = p[n];
    DevPtr dp = dev_alloc(sizeof(int64_t) * (n + 1));
    F3R_HIP(hipMemcpyAsync(dp->p, p.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
    DevPtr oc = dev_alloc(sizeof(int32_t) * std::max<int64_t>(1, nnz));
    DevPtr ov = dev_alloc(sizeof(double) * std::max<int64_t>(1, nnz));
    const int64_t threads = (int64_t)n * kLanes;
    hipLaunchKernelGGL(k_sdainv_compact, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, n, CAP,
                       (const int32_t *)(w ? zlen : wlen)->as<int32_t>(),
                       (const int32_t *)(w ? zc : wc)->as<int32_t>(), (const double *)(w ? zv : wv)->as<double>(),
                       w ? (const double *)d->as<double>() : nullptr, (const int64_t *)dp->as<int64_t>(),
                       oc->as<int32_t>(), ov->as<double>());
    F3R_HIP(hipGetLastError());
    M.n = M.ncols = n;
    M.rptr.resize(n + 1);
    for (int32_t i = 0; i <= n; ++i) M.rptr[i] = (int32_t)p[i];
    M.idx.resize(nnz);
    M.val.resize(nnz);
    if (nnz) {
      F3R_HIP(hipMemcpyAsync(M.idx.data(), oc->p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, st));
      F3R_HIP(hipMemcpyAsync(M.val.data(), ov->p, sizeof(double) * nnz, hipMemcpyDeviceToHost, st));
    }
    F3R_HIP(hipStreamSynchronize(st));
  }
  // Z = (rows z_i / d_i)^T, W^T = rows w_i (ainv.hpp:57-64)
  F.f[0] = csr_transpose(Zr);
  F.f[1] = std::move(Wt);
  return true;
}

}  // namespace

bool sdainv_device(Context *c, const HostCSR &A, const int32_t *d_rptr, const int32_t *d_col, const double *d_val,
                   double tol, double acc, HostFactors &F) {
  hipStream_t st = c->stream;
  const int32_t n = A.n;
  F3R_CHECK(A.ncols == 0 || A.ncols == n, F3R_ERR_ARG, "SDAINV: the matrix must be square");
  if (n == 0) return false;
  std::vector<int32_t> lptr, rows;
  factor_levels(A, true, lptr, rows);
  DevPtr d_rows = upload_i32(c, rows.data(), rows.size());
  // A^T on the device
  DevPtr cnt = dev_alloc(sizeof(int32_t) * n);
  F3R_HIP(hipMemsetAsync(cnt->p, 0, sizeof(int32_t) * n, st));
  hipLaunchKernelGGL(k_t_count, dim3((n + 255) / 256), dim3(256), 0, st, n, d_rptr, d_col, cnt->as<int32_t>());
  std::vector<int32_t> trptr(n + 1, 0);
  F3R_HIP(hipMemcpyAsync(trptr.data() + 1, cnt->p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  F3R_HIP(hipStreamSynchronize(st));
  for (int32_t i = 0; i < n; ++i) trptr[i + 1] += trptr[i];
  const int64_t nnz = A.nnz();
  DevPtr d_trptr = upload_i32(c, trptr.data(), n + 1);
  DevPtr fill = upload_i32(c, trptr.data(), n);
  DevPtr tcol = dev_alloc(sizeof(int32_t) * std::max<int64_t>(1, nnz));
  DevPtr tval = dev_alloc(sizeof(double) * std::max<int64_t>(1, nnz));
  hipLaunchKernelGGL(k_t_fill, dim3((n + 255) / 256), dim3(256), 0, st, n, d_rptr, d_col, d_val,
                     fill->as<int32_t>(), tcol->as<int32_t>(), tval->as<double>());
  hipLaunchKernelGGL(k_t_sort, dim3((unsigned)(((int64_t)n * kLanes + 255) / 256)), dim3(256), 0, st, n,
                     (const int32_t *)d_trptr->as<int32_t>(), tcol->as<int32_t>(), tval->as<double>());
  F3R_HIP(hipGetLastError());
  SdainvArgs a{};
  a.rptr = d_rptr;
  a.col = d_col;
  a.val = d_val;
  a.trptr = d_trptr->as<int32_t>();
  a.tcol = tcol->as<int32_t>();
  a.tval = tval->as<double>();
  a.tol = tol;
  a.acc = acc;
  // slot pools: n * capacity * 24 bytes; larger capacities only while the
  // pools stay within 24 GiB
  const double limit = 24.0 * (1 << 30);
  auto fits = [&](int cap) { return (double)n * cap * 24.0 <= limit; };
  if (fits(64) && sdainv_pass<1>(c, A, lptr, d_rows, a, F)) return true;
  if (fits(256) && sdainv_pass<4>(c, A, lptr, d_rows, a, F)) return true;
  if (fits(1024) && sdainv_pass<16>(c, A, lptr, d_rows, a, F)) return true;
  return false;
}

bool ilu0_device(Context *c, const HostCSR &A, const int32_t *d_arptr, const int32_t *d_acol, const double *d_aval,
                 int32_t blocks, double acc, HostFactors &F) {
  hipStream_t st = c->stream;
  // one block: the whole matrix, already on the device
  HostCSR Rb;
  if (blocks != 1) Rb = ilu0_restrict(A, blocks);
  const HostCSR &R = blocks == 1 ? A : Rb;
  const int32_t n = R.n;
  if (n == 0) return false;
  for (int32_t i = 0; i < n; ++i)
    if (R.rptr[i + 1] - R.rptr[i] > kIluRowCap) return false;
  const std::vector<int32_t> dpos = ilu0_diag_pos(R);
  std::vector<int32_t> lptr, rows;
  factor_levels(R, false, lptr, rows);
  const size_t nnz = R.idx.size();
  DevPtr d_rows = upload_i32(c, rows.data(), rows.size());
  DevPtr d_rptr, d_col, d_val = dev_alloc(sizeof(double) * std::max<size_t>(1, nnz));
  if (blocks == 1) {
    if (nnz) F3R_HIP(hipMemcpyAsync(d_val->p, d_aval, sizeof(double) * nnz, hipMemcpyDeviceToDevice, st));
  } else {
    d_rptr = upload_i32(c, R.rptr.data(), n + 1);
    d_col = upload_i32(c, R.idx.data(), nnz);
    if (nnz) F3R_HIP(hipMemcpyAsync(d_val->p, R.val.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, st));
  }
  const int32_t *rp = blocks == 1 ? d_arptr : d_rptr->as<int32_t>();
  const int32_t *cl = blocks == 1 ? d_acol : d_col->as<int32_t>();
  DevPtr d_dpos = upload_i32(c, dpos.data(), n);
  for (size_t l = 0; l + 1 < lptr.size(); ++l) {
    const int32_t cnt = lptr[l + 1] - lptr[l];
    hipLaunchKernelGGL(k_ilu0_row, dim3(cnt), dim3(kLanes), 0, st, (const int32_t *)d_rows->as<int32_t>() + lptr[l],
                       rp, cl, d_val->as<double>(), (const int32_t *)d_dpos->as<int32_t>(), acc);
  }
  F3R_HIP(hipGetLastError());
  std::vector<double> v(nnz);
  if (nnz) F3R_HIP(hipMemcpyAsync(v.data(), d_val->p, sizeof(double) * nnz, hipMemcpyDeviceToHost, st));
  F3R_HIP(hipStreamSynchronize(st));
  F = ilu0_split(R, v.data());
  return true;
}

}  // namespace f3r
