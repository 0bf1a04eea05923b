"""f3r_amd -- MI355X-native F3R nested mixed-precision solver.

Python mirror of the reference's host interface (suzuki-hpc/F3R,
library/umineko-sparse): CSR / SELL matrices, SDAINV / ILU(0)
preconditioners and the four-level F3R solve, over the C ABI of
libf3r_amd.so (include/f3r_amd.h).  All compute runs in the HIP library;
this module only moves arrays and handles.
"""
import ctypes
from collections import namedtuple

import numpy as np

from . import _capi
from ._capi import BF16, F16, F32, F64, I32, F3RError, check, lib

__all__ = ["BF16", "Context", "vector", "CSR", "CSRDevice", "SELL", "SDAINV", "ILU0", "F3R", "Batch", "algorithm", "io", "solve_res_t",
           "Operator", "AutoRichardson", "FGMRES", "Restarted",
           "F3RError", "F64", "F32", "F16", "device_count"]

# bf16 vectors move as their 16-bit patterns (numpy has no bfloat16)
_NP = {F64: np.float64, F32: np.float32, F16: np.float16, I32: np.int32, BF16: np.uint16}
_DT = {"f64": F64, "f32": F32, "f16": F16, "i32": I32, "bf16": BF16, "double": F64, "float": F32, "half": F16}


def _dtype(d):
    if isinstance(d, int):
        return d
    if isinstance(d, str):
        return _DT[d]
    return {np.dtype(np.float64): F64, np.dtype(np.float32): F32, np.dtype(np.float16): F16,
            np.dtype(np.int32): I32}[np.dtype(d)]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def device_count():
    c = ctypes.c_int(0)
    check(lib.f3r_device_count(ctypes.byref(c)), "device_count")
    return c.value


class Context:
    """One GPU and one HIP stream (replaces exec<device> on the default stream)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib.f3r_ctx_create(device, ctypes.byref(h)), "ctx_create")
        self.handle = h
        self.device = device

    def sync(self):
        check(lib.f3r_ctx_sync(self.handle), "ctx_sync")

    def close(self):
        if self.handle:
            lib.f3r_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class vector:
    """kmm::vector<T, device> (core/tensor.hpp:204-316)."""

    def __init__(self, ctx, dtype, n, _handle=None, _base=None, _borrowed=False):
        self.ctx = ctx
        self._borrowed = _borrowed  # a handle the library owns (an Operator's operands)
        self.dtype = _dtype(dtype)
        self.n = int(n)
        self._base = _base
        if _handle is None:
            h = ctypes.c_void_p()
            check(lib.f3r_vec_create(ctx.handle, self.dtype, self.n, ctypes.byref(h)), "vec_create")
            self.handle = h
        else:
            self.handle = _handle

    @classmethod
    def from_numpy(cls, ctx, arr, dtype=None):
        arr = np.asarray(arr)
        dt = _dtype(dtype) if dtype is not None else _dtype(arr.dtype)
        v = cls(ctx, dt, arr.size)
        if dt == BF16 and arr.dtype != np.uint16:
            # values (not bit patterns): rounded to bf16 on the device from fp64
            v.copy(cls.from_numpy(ctx, arr.astype(np.float64)))
            return v
        v.upload(arr)
        return v

    def view(self, offset, n):
        h = ctypes.c_void_p()
        check(lib.f3r_vec_view(self.handle, offset, n, ctypes.byref(h)), "vec_view")
        return vector(self.ctx, self.dtype, n, _handle=h, _base=self)

    def upload(self, arr):
        a = np.ascontiguousarray(np.asarray(arr).reshape(-1), dtype=_NP[self.dtype])
        if a.size != self.n:
            raise ValueError("upload: size mismatch")
        check(lib.f3r_vec_upload(self.handle, _p(a)), "vec_upload")
        return self

    def download(self):
        out = np.empty(self.n, _NP[self.dtype])
        check(lib.f3r_vec_download(self.handle, _p(out)), "vec_download")
        return out

    def fill(self, v):
        check(lib.f3r_vec_fill(self.handle, float(v)), "vec_fill")
        return self

    def copy(self, src):
        """self = T(src) (vector::copy, tensor.hpp:272-277)"""
        check(lib.f3r_vec_copy(src.handle, self.handle), "vec_copy")
        return self

    def axpy(self, alpha, x):
        """self += alpha * x"""
        check(lib.f3r_vec_axpy(float(alpha), x.handle, self.handle), "vec_axpy")
        return self

    def scale(self, alpha):
        check(lib.f3r_vec_scale(self.handle, float(alpha)), "vec_scale")
        return self

    def __del__(self):
        try:
            if self.handle and not self._borrowed:
                lib.f3r_vec_destroy(self.handle)
            self.handle = None
        except Exception:
            pass


def dot(a, b):
    """DOT (core/tensor.hpp:509-540)"""
    out = ctypes.c_double()
    check(lib.f3r_vec_dot(a.handle, b.handle, ctypes.byref(out)), "vec_dot")
    return out.value


def nrm(a):
    """NRM (core/tensor.hpp:542-571)"""
    out = ctypes.c_double()
    check(lib.f3r_vec_nrm2(a.handle, ctypes.byref(out)), "vec_nrm2")
    return out.value


class CSR:
    """CSR<double, host> (sparse/matrix/csr.hpp:12-50): int32 indices, fp64 values."""

    def __init__(self, n, rptr, idx, val, ncols=None):
        self.n = int(n)
        self.ncols = int(ncols if ncols is not None else n)
        self.rptr = np.ascontiguousarray(rptr, np.int32)
        self.idx = np.ascontiguousarray(idx, np.int32)
        self.val = np.ascontiguousarray(val, np.float64)
        self._dev = {}

    @property
    def nnz(self):
        return int(self.idx.size)

    def nrows(self):
        return self.n

    def diagonal(self):
        """a_ii of every row (0 where the row has no diagonal entry)"""
        d = np.zeros(self.n)
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.rptr))
        on = rows == self.idx
        d[rows[on]] = self.val[on]
        return d

    def factor_levels(self, both=True):
        """(level of every row, number of levels) of the device factorisations:
        row i waits for the rows j < i with a_ij != 0 and, when `both`, a_ji != 0"""
        lev = np.empty(self.n, np.int32)
        nl = ctypes.c_int32()
        check(lib.f3r_factor_levels(self.handle(), 1 if both else 0, _p(lev), ctypes.byref(nl)), "factor_levels")
        return lev, nl.value

    def handle(self, ctx=None):
        """C-ABI matrix handle on ctx (None: host-only)."""
        key = id(ctx) if ctx is not None else None
        if key not in self._dev:
            h = ctypes.c_void_p()
            check(lib.f3r_mat_create_csr(ctx.handle if ctx is not None else None, self.n, self.ncols,
                                         self.nnz, _p(self.rptr), _p(self.idx), _p(self.val),
                                         ctypes.byref(h)), "mat_create_csr")
            self._dev[key] = (h, ctx)
        return self._dev[key][0]

    def invalidate(self):
        for h, _ in self._dev.values():
            lib.f3r_mat_destroy(h)
        self._dev = {}

    def __del__(self):
        try:
            self.invalidate()
        except Exception:
            pass

    # ---- synthetic inputs (host) ----
    @classmethod
    def stencil27(cls, nx, ny=None, nz=None, beta=0.0):
        ny, nz = ny or nx, nz or nx
        nnz = ctypes.c_int64()
        check(lib.f3r_gen_stencil27_nnz(nx, ny, nz, ctypes.byref(nnz)), "gen_stencil27_nnz")
        n = nx * ny * nz
        rptr = np.empty(n + 1, np.int32)
        idx = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float64)
        check(lib.f3r_gen_stencil27(nx, ny, nz, beta, _p(rptr), _p(idx), _p(val)), "gen_stencil27")
        return cls(n, rptr, idx, val)

    @classmethod
    def convdiff7(cls, nx, ny=None, nz=None, sigma=1.5, corr=8.0, peclet=10.0, seed=0):
        ny, nz = ny or nx, nz or nx
        nnz = ctypes.c_int64()
        check(lib.f3r_gen_convdiff7_nnz(nx, ny, nz, ctypes.byref(nnz)), "gen_convdiff7_nnz")
        n = nx * ny * nz
        rptr = np.empty(n + 1, np.int32)
        idx = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float64)
        check(lib.f3r_gen_convdiff7(nx, ny, nz, sigma, corr, peclet, seed, _p(rptr), _p(idx), _p(val)),
              "gen_convdiff7")
        return cls(n, rptr, idx, val)

    @classmethod
    def dominant(cls, npts, nx, ny=None, nz=None, margin=1.0):
        ny, nz = ny or nx, nz or nx
        nnz = ctypes.c_int64()
        check(lib.f3r_gen_dominant_nnz(npts, nx, ny, nz, ctypes.byref(nnz)), "gen_dominant_nnz")
        n = nx * ny * nz
        rptr = np.empty(n + 1, np.int32)
        idx = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float64)
        check(lib.f3r_gen_dominant(npts, nx, ny, nz, margin, _p(rptr), _p(idx), _p(val)), "gen_dominant")
        return cls(n, rptr, idx, val)


def uniform(n, lo=-1.0, hi=1.0, seed=0):
    out = np.empty(n, np.float64)
    check(lib.f3r_gen_uniform(n, lo, hi, seed, _p(out)), "gen_uniform")
    return out


class algorithm:
    """sparse/matrix/algorithm.hpp"""

    @staticmethod
    def scaling(A):
        """a_ij /= sqrt|a_ii| sqrt|a_jj| in place (algorithm.hpp:45-59)"""
        check(lib.f3r_csr_scaling(A.n, _p(A.rptr), _p(A.idx), _p(A.val)), "csr_scaling")
        A.invalidate()
        return A


class io:
    """sparse/matrix/io.hpp"""

    class mm:
        """MatrixMarket files (io::mm, io.hpp:104-217)"""

        @staticmethod
        def read_matrix(path):
            """a coordinate file as a host CSR (io::mm::read_matrix + io::CSR)"""
            h = ctypes.c_void_p()
            check(lib.f3r_mm_read(None, str(path).encode(), ctypes.byref(h)), "mm_read")
            try:
                nr, nc = ctypes.c_int32(), ctypes.c_int32()
                check(lib.f3r_mat_shape(h, ctypes.byref(nr), ctypes.byref(nc)), "mat_shape")
                nnz = ctypes.c_int64()
                check(lib.f3r_mat_info(h, None, ctypes.byref(nnz), None, None, None), "mat_info")
                rptr = np.empty(nr.value + 1, np.int32)
                idx = np.empty(nnz.value, np.int32)
                val = np.empty(nnz.value, np.float64)
                check(lib.f3r_mat_csr_get(h, _p(rptr), _p(idx), _p(val)), "mat_csr_get")
            finally:
                lib.f3r_mat_destroy(h)
            return CSR(nr.value, rptr, idx, val, ncols=nc.value)

        @staticmethod
        def read_vector(path):
            """an "array real general" file of one column"""
            n = ctypes.c_int64()
            check(lib.f3r_mm_read_vector(str(path).encode(), 0, None, ctypes.byref(n)), "mm_read_vector")
            out = np.empty(n.value, np.float64)
            check(lib.f3r_mm_read_vector(str(path).encode(), out.size, _p(out), ctypes.byref(n)), "mm_read_vector")
            return out

        @staticmethod
        def read_system(path, seed=0):
            """(A, b) of io::mm::read_system: b from <name>_b.mtx when that file
            exists, else U[0, 1) from std::mt19937_64(seed), the reference
            drivers' fallback (work/src/f3r.cpp:39)"""
            import os
            A = io.mm.read_matrix(path)
            buf = ctypes.create_string_buffer(len(str(path)) + 8)
            check(lib.f3r_mm_rhs_path(str(path).encode(), buf, len(buf)), "mm_rhs_path")
            bpath = buf.value.decode()
            if os.path.exists(bpath):
                b = io.mm.read_vector(bpath)
                if b.size != A.n:
                    raise F3RError(_capi.ERR_ARG, f"mm.read_system: {bpath} has {b.size} entries, "
                                                  f"the matrix {A.n} rows")
            else:
                b = np.empty(A.n, np.float64)
                check(lib.f3r_gen_uniform_mt(A.n, 0.0, 1.0, seed, _p(b)), "gen_uniform_mt")
            return A, b


def sell_offset_layout(A, C=64, K=4):
    """host chunk-offset layout of CSR A (sigma = 1): dict of cptr, cwidth,
    iptr, idx, coff, val"""
    ns, ni, nc = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    args = (A.n, _p(A.rptr), _p(A.idx), _p(A.val), C, K)
    check(lib.f3r_sell_offset_layout(*args, ctypes.byref(ns), ctypes.byref(ni), ctypes.byref(nc),
                                     None, None, None, None, None, None), "sell_offset_layout")
    nch = (A.n + C - 1) // C
    out = {"cptr": np.empty(nch + 1, np.int32), "cwidth": np.empty(nch, np.int32), "iptr": np.empty(nch, np.int32),
           "idx": np.empty(max(ni.value, 1), np.int32), "coff": np.empty(max(nc.value, 1), np.int32),
           "val": np.empty(max(ns.value, 1), np.float64)}
    check(lib.f3r_sell_offset_layout(*args, ctypes.byref(ns), ctypes.byref(ni), ctypes.byref(nc),
                                     *[_p(out[k]) for k in ("cptr", "cwidth", "iptr", "idx", "coff", "val")]),
          "sell_offset_layout")
    out["idx"], out["coff"], out["val"] = out["idx"][:ni.value], out["coff"][:nc.value], out["val"][:ns.value]
    return out


def sell_layout(A, C=32, sigma=1, K=1):
    """Host SELL-C-sigma layout (cptr, cwidth, idx, val, perm) of a CSR."""
    ns = ctypes.c_int64()
    check(lib.f3r_sell_nslots(A.n, _p(A.rptr), C, sigma, K, ctypes.byref(ns)), "sell_nslots")
    nch = (A.n + C - 1) // C
    cptr = np.empty(nch + 1, np.int32)
    cw = np.empty(max(nch, 1), np.int32)
    sidx = np.empty(ns.value, np.int32)
    sval = np.empty(ns.value, np.float64)
    perm = np.empty(A.n, np.int32)
    check(lib.f3r_sell_layout(A.n, _p(A.rptr), _p(A.idx), _p(A.val), C, sigma, K, _p(cptr), _p(cw),
                              _p(sidx), _p(sval), _p(perm)), "sell_layout")
    return cptr, cw[:nch], sidx, sval, perm


class CSRDevice:
    """CSR<T, device>(CSR<double, host>) (sparse/matrix/csr.hpp:12-50): the CSR
    on a context with values in `dtype`; SpMV with `tpr` lanes per row (a
    power of two up to 64; 0 = by the mean row length; 1 = each row summed in
    CSR order, bit-identical to the SELL kernels)."""

    def __init__(self, ctx, A, dtype="f64", tpr=0):
        self.ctx, self.A, self.dtype, self.tpr = ctx, A, _dtype(dtype), int(tpr)
        h = ctypes.c_void_p()
        check(lib.f3r_mat_cast(A.handle(ctx), self.dtype, ctypes.byref(h)), "mat_cast")
        self.handle = h
        check(lib.f3r_mat_set_threads_per_row(h, self.tpr), "mat_set_threads_per_row")

    def operate(self, x, y):
        check(lib.f3r_spmv(self.handle, x.handle, y.handle), "spmv")

    def compute_residual(self, b, x, r):
        check(lib.f3r_residual(self.handle, b.handle, x.handle, r.handle), "residual")

    def time_spmv(self, x, y, reps=20):
        ms = ctypes.c_double()
        check(lib.f3r_spmv_time(self.handle, x.handle, y.handle, reps, ctypes.byref(ms)), "spmv_time")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_mat_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class SELL:
    """SELL<w, T, device> / SELLsig (sparse/matrix/sell.hpp:35-166).  lanes:
    threads per row of the SpMV, 1 (CSR order) or K (F3R_SELL_TPR)"""

    def __init__(self, ctx, A, dtype="f64", C=64, sigma=1, K=1, offsets=False, device=False, lds=False,
                 lanes=1, _handle=None):
        self.ctx = ctx
        self.A = A
        self.dtype = _dtype(dtype)
        self.C, self.sigma, self.K, self.offsets, self.lds, self.lanes = C, sigma, K, offsets, lds, lanes
        if lanes not in (1, K):
            raise ValueError("SELL: lanes per row must be 1 or K")
        if _handle is None:
            h = ctypes.c_void_p()
            flags = (1 if offsets else 0) | (2 if device else 0) | (4 if lds else 0) | (8 if lanes > 1 else 0)
            check(lib.f3r_mat_to_sell_ex(ctx.handle, A.handle(ctx), self.dtype, C, sigma, K, flags,
                                         ctypes.byref(h)), "mat_to_sell")
            self.handle = h
        else:
            self.handle = _handle

    def cast(self, dtype):
        """cross-precision copy sharing the index arrays (sell.hpp:51-53)"""
        h = ctypes.c_void_p()
        check(lib.f3r_mat_cast(self.handle, _dtype(dtype), ctypes.byref(h)), "mat_cast")
        return SELL(self.ctx, self.A, dtype, self.C, self.sigma, self.K, self.offsets, lanes=self.lanes, _handle=h)

    def nslots(self):
        ns = ctypes.c_int64()
        check(lib.f3r_mat_info(self.handle, None, None, ctypes.byref(ns), None, None), "mat_info")
        return ns.value

    def lds_blocks(self):
        """(blocks of 256 rows, blocks whose x window is staged in LDS)"""
        b, st = ctypes.c_int32(), ctypes.c_int32()
        check(lib.f3r_mat_lds_blocks(self.handle, ctypes.byref(b), ctypes.byref(st)), "mat_lds_blocks")
        return b.value, st.value

    def where(self):
        """where the layout was built, "host" or "device" (a string)"""
        w = ctypes.c_int()
        check(lib.f3r_mat_setup_where(self.handle, ctypes.byref(w)), "mat_setup_where")
        return "device" if w.value == 1 else "host"

    def chunks(self):
        """(chunks, chunks in the chunk-offset format)"""
        c, o = ctypes.c_int32(), ctypes.c_int32()
        check(lib.f3r_mat_chunks(self.handle, ctypes.byref(c), ctypes.byref(o)), "mat_chunks")
        return c.value, o.value

    def operate(self, x, y):
        check(lib.f3r_spmv(self.handle, x.handle, y.handle), "spmv")

    def compute_residual(self, b, x, r):
        check(lib.f3r_residual(self.handle, b.handle, x.handle, r.handle), "residual")

    def time_spmv(self, x, y, reps=20):
        ms = ctypes.c_double()
        check(lib.f3r_spmv_time(self.handle, x.handle, y.handle, reps, ctypes.byref(ms)), "spmv_time")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_mat_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class _Precond:
    SETUP_WHERE = {0: "host", 1: "device", 2: "host (device capacity exceeded)"}

    @property
    def where(self):
        """where the factors were built: "host", "device", or "host (device
        capacity exceeded)" when the device set-up handed the rows back"""
        w = ctypes.c_int()
        check(lib.f3r_prec_setup_where(self.handle, ctypes.byref(w)), "prec_setup_where")
        return self.SETUP_WHERE[w.value]

    def factors(self):
        out = []
        for w in (0, 1):
            nnz = ctypes.c_int64()
            check(lib.f3r_prec_factor_nnz(self.handle, w, ctypes.byref(nnz)), "factor_nnz")
            r = np.empty(self.n + 1, np.int32)
            i = np.empty(nnz.value, np.int32)
            v = np.empty(nnz.value, np.float64)
            check(lib.f3r_prec_factor_get(self.handle, w, _p(r), _p(i), _p(v)), "factor_get")
            out.append((r, i, v))
        return out

    def operate(self, x, y):
        check(lib.f3r_prec_apply(self.handle, x.handle, y.handle), "prec_apply")

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_prec_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class SDAINV(_Precond):
    """SDAINV<T, device>(A, {tol, acc}) (sparse/preconditioner/ainv.hpp:129-147).
    factors() -> (Z, W^T) as CSR triples.  ctx=None builds the host factors only;
    device=True factorises on the GPU (bit-identical factors)."""

    def __init__(self, A, tol=0.1, acc=1.0, dtype="f16", ctx=None, device=False):
        self.n = A.n
        self.kind = 0
        self.dtype = _dtype(dtype)
        h = ctypes.c_void_p()
        check(lib.f3r_sdainv_create_ex(ctx.handle if ctx else None, A.handle(ctx), tol, acc, self.dtype,
                                       1 if device else 0, ctypes.byref(h)), "sdainv_create")
        self.handle = h


class ILU0(_Precond):
    """ILU(0) on `blocks` diagonal blocks + split (ilu.hpp:128-164,
    algorithm.hpp:8-43).  factors() -> (L, U)."""

    def __init__(self, A, blocks=1, acc=1.0, dtype="f16", ctx=None, device=False):
        self.n = A.n
        self.kind = 1
        self.dtype = _dtype(dtype)
        h = ctypes.c_void_p()
        check(lib.f3r_ilu0_create_ex(ctx.handle if ctx else None, A.handle(ctx), blocks, acc, self.dtype,
                                     1 if device else 0, ctypes.byref(h)), "ilu0_create")
        self.handle = h


solve_res_t = namedtuple("solve_res_t", "is_solved res_iter res_nrm2 nrm_b solve_s precond_calls outer_cycles")

VARIANTS = {"half": 0, "float": 1, "double": 2}
# nesting-depth ablations of work/src/depth.cpp:122-167 (0: the F3R nest)
DEPTHS = {"F3R": 0, "F2": 1, "F2H": 2, "F3": 3, "F3H": 4, "F4": 5}


class F3R:
    """Restarted<double>(A, FGMRES<double>(A, FGMRES<float>(A32, FGMRES<float>(A16,
    Richardson(M)))) of work/src/f3r-gpu.cpp:210-266 on one GPU.

    static_weight > 0: the Richardson level with that fixed weight and no probes
    (work/src/static.cpp:99-119).  depth = "F2", "F2H", "F3", "F3H" or "F4": the
    nesting-depth ablations of work/src/depth.cpp:122-167 instead of the F3R
    nest (M applied directly under fp32 / fp16-basis FGMRES levels).  lanes:
    every SpMV with K threads per row (index = 0; the sums are reassociated)."""

    def __init__(self, ctx, A, M, m=(100, 8, 4, 2), c=64, restart=20, eps=1e-8, variant="half",
                 range_scale=0.0, C=0, sigma=0, K=0, use_graph=True, history=True, grid=0, index=1,
                 cap_spmv=False, device_build=False, depth=0, static_weight=0.0, lanes=False):
        self.ctx = ctx
        self.A = A
        self.M = M
        p = _capi.Params()
        check(lib.f3r_params_default(ctypes.byref(p)), "params_default")
        p.m1, p.m2, p.m3, p.m4 = m
        p.c, p.restart, p.eps = c, restart, eps
        p.variant = VARIANTS[variant] if isinstance(variant, str) else int(variant)
        p.range_scale = range_scale
        p.C, p.sigma, p.K = C, sigma, K
        p.use_graph = 1 if use_graph else 0
        p.flags = (1 if history else 0) | (2 if cap_spmv else 0) | (4 if device_build else 0) | \
            (32 if lanes else 0)
        p.grid = grid
        p.index = index
        p.depth = DEPTHS[depth] if isinstance(depth, str) else int(depth)
        p.static_weight = float(static_weight)
        self.params = p
        h = ctypes.c_void_p()
        check(lib.f3r_solver_create(ctx.handle, A.handle(ctx), M.handle, ctypes.byref(p), ctypes.byref(h)),
              "solver_create")
        self.handle = h

    def solve(self, b, x):
        r = _capi.Result()
        check(lib.f3r_solve(self.handle, b.handle, x.handle, ctypes.byref(r)), "solve")
        return solve_res_t(bool(r.is_solved), r.res_iter, r.res_nrm2, r.nrm_b, r.solve_s, r.precond_calls,
                           r.outer_cycles)

    def layouts_where(self):
        """where the matrix layouts were built, "host" or "device" (a string)"""
        w = ctypes.c_int()
        check(lib.f3r_solver_setup_where(self.handle, ctypes.byref(w)), "solver_setup_where")
        return "device" if w.value == 1 else "host"

    def history(self):
        ln = ctypes.c_int32()
        check(lib.f3r_solver_history(self.handle, None, 0, ctypes.byref(ln)), "history")
        out = np.empty(ln.value, np.float64)
        check(lib.f3r_solver_history(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ln.value, ctypes.byref(ln)), "history")
        return out

    def capture(self, max_vectors, every=1):
        """record the innermost level's input vectors of later solves (every
        `every`-th call, up to `max_vectors`; 0 stops)"""
        check(lib.f3r_solver_capture(self.handle, max_vectors, every), "solver_capture")

    def captured(self):
        """the recorded vectors, (count, n) float32 in the matrix's row order"""
        k = ctypes.c_int32()
        check(lib.f3r_solver_captured(self.handle, ctypes.byref(k), None), "solver_captured")
        out = np.empty((k.value, self.A.n), np.float32)
        if k.value:
            check(lib.f3r_solver_captured(self.handle, ctypes.byref(k), _p(out)), "solver_captured")
        return out

    def level_apply(self, level, vin, vout):
        check(lib.f3r_solver_level_apply(self.handle, level, vin.handle, vout.handle), "level_apply")

    def time_kernel(self, which, reps=20, cold=False):
        """mean ms per launch and algorithmic bytes of kernel `which`
        (0 fp16 SpMV, 1 fp16-value/fp32-vector SpMV, 2 fp32 SpMV, 3 fp64 SpMV,
        4 / 5 the level-3 SpMV + CGS dots / CGS update of the last step,
        6 / 7 the same over the solve's mix of steps, 8 the fused Richardson
        step, 9 the innermost entry); cold: each launch alone, after a 512 MiB
        pass that empties the caches"""
        ms, by = ctypes.c_double(), ctypes.c_double()
        w = which | (0x100 if cold else 0)
        check(lib.f3r_solver_time_kernel(self.handle, w, reps, ctypes.byref(ms), ctypes.byref(by)),
              "time_kernel")
        return ms.value, by.value

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_solver_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ------------------------------------------------- composed nests ---
# The reference composes its solvers from objects (sparse/models.hpp:46-267,
# solver/gmres.hpp:56-135, solver/restart.hpp:8-37):
#   Restarted(A, FGMRES(A, FGMRES(A32, FGMRES(A16, AutoRichardson(A16, M, m4, c), m3), m2), 100, True), 20)
# These descriptors build that chain on the device (f3r_nest_create).

class Operator:
    """Operator<L>(shape, lambda) (sparse/models.hpp:46-126): fn(in, out) on
    device vector views of `dtype`; fp16 views carry the solver's power-of-two
    range factor (the operator must then be linear)."""

    def __init__(self, ctx, n, fn, dtype="f32"):
        self.ctx, self.n, self.dtype, self.fn = ctx, int(n), _dtype(dtype), fn
        self.error = None

        def tramp(user, vin, vout):
            try:
                fn(vector(ctx, self.dtype, self.n, _handle=ctypes.c_void_p(vin), _borrowed=True),
                   vector(ctx, self.dtype, self.n, _handle=ctypes.c_void_p(vout), _borrowed=True))
                return 0
            except Exception as e:  # re-raised by Restarted.solve
                self.error = e
                return -1
        self._cb = _capi.APPLY_FN(tramp)
        h = ctypes.c_void_p()
        check(lib.f3r_op_create(ctx.handle, self.n, self.dtype, self._cb, None, ctypes.byref(h)), "op_create")
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_op_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class AutoRichardson:
    """AutoRichardson<T, L>(A, M, {max_iter, period}) (sparse/solver/richardson.hpp:96-156):
    the adaptive fp16 Richardson of work/src/f3r-gpu.cpp:229-260 on the fp16
    matrix with the fp16 preconditioner M; static_weight > 0 fixes the weights
    (work/src/static.cpp:99-119)"""

    def __init__(self, A16, M, max_iter, period, static_weight=0.0):
        self.A16, self.M, self.m4, self.c, self.w = A16, M, int(max_iter), int(period), float(static_weight)


def _mat_dtype(A):
    return A.dtype if isinstance(A, SELL) else _dtype(A)


class FGMRES:
    """FGMRES<T, L>(A, M, {max_iter, normalize}) (sparse/solver/gmres.hpp:56-135).
    T = "f32" (fp32 basis) or "f16" (fp16 storage, fp32 arithmetic) for an
    inner level, "f64" for the outer level; A is a SELL (its precision is the
    level's matrix precision) or a dtype; M another FGMRES, an AutoRichardson,
    a preconditioner (SDAINV, ILU0) or an Operator."""

    def __init__(self, A, M, max_iter, normalize=False, T="f32"):
        self.A, self.M, self.m, self.normalize, self.T = A, M, int(max_iter), bool(normalize), _dtype(T)

    def chain(self):
        """(levels below this one's basis, innermost) with this level first"""
        levels = [] if self.T == F64 else [(self.m, self.T, _mat_dtype(self.A))]
        node = self.M
        while isinstance(node, FGMRES):
            levels.append((node.m, node.T, _mat_dtype(node.A)))
            node = node.M
        return levels, node


class Restarted:
    """Restarted<double, L>(A, S, {max_iter}) (sparse/solver/restart.hpp:8-37) over
    the fp64 outer FGMRES(A, ..., m1, normalize=True); A is the CSR matrix as
    solved.  Layout keywords (C, sigma, K, index, range_scale) as for F3R."""

    def __init__(self, ctx, A, S, max_iter=20, eps=1e-8, C=0, sigma=0, K=0, index=1, range_scale=0.0):
        if not (isinstance(S, FGMRES) and S.T == F64 and S.normalize):
            raise ValueError("Restarted: the outer level is FGMRES(A, ..., T='f64', normalize=True)")
        levels, inner = S.chain()
        self.ctx, self.A, self.keep = ctx, A, (S, inner)
        p = _capi.Params()
        check(lib.f3r_params_default(ctypes.byref(p)), "params_default")
        p.m1, p.restart, p.eps = S.m, int(max_iter), eps
        p.C, p.sigma, p.K, p.index, p.range_scale = C, sigma, K, index, range_scale
        lv = (_capi.Level * len(levels))(*[_capi.Level(m, b, mt) for (m, b, mt) in levels])
        inn = _capi.Inner()
        if isinstance(inner, AutoRichardson):
            inn.kind, inn.P, inn.m4, inn.c, inn.static_weight = 1, inner.M.handle, inner.m4, inner.c, inner.w
        elif isinstance(inner, Operator):
            inn.kind, inn.op = 2, inner.handle
        else:
            inn.kind, inn.P = 0, inner.handle
        self._op = inner if isinstance(inner, Operator) else None
        h = ctypes.c_void_p()
        check(lib.f3r_nest_create(ctx.handle, A.handle(ctx), lv, len(levels), ctypes.byref(inn), ctypes.byref(p),
                                  ctypes.byref(h)), "nest_create")
        self.handle = h

    def solve(self, b, x, cond=None):
        """Solver::solve(b, x, cond) (sparse/models.hpp:191-267); cond(i, |e|) -> bool,
        None: |e| / ||b|| < eps"""
        r = _capi.Result()
        if cond is None:
            st = lib.f3r_solve(self.handle, b.handle, x.handle, ctypes.byref(r))
        else:
            raised = []

            def call(user, i, e):
                # an exception must not cross the C ABI: keep it, stop the solve
                try:
                    return 1 if cond(int(i), float(e)) else 0
                except BaseException as exc:
                    raised.append(exc)
                    return 1
            cb = _capi.COND_FN(call)
            st = lib.f3r_solve_cond(self.handle, b.handle, x.handle, cb, None, ctypes.byref(r))
            if raised:
                raise raised[0]
        if st != 0 and self._op is not None and self._op.error is not None:
            err, self._op.error = self._op.error, None
            raise err
        check(st, "solve")
        return solve_res_t(bool(r.is_solved), r.res_iter, r.res_nrm2, r.nrm_b, r.solve_s, r.precond_calls,
                           r.outer_cycles)

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_solver_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Batch:
    """Independent systems solved in lockstep on one GPU (f3r_batch_*): segment s
    is (As[s], Ps[s]) with its own right-hand side; a system with k right-hand
    sides is listed k times.  Every launch covers all segments; reductions,
    Hessenberg/Givens updates and Richardson weights are per segment.  Ps are
    SDAINV preconditioners (fp16).  Solves the HALF nest of F3R.  use_graph:
    the level-2 body is captured in a HIP graph per probe pattern and
    replayed (the same launches, the same bits).  lanes: the SpMVs with K
    threads per row (needs index = 0 and K >= 2; the sums are reassociated)."""

    def __init__(self, ctx, As, Ps, m=(100, 8, 4, 2), c=64, restart=20, eps=1e-8, C=0, K=0, index=1,
                 range_scale=0.0, use_graph=True, lanes=False):
        self.ctx = ctx
        self.n = [A.n for A in As]
        p = _capi.Params()
        check(lib.f3r_params_default(ctypes.byref(p)), "params_default")
        p.m1, p.m2, p.m3, p.m4 = m
        p.c, p.restart, p.eps = c, restart, eps
        p.C, p.K, p.index, p.range_scale = C, K, index, range_scale
        p.use_graph = 1 if use_graph else 0
        p.flags = (p.flags & ~16) | (16 if lanes else 0)
        ns = len(As)
        ah = (ctypes.c_void_p * ns)(*[A.handle(ctx) for A in As])
        ph = (ctypes.c_void_p * ns)(*[P.handle for P in Ps])
        self._keep = (As, Ps)
        h = ctypes.c_void_p()
        check(lib.f3r_batch_create(ctx.handle, ns, ah, ph, ctypes.byref(p), ctypes.byref(h)), "batch_create")
        self.handle = h

    def info(self):
        """(padded rows of all segments, segments)"""
        rows, segs = ctypes.c_int64(), ctypes.c_int32()
        check(lib.f3r_batch_info(self.handle, ctypes.byref(rows), ctypes.byref(segs)), "batch_info")
        return rows.value, segs.value

    def solve(self, bs, xs):
        """solves every segment: bs, xs lists of fp64 vectors; returns one
        solve_res_t per segment (solve_s: the whole batch)"""
        ns = len(bs)
        bh = (ctypes.c_void_p * ns)(*[b.handle for b in bs])
        xh = (ctypes.c_void_p * ns)(*[x.handle for x in xs])
        res = (_capi.Result * ns)()
        check(lib.f3r_batch_solve(self.handle, bh, xh, res), "batch_solve")
        return [solve_res_t(bool(r.is_solved), r.res_iter, r.res_nrm2, r.nrm_b, r.solve_s, r.precond_calls,
                            r.outer_cycles) for r in res]

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
