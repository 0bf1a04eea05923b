"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests, smoke() and bench.py's
cpu_baseline leg.  The oracle is the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

F64, F32, F16, BF16 = 0, 1, 2, 3
_NP = {F64: np.float64, F32: np.float32, F16: np.float16, BF16: np.uint16}  # bf16 as its bits


class Opts(ctypes.Structure):
    _fields_ = [("m1", ctypes.c_int32), ("m2", ctypes.c_int32), ("m3", ctypes.c_int32),
                ("m4", ctypes.c_int32), ("c", ctypes.c_int32), ("restart", ctypes.c_int32),
                ("eps", ctypes.c_double), ("variant", ctypes.c_int32),
                ("ref_arith", ctypes.c_int32), ("range_scale", ctypes.c_double),
                ("precond", ctypes.c_int32), ("tol", ctypes.c_double), ("acc", ctypes.c_double),
                ("bj_blocks", ctypes.c_int32), ("max_outer", ctypes.c_int32),
                ("nthreads", ctypes.c_int32), ("static_weight", ctypes.c_double),
                ("depth", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("is_solved", ctypes.c_int32), ("res_iter", ctypes.c_int32),
                ("res_nrm2", ctypes.c_double), ("nrm_b", ctypes.c_double),
                ("expl_res", ctypes.c_double), ("solve_s", ctypes.c_double),
                ("setup_s", ctypes.c_double), ("precond_calls", ctypes.c_int64),
                ("nthreads", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_sell_nslots.restype = ctypes.c_int64
        L.orc_factor_nnz.restype = ctypes.c_int64
        L.orc_sdainv.restype = ctypes.c_void_p
        L.orc_ilu0.restype = ctypes.c_void_p
        L.orc_f2h.restype = ctypes.c_uint16
        L.orc_h2f.restype = ctypes.c_float
        L.orc_d2h.restype = ctypes.c_uint16
        L.orc_d2b.restype = ctypes.c_uint16
        L.orc_d2b.argtypes = [ctypes.c_double]
        L.orc_f2b.restype = ctypes.c_uint16
        L.orc_f2b.argtypes = [ctypes.c_float]
        L.orc_b2f.restype = ctypes.c_float
        L.orc_b2f.argtypes = [ctypes.c_uint16]
        L.orc_sdainv.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_ilu0.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_int32, ctypes.c_double]
        L.orc_factor_nnz.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.orc_factor_get.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
        L.orc_factor_free.argtypes = [ctypes.c_void_p]
        L.orc_bj_solve.argtypes = [ctypes.c_void_p] * 4
        L.orc_bj_solve16.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int32]
        _lib = L
    return _lib


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def scaling(n, rptr, idx, val):
    val = np.array(val, dtype=np.float64)
    lib().orc_scaling(n, P(rptr), P(idx), P(val))
    return val


def csr_to_sell(n, rptr, idx, val, C, sigma):
    ns = lib().orc_sell_nslots(n, P(rptr), C, sigma)
    nch = (n + C - 1) // C
    cptr = np.zeros(nch + 1, np.int32)
    sidx = np.zeros(ns, np.int32)
    sval = np.zeros(ns, np.float64)
    perm = np.zeros(n, np.int32)
    lib().orc_csr_to_sell(n, P(rptr), P(idx), P(val), C, sigma, P(cptr), P(sidx), P(sval), P(perm))
    return cptr, sidx, sval, perm


def spmv(n, rptr, idx, val, x, val_prec, x_prec, y_prec, acc_prec, b=None):
    y = np.zeros(n, _NP[y_prec])
    x = np.ascontiguousarray(x, _NP[x_prec])
    if b is not None:
        b = np.ascontiguousarray(b, _NP[y_prec])
    lib().orc_spmv(n, P(rptr), P(idx), P(np.ascontiguousarray(val, np.float64)), val_prec,
                   P(x), x_prec, P(y), y_prec, acc_prec, 1 if b is not None else 0,
                   P(b) if b is not None else None, y_prec)
    return y


def spmv_ld(n, rptr, idx, val, x):
    y = np.zeros(n)
    lib().orc_spmv_ld(n, P(rptr), P(idx), P(val), P(np.ascontiguousarray(x, np.float64)), P(y))
    return y


def _factors(h, n):
    out = []
    for w in (0, 1):
        nnz = lib().orc_factor_nnz(h, w)
        r = np.zeros(n + 1, np.int32)
        i = np.zeros(nnz, np.int32)
        v = np.zeros(nnz, np.float64)
        lib().orc_factor_get(h, w, P(r), P(i), P(v))
        out.append((r, i, v))
    lib().orc_factor_free(h)
    return out


def sdainv(n, rptr, idx, val, tol=0.1, acc=1.0):
    """Returns ((Z rptr, idx, val), (Wt rptr, idx, val))."""
    h = lib().orc_sdainv(n, P(rptr), P(idx), P(val), tol, acc)
    return _factors(h, n)


def ilu0(n, rptr, idx, val, blocks=1, acc=1.0):
    """Returns ((L ...), (U ...)) of ILU(0) on `blocks` diagonal blocks."""
    h = lib().orc_ilu0(n, P(rptr), P(idx), P(val), blocks, acc)
    return _factors(h, n)


def f3r_opts(**kw):
    d = dict(m1=100, m2=8, m3=4, m4=2, c=64, restart=20, eps=1e-8, variant=0, ref_arith=0,
             range_scale=1.0, precond=0, tol=0.1, acc=1.0, bj_blocks=112, max_outer=0,
             nthreads=0, static_weight=0.0, depth=0)
    d.update(kw)
    return Opts(**d)


def f3r_solve(n, rptr, idx, val, b, **kw):
    o = f3r_opts(**kw)
    x = np.zeros(n)
    r = Result()
    hist = np.zeros(o.m1 * o.restart + o.restart + 2)
    lib().orc_f3r_solve(n, P(rptr), P(idx), P(val), P(np.ascontiguousarray(b, np.float64)),
                        P(x), ctypes.byref(o), ctypes.byref(r), P(hist))
    steps = r.res_iter + (1 if r.res_iter == 0 else 0)
    return x, r, hist


def level_apply(n, rptr, idx, val, level, vin, **kw):
    o = f3r_opts(**kw)
    out = np.zeros(n)
    lib().orc_level_apply(n, P(rptr), P(idx), P(val), ctypes.byref(o), level,
                          P(np.ascontiguousarray(vin, np.float64)), P(out))
    return out


def gmres(n, rptr, idx, val, b, m=50, max_iter=10000, tol=1e-12):
    x = np.zeros(n)
    rr = ctypes.c_double()
    it = lib().orc_gmres(n, P(rptr), P(idx), P(val), P(np.ascontiguousarray(b, np.float64)),
                         P(x), m, max_iter, ctypes.c_double(tol), ctypes.byref(rr))
    return x, it, rr.value


def mlp_forward(x, w1, b1, w2, b2, w3, b3):
    rows, nin = x.shape
    h1, h2 = w1.shape[0], w2.shape[0]
    y = np.zeros(rows)
    a = [np.ascontiguousarray(t, np.float64) for t in (x, w1, b1, w2, b2, w3, b3)]
    lib().orc_mlp_forward(rows, nin, h1, h2, *[P(t) for t in a], P(y))
    return y


def mlp_backward(x, w1, b1, w2, b2, w3, b3, g):
    rows, nin = x.shape
    h1, h2 = w1.shape[0], w2.shape[0]
    a = [np.ascontiguousarray(t, np.float64) for t in (x, w1, b1, w2, b2, w3, b3, g)]
    outs = [np.zeros_like(a[k]) for k in (1, 2, 3, 4, 5, 6)]
    lib().orc_mlp_backward(rows, nin, h1, h2, *[P(t) for t in a], *[P(t) for t in outs])
    return outs
