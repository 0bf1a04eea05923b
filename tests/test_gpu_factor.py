"""Device factorisation of SDAINV and ILU(0) through the C ABI: the factors
must equal the oracle's bit for bit (index arrays and fp64 values), on
stencil and random nonsymmetric patterns, with drop tolerances that keep the
rows short and ones that fill them past each device row capacity (64, 256
and 1024 entries), where the device must hand the rows back to the host."""
import numpy as np
import pytest

import oracle_lib as O
import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import f3r_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _matrix(kind):
    import f3r_amd as F
    if kind == "27pt":
        A = F.CSR.stencil27(10, beta=0.5)
    elif kind == "convdiff":
        A = F.CSR.convdiff7(9, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
    elif kind == "lognormal":
        A = F.CSR.convdiff7(9, sigma=2.0, corr=2.0, peclet=0.0, seed=2)
    else:
        n, rptr, idx, val = problems.random_sparse(700, 7, seed=5)
        A = F.CSR(n, rptr, idx, val)
    F.algorithm.scaling(A)
    return A


def _same(got, want):
    for g, w, part in zip(got, want, ("rptr", "idx", "val")):
        assert g.dtype == w.dtype or part != "val"
        assert np.array_equal(g, w), part


@pytest.mark.parametrize("kind,tol,acc", [("27pt", 0.1, 1.0), ("27pt", 0.02, 1.0), ("27pt", 0.005, 1.3),
                                          ("convdiff", 0.1, 1.0), ("convdiff", 0.01, 1.1),
                                          ("lognormal", 0.1, 1.0), ("lognormal", 0.003, 1.0),
                                          ("random", 0.05, 1.0), ("random", 0.005, 1.2)])
def test_sdainv_device_matches_oracle(ctx, kind, tol, acc):
    import f3r_amd as F
    A = _matrix(kind)
    print(f"sdainv device {kind} n={A.n} tol={tol} acc={acc}", flush=True)
    P = F.SDAINV(A, tol, acc, "f16", ctx=ctx, device=True)
    assert P.where == "device"
    (z, w) = P.factors()
    (oz, ow) = O.sdainv(A.n, A.rptr, A.idx, A.val, tol, acc)
    _same(z, oz)
    _same(w, ow)


@pytest.mark.parametrize("tol,lo,hi", [(1e-5, 64, 256), (1e-7, 256, 1024)])
def test_sdainv_device_row_capacities(ctx, tol, lo, hi):
    """rows past 64 (then 256) entries: the build is redone in the next slot size"""
    import f3r_amd as F
    A = _matrix("lognormal")
    print(f"sdainv device lognormal n={A.n} tol={tol}", flush=True)
    P = F.SDAINV(A, tol, 1.0, "f64", ctx=ctx, device=True)
    (z, w) = P.factors()
    longest = max(np.diff(w[0]).max(), np.bincount(z[1], minlength=A.n).max())
    assert lo < longest <= hi
    assert P.where == "device"
    (oz, ow) = O.sdainv(A.n, A.rptr, A.idx, A.val, tol, 1.0)
    _same(z, oz)
    _same(w, ow)


def test_sdainv_device_hands_back_to_host(ctx):
    """no drop on 1331 rows: w_i fills past 1024 entries"""
    import f3r_amd as F
    A = F.CSR.stencil27(11, beta=0.5)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.0, 1.0, "f16", ctx=ctx, device=True)
    assert P.where == "host (device capacity exceeded)"
    (z, w) = P.factors()
    assert np.diff(w[0]).max() > 1024
    (oz, ow) = O.sdainv(A.n, A.rptr, A.idx, A.val, 0.0, 1.0)
    _same(z, oz)
    _same(w, ow)


@pytest.mark.parametrize("kind,blocks,acc", [("27pt", 1, 1.0), ("27pt", 8, 1.2), ("convdiff", 1, 1.0),
                                             ("convdiff", 5, 1.0), ("random", 1, 1.0), ("random", 3, 1.1)])
def test_ilu0_device_matches_oracle(ctx, kind, blocks, acc):
    import f3r_amd as F
    A = _matrix(kind)
    print(f"ilu0 device {kind} n={A.n} blocks={blocks} acc={acc}", flush=True)
    P = F.ILU0(A, blocks, acc, "f16", ctx=ctx, device=True)
    assert P.where == "device"
    (L, U) = P.factors()
    (oL, oU) = O.ilu0(A.n, A.rptr, A.idx, A.val, blocks, acc)
    _same(L, oL)
    _same(U, oU)


@pytest.mark.parametrize("prec", ["sdainv", "ilu0"])
def test_device_factors_give_identical_solves(ctx, prec):
    import f3r_amd as F
    A = F.CSR.convdiff7(12, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
    F.algorithm.scaling(A)
    b = F.vector.from_numpy(ctx, F.uniform(A.n, -1.0, 1.0, seed=9))
    xs = []
    for dev in (False, True):
        P = (F.SDAINV(A, 0.05, 1.0, "f16", ctx=ctx, device=dev) if prec == "sdainv"
             else F.ILU0(A, 4, 1.0, "f16", ctx=ctx, device=dev))
        S = F.F3R(ctx, A, P)
        x = F.vector(ctx, "f64", A.n).fill(0.0)
        r = S.solve(b, x)
        assert r.is_solved
        xs.append((r.res_iter, x.download()))
    assert xs[0][0] == xs[1][0]
    assert np.array_equal(xs[0][1], xs[1][1])
