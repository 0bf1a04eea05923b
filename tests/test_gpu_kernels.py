"""GPU parity of the building blocks through the C ABI: SELL-C-sigma / CSR
SpMV in every precision combination the solve uses (bit-exact against the
oracle, which fuses multiply-adds in the same order), BLAS-1 reductions, and
the SDAINV / ILU(0) preconditioner applies."""
import numpy as np
import pytest

import oracle_lib as O
import problems as PB

pytestmark = pytest.mark.gpu

COMBOS = [  # (value, x, y) precisions; sums in fp64 for fp64 data, else fp32
    ("f64", "f64", "f64", 0, 0, 0, 0),
    ("f32", "f32", "f32", 1, 1, 1, 1),
    ("f16", "f32", "f32", 2, 1, 1, 1),
    ("f16", "f16", "f16", 2, 2, 2, 1),
    ("f16", "f16", "f32", 2, 2, 1, 1),
]
# (C, sigma, K, chunk offsets)
LAYOUTS = [(32, 1, 1, False), (64, 1, 4, False), (64, 256, 2, False), (128, 64, 4, False), (16, 1, 4, False),
           (64, 1, 4, True), (128, 1, 1, True), (256, 1, 2, True), (64, 256, 4, True), (64, 1, 8, True)]
NP = {"f64": np.float64, "f32": np.float32, "f16": np.float16}


@pytest.fixture(scope="module")
def ctx():
    import f3r_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _mats():
    import f3r_amd as F
    A = F.CSR.stencil27(10, 9, 11, beta=0.5)
    F.algorithm.scaling(A)
    yield "27pt", A
    n, rp, ix, v = PB.random_sparse(1000, 40, 5)   # ragged rows, 1..40 entries
    yield "ragged", F.CSR(n, rp, ix, v)
    B = F.CSR.convdiff7(13, sigma=1.5, corr=3.0, peclet=10.0, seed=1)
    F.algorithm.scaling(B)
    yield "convdiff", B


@pytest.mark.parametrize("C,sigma,K,offsets", LAYOUTS)
def test_sell_spmv_bitexact(ctx, C, sigma, K, offsets):
    import f3r_amd as F
    for name, A in _mats():
        x = np.random.default_rng(1).uniform(-1, 1, A.n)
        b = np.random.default_rng(2).uniform(-1, 1, A.n)
        for vd, xd, yd, vp, xp, yp, ap in COMBOS:
            print(f"spmv {name} {vd}/{xd}/{yd} C={C} sigma={sigma} K={K} offsets={offsets}", flush=True)
            S = F.SELL(ctx, A, vd, C=C, sigma=sigma, K=K, offsets=offsets)
            nch, noff = S.chunks()
            if offsets and sigma == 1 and name != "ragged":
                # stencils: every chunk but those near the first and last
                # rows (whose offsets would leave the matrix) uses offsets
                assert noff > 0 and noff >= nch // 2, (name, nch, noff)
            if not offsets:
                assert noff == 0
            xv = F.vector.from_numpy(ctx, x.astype(NP[xd]))
            yv = F.vector(ctx, yd, A.n)
            S.operate(xv, yv)
            got = yv.download()
            ref = O.spmv(A.n, A.rptr, A.idx, A.val, x.astype(NP[xd]), vp, xp, yp, ap)
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), (name, vd, xd, yd)
            if yd == xd:
                bv = F.vector.from_numpy(ctx, b.astype(NP[yd]))
                S.compute_residual(bv, xv, yv)
                ref = O.spmv(A.n, A.rptr, A.idx, A.val, x.astype(NP[xd]), vp, xp, yp, ap,
                             b=b.astype(NP[yd]))
                assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, "resid")


def test_cross_precision_cast_shares_layout(ctx):
    import f3r_amd as F
    A = F.CSR.stencil27(12, beta=0.5)
    F.algorithm.scaling(A)
    S64 = F.SELL(ctx, A, "f64", C=64, sigma=128, K=4)
    S16 = S64.cast("f16")
    assert S16.nslots() == S64.nslots()
    x = np.random.default_rng(0).uniform(-1, 1, A.n).astype(np.float16)
    xv = F.vector.from_numpy(ctx, x)
    yv = F.vector(ctx, "f16", A.n)
    S16.operate(xv, yv)
    ref = O.spmv(A.n, A.rptr, A.idx, A.val, x, 2, 2, 2, 1)
    assert np.array_equal(yv.download().view(np.uint16), ref.view(np.uint16))


def test_fp64_spmv_accuracy_64cube(ctx):
    # fp64 SpMV within 1e-14 relative of a long-double CSR sum, at 64^3
    import f3r_amd as F
    A = F.CSR.stencil27(64, beta=0.5)
    F.algorithm.scaling(A)
    x = np.random.default_rng(3).uniform(-1, 1, A.n)
    S = F.SELL(ctx, A, "f64", C=64, sigma=1, K=4)
    xv = F.vector.from_numpy(ctx, x)
    yv = F.vector(ctx, "f64", A.n)
    S.operate(xv, yv)
    ref = O.spmv_ld(A.n, A.rptr, A.idx, A.val, x)
    got = yv.download()
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    # CSR format on the device too
    Av = F.CSR(A.n, A.rptr, A.idx, A.val)
    from f3r_amd._capi import lib, check
    y2 = F.vector(ctx, "f64", A.n)
    check(lib.f3r_spmv(Av.handle(ctx), xv.handle, y2.handle))
    assert np.max(np.abs(y2.download() - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_blas1(ctx):
    import f3r_amd as F
    rng = np.random.default_rng(7)
    for n in (1, 255, 256, 1000, 100003, 3_000_001):
        a = rng.uniform(-1, 1, n)
        b = rng.uniform(-1, 1, n)
        av = F.vector.from_numpy(ctx, a)
        bv = F.vector.from_numpy(ctx, b)
        d = F.dot(av, bv)
        assert abs(d - a @ b) <= 1e-13 * n
        assert abs(F.nrm(av) - np.linalg.norm(a)) <= 1e-13 * np.sqrt(n)
        a32 = F.vector(ctx, "f32", n).copy(av)
        assert np.array_equal(a32.download(), a.astype(np.float32))
        a16 = F.vector(ctx, "f16", n).copy(av)
        assert np.array_equal(a16.download().view(np.uint16), a.astype(np.float16).view(np.uint16))
        d16 = F.dot(a16, a16)
        r = a.astype(np.float16).astype(np.float64)
        assert abs(d16 - r @ r) <= 1e-5 * (r @ r) + 1e-6
        bv.axpy(0.5, av)
        np.testing.assert_allclose(bv.download(), b + 0.5 * a, rtol=0, atol=1e-15)
        av.scale(3.0)
        np.testing.assert_allclose(av.download(), 3.0 * a, rtol=1e-16)
        av.fill(2.5)
        assert np.all(av.download() == 2.5)
    # views
    v = F.vector.from_numpy(ctx, np.arange(10.0))
    w = v.view(3, 4)
    assert np.array_equal(w.download(), np.arange(3.0, 7.0))


def test_empty_vectors_and_matrix(ctx):
    # zero-length vectors: every BLAS-1 call is a no-op and a dot is 0; the
    # solver refuses an empty matrix with an error status
    import f3r_amd as F
    for dt in ("f64", "f32", "f16", "bf16"):
        print(f"empty vectors {dt}", flush=True)
        a = F.vector(ctx, dt, 0).fill(1.0)
        b = F.vector(ctx, dt, 0).fill(2.0)
        a.axpy(3.0, b)
        assert F.dot(a, b) == 0.0
        assert a.download().size == 0
    A = F.CSR(0, np.array([0], np.int32), np.array([], np.int32), np.array([]))
    P = F.SDAINV(A, 0.1, 1.0)
    with pytest.raises(F.F3RError):
        F.F3R(ctx, A, P)


def test_sdainv_apply_bitexact(ctx):
    import f3r_amd as F
    A = F.CSR.convdiff7(12, sigma=2.0, corr=2.0, peclet=0.0, seed=3)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    (zr, zi, zv), (wr, wi, wv) = P.factors()
    x = np.random.default_rng(4).uniform(-1, 1, A.n).astype(np.float16)
    xv = F.vector.from_numpy(ctx, x)
    yv = F.vector(ctx, "f16", A.n)
    P.operate(xv, yv)
    t = O.spmv(A.n, wr, wi, wv, x, 2, 2, 2, 1)
    ref = O.spmv(A.n, zr, zi, zv, t, 2, 2, 2, 1)
    assert np.array_equal(yv.download().view(np.uint16), ref.view(np.uint16))


def test_ilu0_apply_matches_oracle(ctx):
    import f3r_amd as F
    A = F.CSR.stencil27(12, beta=0.5)
    F.algorithm.scaling(A)
    P = F.ILU0(A, blocks=4, acc=1.0, dtype="f64", ctx=ctx)
    x = np.random.default_rng(5).uniform(-1, 1, A.n)
    xv = F.vector.from_numpy(ctx, x)
    yv = F.vector(ctx, "f64", A.n)
    P.operate(xv, yv)
    O.lib().orc_ilu0.restype = __import__("ctypes").c_void_p
    h = O.lib().orc_ilu0(A.n, O.P(A.rptr), O.P(A.idx), O.P(A.val), 4, 1.0)
    out, tmp = np.zeros(A.n), np.zeros(A.n)
    O.lib().orc_bj_solve(h, O.P(x), O.P(out), O.P(tmp))
    O.lib().orc_factor_free(h)
    assert np.array_equal(yv.download(), out)
    # fp16 factors: close to the fp64 apply
    P16 = F.ILU0(A, blocks=4, acc=1.0, dtype="f16", ctx=ctx)
    x16 = F.vector.from_numpy(ctx, x.astype(np.float16))
    y16 = F.vector(ctx, "f16", A.n)
    P16.operate(x16, y16)
    err = np.linalg.norm(y16.download().astype(np.float64) - out) / np.linalg.norm(out)
    assert err < 5e-3
    # bit-exact against the oracle's fp16 block-Jacobi solve with fp32 sums
    for kind in ("27pt", "convdiff"):
        if kind == "convdiff":
            A = F.CSR.convdiff7(14, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
            F.algorithm.scaling(A)
        P16 = F.ILU0(A, blocks=8, acc=1.0, dtype="f16", ctx=ctx)
        xh = np.random.default_rng(6).uniform(-1, 1, A.n).astype(np.float16)
        x16 = F.vector.from_numpy(ctx, xh)
        y16 = F.vector(ctx, "f16", A.n)
        P16.operate(x16, y16)
        h = O.lib().orc_ilu0(A.n, O.P(A.rptr), O.P(A.idx), O.P(A.val), 8, 1.0)
        o16, t16 = np.zeros(A.n, np.uint16), np.zeros(A.n, np.uint16)
        O.lib().orc_bj_solve16(h, O.P(xh.view(np.uint16)), O.P(o16), O.P(t16), 0)
        O.lib().orc_factor_free(h)
        got = y16.download().view(np.uint16)
        bad = np.flatnonzero(got != o16)
        print(kind, "ilu16 mismatches", bad.size, bad[:8], got[bad[:8]].view(np.float16), o16[bad[:8]].view(np.float16))
        assert bad.size == 0, kind


@pytest.mark.parametrize("blocks", [31, 32, 64, 100, 128])
@pytest.mark.parametrize("dtype", ["f64", "f16"])
def test_bjilu0_block_per_workgroup_matches_oracle(ctx, blocks, dtype):
    """32 or more diagonal blocks: one workgroup per block walks its levels"""
    import f3r_amd as F
    A = F.CSR.convdiff7(16, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
    F.algorithm.scaling(A)
    print(f"bjilu0 blocks={blocks} {dtype}", flush=True)
    P = F.ILU0(A, blocks=blocks, acc=1.0, dtype=dtype, ctx=ctx)
    x = np.random.default_rng(7).uniform(-1, 1, A.n)
    h = O.lib().orc_ilu0(A.n, O.P(A.rptr), O.P(A.idx), O.P(A.val), blocks, 1.0)
    if dtype == "f64":
        xv = F.vector.from_numpy(ctx, x)
        yv = F.vector(ctx, "f64", A.n)
        P.operate(xv, yv)
        out, tmp = np.zeros(A.n), np.zeros(A.n)
        O.lib().orc_bj_solve(h, O.P(x), O.P(out), O.P(tmp))
        got = yv.download()
    else:
        xh = x.astype(np.float16)
        xv = F.vector.from_numpy(ctx, xh)
        yv = F.vector(ctx, "f16", A.n)
        P.operate(xv, yv)
        out, tmp = np.zeros(A.n, np.uint16), np.zeros(A.n, np.uint16)
        O.lib().orc_bj_solve16(h, O.P(xh.view(np.uint16)), O.P(out), O.P(tmp), 0)
        got = yv.download().view(np.uint16)
    O.lib().orc_factor_free(h)
    assert np.array_equal(got, out)


def _device_fits(A, C, sigma, offsets):
    # the device build's on-chip sorts: sigma windows of at most 4096 rows and,
    # in the offset format, chunks of at most 4096 entries (else the host builds)
    if sigma > 4096:
        return False
    if not offsets or C < 64:
        return True
    lens = np.diff(np.asarray(A.rptr))
    perm = np.concatenate([w0 + np.argsort(-lens[w0:w0 + sigma], kind="stable")
                           for w0 in range(0, A.n, sigma)])
    sl = lens[perm]
    return max(int(sl[c0:c0 + C].sum()) for c0 in range(0, A.n, C)) <= 4096


@pytest.mark.parametrize("C,sigma,K,offsets", [(64, 1, 4, True), (128, 1, 8, True), (64, 1, 4, False),
                                               (256, 1, 2, True), (64, 16, 4, False), (32, 256, 2, False),
                                               (128, 64, 8, True), (64, 4096, 1, False)])
def test_device_sell_build_matches_host(ctx, C, sigma, K, offsets):
    # the device-built layout equals the host-built one: same slot count and
    # chunk split, and bit-identical SpMV results in fp64 and fp16 (the rows
    # sorted inside sigma windows on the device as on the host: y comes back
    # through the permutation)
    import f3r_amd as F
    for name, A in _mats():
        print(f"device build {name} C={C} sigma={sigma} K={K} offsets={offsets}", flush=True)
        Hs = F.SELL(ctx, A, "f64", C=C, sigma=sigma, K=K, offsets=offsets)
        Ds = F.SELL(ctx, A, "f64", C=C, sigma=sigma, K=K, offsets=offsets, device=True)
        assert Hs.where() == "host" and Ds.where() == ("device" if _device_fits(A, C, sigma, offsets) else "host"), name
        assert Ds.nslots() == Hs.nslots() and Ds.chunks() == Hs.chunks(), name
        x = np.random.default_rng(3).uniform(-1, 1, A.n)
        for dt, np_t in (("f64", np.float64), ("f16", np.float16)):
            Hd, Dd = (Hs, Ds) if dt == "f64" else (Hs.cast(dt), Ds.cast(dt))
            xv = F.vector.from_numpy(ctx, x.astype(np_t))
            y1, y2 = F.vector(ctx, dt, A.n), F.vector(ctx, dt, A.n)
            Hd.operate(xv, y1)
            Dd.operate(xv, y2)
            assert np.array_equal(y1.download().view(np.uint8), y2.download().view(np.uint8)), (name, dt)


# ----------------------------------------------------------- CSR SpMV ----
CSR_COMBOS = COMBOS + [("f64", "f32", "f32", 0, 1, 1, 0)]


@pytest.mark.parametrize("vd,xd,yd,vp,xp,yp,ap", CSR_COMBOS)
def test_csr_spmv_one_lane_per_row_bitexact(ctx, vd, xd, yd, vp, xp, yp, ap):
    # one lane per row sums in CSR order: the oracle's bits, residual too
    import f3r_amd as F
    for name, A in _mats():
        x = np.random.default_rng(1).uniform(-1, 1, A.n).astype(NP[xd])
        b = np.random.default_rng(2).uniform(-1, 1, A.n).astype(NP[yd])
        print(f"csr spmv {name} {vd}/{xd}/{yd} tpr=1", flush=True)
        M = F.CSRDevice(ctx, A, vd, tpr=1)
        xv = F.vector.from_numpy(ctx, x)
        yv = F.vector(ctx, yd, A.n)
        M.operate(xv, yv)

        def oracle(**kw):
            if (vd, xd, yd) == ("f64", "f32", "f32"):
                # fp64 sums of fp32 data rounded once to fp32: the all-fp64 oracle on the
                # (exact) fp64 images of x and b, then rounded to nearest
                kw = {k: v.astype(np.float64) for k, v in kw.items()}
                return O.spmv(A.n, A.rptr, A.idx, A.val, x.astype(np.float64), 0, 0, 0, 0, **kw).astype(np.float32)
            return O.spmv(A.n, A.rptr, A.idx, A.val, x, vp, xp, yp, ap, **kw)
        ref = oracle()
        assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, vd, xd, yd)
        bv = F.vector.from_numpy(ctx, b)
        M.compute_residual(bv, xv, yv)
        ref = oracle(b=b)
        assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, "resid")


@pytest.mark.parametrize("tpr", [0, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("vd,xd,yd", [("f64", "f64", "f64"), ("f32", "f32", "f32"), ("f16", "f16", "f16"),
                                      ("f16", "f32", "f32")])
def test_csr_spmv_lanes_per_row(ctx, tpr, vd, xd, yd):
    # several lanes per row reassociate the sum: within the accumulation's
    # rounding of sum |a_ij x_j| (fp64 sums for fp64 values, else fp32) plus
    # the output's rounding
    import f3r_amd as F
    for name, A in _mats():
        x = np.random.default_rng(3).uniform(-1, 1, A.n).astype(NP[xd])
        vals = A.val.astype(NP[vd]).astype(np.float64)
        xs = x.astype(np.float64)
        rows = np.repeat(np.arange(A.n), np.diff(A.rptr))
        exact = np.bincount(rows, vals * xs[A.idx], minlength=A.n)
        mag = np.bincount(rows, np.abs(vals * xs[A.idx]), minlength=A.n)
        print(f"csr spmv {name} {vd}/{xd}/{yd} tpr={tpr}", flush=True)
        M = F.CSRDevice(ctx, A, vd, tpr=tpr)
        xv = F.vector.from_numpy(ctx, x)
        yv = F.vector(ctx, yd, A.n)
        M.operate(xv, yv)
        got = yv.download().astype(np.float64)
        nmax = int(np.diff(A.rptr).max())
        u_acc = 2.0 ** -53 if vd == "f64" else 2.0 ** -24
        u_out = {"f64": 2.0 ** -53, "f32": 2.0 ** -24, "f16": 2.0 ** -11}[yd]
        tol = 2 * nmax * u_acc * mag + u_out * np.abs(exact) + (6e-8 if yd == "f16" else 0.0)
        assert np.all(np.abs(got - exact) <= tol), (name, float(np.max(np.abs(got - exact) - tol)))


@pytest.mark.parametrize("C,sigma,K", [(64, 1, 2), (32, 256, 4), (128, 1, 8)])
@pytest.mark.parametrize("vd,xd,yd", [("f64", "f64", "f64"), ("f32", "f32", "f32"), ("f16", "f16", "f16"),
                                      ("f16", "f32", "f32")])
def test_sell_threads_per_row(ctx, C, sigma, K, vd, xd, yd):
    # K threads per row reassociate the sum: within the accumulation's
    # rounding of sum |a_ij x_j| plus the output's rounding, as CSR with
    # several lanes per row; the residual form likewise
    import f3r_amd as F
    for name, A in _mats():
        x = np.random.default_rng(3).uniform(-1, 1, A.n).astype(NP[xd])
        b = np.random.default_rng(4).uniform(-1, 1, A.n).astype(NP[yd])
        vals = A.val.astype(NP[vd]).astype(np.float64)
        xs = x.astype(np.float64)
        rows = np.repeat(np.arange(A.n), np.diff(A.rptr))
        exact = np.bincount(rows, vals * xs[A.idx], minlength=A.n)
        mag = np.bincount(rows, np.abs(vals * xs[A.idx]), minlength=A.n)
        print(f"sell {K} threads per row {name} {vd}/{xd}/{yd} C={C} sigma={sigma}", flush=True)
        M = F.SELL(ctx, A, vd, C=C, sigma=sigma, K=K, lanes=K)
        xv = F.vector.from_numpy(ctx, x)
        yv, rv = F.vector(ctx, yd, A.n), F.vector(ctx, yd, A.n)
        M.operate(xv, yv)
        M.compute_residual(F.vector.from_numpy(ctx, b), xv, rv)
        nmax = int(np.diff(A.rptr).max())
        u_acc = 2.0 ** -53 if vd == "f64" else 2.0 ** -24
        u_out = {"f64": 2.0 ** -53, "f32": 2.0 ** -24, "f16": 2.0 ** -11}[yd]
        for got, ref in ((yv.download().astype(np.float64), exact),
                         (rv.download().astype(np.float64), b.astype(np.float64) - exact)):
            tol = 2 * nmax * u_acc * mag + u_out * np.abs(ref) + (6e-8 if yd == "f16" else 0.0)
            assert np.all(np.abs(got - ref) <= tol), (name, float(np.max(np.abs(got - ref) - tol)))


@pytest.mark.parametrize("n", [1, 5, 63, 65, 4097])
def test_device_build_and_threads_per_row_edge_sizes(ctx, n):
    # sizes around one chunk, one wavefront and one sigma window (4096 rows):
    # the device build equals the host build (sigma 4096 and 16, K 4 and 2),
    # and K threads per row stay within the rounding bound of the fp64 sum
    import f3r_amd as F
    A = F.CSR(*PB.random_sparse(n, 12, 100 + n))
    x = np.random.default_rng(n).uniform(-1, 1, n)
    xv = F.vector.from_numpy(ctx, x)
    rows = np.repeat(np.arange(n), np.diff(A.rptr))
    exact = np.bincount(rows, A.val * x[A.idx], minlength=n)
    mag = np.bincount(rows, np.abs(A.val * x[A.idx]), minlength=n)
    nmax = int(np.diff(A.rptr).max())
    for C, sigma, K in ((64, 4096, 4), (32, 16, 2)):
        print(f"edge n={n} C={C} sigma={sigma} K={K}", flush=True)
        Hs = F.SELL(ctx, A, "f64", C=C, sigma=sigma, K=K)
        Ds = F.SELL(ctx, A, "f64", C=C, sigma=sigma, K=K, device=True)
        assert Ds.where() == "device" and Ds.nslots() == Hs.nslots()
        y1, y2 = F.vector(ctx, "f64", n), F.vector(ctx, "f64", n)
        Hs.operate(xv, y1)
        Ds.operate(xv, y2)
        assert np.array_equal(y1.download(), y2.download())
        Ls = F.SELL(ctx, A, "f64", C=C, sigma=sigma, K=K, lanes=K, device=True)
        y3 = F.vector(ctx, "f64", n)
        Ls.operate(xv, y3)
        tol = 2 * nmax * 2.0 ** -53 * mag + 2.0 ** -53 * np.abs(exact)
        assert np.all(np.abs(y3.download() - exact) <= tol)


def test_sell_threads_per_row_rejected_combinations(ctx):
    import f3r_amd as F
    A = next(_mats())[1]
    for kw in ({"K": 1, "lanes": 1}, {"K": 4, "lanes": 4, "offsets": True}, {"K": 4, "lanes": 4, "lds": True}):
        if kw.get("lanes", 1) == 1:
            continue
        with pytest.raises(F.F3RError):
            F.SELL(ctx, A, "f64", C=64, **kw)
    with pytest.raises(ValueError):
        F.SELL(ctx, A, "f64", C=64, K=4, lanes=2)


# ----------------------------------------------- LDS-staged x gathers ----
@pytest.mark.parametrize("C,sigma,K", [(64, 1, 4), (32, 1, 1), (128, 256, 2), (64, 1, 8)])
def test_sell_lds_gathers_bitexact(ctx, C, sigma, K):
    # the same slots and sums as the global gathers: the oracle's bits, for
    # banded matrices (every block staged) and a wide random one (none staged)
    import f3r_amd as F
    mats = [("band2d", F.CSR(*PB.stencil7(256, 48, 1))), ("tridiag", F.CSR(*PB.stencil7(5000, 1, 1))),
            ("wide", F.CSR(*PB.random_sparse(40000, 12, 7)))]
    for name, A in mats:
        x = np.random.default_rng(1).uniform(-1, 1, A.n)
        b = np.random.default_rng(2).uniform(-1, 1, A.n)
        for vd, xd, yd, vp, xp, yp, ap in COMBOS:
            print(f"lds spmv {name} {vd}/{xd}/{yd} C={C} sigma={sigma} K={K}", flush=True)
            S = F.SELL(ctx, A, vd, C=C, sigma=sigma, K=K, lds=True)
            nb, staged = S.lds_blocks()
            if name == "wide":
                assert staged < nb // 2, (nb, staged)
            elif sigma == 1:
                assert staged == nb, (name, nb, staged)
            xv = F.vector.from_numpy(ctx, x.astype(NP[xd]))
            yv = F.vector(ctx, yd, A.n)
            S.operate(xv, yv)
            ref = O.spmv(A.n, A.rptr, A.idx, A.val, x.astype(NP[xd]), vp, xp, yp, ap)
            assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, vd, xd, yd)
            if yd == xd:
                bv = F.vector.from_numpy(ctx, b.astype(NP[yd]))
                S.compute_residual(bv, xv, yv)
                ref = O.spmv(A.n, A.rptr, A.idx, A.val, x.astype(NP[xd]), vp, xp, yp, ap, b=b.astype(NP[yd]))
                assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, "resid")


def test_sell_lds_rejects_offsets(ctx):
    import f3r_amd as F
    A = F.CSR(*PB.stencil7(16, 16, 1))
    with pytest.raises(F.F3RError):
        F.SELL(ctx, A, "f16", C=64, sigma=1, K=4, offsets=True, lds=True)


# ----------------------------------------------------------- bfloat16 ----
def test_bf16_vector_casts_and_blas1(ctx):
    # device conversions into bf16 (from fp64 correctly rounded, from fp32 and
    # fp16 nearest even) and back, against the oracle's pinned conversions;
    # axpy, scale and dot in fp32 arithmetic
    import f3r_amd as F
    L = O.lib()
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(-1, 1, 5000) * 2.0 ** rng.integers(-135, 127, 5000),
                        [0.0, -0.0, 3.3961e38, 3.4e38, 2.0 ** -134, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8]])
    v64 = F.vector.from_numpy(ctx, x)
    vb = F.vector(ctx, "bf16", x.size).copy(v64)
    want = np.array([L.orc_d2b(float(t)) for t in x], np.uint16)
    assert np.array_equal(vb.download(), want)
    xf = x.astype(np.float32)
    vb.copy(F.vector.from_numpy(ctx, xf))
    assert np.array_equal(vb.download(), np.array([L.orc_f2b(float(t)) for t in xf], np.uint16))
    back = F.vector(ctx, "f32", x.size).copy(vb)
    assert np.array_equal(back.download(), np.array([L.orc_b2f(int(b)) for b in vb.download()], np.float32))
    # BLAS-1 on bf16 data: fp32 arithmetic, one rounding to bf16 per element
    a = rng.uniform(-1, 1, 4096)
    b = rng.uniform(-1, 1, 4096)
    va, vbb = F.vector.from_numpy(ctx, a, "bf16"), F.vector.from_numpy(ctx, b, "bf16")
    af = np.array([L.orc_b2f(int(t)) for t in va.download()], np.float32)
    bf = np.array([L.orc_b2f(int(t)) for t in vbb.download()], np.float32)
    d = F.dot(va, vbb)
    assert abs(d - float(af.astype(np.float64) @ bf.astype(np.float64))) <= 4096 * 2.0 ** -24 * float(np.abs(af) @ np.abs(bf))
    vbb.axpy(0.5, va)
    exp = np.array([L.orc_f2b(float(np.float32(np.float64(0.5) * ai + bi))) for ai, bi in zip(af, bf)], np.uint16)
    # fma(0.5, a, b) in fp32 then bf16: 0.5 * a is exact, so fma = fp32(a/2 + b)
    assert np.array_equal(vbb.download(), exp)


@pytest.mark.parametrize("fmt", ["sell-int32", "sell-offsets", "csr"])
def test_bf16_spmv_bitexact(ctx, fmt):
    import f3r_amd as F
    for name, A in _mats():
        if fmt == "sell-offsets" and name == "ragged":
            continue
        x = np.random.default_rng(4).uniform(-1, 1, A.n)
        b = np.random.default_rng(5).uniform(-1, 1, A.n)
        for xd, yd, xp, yp in (("bf16", "bf16", 3, 3), ("f32", "f32", 1, 1)):
            print(f"bf16 spmv {name} {fmt} x {xd}", flush=True)
            if fmt == "csr":
                M = F.CSRDevice(ctx, A, "bf16", tpr=1)
            else:
                M = F.SELL(ctx, A, "bf16", C=64, sigma=1, K=4, offsets=fmt == "sell-offsets")
            xv = F.vector.from_numpy(ctx, x, xd)
            xh = xv.download()
            yv = F.vector(ctx, yd, A.n)
            M.operate(xv, yv)
            ref = O.spmv(A.n, A.rptr, A.idx, A.val, xh, 3, xp, yp, 1)
            assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, fmt, xd)
            bv = F.vector.from_numpy(ctx, b, yd)
            M.compute_residual(bv, xv, yv)
            ref = O.spmv(A.n, A.rptr, A.idx, A.val, xh, 3, xp, yp, 1, b=bv.download())
            assert np.array_equal(yv.download().view(np.uint8), ref.view(np.uint8)), (name, fmt, "resid")
