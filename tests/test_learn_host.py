"""Host logic of the learned SDAINV correction (no GPU): the row features
against their definition and closed forms, the rescaled factors, parameter
packing.  The learned correction has no counterpart in the reference."""
import numpy as np
import pytest


def _scaled(A):
    import f3r_amd as F
    d = A.diagonal()
    F.algorithm.scaling(A)
    return d


def test_features_closed_form_27pt():
    # scaled 27-point, diagonal 26 / neighbours -1: every entry / 26; SDAINV
    # at tol 0.1 gives W = I and Z = I / acc
    import f3r_amd as F
    from f3r_amd import learn
    nx = 6
    A = F.CSR.stencil27(nx)
    d = _scaled(A)
    acc = 2.0
    P = F.SDAINV(A, 0.1, acc, "f16", ctx=None)
    f = learn.features(A, P, d)
    assert f.shape == (A.n, learn.NIN) and f.dtype == np.float32
    assert np.all(f[:, learn.NFEAT:] == 0)
    assert np.allclose(f[:, 0], np.log(26.0), rtol=1e-6)
    interior = (1 + nx + nx * nx) * 2  # a cell with all 26 neighbours
    assert np.allclose(f[interior, 1:7], -1 / 26, rtol=1e-6)
    assert f[interior, 7] == pytest.approx(1.0, rel=1e-6)
    assert abs(f[interior, 8]) < 1e-6
    # corner (0,0,0): its 7 neighbours all lie above the diagonal
    assert np.all(f[0, 1:4] == 0)
    assert np.allclose(f[0, 4:7], -1 / 26, rtol=1e-6)
    assert f[0, 7] == pytest.approx(7 / 26, rel=1e-6)
    assert f[0, 8] == pytest.approx(1 - 7 / 26, rel=1e-6)
    assert np.allclose(f[:, 9], 1 / acc)
    assert np.all(f[:, 10] == 0) and np.all(f[:, 11] == 0)


def test_features_definition_7pt():
    # lognormal 7-point: features 1..6 are the three nearest entries below and
    # above the diagonal in column order; the sums and factor features as defined
    import f3r_amd as F
    from f3r_amd import learn
    nx = 7
    A = F.CSR.convdiff7(nx, sigma=1.5, corr=3.0, peclet=0.0, seed=5)
    d = _scaled(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=None)
    f = learn.features(A, P, d)
    n = A.n
    dense = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(A.rptr))
    dense[rows, A.idx] = A.val
    for i in range(n):
        cols = A.idx[A.rptr[i]:A.rptr[i + 1]]
        lo = [dense[i, j] for j in sorted(cols[cols < i])[-3:]]
        hi = [dense[i, j] for j in sorted(cols[cols > i])[:3]]
        want = [0.0] * (3 - len(lo)) + lo + hi + [0.0] * (3 - len(hi))
        assert np.array_equal(f[i, 1:7], np.float32(want)), i
    # an interior row: -z, -y, -x, +x, +y, +z
    c = nx * nx * 3 + nx * 3 + 3
    assert np.array_equal(f[c, 1:7], np.float32([dense[c, c + o] for o in (-nx * nx, -nx, -1, 1, nx, nx * nx)]))
    off = dense - np.diag(np.diag(dense))
    assert np.allclose(f[:, 7], np.abs(off).sum(1), rtol=1e-5)
    assert np.allclose(f[:, 8], dense.sum(1), rtol=1e-4, atol=1e-6)
    assert np.allclose(f[:, 0], np.log(np.abs(d)), rtol=1e-6)
    (zr, zi, zv), (wr, wi, wv) = P.factors()
    Z = np.zeros((n, n))
    Z[np.repeat(np.arange(n), np.diff(zr)), zi] = zv
    W = np.zeros((n, n))
    W[np.repeat(np.arange(n), np.diff(wr)), wi] = wv
    assert np.allclose(f[:, 9], np.diag(Z), rtol=1e-6)
    assert np.allclose(f[:, 10], np.abs(W - np.diag(np.diag(W))).sum(1), rtol=1e-5, atol=1e-7)
    assert np.allclose(f[:, 11], np.abs(Z - np.diag(np.diag(Z))).sum(1), rtol=1e-5, atol=1e-7)
    assert np.any(f[:, 10] > 0) and np.any(f[:, 11] > 0)  # the 7-point pattern keeps fill


def test_features_without_diagonal():
    import f3r_amd as F
    from f3r_amd import learn
    A = F.CSR.stencil27(4)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=None)
    assert np.all(learn.features(A, P, None)[:, 0] == 0)


def test_rescale_scales_columns_of_z():
    import f3r_amd as F
    from f3r_amd import learn
    A = F.CSR.convdiff7(6, sigma=1.0, corr=2.0, peclet=0.0, seed=2)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=None)
    s = np.random.default_rng(0).uniform(0.5, 2.0, A.n).astype(np.float32)
    Q = learn.Rescaled(P, s, ctx=None)
    (zr, zi, zv), (wr, wi, wv) = P.factors()
    (qr, qi, qv), (vr, vi, vv) = Q.factors()
    assert np.array_equal(zr, qr) and np.array_equal(zi, qi)
    assert np.array_equal(qv, zv * s[zi].astype(np.float64))
    assert np.array_equal(wr, vr) and np.array_equal(wi, vi) and np.array_equal(wv, vv)


def test_rescale_rejects_bad_scales():
    import f3r_amd as F
    from f3r_amd import learn
    A = F.CSR.stencil27(3)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=None)
    s = np.ones(A.n, np.float32)
    s[3] = 0.0
    with pytest.raises(F.F3RError):
        learn.Rescaled(P, s, ctx=None)
    s[3] = np.nan
    with pytest.raises(F.F3RError):
        learn.Rescaled(P, s, ctx=None)
    I = F.ILU0(A, blocks=1, acc=1.0, dtype="f16", ctx=None)
    with pytest.raises(F.F3RError):
        learn.Rescaled(I, np.ones(A.n, np.float32), ctx=None)


def test_pack_unpack_roundtrip():
    from f3r_amd import learn
    assert learn.NPARAMS == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1
    p = np.arange(learn.NPARAMS, dtype=np.float32)
    d = learn.unpack(p)
    assert d["w1"].shape == (64, 16) and d["w2"].shape == (64, 64) and d["b3"].shape == (1,)
    assert np.array_equal(learn.pack(d), p)


def test_family_params_in_range():
    from f3r_amd import learn
    for seed in range(20):
        s, c = learn.family_params(seed)
        assert 0.5 <= s <= 2.0 and 2.0 <= c <= 16.0
    assert learn.family_params(3) == learn.family_params(3)
    assert learn.family_params(3) != learn.family_params(4)


def test_batch_family():
    from f3r_amd import batch
    fam = batch.family(0, 64)
    assert len(fam) == 64
    assert sum(1 for f in fam if f[1] == 7) == 32 and sum(1 for f in fam if f[1] == 27) == 32
    assert all(16 <= f[2] <= 64 and 0.5 <= f[3] <= 2.0 for f in fam)
    assert len({f[2] for f in fam}) > 10
    assert batch.family(0, 64) == fam and batch.family(1, 64) != fam
