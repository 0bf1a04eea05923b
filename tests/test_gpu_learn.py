"""GPU parity of the learned SDAINV correction (no counterpart in the
reference; the checks are the oracle's fp64 MLP, a numpy emulation of the
kernels' bf16 roundings, and scipy sparse products in fp64).

Tolerances: the MFMA layers take bf16 operands and accumulate in fp32.
Against an emulation that rounds the same operands to bf16 the difference is
fp32 summation order (and, rarely, one bf16 step of a hidden unit whose
pre-activation lies on a rounding boundary): 2e-3 of the output scale.
Against the fp64 oracle on the same bf16 inputs and weights (hidden layer
not rounded): 1e-2 of each row's magnitude and a gradient cosine above 0.995.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import f3r_amd as F
    c = F.Context(0)
    yield c
    c.close()


def bf16(a):
    """round to bfloat16 (nearest even), returned as float64"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _params(rng, scale3=0.3):
    from f3r_amd import learn
    d = {"w1": rng.uniform(-0.6, 0.6, (64, 16)), "b1": rng.uniform(-0.1, 0.1, 64),
         "w2": rng.uniform(-0.3, 0.3, (64, 64)), "b2": rng.uniform(-0.1, 0.1, 64),
         "w3": rng.uniform(-scale3, scale3, 64), "b3": rng.uniform(-0.1, 0.1, 1)}
    d["w1"][:, 12:] = 0.0
    return learn.pack(d)


def _emulate(p, x):
    """the kernels' arithmetic in float64 with their bf16 roundings"""
    from f3r_amd import learn
    d = learn.unpack(p.astype(np.float32))
    xb, w1, w2 = bf16(x), bf16(d["w1"]), bf16(d["w2"])
    z1 = xb @ w1.T + d["b1"]
    h1 = bf16(np.maximum(z1, 0))
    z2 = h1 @ w2.T + d["b2"]
    y = np.maximum(z2, 0) @ d["w3"].astype(np.float64) + d["b3"][0]
    return y, (xb, w1, w2, h1, z2, d)


def _emulate_backward(p, x, g):
    y, (xb, w1, w2, h1, z2, d) = _emulate(p, x)
    w3 = d["w3"].astype(np.float64)
    dz2 = np.where(z2 > 0, g[:, None] * w3[None, :], 0.0)
    gw3 = (g[:, None] * np.maximum(z2, 0)).sum(0)
    gb2 = dz2.sum(0)
    gw2 = bf16(dz2).T @ h1
    dh1 = bf16(dz2) @ w2
    dh1 = np.where(h1 > 0, dh1, 0.0)
    gb1 = dh1.sum(0)
    gw1 = bf16(dh1).T @ xb
    gw1[:, 12:] = 0.0
    from f3r_amd import learn
    return learn.pack({"w1": gw1, "b1": gb1, "w2": gw2, "b2": gb2, "w3": gw3, "b3": [g.sum()]})


def _feat(rng, n):
    x = np.zeros((n, 16), np.float32)
    x[:, :12] = rng.normal(0, 1, (n, 12))
    return x


@pytest.mark.parametrize("n", [1, 63, 1000, 70001])
def test_mlp_forward(ctx, n):
    from f3r_amd import learn
    rng = np.random.default_rng(n)
    p = _params(rng)
    x = _feat(rng, n)
    m = learn.MLP(ctx, seed=3, params=p)
    assert np.array_equal(m.get(), p.astype(np.float32))
    y = m.forward(x).astype(np.float64)
    ye, _ = _emulate(p, x)
    scale = max(np.abs(ye).max(), 1e-3)
    err = np.abs(y - ye).max() / scale
    print(f"forward n={n}: max |gpu - emulation| / scale = {err:.2e}")
    assert err < 2e-3
    # fp64 oracle on the same bf16 inputs and weights (no rounding of the
    # hidden layer): within 1e-2 of each row's magnitude sum_k |w3_k| h2_k + |b3|
    d = learn.unpack(p.astype(np.float32))
    yo = O.mlp_forward(bf16(x)[:, :12], bf16(d["w1"])[:, :12], d["b1"], bf16(d["w2"]), d["b2"],
                       d["w3"].reshape(1, -1), d["b3"])
    _, (xb, w1, w2, h1, z2, _) = _emulate(p, x)
    rowscale = np.maximum(z2, 0) @ np.abs(d["w3"].astype(np.float64)) + abs(float(d["b3"][0]))
    err_o = (np.abs(y - yo) / rowscale).max()
    print(f"forward n={n}: max |gpu - fp64 oracle| / row magnitude = {err_o:.2e}")
    assert err_o < 1e-2


def test_mlp_initial_output_is_zero(ctx):
    # the last layer starts at zero: y = 0, s = 1, plain SDAINV
    from f3r_amd import learn
    m = learn.MLP(ctx, seed=11)
    y = m.forward(_feat(np.random.default_rng(0), 500))
    assert np.all(y == 0)
    p = m.get()
    assert np.all(learn.unpack(p)["w1"][:, 12:] == 0)


@pytest.mark.parametrize("n", [64, 5000, 100003])
def test_mlp_backward(ctx, n):
    from f3r_amd import learn
    rng = np.random.default_rng(100 + n)
    p = _params(rng)
    x = _feat(rng, n)
    g = rng.normal(0, 1, n).astype(np.float32)
    m = learn.MLP(ctx, seed=0, params=p)
    m.grad_zero()
    m.backward(x, g)
    gg = m.grad().astype(np.float64)
    ge = _emulate_backward(p, x, g.astype(np.float64))
    gp, ep = learn.unpack(gg), learn.unpack(ge)
    for k in gp:
        scale = max(np.abs(ep[k]).max(), 1e-6)
        err = np.abs(gp[k] - ep[k]).max() / scale
        print(f"backward n={n} {k}: max rel err vs emulation {err:.2e}")
        assert err < 2e-3, k
    d = learn.unpack(p.astype(np.float32))
    go = O.mlp_backward(bf16(x)[:, :12], bf16(d["w1"])[:, :12], d["b1"], bf16(d["w2"]), d["b2"],
                        d["w3"].reshape(1, -1), d["b3"], g.astype(np.float64))
    gw1 = np.zeros((64, 16))
    gw1[:, :12] = go[0]
    ref = learn.pack({"w1": gw1, "b1": go[1], "w2": go[2], "b2": go[3], "w3": go[4].reshape(-1),
                      "b3": go[5]}).astype(np.float64)
    cos = gg @ ref / np.linalg.norm(gg) / np.linalg.norm(ref)
    print(f"backward n={n}: cosine vs fp64 oracle {cos:.6f}")
    assert cos > 0.995
    # accumulation: a second backward adds the same gradient
    m.backward(x, g)
    assert np.allclose(m.grad(), 2 * gg, rtol=1e-5, atol=1e-6 * np.abs(gg).max())


def test_adam_step(ctx):
    # one Adam step from zero moments moves every parameter with a non-zero
    # gradient by lr * sign(g) (bias-corrected m/sqrt(v) = g/|g|)
    from f3r_amd import learn
    rng = np.random.default_rng(5)
    p = _params(rng)
    m = learn.MLP(ctx, seed=0, params=p)
    x = _feat(rng, 777)
    m.grad_zero()
    m.backward(x, rng.normal(0, 1, 777).astype(np.float32))
    g = m.grad()
    m.adam(lr=1e-2, eps=1e-12)
    q = m.get()
    live = np.abs(g) > 1e-6
    live[:1024] &= (np.arange(1024) % 16) < 12
    assert np.allclose(q[live] - p[live], -1e-2 * np.sign(g[live]), rtol=1e-3, atol=1e-6)
    assert np.all(learn.unpack(q)["w1"][:, 12:] == 0)


def _system(F, m=12, seed=7):
    A = F.CSR.convdiff7(m, sigma=1.2, corr=3.0, peclet=0.0, seed=seed)
    d = A.diagonal()
    F.algorithm.scaling(A)
    return A, d


def _dense(n, rptr, idx, val):
    import scipy.sparse as sp
    return sp.csr_matrix((val, idx, rptr), shape=(n, n))


def test_lsys_loss_and_gradient(ctx):
    import f3r_amd as F
    from f3r_amd import learn
    A, d = _system(F)
    n = A.n
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    feat = learn.features(A, P, d)
    L = learn.LearnSystem(ctx, A, P, feat)
    rng = np.random.default_rng(9)
    Zp = rng.normal(0, 1, (3, n)).astype(np.float32)
    L.set_probes(Zp)
    p = _params(rng, scale3=0.05)
    m = learn.MLP(ctx, seed=0, params=p)
    y = m.forward(feat).astype(np.float64)
    assert np.abs(y).max() < 2.5  # away from the clamp
    s_gpu = L.scales(m, normalize=False)
    assert np.allclose(s_gpu, np.exp(y), rtol=1e-5)
    s_n = L.scales(m)
    assert np.allclose(s_n, np.exp(y - y.mean()), rtol=1e-4)
    # fp64 reference of both losses and of dL/dy
    Am = _dense(n, A.rptr, A.idx, A.val)
    (zr, zi, zv), (wr, wi, wv) = P.factors()
    Z, Wt = _dense(n, zr, zi, zv), _dense(n, wr, wi, wv)
    s = s_gpu.astype(np.float64)
    ref = {0: [0.0, np.zeros(n)], 1: [0.0, np.zeros(n)]}
    for z in Zp.astype(np.float64):
        u = Wt @ z
        q = Am @ (Z @ (s * u))
        zz, a, b = z @ z, z @ q, q @ q
        r = z - q
        ref[0][0] += (r @ r) / zz
        ref[0][1] += -2.0 / zz * (Z.T @ (Am.T @ r)) * u * s
        ref[1][0] += 1.0 - a * a / (zz * b)
        ref[1][1] += -2.0 * a / (zz * b * b) * (Z.T @ (Am.T @ (b * z - a * q))) * u * s
    for kind in (0, 1):
        loss, gy = ref[kind][0] / len(Zp), ref[kind][1] / len(Zp)
        m.grad_zero()
        lg = L.loss(m, grad=False, kind=kind)
        print(f"kind {kind}: loss gpu {lg:.8f} fp64 {loss:.8f}")
        assert lg == pytest.approx(loss, rel=1e-4)
        assert np.all(m.grad() == 0)
        lg2 = L.loss(m, grad=True, kind=kind)
        assert lg2 == pytest.approx(lg, rel=1e-6)
        gg = m.grad().astype(np.float64)
        ge = _emulate_backward(p, feat, gy)
        cos = gg @ ge / np.linalg.norm(gg) / np.linalg.norm(ge)
        rel = np.linalg.norm(gg - ge) / np.linalg.norm(ge)
        print(f"kind {kind}: loss gradient cosine {cos:.6f}, rel err {rel:.2e}")
        assert cos > 0.999 and rel < 2e-2


def test_capture_first_vector_is_rhs(ctx):
    # the first innermost input of a level-2 application is fp32(b)
    import f3r_amd as F
    A, _ = _system(F, m=10)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    S = F.F3R(ctx, A, P, use_graph=True)
    S.capture(5, 3)
    b = np.random.default_rng(2).uniform(-1, 1, A.n)
    vb, out = F.vector.from_numpy(ctx, b), F.vector(ctx, "f64", A.n)
    S.level_apply(2, vb, out)
    Zc = S.captured()
    assert Zc.shape == (5, A.n)
    assert np.array_equal(Zc[0], b.astype(np.float32))
    assert np.all(np.isfinite(Zc)) and np.all(np.linalg.norm(Zc, axis=1) > 0)
    # a full solve while capturing (level 2 eager) matches one with graphs
    S.capture(0)
    x1 = F.vector(ctx, "f64", A.n).fill(0.0)
    r1 = S.solve(vb, x1)
    S.capture(8, 5)
    x2 = F.vector(ctx, "f64", A.n).fill(0.0)
    r2 = S.solve(vb, x2)
    assert r1.res_iter == r2.res_iter and np.array_equal(x1.download(), x2.download())
    assert S.captured().shape == (8, A.n)


def test_training_lowers_loss_and_learned_solve_converges(ctx):
    import f3r_amd as F
    from f3r_amd import learn
    systems, hs = [], []
    for seed in (1, 2, 3):
        A, d = _system(F, m=14, seed=seed)
        P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=None)
        feat = learn.features(A, P, d)
        L = learn.LearnSystem(ctx, A, P, feat)
        L.set_probes(learn.solver_probes(ctx, A, P, k=6, every=3, seed=seed))
        systems.append(L)
        hs.append({"A": A, "P": P, "feat": feat, "seed": seed})
    m0 = learn.MLP(ctx, seed=0)
    base = np.mean([L.loss(m0, kind=0) for L in systems])
    mlp, hist = learn.train(ctx, systems, epochs=25, lr=3e-3, seed=0, kind=0)
    final = np.mean([L.loss(mlp, kind=0) for L in systems])
    print(f"training loss: plain SDAINV {base:.5f} -> learned {final:.5f}; per epoch {np.round(hist, 5)}")
    assert hist[0] == pytest.approx(base, rel=0.2)
    assert final < base
    r = learn.evaluate(ctx, mlp, hs[0], normalize=False)
    print(f"held-in system: {r}")
    for k in ("plain", "learned"):
        assert r[k]["solved"] and r[k]["true_rel_res"] <= 1e-8
