"""GPU parity of the ablation nests through the C ABI against the oracle:
the nesting-depth variants F2, F2H, F3, F3H, F4 (work/src/depth.cpp:122-167)
and the static-weight Richardson (work/src/static.cpp:99-119).  Bar: a true
fp64 relative residual <= 1e-8 and the oracle's outer-iteration count within
+-10% (at least +-1); the static weight 1 equals the adaptive solve with no
probe bit for bit; one application of each nest matches the oracle's to
fp32 / fp16 accuracy."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

DEPTH = {"F2": 1, "F2H": 2, "F3": 3, "F3H": 4, "F4": 5}


@pytest.fixture(scope="module")
def ctx():
    import f3r_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _problem(kind, m):
    import f3r_amd as F
    if kind == "27pt":
        A = F.CSR.stencil27(m, beta=0.5)
    elif kind == "lognormal":
        A = F.CSR.convdiff7(m, sigma=2.0, corr=4.0, peclet=0.0, seed=5)
    else:
        A = F.CSR.convdiff7(m, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
    F.algorithm.scaling(A)
    return A


def _scale(n):
    return float(2.0 ** np.floor(np.log2(np.sqrt(n) / 4.0)))


def _prec(F, A, ctx, prec):
    if prec == "ilu0":
        return F.ILU0(A, blocks=8, acc=1.0, dtype="f16", ctx=ctx)
    return F.SDAINV(A, 0.05, 1.0, "f16", ctx=ctx)


def _true_res(F, ctx, A, x, b):
    import scipy.sparse as sp
    M = sp.csr_matrix((A.val, A.idx, A.rptr), shape=(A.n, A.n))
    return np.linalg.norm(b - M @ x) / np.linalg.norm(b)


@pytest.mark.parametrize("depth", ["F2", "F2H", "F3", "F3H", "F4"])
@pytest.mark.parametrize("kind,m,prec", [("convdiff", 20, "ilu0"), ("lognormal", 20, "sdainv")])
def test_depth_solve_matches_oracle(ctx, depth, kind, m, prec):
    import f3r_amd as F
    A = _problem(kind, m)
    s = _scale(A.n)
    b = F.uniform(A.n, -1.0, 1.0, seed=11)
    print(f"depth {depth} {kind} {m}^3 {prec}", flush=True)
    P = _prec(F, A, ctx, prec)
    S = F.F3R(ctx, A, P, depth=depth, range_scale=s)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    r = S.solve(bv, xv)
    x = xv.download()
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, depth=DEPTH[depth], range_scale=s,
                           precond=1 if prec == "ilu0" else 0, tol=0.05, bj_blocks=8)
    assert r.is_solved and ro.is_solved
    assert _true_res(F, ctx, A, x, b) <= 1e-8
    assert abs(r.res_iter - ro.res_iter) <= max(1, round(0.1 * ro.res_iter)), (r.res_iter, ro.res_iter)
    assert r.precond_calls == r.res_iter * 64


@pytest.mark.parametrize("depth", ["F2", "F2H", "F3", "F3H", "F4"])
def test_depth_one_application_matches_oracle(ctx, depth):
    import f3r_amd as F
    A = _problem("convdiff", 12)
    s = _scale(A.n)
    P = F.SDAINV(A, 0.05, 1.0, "f16", ctx=ctx)
    S = F.F3R(ctx, A, P, depth=depth, range_scale=s)
    v = np.random.default_rng(4).standard_normal(A.n)
    v /= np.linalg.norm(v)
    vi = F.vector.from_numpy(ctx, v)
    vo = F.vector(ctx, "f64", A.n)
    S.level_apply(2, vi, vo)
    got = vo.download()
    want = O.level_apply(A.n, A.rptr, A.idx, A.val, 2, v, depth=DEPTH[depth], range_scale=s, tol=0.05)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    tol = 2e-2 if depth.endswith("H") or depth == "F4" else 2e-3
    assert rel < tol, rel


@pytest.mark.parametrize("prec", ["sdainv", "ilu0"])
def test_static_weight_solve_matches_oracle(ctx, prec):
    import f3r_amd as F
    A = _problem("convdiff", 20)
    s = _scale(A.n)
    b = F.uniform(A.n, -1.0, 1.0, seed=12)
    for w in (1.0, 0.8):
        P = _prec(F, A, ctx, prec)
        S = F.F3R(ctx, A, P, m=(100, 8, 4, 3), static_weight=w, range_scale=s)
        bv = F.vector.from_numpy(ctx, b)
        xv = F.vector(ctx, "f64", A.n).fill(0.0)
        r = S.solve(bv, xv)
        _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, m4=3, static_weight=w, range_scale=s,
                               precond=1 if prec == "ilu0" else 0, tol=0.05, bj_blocks=8)
        assert r.is_solved and _true_res(F, ctx, A, xv.download(), b) <= 1e-8
        assert abs(r.res_iter - ro.res_iter) <= max(1, round(0.1 * ro.res_iter)), (w, r.res_iter, ro.res_iter)


def test_static_one_equals_adaptive_without_probes(ctx):
    import f3r_amd as F
    A = _problem("convdiff", 16)
    s = _scale(A.n)
    b = F.vector.from_numpy(ctx, F.uniform(A.n, -1.0, 1.0, seed=13))
    P = F.SDAINV(A, 0.05, 1.0, "f16", ctx=ctx)
    xs = []
    for kw in (dict(static_weight=1.0), dict(c=1 << 30)):
        S = F.F3R(ctx, A, P, range_scale=s, **kw)
        x = F.vector(ctx, "f64", A.n).fill(0.0)
        r = S.solve(b, x)
        xs.append((r.res_iter, x.download()))
    assert xs[0][0] == xs[1][0] and np.array_equal(xs[0][1], xs[1][1])


def test_depth_rejects_other_variants(ctx):
    import f3r_amd as F
    A = _problem("convdiff", 6)
    P = F.SDAINV(A, 0.05, 1.0, "f32", ctx=ctx)
    with pytest.raises(Exception):
        F.F3R(ctx, A, P, depth="F3", variant="float")


# ------------------------------------------- against the reference's drivers --
# The CSV lines of work/src/depth.cpp and work/src/static.cpp on the
# reference generator's 32^3 HPCG matrix (BJILU0 on 112 blocks, restart 3;
# tests/golden/ablation_hpcg32.npz).  Outer steps agree to one step where the
# levels compute in fp32 as here; the reference's fp16-arithmetic levels
# (F3H, F2H) need at least as many steps as this build's fp32 arithmetic on
# fp16 storage; at weight 1.2 both stall above 1e-8 for all 3 x 100 steps.
REF_NESTS = {"f4": "F4", "f3": "F3", "f3h": "F3H", "f2": "F2", "f2h": "F2H"}


def _hpcg32(F, golden, ctx):
    import problems as PB
    g = golden("ablation_hpcg32")
    n, rp, ix, v = PB.stencil27(32)
    A = F.CSR(n, rp, ix, v)
    F.algorithm.scaling(A)
    b = np.random.default_rng(int(g["b_seed"])).uniform(-1, 1, n)
    P = F.ILU0(A, blocks=112, acc=1.0, dtype="f16", ctx=ctx)
    return g, A, b, P


@pytest.mark.parametrize("nest", sorted(REF_NESTS))
def test_depth_against_reference_drivers(ctx, golden, nest):
    import f3r_amd as F
    g, A, b, P = _hpcg32(F, golden, ctx)
    ref_it = int(g[f"{nest}_csv"][0]) // 64
    S = F.F3R(ctx, A, P, depth=REF_NESTS[nest], restart=3)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    print(f"depth {nest} vs the reference's driver", flush=True)
    r = S.solve(bv, xv)
    rel = _true_res(F, ctx, A, xv.download(), b)
    print(f"  gpu {r.res_iter} steps, reference {ref_it}, true residual {rel:.2e}", flush=True)
    assert r.is_solved and rel <= 1e-8
    if nest in ("f3h", "f2h"):
        assert r.res_iter <= ref_it
    else:
        assert abs(r.res_iter - ref_it) <= 1


@pytest.mark.parametrize("w", [0.8, 1.0, 1.2])
def test_static_against_reference_driver(ctx, golden, w):
    import f3r_amd as F
    g, A, b, P = _hpcg32(F, golden, ctx)
    it, impl, _ = g[f"static{w:.1f}_csv"]
    ref_it = int(it) // 64
    S = F.F3R(ctx, A, P, static_weight=w, restart=3)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    print(f"static weight {w} vs the reference's driver", flush=True)
    r = S.solve(bv, xv)
    print(f"  gpu {r.res_iter} steps solved={r.is_solved}, reference {ref_it} ({impl:.2e})", flush=True)
    if impl <= 1e-8:
        assert r.is_solved and abs(r.res_iter - ref_it) <= 1
        assert _true_res(F, ctx, A, xv.download(), b) <= 1e-8
    else:
        assert ref_it == 300 and not r.is_solved and r.res_iter == 300
