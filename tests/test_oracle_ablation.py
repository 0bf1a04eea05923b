"""Pins the oracle's ablation paths (work/src/depth.cpp:122-167,
work/src/static.cpp:99-119) before the GPU is checked against them:
  - F3 (M applied directly under two fp32-basis levels) is the F3R nest with
    one Richardson step of static weight 1, bit for bit;
  - a static weight of 1 is the adaptive scheme when no probe ever fires,
    and a static weight of 1/2 halves a one-step Richardson exactly;
  - an fp16-basis level computes the operator of its fp32-basis twin to
    fp16 accuracy, and its power-of-two storage scale changes nothing but
    fp16 range effects;
  - every depth solves to a true fp64 residual <= 1e-8."""
import numpy as np
import pytest

import oracle_lib as O


def _problem(m=12, kind="convdiff"):
    import f3r_amd as F
    if kind == "convdiff":
        A = F.CSR.convdiff7(m, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
    else:
        A = F.CSR.stencil27(m, beta=0.5)
    F.algorithm.scaling(A)
    b = F.uniform(A.n, -1.0, 1.0, seed=9)
    s = float(2.0 ** np.floor(np.log2(np.sqrt(A.n) / 4.0)))
    return A, b, s


@pytest.mark.parametrize("precond", [0, 1])
def test_f3_is_f3r_with_one_static_step(precond):
    A, b, s = _problem(14)
    x1, r1, h1 = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, depth=3, precond=precond, bj_blocks=8,
                             range_scale=s)
    x2, r2, h2 = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, depth=0, m2=8, m3=8, m4=1, static_weight=1.0,
                             precond=precond, bj_blocks=8, range_scale=s)
    assert r1.res_iter == r2.res_iter and r1.precond_calls == r2.precond_calls
    assert np.array_equal(x1, x2)
    assert np.array_equal(h1, h2)


@pytest.mark.parametrize("variant", [0, 1])
def test_static_one_is_adaptive_without_probes(variant):
    A, b, s = _problem(12)
    kw = dict(variant=variant, range_scale=s, m4=3)
    x1, r1, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, static_weight=1.0, **kw)
    x2, r2, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, c=1 << 30, **kw)
    assert r1.res_iter == r2.res_iter and np.array_equal(x1, x2)
    # with probes the adaptive weights move away from 1 and the solves differ
    x3, r3, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, c=4, **kw)
    assert not np.array_equal(x1, x3)


def test_static_half_weight_halves_one_step():
    A, b, s = _problem(10)
    v = np.random.default_rng(1).uniform(0.5, 1.0, A.n)
    one = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, m4=1, static_weight=1.0)
    half = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, m4=1, static_weight=0.5)
    big = np.abs(one) > 1e-3  # away from the fp16 subnormal range
    assert big.mean() > 0.9
    assert np.array_equal(half[big], one[big] * 0.5)
    # two steps: w M v + w M (v - A w M v) is not linear in w
    two1 = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, m4=2, static_weight=1.0)
    two5 = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, m4=2, static_weight=0.5)
    assert not np.allclose(two5, 0.5 * two1)


@pytest.mark.parametrize("h,f", [(4, 3), (2, 1)])
def test_fp16_level_matches_fp32_twin(h, f):
    """F3H's and F2H's fp16 levels against F3's and F2's fp32 levels"""
    A, b, s = _problem(10)
    v = np.random.default_rng(2).standard_normal(A.n)
    v /= np.linalg.norm(v)
    oh = O.level_apply(A.n, A.rptr, A.idx, A.val, 2, v, depth=h, range_scale=s)
    of = O.level_apply(A.n, A.rptr, A.idx, A.val, 2, v, depth=f, range_scale=s)
    rel = np.linalg.norm(oh - of) / np.linalg.norm(of)
    assert 1e-7 < rel < 2e-2, rel


def test_fp16_level_storage_scale():
    """the fp16 storage scale only moves values relative to fp16's range"""
    A, b, s = _problem(10)
    v = np.random.default_rng(3).standard_normal(A.n)
    v /= np.linalg.norm(v)
    o1 = O.level_apply(A.n, A.rptr, A.idx, A.val, 2, v, depth=2, range_scale=s)
    o4 = O.level_apply(A.n, A.rptr, A.idx, A.val, 2, v, depth=2, range_scale=4 * s)
    assert np.linalg.norm(o4 - o1) / np.linalg.norm(o1) < 1e-2
    # far too small a scale pushes the basis into fp16 subnormals
    o0 = O.level_apply(A.n, A.rptr, A.idx, A.val, 2, v, depth=2, range_scale=2.0 ** -12)
    assert np.linalg.norm(o0 - o1) / np.linalg.norm(o1) > np.linalg.norm(o4 - o1) / np.linalg.norm(o1)


@pytest.mark.parametrize("depth", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("precond", [0, 1])
def test_depth_variants_solve(depth, precond):
    A, b, s = _problem(16, "27pt")
    x, r, h = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, depth=depth, precond=precond, bj_blocks=8,
                          range_scale=s)
    assert r.is_solved and r.expl_res <= 1e-8
    inner = {1: 64, 2: 64, 3: 64, 4: 64, 5: 64}[depth]
    assert r.precond_calls == r.res_iter * inner


def test_depth_rejects_reference_arithmetic():
    A, b, s = _problem(6)
    o = O.f3r_opts(depth=3, ref_arith=1)
    x = np.zeros(A.n)
    r = O.Result()
    assert O.lib().orc_f3r_solve(A.n, O.P(A.rptr), O.P(A.idx), O.P(A.val), O.P(b), O.P(x),
                                 O.ctypes.byref(o), O.ctypes.byref(r), None) == -1


# ------------------------------------------- against the reference's drivers --
# tests/golden/ablation_hpcg32.npz: the CSV lines of work/src/depth.cpp (F4,
# F3, F3H, F2, F2H) and work/src/static.cpp (weights 0.8, 1.0, 1.2) on the
# reference generator's 32^3 HPCG matrix, BJILU0 on 112 blocks, restart 3.
# The reference computes its fp16-basis levels in _Float16 (dots, norms and
# Givens rotations included); this build stores them in fp16 and computes in
# fp32.  Where the arithmetic is alike (fp32-basis levels), the outer-step
# counts agree to one step (+-10% of 2-3 steps rounds to zero; the fp16
# triangular solves round differently); where the reference computes in
# fp16, fp32 arithmetic needs no more steps.
NESTS = {"f4": 5, "f3": 3, "f3h": 4, "f2": 1, "f2h": 2}


def _hpcg32(golden):
    import problems as PB
    g = golden("ablation_hpcg32")
    n, rp, ix, v = PB.stencil27(32)
    b = np.random.default_rng(int(g["b_seed"])).uniform(-1, 1, n)
    return g, n, rp, ix, O.scaling(n, rp, ix, v), b


@pytest.mark.parametrize("nest", sorted(NESTS))
def test_depth_nests_against_reference_drivers(golden, nest):
    g, n, rp, ix, vs, b = _hpcg32(golden)
    ref_it = int(g[f"{nest}_csv"][0]) // 64
    _, r, _ = O.f3r_solve(n, rp, ix, vs, b, depth=NESTS[nest], precond=1, bj_blocks=112, restart=3)
    print(nest, "oracle", r.res_iter, "reference", ref_it)
    assert r.is_solved and r.expl_res <= 1e-8
    if nest in ("f3h", "f2h"):
        assert r.res_iter <= ref_it
    else:
        assert abs(r.res_iter - ref_it) <= 1


@pytest.mark.parametrize("w", [0.8, 1.0, 1.2])
@pytest.mark.parametrize("ref_arith", [0, 1])
def test_static_weights_against_reference_driver(golden, w, ref_arith):
    g, n, rp, ix, vs, b = _hpcg32(golden)
    it, impl, _ = g[f"static{w:.1f}_csv"]
    ref_it = int(it) // 64  # Iter = outer steps x m2 m3 m4 = 8 x 4 x 2
    _, r, _ = O.f3r_solve(n, rp, ix, vs, b, static_weight=w, precond=1, bj_blocks=112, restart=3,
                          ref_arith=ref_arith)
    print(w, ref_arith, "oracle", r.res_iter, r.is_solved, "reference", ref_it, impl)
    if impl <= 1e-8:
        assert r.is_solved and abs(r.res_iter - ref_it) <= 1
    else:
        # the reference stalls above 1e-8 for all 3 x 100 outer steps, and so does the oracle
        assert ref_it == 300 and not r.is_solved and r.res_iter == 300
