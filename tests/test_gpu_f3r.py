"""GPU parity of the nested solve through the C ABI, against the oracle in
this build's arithmetic (fp32 sums of fp16 data, fused multiply-adds):
  - one innermost (Richardson) application: bit-exact, for diagonal SDAINV,
    general SDAINV and ILU(0);
  - one level-3 application: within fp32 rounding of the reductions;
  - full solves: true fp64 relative residual <= 1e-8 and outer iteration
    count within +-10% of the oracle; graph replay identical to eager launch.
"""
import numpy as np
import pytest

import oracle_lib as O
import problems as PB

pytestmark = pytest.mark.gpu


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
    elif kind == "27sym":
        A = F.CSR.stencil27(m)
    elif kind == "lognormal":
        A = F.CSR.convdiff7(m, sigma=1.5, corr=3.0, peclet=0.0, seed=7)
    else:
        A = F.CSR.convdiff7(m, sigma=1.5, corr=4.0, peclet=10.0, seed=3)
    F.algorithm.scaling(A)
    return A


def _scale(n):
    return float(2.0 ** np.floor(np.log2(np.sqrt(n) / 4.0))) if n >= 16 else 1.0


def _prec(F, A, ctx, kind):
    if kind == "ilu0":
        return F.ILU0(A, blocks=8, acc=1.0, dtype="f16", ctx=ctx)
    return F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)


@pytest.mark.parametrize("kind,prec", [("27pt", "sdainv"), ("lognormal", "sdainv"), ("convdiff", "ilu0"),
                                       ("27pt", "ilu0")])
def test_level4_bitexact(ctx, kind, prec):
    import f3r_amd as F
    A = _problem(kind, 14)
    P = _prec(F, A, ctx, prec)
    s = _scale(A.n)
    S = F.F3R(ctx, A, P, range_scale=s, use_graph=False)
    v = np.random.default_rng(1).uniform(-1, 1, A.n)
    v /= np.linalg.norm(v)
    vin = F.vector.from_numpy(ctx, v)
    vout = F.vector(ctx, "f64", A.n)
    print(f"level4 {kind} {prec}", flush=True)
    S.level_apply(4, vin, vout)
    ref = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, variant=0, ref_arith=0, range_scale=s,
                        precond=1 if prec == "ilu0" else 0, bj_blocks=8)
    got = vout.download()
    bad = np.flatnonzero(got != ref)
    print("level4 mismatches", bad.size, bad[:8], got[bad[:8]], ref[bad[:8]], flush=True)
    assert bad.size == 0


@pytest.mark.parametrize("kind,prec,m4", [("27pt", "sdainv", 1), ("27pt", "sdainv", 3), ("convdiff", "ilu0", 1),
                                          ("convdiff", "ilu0", 3)])
def test_level4_richardson_steps_bitexact(ctx, kind, prec, m4):
    # one (no residual step) and three Richardson steps; the result is h / s
    # in every case
    import f3r_amd as F
    A = _problem(kind, 12)
    P = _prec(F, A, ctx, prec)
    s = _scale(A.n)
    S = F.F3R(ctx, A, P, m=(100, 8, 4, m4), range_scale=s, use_graph=False)
    v = np.random.default_rng(4).uniform(-1, 1, A.n)
    v /= np.linalg.norm(v)
    vin = F.vector.from_numpy(ctx, v)
    vout = F.vector(ctx, "f64", A.n)
    print(f"level4 {kind} {prec} m4={m4}", flush=True)
    S.level_apply(4, vin, vout)
    ref = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, variant=0, ref_arith=0, range_scale=s, m4=m4,
                        precond=1 if prec == "ilu0" else 0, bj_blocks=8)
    assert np.array_equal(vout.download(), ref)


def test_non_power_of_two_range_scale(ctx):
    # level 3 keeps an fp32 Z when Z = h / s would not be exact in fp16
    # storage; level 4 stays bit-exact and the solve matches the oracle's
    import f3r_amd as F
    A = _problem("27pt", 14)
    P = _prec(F, A, ctx, "sdainv")
    S = F.F3R(ctx, A, P, range_scale=3.0, use_graph=False)
    v = np.random.default_rng(6).uniform(-1, 1, A.n)
    v /= np.linalg.norm(v)
    vin, vout = F.vector.from_numpy(ctx, v), F.vector(ctx, "f64", A.n)
    S.level_apply(4, vin, vout)
    ref = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, variant=0, ref_arith=0, range_scale=3.0)
    assert np.array_equal(vout.download(), ref)
    b = F.uniform(A.n, -1.0, 1.0, seed=8)
    S2 = F.F3R(ctx, A, P, range_scale=3.0)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    res = S2.solve(F.vector.from_numpy(ctx, b), xv)
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=3.0)
    assert res.is_solved and _true_res(A, xv.download(), b) <= 1e-8
    assert abs(res.res_iter - ro.res_iter) <= max(1, int(np.ceil(0.1 * ro.res_iter)))


@pytest.mark.parametrize("kind", ["27pt", "lognormal"])
def test_level3_close(ctx, kind):
    import f3r_amd as F
    A = _problem(kind, 16)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    s = _scale(A.n)
    S = F.F3R(ctx, A, P, range_scale=s, use_graph=False)
    v = np.random.default_rng(2).uniform(-1, 1, A.n)
    v /= np.linalg.norm(v)
    vin = F.vector.from_numpy(ctx, v)
    vout = F.vector(ctx, "f64", A.n)
    S.level_apply(3, vin, vout)
    ref = O.level_apply(A.n, A.rptr, A.idx, A.val, 3, v, variant=0, ref_arith=0, range_scale=s)
    err = np.linalg.norm(vout.download() - ref) / np.linalg.norm(ref)
    assert err < 1e-3, err


def _true_res(A, x, b):
    r = b - O.spmv_ld(A.n, A.rptr, A.idx, A.val, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.mark.parametrize("kind,m,prec", [("27pt", 24, "sdainv"), ("27sym", 20, "sdainv"),
                                         ("lognormal", 16, "sdainv"), ("convdiff", 16, "sdainv"),
                                         ("convdiff", 16, "ilu0")])
def test_solve_matches_oracle(ctx, kind, m, prec):
    import f3r_amd as F
    A = _problem(kind, m)
    P = _prec(F, A, ctx, prec)
    s = _scale(A.n)
    b = F.uniform(A.n, -1.0, 1.0, seed=5)
    S = F.F3R(ctx, A, P, range_scale=s)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    print(f"solve {kind} {m} {prec}", flush=True)
    res = S.solve(bv, xv)
    x = xv.download()
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=s,
                           precond=1 if prec == "ilu0" else 0, bj_blocks=8)
    assert res.is_solved and ro.is_solved
    assert _true_res(A, x, b) <= 1e-8
    assert abs(res.res_iter - ro.res_iter) <= max(1, int(np.ceil(0.1 * ro.res_iter))), (res.res_iter, ro.res_iter)
    h = S.history()
    assert len(h) >= res.res_iter and h[-1] < 1e-8


@pytest.mark.parametrize("sigma", [1, 256])
def test_solve_irregular_matrix(ctx, sigma):
    # random columns and row lengths (1..24 entries, weak diagonal dominance):
    # int32-index chunks, general SDAINV factors (CSR applies, no diagonal
    # fast path), rows sorted inside sigma windows and the matrix renumbered
    import f3r_amd as F
    A = F.CSR(*PB.random_sparse(12000, 24, 29, diag_dom=1.02))
    F.algorithm.scaling(A)
    P = _prec(F, A, ctx, "sdainv")
    s = _scale(A.n)
    b = F.uniform(A.n, -1.0, 1.0, seed=6)
    S = F.F3R(ctx, A, P, range_scale=s, sigma=sigma, C=64)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    print(f"solve irregular sigma={sigma}", flush=True)
    res = S.solve(bv, xv)
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=s)
    print(f"  gpu {res.res_iter} steps, oracle {ro.res_iter}", flush=True)
    assert res.is_solved and ro.is_solved
    assert _true_res(A, xv.download(), b) <= 1e-8
    assert abs(res.res_iter - ro.res_iter) <= max(1, int(np.ceil(0.1 * ro.res_iter))), (res.res_iter, ro.res_iter)


@pytest.mark.parametrize("kind", ["irregular", "27pt", "lognormal"])
def test_solve_threads_per_row(ctx, kind):
    # every SpMV of the nest with K threads per row (int32-index layouts, the
    # fused epilogues on the row's first lane): within +-10 % (at least one)
    # of the oracle's outer steps, to 1e-8, and graph replay equal to eager
    import f3r_amd as F
    if kind == "irregular":
        A = F.CSR(*PB.random_sparse(12000, 24, 29, diag_dom=1.02))
        F.algorithm.scaling(A)
    else:
        A = _problem(kind, 20 if kind == "27pt" else 16)
    P = _prec(F, A, ctx, "sdainv")
    s = _scale(A.n)
    b = F.uniform(A.n, -1.0, 1.0, seed=9)
    bv = F.vector.from_numpy(ctx, b)
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=s)
    xs, its = [], []
    for g in (True, False):
        S = F.F3R(ctx, A, P, range_scale=s, index=0, lanes=True, use_graph=g)
        xv = F.vector(ctx, "f64", A.n).fill(0.0)
        res = S.solve(bv, xv)
        print(f"{kind} threads per row graph={g}: {res.res_iter} steps, oracle {ro.res_iter}", flush=True)
        assert res.is_solved and _true_res(A, xv.download(), b) <= 1e-8
        assert abs(res.res_iter - ro.res_iter) <= max(1, int(np.ceil(0.1 * ro.res_iter)))
        xs.append(xv.download())
        its.append(res.res_iter)
    assert its[0] == its[1] and np.array_equal(xs[0], xs[1])
    with pytest.raises(F.F3RError):
        F.F3R(ctx, A, P, index=1, lanes=True)


def test_graph_replay_identical_to_eager(ctx):
    import f3r_amd as F
    A = _problem("27pt", 20)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    b = F.uniform(A.n, -1.0, 1.0, seed=9)
    bv = F.vector.from_numpy(ctx, b)
    outs = []
    for g in (False, True, True):
        S = F.F3R(ctx, A, P, use_graph=g, c=8)   # c = 8: probes inside the captured bodies
        xv = F.vector(ctx, "f64", A.n).fill(0.0)
        r = S.solve(bv, xv)
        outs.append((r.res_iter, xv.download()))
        # a second solve with the same object restarts the weights: same answer
        xv.fill(0.0)
        r2 = S.solve(bv, xv)
        assert r2.res_iter == r.res_iter and np.array_equal(xv.download(), outs[-1][1])
    assert outs[0][0] == outs[1][0] == outs[2][0]
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[1][1], outs[2][1])


@pytest.mark.parametrize("kind", ["27pt", "lognormal"])
def test_layouts_give_identical_solves(ctx, kind):
    # The chunk-offset format sums every row's nonzeros in CSR order (its
    # extra slots add +0), and K only regroups the same fma chain.  The
    # whole solve is therefore bit-identical across index formats and K.
    import f3r_amd as F
    A = _problem(kind, 20)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    bv = F.vector.from_numpy(ctx, F.uniform(A.n, -1.0, 1.0, seed=4))
    outs = []
    for index, K in ((0, 4), (1, 4), (1, 8), (0, 8), (1, 0)):
        S = F.F3R(ctx, A, P, index=index, K=K, C=64)
        xv = F.vector(ctx, "f64", A.n).fill(0.0)
        r = S.solve(bv, xv)
        print(f"{kind} index={index} K={K}: {r.res_iter} outer steps")
        outs.append((r.res_iter, xv.download()))
    for it, x in outs[1:]:
        assert it == outs[0][0] and np.array_equal(x, outs[0][1])


@pytest.mark.parametrize("sigma", [1, 64])
def test_device_built_layouts_give_identical_solves(ctx, sigma):
    # sigma 64: the boundary rows (fewer entries) move inside their windows,
    # the matrix is renumbered, b and x are permuted on the way in and out
    import f3r_amd as F
    A = _problem("lognormal", 18)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    bv = F.vector.from_numpy(ctx, F.uniform(A.n, -1.0, 1.0, seed=8))
    outs = []
    for dev in (False, True):
        S = F.F3R(ctx, A, P, device_build=dev, sigma=sigma)
        assert S.layouts_where() == ("device" if dev else "host")
        xv = F.vector(ctx, "f64", A.n).fill(0.0)
        r = S.solve(bv, xv)
        outs.append((r.res_iter, xv.download()))
    assert outs[0][0] == outs[1][0] and np.array_equal(outs[0][1], outs[1][1])


def test_reference_driver_csv_line():
    # the C++ driver on the C ABI prints the reference's CSV line
    # (work/src/f3r-gpu.cpp:97-102): name,F3R,__half,m2,m3,m4,c,SDAINV,acc,Time,Iter,ImplRes,ExplRes
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "f3r_amd", "f3r_gpu")
    assert os.path.exists(exe), "f3r_amd/f3r_gpu is not built"
    out = subprocess.run([exe, "gen:27:32:0.5", "1.0", "2", "8", "4", "2", "64", "20"], capture_output=True,
                         text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    f = out.stdout.strip().split(",")
    assert f[:8] == ["stencil27_32", "F3R", "__half", "8", "4", "2", "64", "SDAINV"]
    t, it, impl, expl = float(f[9]), int(f[10]), float(f[11]), float(f[12])
    assert t > 0 and it > 0 and impl <= 1e-8 and expl <= 1e-8


def test_solve_edge_cases(ctx):
    import f3r_amd as F
    A = _problem("27pt", 9)    # n = 729: not a multiple of the chunk height
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    b = F.uniform(A.n, -1.0, 1.0, seed=1)
    bv = F.vector.from_numpy(ctx, b)
    S = F.F3R(ctx, A, P, C=64, sigma=64, K=4)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    r = S.solve(bv, xv)
    x = xv.download()
    assert r.is_solved and _true_res(A, x, b) <= 1e-8
    # starting from the solution: solved before the first Arnoldi step
    r0 = S.solve(bv, xv)
    assert r0.is_solved and r0.res_iter == 0
    # outer budget exhausted: reported, not an error
    S1 = F.F3R(ctx, A, P, m=(1, 1, 1, 1), restart=1)
    xv.fill(0.0)
    r1 = S1.solve(bv, xv)
    assert not r1.is_solved and r1.res_iter == 1
    # a zero right-hand side from x = 0: solved with no step, x stays 0
    zv = F.vector(ctx, "f64", A.n).fill(0.0)
    xv.fill(0.0)
    rz = S.solve(zv, xv)
    assert rz.is_solved and rz.res_iter == 0 and rz.nrm_b == 0.0 and not np.any(xv.download())


def test_c2_shape_64cube(ctx):
    # the 27-point nonsymmetric problem of the benchmark, at 64^3
    import f3r_amd as F
    A = _problem("27pt", 64)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    b = F.uniform(A.n, -1.0, 1.0, seed=0)
    S = F.F3R(ctx, A, P)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    r = S.solve(bv, xv)
    assert r.is_solved and _true_res(A, xv.download(), b) <= 1e-8
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=_scale(A.n))
    assert abs(r.res_iter - ro.res_iter) <= max(1, int(np.ceil(0.1 * ro.res_iter)))
    # the kernel timings of the bench, back to back and with the caches emptied first
    for which in range(4):
        ms, by = S.time_kernel(which, reps=3)
        msc, byc = S.time_kernel(which, reps=2, cold=True)
        assert ms > 0 and msc > 0 and by == byc and by > 0, (which, ms, msc, by, byc)
    # and the solve after them is unchanged
    xv.fill(0.0)
    r2 = S.solve(bv, xv)
    assert r2.res_iter == r.res_iter and r2.res_nrm2 == r.res_nrm2


@pytest.mark.parametrize("variant,prec", [("float", "sdainv"), ("double", "sdainv"), ("float", "ilu0"),
                                          ("double", "ilu0")])
def test_level4_variants_bitexact(ctx, variant, prec):
    # the FLOAT / DOUBLE builds: Richardson in fp32 / fp64 with the fp64 matrix
    import f3r_amd as F
    A = _problem("lognormal", 13)
    P = _prec(F, A, ctx, prec)
    S = F.F3R(ctx, A, P, variant=variant, use_graph=False)
    v = np.random.default_rng(3).uniform(-1, 1, A.n)
    v /= np.linalg.norm(v)
    vin = F.vector.from_numpy(ctx, v)
    vout = F.vector(ctx, "f64", A.n)
    S.level_apply(4, vin, vout)
    ref = O.level_apply(A.n, A.rptr, A.idx, A.val, 4, v, variant={"float": 1, "double": 2}[variant],
                        ref_arith=0, precond=1 if prec == "ilu0" else 0, bj_blocks=8)
    got = vout.download()
    bad = np.flatnonzero(got != ref)
    print(variant, prec, "level4 mismatches", bad.size, flush=True)
    assert bad.size == 0


@pytest.mark.parametrize("variant", ["float", "double"])
@pytest.mark.parametrize("kind", ["27pt", "convdiff"])
def test_solve_variants_match_oracle(ctx, variant, kind):
    import f3r_amd as F
    A = _problem(kind, 18)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    b = F.uniform(A.n, -1.0, 1.0, seed=8)
    S = F.F3R(ctx, A, P, variant=variant)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    res = S.solve(bv, xv)
    _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant={"float": 1, "double": 2}[variant],
                           ref_arith=0)
    assert res.is_solved and ro.is_solved
    assert _true_res(A, xv.download(), b) <= 1e-8
    assert abs(res.res_iter - ro.res_iter) <= max(1, int(np.ceil(0.1 * ro.res_iter))), (res.res_iter, ro.res_iter)
