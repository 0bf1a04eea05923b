"""Batched solves of the seeded family (independent systems on concurrent
streams, autotuned SELL layouts): every system and right-hand side reaches a
true fp64 relative residual of 1e-8."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_small():
    from f3r_amd import batch
    out = batch.run(count=6, nrhs=2, streams=3, seed=1, tune=True, log=print,
                    cands={"Cs": (64, 128), "sigmas": (1,), "Ks": (2, 4), "reps": 5})
    assert out["all_solved"] and out["worst_true_rel_res"] <= 1e-8
    assert len(out["results"]) == 6 and all(len(r["outer"]) == 2 for r in out["results"])
    for r in out["results"]:
        assert r["layout"]["C"] in (64, 128) and r["layout"]["K"] in (2, 4)


def test_autotune_candidates_agree_and_the_tuned_solve_is_unchanged():
    # every candidate computes the same SpMV (SELL layouts and one-lane CSR
    # sum in CSR order: the same bits; several lanes per row within fp32
    # rounding), the winner is the fastest, and a solve with the tuned layout
    # gives the default layout's solution bit for bit
    import f3r_amd as F
    from f3r_amd.tune import autotune, solver_layout
    ctx = F.Context(0)
    A = F.CSR.stencil27(40, beta=0.5)
    F.algorithm.scaling(A)
    best, table = autotune(ctx, A, "f16", Cs=(32, 64, 256), sigmas=(1, 256), Ks=(1, 4, 8), offsets=(False, True),
                           lds=(False, True), tprs=(1, 4, 16), reps=5, check=True)
    # SELL: 3 C x 2 sigma x 3 K with one column index per slot (sigma 256 >= every C), with and
    # without LDS-staged gathers, K threads per row for those of them with K >= 2 and no staging
    # (3 C x 2 sigma x 2 K), and the chunk-offset format for sigma 1 and C >= 64 (2 C x 3 K);
    # CSR: 3 lane counts
    sells = [e for e in table if e["format"] == "sell"]
    assert len(sells) == 2 * 18 + 12 + 6 and len(table) == len(sells) + 3
    assert sum(e["lanes"] > 1 for e in sells) == 12
    assert best["ms"] == min(e["ms"] for e in table) > 0
    ref = next(e for e in sells if e["lanes"] == 1)["y"]
    for e in table:
        if (e["format"] == "sell" and e["lanes"] == 1) or (e["format"] == "csr" and e["tpr"] == 1):
            assert np.array_equal(e["y"].view(np.uint16), ref.view(np.uint16)), e
        else:
            assert np.allclose(e["y"].astype(np.float64), ref.astype(np.float64), rtol=2e-3, atol=1e-4), e
    b = F.uniform(A.n, -1.0, 1.0, seed=3)
    bv = F.vector.from_numpy(ctx, b)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    xs, its = [], []
    for kw in ({}, solver_layout(best)):
        x = F.vector(ctx, "f64", A.n).fill(0.0)
        r = F.F3R(ctx, A, P, **kw).solve(bv, x)
        assert r.is_solved
        xs.append(x.download())
        its.append(r.res_iter)
    print("tuned", best, "outer steps", its, flush=True)
    if solver_layout(best)["sigma"] == 1:  # the same row order: the same reductions, the same bits
        assert np.array_equal(xs[0], xs[1])
    else:  # a sorted row order changes the reductions' order
        assert abs(its[0] - its[1]) <= 1
    ctx.close()


# ------------------------------------------------ lockstep batch (f3r_batch) --
def _family(F, count, seed):
    from f3r_amd import batch
    fam = batch.family(seed, count)
    out = []
    for (i, npts, nx, delta) in fam:
        nx = min(nx, 20)  # small systems: the test is about the per-segment bookkeeping
        out.append((i, npts, nx, delta, batch.make_system(npts, nx, delta)))
    return out


def _true_res(A, x, b):
    import numpy as np
    import scipy.sparse as sp
    M = sp.csr_matrix((A.val, A.idx, A.rptr), shape=(A.n, A.n))
    return np.linalg.norm(b - M @ x) / np.linalg.norm(b)


def test_lockstep_batch_matches_single_solves():
    """every segment: true residual <= 1e-8, outer steps within +-1 of its own
    single-system solve and +-10% (at least +-1) of the oracle's"""
    import numpy as np
    import f3r_amd as F
    import oracle_lib as O
    ctx = F.Context(0)
    fam = _family(F, 6, seed=3)
    As, Ps, bs = [], [], []
    for (i, npts, nx, delta, A) in fam:
        P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
        for r in range(2):
            As.append(A)
            Ps.append(P)
            bs.append(F.uniform(A.n, -1.0, 1.0, seed=1000 * i + r))
    B = F.Batch(ctx, As, Ps)
    rows, segs = B.info()
    assert segs == len(As) and rows % 2048 == 0 and rows >= sum(A.n for A in As)
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    xv = [F.vector(ctx, "f64", A.n).fill(0.0) for A in As]
    res = B.solve(bv, xv)
    for s, (A, P, b) in enumerate(zip(As, Ps, bs)):
        x = xv[s].download()
        print(f"segment {s}: n={A.n} outer={res[s].res_iter} solved={res[s].is_solved}", flush=True)
        assert res[s].is_solved and _true_res(A, x, b) <= 1e-8
        S = F.F3R(ctx, A, P)
        x1 = F.vector(ctx, "f64", A.n).fill(0.0)
        r1 = S.solve(bv[s], x1)
        assert abs(r1.res_iter - res[s].res_iter) <= 1, (s, r1.res_iter, res[s].res_iter)
        if s % 4 == 0:
            scale = float(2.0 ** np.floor(np.log2(max(1.0, np.sqrt(A.n) / 4.0))))
            _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, range_scale=scale)
            assert abs(ro.res_iter - res[s].res_iter) <= max(1, round(0.1 * ro.res_iter))
    ctx.close()


def test_lockstep_batch_non_power_of_two_scale():
    """a range scale that is not a power of two keeps level 3's Z in fp32 (its
    fp16 storage is exact only for powers of two); the solves still match
    single-system solves with the same scale"""
    import f3r_amd as F
    ctx = F.Context(0)
    As, Ps, bs = [], [], []
    for (i, npts, nx, delta, A) in _family(F, 2, seed=5):
        P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
        for r in range(2):
            As.append(A)
            Ps.append(P)
            bs.append(F.uniform(A.n, -1.0, 1.0, seed=77 * i + r))
    B = F.Batch(ctx, As, Ps, range_scale=3.0)
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    xv = [F.vector(ctx, "f64", A.n).fill(0.0) for A in As]
    res = B.solve(bv, xv)
    for s, (A, P, b) in enumerate(zip(As, Ps, bs)):
        print(f"segment {s}: n={A.n} outer={res[s].res_iter}", flush=True)
        assert res[s].is_solved and _true_res(A, xv[s].download(), b) <= 1e-8
        S = F.F3R(ctx, A, P, range_scale=3.0)
        x1 = F.vector(ctx, "f64", A.n).fill(0.0)
        r1 = S.solve(bv[s], x1)
        assert abs(r1.res_iter - res[s].res_iter) <= 1, (s, r1.res_iter, res[s].res_iter)
    ctx.close()


def test_lockstep_batch_segments_are_independent():
    """a zero right-hand side and a solved start finish at step 0 without
    disturbing the others; a segment's result does not depend on its company"""
    import numpy as np
    import f3r_amd as F
    ctx = F.Context(0)
    A1 = F.CSR.dominant(7, 12, margin=0.7)
    F.algorithm.scaling(A1)
    A2 = F.CSR.dominant(27, 9, margin=1.2)
    F.algorithm.scaling(A2)
    P1 = F.SDAINV(A1, 0.1, 1.0, "f16", ctx=ctx)
    P2 = F.SDAINV(A2, 0.1, 1.0, "f16", ctx=ctx)
    b1 = F.uniform(A1.n, -1.0, 1.0, seed=5)
    b2 = F.uniform(A2.n, -1.0, 1.0, seed=6)
    # alone
    B = F.Batch(ctx, [A1], [P1])
    x = F.vector(ctx, "f64", A1.n).fill(0.0)
    alone = B.solve([F.vector.from_numpy(ctx, b1)], [x])[0]
    x_alone = x.download()
    # in company: A2 with b2, A1 with 0, A1 with b1
    B3 = F.Batch(ctx, [A2, A1, A1], [P2, P1, P1])
    bv = [F.vector.from_numpy(ctx, b2), F.vector(ctx, "f64", A1.n).fill(0.0), F.vector.from_numpy(ctx, b1)]
    xv = [F.vector(ctx, "f64", A2.n).fill(0.0), F.vector(ctx, "f64", A1.n).fill(0.0),
          F.vector(ctx, "f64", A1.n).fill(0.0)]
    res = B3.solve(bv, xv)
    assert res[1].is_solved and res[1].res_iter == 0 and not np.any(xv[1].download())
    assert res[2].res_iter == alone.res_iter and np.array_equal(xv[2].download(), x_alone)
    assert res[0].is_solved and _true_res(A2, xv[0].download(), b2) <= 1e-8
    # warm start from the solution: solved at step 0
    res2 = B3.solve(bv, xv)
    assert all(r.is_solved and r.res_iter == 0 for r in res2)
    ctx.close()


def test_lockstep_batch_shared_matrix_copies_identical():
    """four right-hand sides per system: one stored matrix serving all four
    copies gives the same solutions, bit for bit, as one matrix copy per
    segment (the same segments interleaved, so no system repeats in a row)"""
    import numpy as np
    import f3r_amd as F
    ctx = F.Context(0)
    A1 = F.CSR.dominant(7, 13, margin=0.6)
    F.algorithm.scaling(A1)
    A2 = F.CSR.dominant(27, 10, margin=0.9)
    F.algorithm.scaling(A2)
    P1 = F.SDAINV(A1, 0.1, 1.0, "f16", ctx=ctx)
    P2 = F.SDAINV(A2, 0.1, 1.0, "f16", ctx=ctx)
    rhs = {(k, r): F.uniform(A.n, -1.0, 1.0, seed=100 * k + r) for k, A in ((1, A1), (2, A2)) for r in range(4)}
    grouped = [(1, r) for r in range(4)] + [(2, r) for r in range(4)]
    mixed = [(k, r) for r in range(4) for k in (1, 2)]
    sols = {}
    for order in (grouped, mixed):
        As = [A1 if k == 1 else A2 for (k, r) in order]
        Ps = [P1 if k == 1 else P2 for (k, r) in order]
        B = F.Batch(ctx, As, Ps)
        bv = [F.vector.from_numpy(ctx, rhs[key]) for key in order]
        xv = [F.vector(ctx, "f64", A.n).fill(0.0) for A in As]
        res = B.solve(bv, xv)
        assert all(r.is_solved for r in res)
        for key, x, r in zip(order, xv, res):
            sols.setdefault(key, []).append((r.res_iter, x.download()))
    for key, ((i0, x0), (i1, x1)) in sols.items():
        assert i0 == i1 and np.array_equal(x0, x1), key
    ctx.close()


def test_lockstep_rejects_bad_parameters():
    # every invalid combination is refused before any launch
    import ctypes

    import f3r_amd as F
    from f3r_amd import _capi
    ctx = F.Context(0)
    A = F.CSR.stencil27(8)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    for kw in ({"C": 16}, {"C": 32}, {"C": 96}, {"C": 4096}, {"K": 3}, {"K": 16}):
        with pytest.raises(F.F3RError) as e:
            F.Batch(ctx, [A, A], [P, P], **kw)
        assert e.value.status == -1, kw
    for field, value in (("variant", 1), ("static_weight", 1.0), ("depth", 3)):
        p = _capi.Params()
        _capi.lib.f3r_params_default(ctypes.byref(p))
        setattr(p, field, value)
        ah = (ctypes.c_void_p * 1)(A.handle(ctx))
        ph = (ctypes.c_void_p * 1)(P.handle)
        h = ctypes.c_void_p()
        assert _capi.lib.f3r_batch_create(ctx.handle, 1, ah, ph, ctypes.byref(p), ctypes.byref(h)) == -1, field
    # a valid one still builds
    B = F.Batch(ctx, [A, A], [P, P], C=128, K=8)
    assert B.info()[1] == 2
    del B
    ctx.close()


@pytest.mark.parametrize("c", [64, 8])
def test_lockstep_graph_replay_equals_eager(c):
    """the level-2 body captured per probe pattern and replayed gives the
    eager launches' bits, over two solves in a row (the second replays the
    graphs the first captured, with host reads and copies between replays);
    c = 8 puts probe calls inside the captured body"""
    import numpy as np
    import f3r_amd as F
    ctx = F.Context(0)
    As, Ps, bs = [], [], []
    for (i, npts, nx, delta, A) in _family(F, 3, seed=11):
        P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
        for r in range(2):
            As.append(A)
            Ps.append(P)
            bs.append(F.uniform(A.n, -1.0, 1.0, seed=31 * i + r))
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    out = {}
    for g in (False, True):
        B = F.Batch(ctx, As, Ps, c=c, use_graph=g)
        runs = []
        for rep in range(2):
            xv = [F.vector(ctx, "f64", A.n).fill(0.0) for A in As]
            res = B.solve(bv, xv)
            runs.append(([r.res_iter for r in res], [r.res_nrm2 for r in res], [x.download() for x in xv]))
        out[g] = runs
        print(f"c={c} graph={g}: outer steps {runs[0][0]}", flush=True)
    for rep in range(2):
        it0, n0, x0 = out[False][rep]
        it1, n1, x1 = out[True][rep]
        assert it0 == it1 and n0 == n1
        assert all(np.array_equal(a, b) for a, b in zip(x0, x1))
    for s, (A, b) in enumerate(zip(As, bs)):
        assert _true_res(A, out[True][1][2][s], b) <= 1e-8
    ctx.close()


def test_lockstep_many_outer_steps_graph_cap():
    """more outer steps than the solver keeps whole-step graphs for (64): the
    steps past the cap replay only the level-2 bodies; every step's bits
    match the eager launches"""
    import numpy as np
    import f3r_amd as F
    ctx = F.Context(0)
    A = F.CSR.dominant(7, 20, margin=0.005)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    bs = [F.uniform(A.n, -1.0, 1.0, seed=5 + r) for r in range(2)]
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    out = {}
    for g in (False, True):
        B = F.Batch(ctx, [A, A], [P, P], m=(100, 1, 1, 1), eps=1e-10, use_graph=g)
        xv = [F.vector(ctx, "f64", A.n).fill(0.0) for _ in bs]
        res = B.solve(bv, xv)
        out[g] = ([r.res_iter for r in res], [r.res_nrm2 for r in res], [x.download() for x in xv],
                  [r.is_solved for r in res])
        print(f"graph={g}: outer steps {out[g][0]}", flush=True)
    assert out[False][0] == out[True][0] and out[False][1] == out[True][1]
    assert all(np.array_equal(a, b) for a, b in zip(out[False][2], out[True][2]))
    assert all(out[True][3]) and min(out[True][0]) > 64
    for x, b in zip(out[True][2], bs):
        assert _true_res(A, x, b) <= 1e-9
    ctx.close()


def test_lockstep_restart_cycles_graph_equals_eager():
    """outer cycles of m1 = 5 steps: the whole-step graphs of steps 0..4 are
    replayed in every restart cycle; the bits match the eager launches"""
    import numpy as np
    import f3r_amd as F
    ctx = F.Context(0)
    A = F.CSR.dominant(7, 16, margin=0.05)
    F.algorithm.scaling(A)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    bs = [F.uniform(A.n, -1.0, 1.0, seed=21 + r) for r in range(2)]
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    out = {}
    for g in (False, True):
        B = F.Batch(ctx, [A, A], [P, P], m=(5, 1, 1, 1), restart=20, use_graph=g)
        xv = [F.vector(ctx, "f64", A.n).fill(0.0) for _ in bs]
        res = B.solve(bv, xv)
        out[g] = ([r.res_iter for r in res], [r.outer_cycles for r in res], [x.download() for x in xv],
                  [r.is_solved for r in res])
        print(f"graph={g}: outer steps {out[g][0]} cycles {out[g][1]}", flush=True)
    assert out[False][0] == out[True][0] and out[False][1] == out[True][1]
    assert all(np.array_equal(a, b) for a, b in zip(out[False][2], out[True][2]))
    assert all(out[True][3]) and min(out[True][1]) >= 3
    for x, b in zip(out[True][2], bs):
        assert _true_res(A, x, b) <= 1e-8
    ctx.close()


@pytest.mark.parametrize("nrhs", [1, 2, 4])
def test_lockstep_threads_per_row(nrhs):
    """the lockstep SpMVs with K threads per row (int32-index layout): every
    segment solved to 1e-8 within one outer step of the one-thread-per-row
    batch, and graph replay bit-identical to eager launches; nrhs = 1 runs
    the single-copy kernels, 2 and 4 the multi-copy ones"""
    import numpy as np
    import f3r_amd as F
    ctx = F.Context(0)
    As, Ps, bs = [], [], []
    for (i, npts, nx, delta, A) in _family(F, 4, seed=5):
        P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
        for r in range(nrhs):
            As.append(A)
            Ps.append(P)
            bs.append(F.uniform(A.n, -1.0, 1.0, seed=41 * i + r))
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    out = {}
    for name, kw in (("one", dict(index=0)), ("lanes", dict(index=0, lanes=True)),
                     ("lanes eager", dict(index=0, lanes=True, use_graph=False))):
        B = F.Batch(ctx, As, Ps, **kw)
        xv = [F.vector(ctx, "f64", A.n).fill(0.0) for A in As]
        res = B.solve(bv, xv)
        out[name] = ([r.res_iter for r in res], [x.download() for x in xv], [r.is_solved for r in res])
        print(f"nrhs={nrhs} {name}: outer steps {out[name][0]}", flush=True)
    assert all(out["lanes"][2])
    assert all(abs(a - b) <= 1 for a, b in zip(out["one"][0], out["lanes"][0]))
    assert out["lanes"][0] == out["lanes eager"][0]
    assert all(np.array_equal(a, b) for a, b in zip(out["lanes"][1], out["lanes eager"][1]))
    for s, (A, b) in enumerate(zip(As, bs)):
        assert _true_res(A, out["lanes"][1][s], b) <= 1e-8
    # K threads per row need the int32-index layout and K >= 2
    for kw in (dict(index=1, lanes=True), dict(index=0, K=1, lanes=True)):
        with pytest.raises(F.F3RError):
            F.Batch(ctx, As, Ps, **kw)
    ctx.close()


def test_lockstep_many_small_segments():
    """1024 segments (256 systems of 8^3 to 12^3 rows, 4 right-hand sides
    each), each padded to a 2048-row tile: every segment's reductions,
    rotations and weights stay its own, and every solve reaches 1e-8"""
    import numpy as np
    import f3r_amd as F
    from f3r_amd import batch
    ctx = F.Context(0)
    rng = np.random.default_rng(17)
    As, Ps, bs, uniq = [], [], [], []
    for i in range(256):
        npts = 7 if i % 2 == 0 else 27
        A = batch.make_system(npts, int(rng.integers(8, 13)), float(rng.uniform(0.5, 2.0)))
        P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
        uniq.append(A)
        for r in range(4):
            As.append(A)
            Ps.append(P)
            bs.append(F.uniform(A.n, -1.0, 1.0, seed=1000 * i + r))
    B = F.Batch(ctx, As, Ps)
    rows, segs = B.info()
    assert segs == 1024 and rows == 1024 * 2048
    bv = [F.vector.from_numpy(ctx, b) for b in bs]
    xv = [F.vector(ctx, "f64", A.n).fill(0.0) for A in As]
    res = B.solve(bv, xv)
    print(f"1024 segments: outer steps {min(r.res_iter for r in res)}..{max(r.res_iter for r in res)}", flush=True)
    assert all(r.is_solved for r in res)
    worst = max(_true_res(A, x.download(), b) for A, x, b in zip(As, xv, bs))
    assert worst <= 1e-8, worst
    ctx.close()
