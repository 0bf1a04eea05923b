"""The five workload configurations at their stated sizes, one named test each:

  C1  plain F3R, symmetric 27-point 64^3, b = A*1, restart 3 (the CPU-path
      configuration): outer steps equal to the reference's own run
      (tests/golden/f3r_ref_steps.npz) and to the oracle, true residual <= 1e-8;
  C2  plain F3R, nonsymmetric 27-point 128^3 (beta = 0.5), b ~ U[-1, 1]:
      outer steps within +-10% of a full oracle solve in this build's
      arithmetic and of one in the reference's arithmetic, true residual <= 1e-8;
  C3  the precision study on the 7-point lognormal convection-diffusion
      system at 192^3: every variant solves to 1e-8, the per-level operator
      errors are ordered by precision, the solutions agree with the fp64 x*;
  C4  64 systems x 4 right-hand sides in lockstep: every segment <= 1e-8,
      sampled segments within +-10% (at least +-1) of the oracle's steps;
  C5  the committed learned SDAINV correction on 16 held-out 96^3 systems:
      every solve <= 1e-8, fewer outer steps in total and less mean solve
      time than plain SDAINV.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import problems as PB

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ctx():
    import f3r_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _scale(n):
    return float(2.0 ** np.floor(np.log2(np.sqrt(n) / 4.0)))


def _true_res(F, ctx, A, bv, xv):
    S64 = F.SELL(ctx, A, "f64", C=64, sigma=1, K=4)
    r = F.vector(ctx, "f64", A.n)
    S64.compute_residual(bv, xv, r)
    return F.nrm(r) / F.nrm(bv)


def test_c1_symmetric_27pt_64_rowsum_rhs(ctx, golden):
    import f3r_amd as F
    g = golden("f3r_ref_steps")
    key = "c1s27_64_m842"
    d = json.loads(str(g[f"{key}_desc"]))
    n, rp, ix, v, b = PB.from_desc(d)
    A = F.CSR(n, rp, ix, v.copy())
    F.algorithm.scaling(A)
    print(f"C1 n={n} nnz={A.nnz}", flush=True)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=True)
    S = F.F3R(ctx, A, P, restart=3)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", n).fill(0.0)
    res = S.solve(bv, xv)
    _, ro, _ = O.f3r_solve(n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=_scale(n), restart=3)
    ref = int(g[f"{key}_p16_res"][1])
    rel = _true_res(F, ctx, A, bv, xv)
    print(f"  gpu {res.res_iter} steps, oracle {ro.res_iter}, reference {ref}, true residual {rel:.3e}", flush=True)
    assert res.is_solved and rel <= 1e-8
    assert res.res_iter == ref == ro.res_iter


def test_c2_27pt_nonsymmetric_128(ctx):
    import f3r_amd as F
    A = F.CSR.stencil27(128, beta=0.5)
    F.algorithm.scaling(A)
    b = F.uniform(A.n, -1.0, 1.0, seed=2024)
    print(f"C2 n={A.n} nnz={A.nnz}", flush=True)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=True)
    S = F.F3R(ctx, A, P, device_build=True)
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n).fill(0.0)
    res = S.solve(bv, xv)
    rel = _true_res(F, ctx, A, bv, xv)
    print(f"  gpu {res.res_iter} steps in {res.solve_s * 1e3:.2f} ms, true residual {rel:.3e}", flush=True)
    _, rb, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=_scale(A.n))
    print(f"  oracle (this build's arithmetic) {rb.res_iter} steps", flush=True)
    _, rr, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=1)
    print(f"  oracle (the reference's arithmetic) {rr.res_iter} steps", flush=True)
    assert res.is_solved and rel <= 1e-8
    for ref in (rb.res_iter, rr.res_iter):
        assert abs(res.res_iter - ref) <= int(0.1 * ref), (res.res_iter, ref)


def test_c3_precision_study_192(ctx):
    from f3r_amd import precision, uniform
    A = precision.convdiff_system(192)
    b = uniform(A.n, -1.0, 1.0, seed=0)
    print(f"C3 n={A.n} nnz={A.nnz}", flush=True)
    out = precision.study(ctx, A, b, log=lambda m: print("  " + m, flush=True))
    assert out["reference"]["solved"] and out["reference"]["true_rel_res"] <= 1e-11
    steps = [out["solves"][v]["outer"] for v in ("half", "float", "double")]
    for v in ("half", "float", "double"):
        s = out["solves"][v]
        assert s["solved"] and s["true_rel_res"] <= 1e-8, (v, s)
        # a 1e-8 residual bounds the error by the conditioning: far below 1e-4
        assert s["err_vs_fp64"] < 1e-4, (v, s)
    assert max(steps) - min(steps) <= 1, steps
    for k in (2, 3, 4):
        row = out["levels"][k]
        assert row["double"]["op_err"] == 0.0
        # what precision costs: fp32 levels near fp32 rounding, fp16 near fp16
        assert row["float"]["op_err"] < 1e-5 and row["half"]["op_err"] < 1e-2, (k, row)
        assert row["half"]["op_err"] > row["float"]["op_err"]
        # every precision is as good an inverse to 1%
        ir = [row[v]["inv_res"] for v in ("half", "float", "double")]
        assert max(ir) <= 1.01 * min(ir), (k, ir)


def test_c4_batch_64x4_lockstep(ctx):
    import f3r_amd as F
    from f3r_amd import batch
    out = batch.run_lockstep(64, 4, seed=0, log=lambda m: print("  " + m, flush=True))
    assert out["solves"] == 256 and out["all_solved"]
    assert out["worst_true_rel_res"] <= 1e-8
    # sampled segments against the oracle: the three smallest systems of each
    # stencil, first right-hand side
    fam = batch.family(0, 64)
    picks = sorted(fam, key=lambda t: t[2])
    picks = [t for t in picks if t[1] == 7][:2] + [t for t in picks if t[1] == 27][:2]
    steps = out["outer_steps_per_segment"]
    for (i, npts, nx, delta) in picks:
        A = batch.make_system(npts, nx, delta)
        b = F.uniform(A.n, -1.0, 1.0, seed=16 * i)
        _, ro, _ = O.f3r_solve(A.n, A.rptr, A.idx, A.val, b, variant=0, ref_arith=0, range_scale=_scale(A.n))
        got = steps[4 * i]
        print(f"  system {i} ({npts}-point {nx}^3): lockstep {got} steps, oracle {ro.res_iter}", flush=True)
        assert ro.is_solved and abs(got - ro.res_iter) <= max(1, int(0.1 * ro.res_iter))


def test_c5_learned_vs_plain_96_heldout(ctx):
    from f3r_amd import learn
    path = os.path.join(ROOT, "f3r_amd", "data", "learned_sdainv_96.npz")
    d = np.load(path, allow_pickle=False)
    mlp = learn.MLP(ctx, params=d["params"])
    normalize = int(d["kind"]) == 1 if "kind" in d.files else True
    seeds = list(range(100000, 100016))
    plain_s, learned_s, plain_it, learned_it = [], [], [], []
    for hs in learn.host_systems(96, seeds, threads=8):
        r = learn.evaluate(ctx, mlp, hs, reps=3, normalize=normalize)
        print(f"  seed {hs['seed']}: plain {r['plain']['outer']} steps {r['plain']['solve_s'] * 1e3:.2f} ms, "
              f"learned {r['learned']['outer']} steps {r['learned']['solve_s'] * 1e3:.2f} ms", flush=True)
        for k in ("plain", "learned"):
            assert r[k]["solved"] and r[k]["true_rel_res"] <= 1e-8, (hs["seed"], k, r[k])
        plain_s.append(r["plain"]["solve_s"])
        learned_s.append(r["learned"]["solve_s"])
        plain_it.append(r["plain"]["outer"])
        learned_it.append(r["learned"]["outer"])
    print(f"  mean plain {np.mean(plain_s) * 1e3:.3f} ms ({sum(plain_it)} steps), learned "
          f"{np.mean(learned_s) * 1e3:.3f} ms ({sum(learned_it)} steps)", flush=True)
    assert sum(learned_it) < sum(plain_it)
    assert np.mean(learned_s) < np.mean(plain_s)
