"""Batched solves of the seeded family (independent systems on concurrent
streams, autotuned SELL layouts): every system and right-hand side reaches a
true fp64 relative residual of 1e-8."""
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


def test_autotune_picks_the_fastest():
    import f3r_amd as F
    from f3r_amd.tune import autotune
    ctx = F.Context(0)
    A = F.CSR.stencil27(48, beta=0.5)
    F.algorithm.scaling(A)
    best, table = autotune(ctx, A, "f16", Cs=(32, 64, 256), sigmas=(1, 256), Ks=(1, 4), reps=10)
    assert best["ms"] == min(t[0] for t in table) and best["ms"] > 0
    assert len(table) == 3 * 2 * 2   # every (C, sigma, K) with sigma = 1 or sigma >= C
    ctx.close()
