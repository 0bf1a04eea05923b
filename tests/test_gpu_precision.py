"""Precision study on the GPU (7-point lognormal convection-diffusion):
the fp64 reference solution against the oracle's fp64 GMRES, and the error
each level adds per variant."""
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


def test_precision_study_small(ctx):
    from f3r_amd import precision, uniform
    A = precision.convdiff_system(16, sigma=1.5, corr=4.0, peclet=10.0, seed=1)
    b = uniform(A.n, -1.0, 1.0, seed=1)
    out = precision.study(ctx, A, b, log=print)
    # the fp64 reference solution agrees with the oracle's fp64 GMRES(50)
    xo, it, rr = O.gmres(A.n, A.rptr, A.idx, A.val, b, m=50, tol=1e-13)
    assert rr <= 1e-12
    assert out["reference"]["true_rel_res"] <= 1e-11
    import f3r_amd as F
    ctx2 = ctx
    S = F.F3R(ctx2, A, F.SDAINV(A, 0.1, 1.0, "f64", ctx=ctx2), variant="double", eps=1e-12, restart=50)
    x = F.vector(ctx2, "f64", A.n).fill(0.0)
    S.solve(F.vector.from_numpy(ctx2, b), x)
    assert np.linalg.norm(x.download() - xo) / np.linalg.norm(xo) < 1e-9
    for k in (2, 3, 4):
        lv = out["levels"][k]
        assert lv["double"]["op_err"] == 0.0
        assert lv["float"]["op_err"] < 1e-4, k
        assert lv["half"]["op_err"] < 5e-2, k
        # every level is a contraction of the residual
        for v in ("half", "float", "double"):
            assert lv[v]["inv_res"] < 1.0, (k, v)
    # deeper levels are weaker inverses (fewer inner steps)
    assert out["levels"][2]["double"]["inv_res"] < out["levels"][3]["double"]["inv_res"] \
        < out["levels"][4]["double"]["inv_res"]
    for v, r in out["solves"].items():
        assert r["solved"] and r["true_rel_res"] <= 1e-8, v
        assert r["err_vs_fp64"] < 1e-5, v
