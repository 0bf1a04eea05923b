"""Composed nests (f3r_nest_create; the C++ layer include/f3r_amd.hpp and the
Python descriptors FGMRES / AutoRichardson / Operator / Restarted), against
the solvers they must equal:

  - the F3R nest composed from objects = f3r_solver_create, bit for bit
    (the C++ driver f3r_amd/f3r_nest, written in the structure of
    work/src/f3r-gpu.cpp:210-266, and the Python descriptors);
  - M applied directly under two levels = the F3 depth ablation, and the same
    nest with M inside a caller's Operator = M applied directly, bit for bit;
  - a caller's stopping rule (Solver::solve's cond) and a failing operator.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ctx():
    import f3r_amd as F
    c = F.Context(0)
    yield c
    c.close()


def test_cpp_layer_driver_checks():
    exe = os.path.join(ROOT, "f3r_amd", "f3r_nest")
    print("f3r_nest gen:27:24:0.5 1.0 1 8 4 2 64 20 check", flush=True)
    out = subprocess.run([exe, "gen:27:24:0.5", "1.0", "1", "8", "4", "2", "64", "20", "check"],
                         capture_output=True, text=True, timeout=180)
    print(out.stdout, out.stderr, flush=True)
    lines = out.stdout.strip().split("\n")
    f = lines[0].split(",")
    # Problem,Method,Prec,M2,M3,M4,W,Precond,ACC,Time,Iter,ImplRes,ExplRes
    assert f[0] == "stencil27_24" and float(f[11]) < 1e-8 and float(f[12]) < 1e-8
    checks = [ln for ln in lines if ln.startswith("check ")]
    assert len(checks) == 6 and all(": ok" in ln for ln in checks), checks
    assert out.returncode == 0


def _system(F, kind, m):
    if kind == "lognormal":
        A = F.CSR.convdiff7(m, sigma=1.5, corr=3.0, peclet=0.0, seed=7)
    else:
        A = F.CSR.stencil27(m, beta=0.5)
    F.algorithm.scaling(A)
    return A


@pytest.mark.parametrize("kind", ["lognormal", "27pt"])
def test_python_f3r_nest_equals_solver_create(ctx, kind):
    import f3r_amd as F
    A = _system(F, kind, 20)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    b = F.uniform(A.n, -1.0, 1.0, seed=3)
    bv = F.vector.from_numpy(ctx, b)
    x1 = F.vector(ctx, "f64", A.n).fill(0.0)
    r1 = F.F3R(ctx, A, P).solve(bv, x1)
    R = F.AutoRichardson("f16", P, 2, 64)
    nest = F.FGMRES("f64", F.FGMRES("f32", F.FGMRES("f16", R, 4), 8), 100, normalize=True, T="f64")
    S = F.Restarted(ctx, A, nest, 20)
    x2 = F.vector(ctx, "f64", A.n).fill(0.0)
    nb = np.linalg.norm(b)
    r2 = S.solve(bv, x2, cond=lambda i, e: e / nb < 1e-8)
    print(f"{kind}: {r1.res_iter} / {r2.res_iter} outer steps", flush=True)
    assert r1.res_iter == r2.res_iter and r1.res_nrm2 == r2.res_nrm2
    assert np.array_equal(x1.download(), x2.download())


def test_python_operator_innermost_equals_prec(ctx):
    # non-diagonal SDAINV factors (7-point lognormal): the operator applies the
    # caller's own preconditioner handle on the solver's fp16 operands
    import f3r_amd as F
    A = _system(F, "lognormal", 18)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    b = F.uniform(A.n, -1.0, 1.0, seed=4)
    bv = F.vector.from_numpy(ctx, b)
    calls = []

    def apply(vin, vout):
        calls.append(1)
        P.operate(vin, vout)

    xs = []
    for inner in (P, F.Operator(ctx, A.n, apply, "f16")):
        nest = F.FGMRES("f64", F.FGMRES("f32", F.FGMRES("f16", inner, 8), 8), 100, normalize=True, T="f64")
        x = F.vector(ctx, "f64", A.n).fill(0.0)
        r = F.Restarted(ctx, A, nest, 20).solve(bv, x)
        assert r.is_solved
        xs.append(x.download())
    assert len(calls) > 0
    assert np.array_equal(xs[0], xs[1])
    rel = np.linalg.norm(b - O.spmv_ld(A.n, A.rptr, A.idx, A.val, xs[1])) / np.linalg.norm(b)
    assert rel <= 1e-8


def test_python_cond_and_operator_error(ctx):
    import f3r_amd as F
    A = _system(F, "27pt", 16)
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx)
    bv = F.vector.from_numpy(ctx, F.uniform(A.n, -1.0, 1.0, seed=5))
    nest = F.FGMRES("f64", F.FGMRES("f32", F.FGMRES("f16", F.AutoRichardson("f16", P, 2, 64), 4), 8), 100,
                    normalize=True, T="f64")
    S = F.Restarted(ctx, A, nest, 20)
    x = F.vector(ctx, "f64", A.n).fill(0.0)
    seen = []
    r = S.solve(bv, x, cond=lambda i, e: seen.append(i) or i >= 1)
    assert r.res_iter == 1 and r.is_solved and seen == [0, 1]

    # an exception in the stopping rule stops the solve and reaches the caller
    class Stop(Exception):
        pass

    def stop_rule(i, e):
        if i >= 1:
            raise Stop("stop rule failed")
        return False
    with pytest.raises(Stop):
        S.solve(bv, F.vector(ctx, "f64", A.n).fill(0.0), cond=stop_rule)

    class Boom(Exception):
        pass

    def bad(vin, vout):
        raise Boom("operator failed")
    nest = F.FGMRES("f64", F.FGMRES("f32", F.Operator(ctx, A.n, bad), 4), 100, normalize=True, T="f64")
    S = F.Restarted(ctx, A, nest, 20)
    with pytest.raises(Boom):
        S.solve(bv, F.vector(ctx, "f64", A.n).fill(0.0))
    # a Richardson under an fp16-basis level is refused at creation
    nest = F.FGMRES("f64", F.FGMRES("f16", F.AutoRichardson("f16", P, 2, 64), 4, T="f16"), 100, normalize=True,
                    T="f64")
    with pytest.raises(F.F3RError):
        F.Restarted(ctx, A, nest, 20)
