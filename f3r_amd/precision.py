"""Precision study of the nested solve: the error each level adds, per
variant, against fp64.

For one system A x = b (scaled as the reference driver scales it,
algorithm.hpp:45-59) the study reports

  * x*, the fp64 reference solution: the DOUBLE variant (every level fp64,
    src/f3r.cpp:105-133) solved to a 1e-12 relative residual;
  * for each level k = 2, 3, 4 (level 2 = fp32 FGMRES(m2), 3 = FGMRES(m3)
    on fp16 matrix values, 4 = adaptive Richardson on fp16 data) and each
    variant v in {half, float, double}, one application z = L_k^v(v) to the
    unit vector v = b / ||b|| (inner levels take unit-norm Krylov vectors:
    their Arnoldi starts from the input unnormalised, solver/gmres.hpp:97-102),
    from the same fresh Richardson state:
        op_err  = ||z - L_k^double(v)|| / ||L_k^double(v)||   (what precision costs)
        inv_res = ||v - A z||                                  (how good an inverse the level is)
  * for each variant the full solve at eps = 1e-8: outer steps, solve time,
    true residual and ||x - x*|| / ||x*||.

python -m f3r_amd.precision --nx 192   (the 7-point lognormal / Peclet-10 system)
"""
import json
import time

import numpy as np

from . import CSR, F3R, SDAINV, SELL, Context, algorithm, nrm, uniform, vector


def convdiff_system(nx, sigma=1.5, corr=8.0, peclet=10.0, seed=0):
    A = CSR.convdiff7(nx, sigma=sigma, corr=corr, peclet=peclet, seed=seed)
    algorithm.scaling(A)
    return A


def study(ctx, A, b, m=(100, 8, 4, 2), c=64, tol=0.1, acc=1.0, log=None):
    n = A.n
    S64 = SELL(ctx, A, "f64")
    vb = vector.from_numpy(ctx, b)
    vu = vector.from_numpy(ctx, b / np.linalg.norm(b))
    r = vector(ctx, "f64", n)

    def resid(x, rhs=vb):
        S64.compute_residual(rhs, x, r)
        return nrm(r) / nrm(rhs)

    solvers = {}
    for v, dt in (("half", "f16"), ("float", "f32"), ("double", "f64")):
        P = SDAINV(A, tol, acc, dt, ctx=ctx)
        solvers[v] = (F3R(ctx, A, P, m=m, c=c, variant=v), P)
    # fp64 reference solution
    t0 = time.time()
    Sref = F3R(ctx, A, solvers["double"][1], m=m, c=c, variant="double", eps=1e-12, restart=50)
    xs = vector(ctx, "f64", n).fill(0.0)
    rs = Sref.solve(vb, xs)
    xstar = xs.download()
    out = {"n": n, "nnz": A.nnz, "reference": {"outer": rs.res_iter, "solved": rs.is_solved,
                                                 "true_rel_res": resid(xs)}}
    if log:
        log(f"fp64 reference: {rs.res_iter} outer steps, true rel. res {out['reference']['true_rel_res']:.2e} "
            f"({time.time() - t0:.1f} s)")
    # one application of every level
    z = vector(ctx, "f64", n)
    levels = {}
    for k in (2, 3, 4):
        apps = {}
        for v in ("double", "float", "half"):
            solvers[v][0].level_apply(k, vu, z)
            apps[v] = z.download()
        ref = apps["double"]
        row = {}
        for v, zz in apps.items():
            zv = vector.from_numpy(ctx, zz)
            row[v] = {"op_err": float(np.linalg.norm(zz - ref) / np.linalg.norm(ref)), "inv_res": resid(zv, vu)}
        levels[k] = row
        if log:
            log(f"level {k}: " + ", ".join(f"{v} op_err {row[v]['op_err']:.2e} inv_res {row[v]['inv_res']:.3e}"
                                             for v in ("half", "float", "double")))
    out["levels"] = levels
    # full solves
    full = {}
    ns = np.linalg.norm(xstar)
    for v in ("half", "float", "double"):
        x = vector(ctx, "f64", n).fill(0.0)
        res = solvers[v][0].solve(vb, x)
        xv = x.download()
        full[v] = {"outer": res.res_iter, "solved": res.is_solved, "solve_s": res.solve_s,
                   "true_rel_res": resid(x), "err_vs_fp64": float(np.linalg.norm(xv - xstar) / ns)}
        if log:
            log(f"solve {v}: {res.res_iter} outer steps, {res.solve_s * 1e3:.1f} ms, "
                f"true rel. res {full[v]['true_rel_res']:.2e}, error vs fp64 {full[v]['err_vs_fp64']:.2e}")
    out["solves"] = full
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="per-level precision study (7-point lognormal convection-diffusion)")
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--sigma", type=float, default=1.5)
    ap.add_argument("--corr", type=float, default=8.0)
    ap.add_argument("--peclet", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.time()

    def log(msg):
        print(f"[precision {time.time() - t0:6.1f}s] {msg}", flush=True)

    A = convdiff_system(a.nx, a.sigma, a.corr, a.peclet, a.seed)
    b = uniform(A.n, -1.0, 1.0, seed=a.seed)
    log(f"matrix {a.nx}^3: n={A.n} nnz={A.nnz}")
    ctx = Context(0)
    out = study(ctx, A, b, log=log)
    out["config"] = {"nx": a.nx, "sigma": a.sigma, "corr": a.corr, "peclet": a.peclet, "seed": a.seed}
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
