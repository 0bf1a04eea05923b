"""Ablation study of the nest: nesting depth and the Richardson weights.

For one system A x = b (scaled as the reference driver scales it,
algorithm.hpp:45-59; b uniform in [-1, 1]) every nest below solves to the
fp64 relative residual eps with the same fp16 preconditioner M:

  * F3R: fp64 FGMRES(100) > fp32 FGMRES(8) > fp16-value FGMRES(4) > adaptive
    Richardson(2) with M (work/src/f3r-gpu.cpp:210-266);
  * F3R-Static(w): the same with every Richardson weight fixed at w and no
    probes (work/src/static.cpp:99-119);
  * F2, F2H, F3, F3H, F4: M applied directly under one to three FGMRES levels
    (work/src/depth.cpp:122-167); "H" levels store their basis in fp16.

Reported per nest: solve time (median of `reps`), outer steps, innermost
applications, the implicit and the true fp64 relative residual.  The CSV
lines follow the reference drivers' format (name, solver, type, m2, m3, m4,
c or weight, precond, acc, time, innermost applications, implicit residual,
true residual; work/src/depth.cpp:169-187, static.cpp:121-134).

python -m f3r_amd.ablation --nx 128 --matrix 27pt --precond sdainv
"""
import json
import time

import numpy as np

from . import CSR, F3R, ILU0, SDAINV, SELL, Context, algorithm, nrm, uniform, vector

NESTS = [("F3R", {}), ("F3R-Static", {"static_weight": 0.8}), ("F3R-Static", {"static_weight": 1.0}),
         ("F3R-Static", {"static_weight": 1.2}), ("F2", {"depth": "F2"}), ("F2H", {"depth": "F2H"}),
         ("F3", {"depth": "F3"}), ("F3H", {"depth": "F3H"}), ("F4", {"depth": "F4"})]


def system(matrix, nx):
    if matrix == "27pt":
        A = CSR.stencil27(nx, beta=0.5)
    else:
        A = CSR.convdiff7(nx, sigma=1.5, corr=8.0, peclet=10.0, seed=0)
    algorithm.scaling(A)
    return A


def run(ctx, A, b, precond="sdainv", reps=3, m=(100, 8, 4, 2), restart=20, eps=1e-8, log=None, name="system"):
    n = A.n
    if precond == "sdainv":
        P = SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=True)
        ptag = "SDAINV,0.1,"
    else:
        P = ILU0(A, 112, 1.0, "f16", ctx=ctx, device=True)
        ptag = "BJILU0,1.0,"
    S64 = SELL(ctx, A, "f64")
    vb = vector.from_numpy(ctx, b)
    r = vector(ctx, "f64", n)
    rows, csv = [], []
    for label, kw in NESTS:
        S = F3R(ctx, A, P, m=m, restart=restart, eps=eps, device_build=True, **kw)
        x = vector(ctx, "f64", n)
        times, res = [], None
        x.fill(0.0)
        S.solve(vb, x)  # warm-up: the inner levels' graphs are captured here
        for _ in range(reps):
            x.fill(0.0)
            ctx.sync()
            res = S.solve(vb, x)
            times.append(res.solve_s)
        S64.compute_residual(vb, x, r)
        true_rel = nrm(r) / nrm(vb)
        t = float(np.median(times))
        w = kw.get("static_weight")
        row = {"nest": label if w is None else f"{label}({w})", "solved": bool(res.is_solved),
               "outer": res.res_iter, "innermost": int(res.precond_calls), "solve_ms": round(t * 1e3, 2),
               "implicit_rel_res": res.res_nrm2 / res.nrm_b, "true_rel_res": true_rel}
        rows.append(row)
        m2, m3, m4 = {"F2": (64, 0, 0), "F2H": (64, 0, 0), "F3": (8, 8, 0), "F3H": (8, 8, 0),
                      "F4": (8, 4, 2)}.get(label, m[1:])
        csv.append(f"{name},{label},half,{m2},{m3},{m4},{w if w is not None else 64},{ptag}"
                   f"{t:e},{int(res.precond_calls)},{row['implicit_rel_res']:e},{true_rel:e}")
        if log:
            log(f"{row['nest']:>16s}: solved {row['solved']} outer {row['outer']:4d} innermost "
                f"{row['innermost']:6d} {row['solve_ms']:9.2f} ms true rel. res {true_rel:.2e}")
        del S
    return rows, csv


def markdown(rows, title):
    out = [f"### {title}", "", "| nest | solved | outer steps | innermost applications | solve ms | true rel. res |",
           "|---|---|---:|---:|---:|---:|"]
    for r in rows:
        out.append(f"| {r['nest']} | {r['solved']} | {r['outer']} | {r['innermost']} | {r['solve_ms']} | "
                   f"{r['true_rel_res']:.2e} |")
    return "\n".join(out) + "\n"


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="nesting-depth and static-weight ablations of the F3R nest")
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--matrix", choices=["27pt", "lognormal"], default="27pt")
    ap.add_argument("--precond", choices=["sdainv", "bjilu0"], default="sdainv")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.time()

    def log(msg):
        print(f"[ablation {time.time() - t0:6.1f}s] {msg}", flush=True)

    A = system(a.matrix, a.nx)
    b = uniform(A.n, -1.0, 1.0, seed=2024)
    log(f"{a.matrix} {a.nx}^3: n={A.n} nnz={A.nnz}, precond {a.precond}")
    ctx = Context(0)
    rows, csv = run(ctx, A, b, a.precond, a.reps, log=log, name=f"{a.matrix}{a.nx}")
    for line in csv:
        print(line, flush=True)
    out = {"matrix": a.matrix, "nx": a.nx, "n": A.n, "nnz": A.nnz, "precond": a.precond, "rows": rows,
           "csv": csv, "markdown": markdown(rows, f"{a.matrix} {a.nx}^3, {a.precond}")}
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k not in ("markdown", "csv")}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
