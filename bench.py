"""Benchmark: plain F3R solve time to a 1e-8 fp64 relative residual on one
MI355X, and the fp16 SpMV HBM bandwidth of the solver's matrix layout.

Workload (one step = one full solve from x0 = 0): the 27-point
nonsymmetric stencil on a 128^3 grid (2,097,152 rows, 55,742,968 nonzeros),
diagonal 26, off-diagonals -1 + 0.5*sign(dx), diagonally scaled as the
reference driver does; b uniform in [-1, 1) from a seed; SDAINV (tol 0.1,
acc 1.0); m = (100; 8, 4, 2), c = 64, restart 20.  All inputs are resident in
HBM before the timed region.

Learned vs plain (the second half of the metric): the per-row MLP correction
of SDAINV, trained on the 7-point lognormal-diffusion family at 96^3
(f3r_amd/learn.py; parameters in f3r_amd/data/), against plain SDAINV F3R on
held-out systems of that family, reported under "learned".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 128] [--no-cpu] [--learned N] [--batch 0|1]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solve time to 1e-8 fp64 rel. residual, learned vs plain F3R; fp16 SpMV HBM GB/s"
PEAK_HBM_GBS = 8000.0
# the solve's hot kernels (f3r_solver_time_kernel modes): the metric's fp16
# SpMV, then the four kernels of every level-3 Arnoldi step
KERNELS = (("spmv_f16", 0), ("l3_spmv_cgs_dots", 6), ("l4_richardson_step", 8), ("l3_cgs_update_norm", 7),
           ("l4_entry", 9))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(F, A, b, reps=3):
    """The oracle's F3R (the C++/OpenMP restatement of the reference's nest,
    fp16 emulated through F16C conversions) timed on the host's cores: full
    solves, solve only (set-up excluded, as work/src/f3r.cpp:72-74), mean of
    `reps` repetitions (work/src/f3r.cpp:261-272), on
      C2: the bench's own 128^3 nonsymmetric system, this build's arithmetic
          (fp32 sums of fp16 data; the reference's fp16 arithmetic emulated
          on the host is slower still, measured on C1);
      C1: the symmetric 27-point 64^3 system with b = A*1 and restart 3 (the
          CPU-path configuration), in both arithmetics."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as O

    def timed(Ab, bb, **kw):
        ts, r = [], None
        for _ in range(reps):
            _, r, _ = O.f3r_solve(Ab.n, Ab.rptr, Ab.idx, Ab.val, bb, **kw)
            ts.append(r.solve_s)
        return {"mean_s": round(float(np.mean(ts)), 4), "reps_s": [round(t, 4) for t in ts],
                "outer_steps": int(r.res_iter), "true_rel_res": float(r.expl_res), "solved": bool(r.is_solved)}, r

    scale2 = float(2.0 ** np.floor(np.log2(np.sqrt(A.n) / 4.0)))
    c2, r2 = timed(A, b, variant=0, ref_arith=0, range_scale=scale2)
    log(f"[bench] cpu C2: {c2}")
    A1 = F.CSR.stencil27(64)
    b1 = np.add.reduceat(A1.val, A1.rptr[:-1])  # b = A * 1 of the unscaled matrix
    F.algorithm.scaling(A1)
    scale1 = float(2.0 ** np.floor(np.log2(np.sqrt(A1.n) / 4.0)))
    c1, _ = timed(A1, b1, variant=0, ref_arith=0, range_scale=scale1, restart=3)
    c1r, _ = timed(A1, b1, variant=0, ref_arith=1, restart=3)
    log(f"[bench] cpu C1: {c1}, reference arithmetic {c1r}")
    return {"value": c2["mean_s"], "unit": "s", "cores": int(r2.nthreads), "kind": "port",
            "cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
            "sample": f"full F3R solves of the 128^3 bench system (C2), mean of {reps}, solve only; "
                      f"{int(r2.nthreads)} OpenMP threads on {cpu_model()}",
            "c2_128": c2,
            "c1_64_symmetric_rowsum_rhs": {"build_arithmetic": c1, "reference_arithmetic": c1r}}


def measured_traffic(kernel):
    """bytes per launch of one of the solve's kernels from the committed
    hardware-counter summary (two rocprofv3 --pmc runs of `bench.py
    --kernel-only`, FETCH_SIZE and WRITE_SIZE, each scaled by what the same
    counter reads for a copy kernel of known size); None when absent"""
    path = os.path.join(ROOT, "profiles", "kernels_c2_128_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if kernel not in t.get("kernels", {}):
        return None
    k = t["kernels"][kernel]
    return {"bytes": k["bytes_per_launch"], "fetch_bytes": k["fetch_bytes"], "write_bytes": k["write_bytes"],
            "fetch_scale": t["fetch_scale"], "write_scale": t["write_scale"],
            "source": "profiles/kernels_c2_128_traffic.md"}


def learned_vs_plain(F, ctx, count, nx=96, reps=3):
    """plain vs learned F3R on `count` held-out systems of the learned
    correction's family (seeds 100000...), best of `reps` solves each"""
    import numpy as np
    from f3r_amd import learn
    path = os.path.join(ROOT, "f3r_amd", "data", "learned_sdainv_96.npz")
    if not os.path.exists(path):
        return {"status": "no trained parameters at f3r_amd/data/learned_sdainv_96.npz"}
    d = np.load(path, allow_pickle=False)
    mlp = learn.MLP(ctx, params=d["params"])
    # a model trained with the scale-blind loss (kind 1) leaves the global
    # scale free: its scales are divided by their geometric mean
    kind = int(d["kind"]) if "kind" in d.files else 1
    seeds = list(range(100000, 100000 + count))
    rows = []
    for hs in learn.host_systems(nx, seeds, threads=min(8, count)):
        r = learn.evaluate(ctx, mlp, hs, reps=reps, normalize=kind == 1)
        log(f"[bench] learned: seed {hs['seed']} plain {r['plain']['outer']} steps {r['plain']['solve_s'] * 1e3:.2f} ms, "
            f"learned {r['learned']['outer']} steps {r['learned']['solve_s'] * 1e3:.2f} ms")
        rows.append({"seed": hs["seed"], "sigma": round(hs["sigma"], 3), "corr": round(hs["corr"], 2),
                     "plain_s": r["plain"]["solve_s"], "learned_s": r["learned"]["solve_s"],
                     "plain_outer": r["plain"]["outer"], "learned_outer": r["learned"]["outer"],
                     "plain_res": r["plain"]["true_rel_res"], "learned_res": r["learned"]["true_rel_res"]})
    ps = float(np.mean([r["plain_s"] for r in rows]))
    ls = float(np.mean([r["learned_s"] for r in rows]))
    return {"workload": f"7-point lognormal diffusion {nx}^3 (sigma ~ U[0.5,2], corr ~ U[2,16]), held-out seeds "
                        f"{seeds[0]}..{seeds[-1]}; MLP 12-64-64-1 bf16 trained on {int(d['train'])} systems",
            "plain_s": round(ps, 6), "learned_s": round(ls, 6), "speedup": round(ps / ls, 4),
            "all_solved": all(r["plain_res"] <= 1e-8 and r["learned_res"] <= 1e-8 for r in rows),
            "systems": rows}


def setup_timings(F, ctx, A, m):
    """wall time of each set-up step on the device and on the host (same
    factors: checked), for the bench matrix"""
    import numpy as np
    out = {"workload": f"set-up of the 27-point nonsymmetric {m}^3 matrix (n={A.n}, nnz={A.nnz})"}
    t = time.time()
    A.handle(ctx)
    ctx.sync()
    out["csr_upload_s"] = round(time.time() - t, 3)
    t = time.time()
    _, nl = A.factor_levels(True)
    out["levels"] = nl
    out["level_sets_s"] = round(time.time() - t, 3)
    for name, make in (("sdainv", lambda dev: F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=dev)),
                       ("ilu0", lambda dev: F.ILU0(A, 1, 1.0, "f16", ctx=ctx, device=dev))):
        facs = {}
        for dev in (True, False):
            t = time.time()
            P = make(dev)
            ctx.sync()
            out[f"{name}_{'device' if dev else 'host'}_s"] = round(time.time() - t, 3)
            out[f"{name}_{'device' if dev else 'host'}_where"] = P.where
            facs[dev] = P.factors()
            del P
        out[f"{name}_factors_identical"] = all(np.array_equal(a, b) for fa, fb in zip(facs[True], facs[False])
                                               for a, b in zip(fa, fb))
        log(f"[bench] {name}: device {out[f'{name}_device_s']} s, host {out[f'{name}_host_s']} s, "
            f"identical {out[f'{name}_factors_identical']}")
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=True)
    for dev in (True, False):
        t = time.time()
        S = F.F3R(ctx, A, P, device_build=dev)
        ctx.sync()
        out[f"solver_layouts_{'device' if dev else 'host'}_s"] = round(time.time() - t, 3)
        del S
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--blocks", type=int, default=0, help="blocks of the streaming reduction kernels (0 = auto)")
    ap.add_argument("--cap-spmv", action="store_true", help="run the fused SpMV+reduction kernels on that grid too")
    ap.add_argument("--host-build", action="store_true",
                    help="factorise SDAINV and build the matrix layouts on the host (default: device)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--batch", type=int, default=1, help="1: also time the 64 x 4 lockstep batch (0: skip)")
    ap.add_argument("--learned", type=int, default=16, metavar="N",
                    help="held-out 96^3 lognormal systems for learned vs plain F3R (0: skip)")
    ap.add_argument("--setup-only", action="store_true",
                    help="no solves: time the set-up steps (SDAINV and ILU(0) factors, solver layouts) on the "
                         "device and on the host, check the factors agree, print one JSON line and exit")
    ap.add_argument("--kernel-only", type=int, default=0, metavar="REPS",
                    help="no solves: launch each hot kernel of the solve (KERNELS) REPS rounds and a plain copy "
                         "of a known size REPS times (for hardware-counter runs), print their timings and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the solve is not sharded across GPUs "
                                   "(row blocks with halo exchange would shard it); run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import f3r_amd as F

    m = args.grid
    t0 = time.time()
    A = F.CSR.stencil27(m, beta=0.5)
    F.algorithm.scaling(A)
    b = F.uniform(A.n, -1.0, 1.0, seed=2024)
    log(f"[bench] matrix {m}^3: n={A.n} nnz={A.nnz} ({time.time() - t0:.1f} s)")
    ctx = F.Context(0)
    if args.setup_only:
        setup_timings(F, ctx, A, m)
        return
    t0 = time.time()
    P = F.SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=not args.host_build)
    factor_s = time.time() - t0
    S = F.F3R(ctx, A, P, use_graph=not args.no_graph, grid=args.blocks, cap_spmv=args.cap_spmv,
              device_build=not args.host_build)
    setup_s = time.time() - t0
    log(f"[bench] setup {setup_s:.2f} s (SDAINV factors on the {P.where}: {factor_s:.2f} s)")
    if args.kernel_only:
        # for hardware-counter runs: each kernel back to back `REPS` rounds,
        # in this order, then a copy of known size for calibration
        out = {"reps": args.kernel_only, "kernels": {}}
        for name, w in KERNELS:
            ms, by = S.time_kernel(w, reps=args.kernel_only)
            out["kernels"][name] = {"which": w, "us": round(ms * 1e3, 2), "algorithmic_bytes": by,
                                    "launches": (args.kernel_only + 1) * (S.params.m3 if w in (6, 7) else 1)}
        # calibration copy: 2^27 fp32 elements (512 MiB read, 512 MiB written),
        # far larger than the L2s and the last-level cache
        nc = 1 << 27
        xc = F.vector(ctx, "f32", nc).fill(1.0)
        yc = F.vector(ctx, "f32", nc)
        ctx.sync()
        tc = time.time()
        for _ in range(args.kernel_only):
            yc.copy(xc)
        ctx.sync()
        out.update({"copy_elements": nc, "copy_bytes_read": 4 * nc, "copy_bytes_written": 4 * nc,
                    "copy_launches": args.kernel_only,
                    "copy_us_wall": round((time.time() - tc) / args.kernel_only * 1e6, 2)})
        print(json.dumps(out), flush=True)
        return
    bv = F.vector.from_numpy(ctx, b)
    xv = F.vector(ctx, "f64", A.n)
    for w in range(args.warmup):
        xv.fill(0.0)
        r = S.solve(bv, xv)
        log(f"[bench] warmup {w}: outer={r.res_iter} solved={r.is_solved} {r.solve_s * 1e3:.1f} ms")
    times = []
    res = None
    for k in range(args.steps):
        xv.fill(0.0)
        ctx.sync()
        res = S.solve(bv, xv)   # synchronises before it returns
        times.append(res.solve_s)
        log(f"[bench] step {k}: outer={res.res_iter} solved={res.is_solved} {res.solve_s * 1e3:.2f} ms")
    # true residual of the last solve, fp64 on the device
    A64 = F.SELL(ctx, A, "f64", C=64, sigma=1, K=4)
    rv = F.vector(ctx, "f64", A.n)
    A64.compute_residual(bv, xv, rv)
    true_rel = F.nrm(rv) / F.nrm(bv)
    # every hot kernel of the solve, timed with HIP events on the solver's
    # stream (back-to-back launches in the solve's own mix of Arnoldi steps),
    # with its algorithmic bytes per launch
    kern = {}
    for name, w in KERNELS + (("spmv_f16val_f32vec", 1), ("spmv_f32", 2), ("spmv_f64", 3)):
        ms, by = S.time_kernel(w, reps=20)
        kern[name] = {"us": round(ms * 1e3, 2), "algorithmic_bytes": by, "GB/s": round(by / ms / 1e6, 1),
                      "frac": round(by / ms / 1e6 / PEAK_HBM_GBS, 4)}
        log(f"[bench] {name}: {kern[name]}")
    # the metric's fp16 SpMV with cold caches (each launch after a 512 MiB pass)
    ms16c, by16 = S.time_kernel(0, reps=10, cold=True)
    kern["spmv_f16"]["us_cold"] = round(ms16c * 1e3, 2)
    kern["spmv_f16"]["frac_cold"] = round(by16 / ms16c / 1e6 / PEAK_HBM_GBS, 4)
    # the same fp16 SpMV in the int32-index layout (a column index per slot),
    # the layout the 70%-of-peak target is stated for
    xi = F.vector.from_numpy(ctx, np.ones(A.n), "f16")
    yi = F.vector(ctx, "f16", A.n)
    S32 = F.SELL(ctx, A, "f16", C=64, sigma=1, K=4, offsets=False)
    ms_i = S32.time_spmv(xi, yi, reps=50)
    by_i = 6.0 * A.nnz + 4.0 * A.n
    kern["spmv_f16_int32_index"] = {"us": round(ms_i * 1e3, 2), "algorithmic_bytes": by_i,
                                    "GB/s": round(by_i / ms_i / 1e6, 1),
                                    "frac": round(by_i / ms_i / 1e6 / PEAK_HBM_GBS, 4)}
    del S32
    # roofline: the kernel that takes the most time in the solve, the level-3
    # fp16 SpMV with the fused CGS dots
    head = kern["l3_spmv_cgs_dots"]
    ach = head["algorithmic_bytes"] / (head["us"] * 1e-6) / 1e9
    mean_s = float(np.mean(times))
    out = {
        "metric": METRIC,
        "value": round(mean_s, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f16/f32/f64",
        "data": "synthetic",
        "config": {"workload": f"plain F3R, 27-point nonsymmetric {m}^3 (beta=0.5 along x), SDAINV tol 0.1, "
                               f"m=(100;8,4,2), c=64, restart 20, eps 1e-8",
                   "n": A.n, "nnz": A.nnz, "outer_iters": res.res_iter, "solved": bool(res.is_solved),
                   "true_rel_res": float(true_rel), "precond_calls": res.precond_calls,
                   "setup_s": round(setup_s, 2), "layout_build": S.layouts_where(),
                   "sdainv_factors": P.where, "sdainv_s": round(factor_s, 3),
                   "graph": not args.no_graph},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None, "traffic_detail": None,
                     "kernel": "k_sell<half,half,float,8,EpiDots<float,4>>: level-3 fp16 SpMV w = A16 z_j (fp16 "
                               "values, fp16 z, fp32 sums) with the CGS dots <w, v_k>, k <= j, fused; the solve's "
                               "mix j = 0..3, back to back; algorithmic bytes per launch = 2 nnz (values, unpadded; "
                               "no column indices in the chunk-offset layout) + 4 B per nonzero of the int32-index "
                               "chunks + n (2 z + 4 w + 4 (j+1) v), averaged over j",
                     "us_per_launch": head["us"], "algorithmic_bytes": head["algorithmic_bytes"],
                     "metric_spmv_f16": kern["spmv_f16"]},
        "kernels": kern,
    }
    tr = measured_traffic("l3_spmv_cgs_dots")
    if tr:
        out["roofline"]["traffic"] = tr["bytes"]
        out["roofline"]["traffic_detail"] = tr
    if args.learned > 0:
        out["learned"] = learned_vs_plain(F, ctx, args.learned)
    if args.batch:
        from f3r_amd import batch
        del S
        bo = batch.run_lockstep(64, 4, seed=0, log=log, tune=True)
        out["batch"] = {"workload": "64 systems (16^3..64^3, 7- and 27-point, diagonal margin U[0.5,2]) x 4 "
                                    "right-hand sides, lockstep F3R (one set of launches for all 256 solves)",
                        **{k: bo[k] for k in ("solves", "padded_rows", "rows_total", "wall_s", "walls_s",
                                              "all_solved", "outer_steps_max", "outer_steps_mean",
                                              "worst_true_rel_res", "setup_s", "layout")}}
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, full solves) ...")
        out["cpu_baseline"] = cpu_baseline(F, A, b)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
