"""Batched solves: many independent systems, several right-hand sides each.

The family (seeded): edge lengths uniform integers in [16, 64]; even-indexed
systems are 7-point, odd-indexed 27-point stencils with unit off-diagonals and
diagonal (1 + delta) * sum|off|, delta ~ U[0.5, 2] (CSR.dominant); every
system is scaled as the reference driver scales (algorithm.hpp:45-59) and
gets `nrhs` right-hand sides uniform in [-1, 1].

Every system has its own F3R solver (SDAINV, m = (100; 8, 4, 2)) whose SELL
layout (C, sigma, K, index format, threads per row) is autotuned on that
system's fp16 matrix (tune.autotune).  The systems are spread over `streams` contexts, one HIP
stream each, each driven by its own host thread; the HIP library releases the
GIL, so the streams run concurrently and small systems overlap on the GPU.
The reference solves one system per run (work/src/f3r-gpu.cpp:44-288).
"""
import threading
import time

import numpy as np

from . import CSR, F3R, SDAINV, SELL, Context, algorithm, nrm, uniform, vector
from .tune import autotune, solver_layout


def family(seed=0, count=64):
    """[(index, npts, nx, delta)] of the batch"""
    u = uniform(2 * count, 0.0, 1.0, seed=0xBA7C0000 + seed)
    out = []
    for i in range(count):
        nx = 16 + min(48, int(u[2 * i] * 49))
        out.append((i, 7 if i % 2 == 0 else 27, nx, 0.5 + 1.5 * float(u[2 * i + 1])))
    return out


def make_system(npts, nx, delta):
    A = CSR.dominant(npts, nx, margin=delta)
    algorithm.scaling(A)
    return A


class _Worker:
    def __init__(self, device, jobs, nrhs, tune, seed, cands):
        self.ctx = Context(device)
        self.jobs, self.nrhs, self.tune, self.seed, self.cands = jobs, nrhs, tune, seed, cands
        self.results, self.error = [], None
        self.ready = []

    def setup(self):
        try:
            for (i, npts, nx, delta) in self.jobs:
                t0 = time.time()
                A = make_system(npts, nx, delta)
                if self.tune:
                    best, _ = autotune(self.ctx, A, "f16", tprs=(), lds=(False,), **self.cands)
                    L = solver_layout(best)
                    lay = (L["C"], L["sigma"], L["K"], L["index"], L.get("lanes", False))
                else:
                    lay = (64, 1, 4, 1, False)
                P = SDAINV(A, 0.1, 1.0, "f16", ctx=self.ctx)
                S = F3R(self.ctx, A, P, C=lay[0], sigma=lay[1], K=lay[2], index=lay[3], lanes=lay[4])
                bs = [vector.from_numpy(self.ctx, uniform(A.n, -1.0, 1.0, seed=(self.seed << 20) + 16 * i + r))
                      for r in range(self.nrhs)]
                xs = [vector(self.ctx, "f64", A.n) for _ in range(self.nrhs)]
                self.ready.append((i, npts, nx, delta, A, P, S, bs, xs, lay, time.time() - t0))
        except Exception as e:  # reported by the caller
            self.error = e

    def solve(self):
        try:
            for (i, npts, nx, delta, A, P, S, bs, xs, lay, setup_s) in self.ready:
                outs = []
                for b, x in zip(bs, xs):
                    x.fill(0.0)
                    r = S.solve(b, x)
                    outs.append((r.res_iter, r.is_solved, r.solve_s))
                self.results.append({"system": i, "npts": npts, "nx": nx, "n": A.n, "delta": round(delta, 4),
                                     "layout": {"C": lay[0], "sigma": lay[1], "K": lay[2], "index": lay[3],
                                                "lanes": lay[4]},
                                     "setup_s": round(setup_s, 3),
                                     "outer": [o[0] for o in outs], "solved": all(o[1] for o in outs),
                                     "solve_s": [o[2] for o in outs]})
        except Exception as e:
            self.error = e

    def check(self):
        """true relative residual of every solution (fp64 SELL on the device)"""
        worst = 0.0
        for (i, npts, nx, delta, A, P, S, bs, xs, lay, _) in self.ready:
            S64 = SELL(self.ctx, A, "f64")
            r = vector(self.ctx, "f64", A.n)
            for b, x in zip(bs, xs):
                S64.compute_residual(b, x, r)
                worst = max(worst, nrm(r) / nrm(b))
        return worst


def run(count=64, nrhs=4, streams=4, seed=0, tune=True, device=0,
        cands=None, log=None):
    """Solves the batch; returns a summary dict (wall time of the solve phase
    over all streams, per-system results, the worst true residual)."""
    cands = cands or {"Cs": (64, 128, 256), "sigmas": (1,), "Ks": (1, 2, 4), "offsets": (False, True), "reps": 10}
    fam = family(seed, count)
    # largest systems first, dealt round-robin, so the streams finish together
    order = sorted(fam, key=lambda t: -(t[2] ** 3) * (27 if t[1] == 27 else 7))
    workers = [_Worker(device, order[w::streams], nrhs, tune, seed, cands) for w in range(streams)]
    t0 = time.time()
    th = [threading.Thread(target=w.setup) for w in workers]
    [t.start() for t in th]
    [t.join() for t in th]
    for w in workers:
        if w.error:
            raise w.error
    setup_s = time.time() - t0
    if log:
        log(f"batch setup {setup_s:.1f} s ({count} systems, {streams} streams)")
    for w in workers:
        w.ctx.sync()
    t0 = time.time()
    th = [threading.Thread(target=w.solve) for w in workers]
    [t.start() for t in th]
    [t.join() for t in th]
    for w in workers:
        w.ctx.sync()
    wall = time.time() - t0
    for w in workers:
        if w.error:
            raise w.error
    res = sorted([r for w in workers for r in w.results], key=lambda r: r["system"])
    worst = max(w.check() for w in workers)
    rows = sum(r["n"] for r in res)
    out = {"systems": count, "rhs": nrhs, "streams": streams, "rows_total": rows, "solves": count * nrhs,
           "wall_s": wall, "setup_s": setup_s, "all_solved": all(r["solved"] for r in res),
           "worst_true_rel_res": worst, "sum_solve_s": float(sum(sum(r["solve_s"]) for r in res)),
           "results": res}
    if log:
        log(f"batch: {count * nrhs} solves in {wall * 1e3:.1f} ms wall over {streams} streams "
            f"(sum of per-solve times {out['sum_solve_s'] * 1e3:.1f} ms), all solved {out['all_solved']}, "
            f"worst true residual {worst:.2e}")
    return out


def tune_lockstep(ctx, fam, log=None):
    """Batch layout keywords (C, K, index, lanes) autotuned on the family's
    largest 27-point system, the lockstep SpMVs' dominant cost: chunk heights
    64 to 256, K, the index format and K threads per row (sigma 1: the
    lockstep layout keeps each segment's rows in order)"""
    i, npts, nx, delta = max((t for t in fam if t[1] == 27), key=lambda t: t[2])
    A = make_system(npts, nx, delta)
    best, table = autotune(ctx, A, "f16", Cs=(64, 128, 256), sigmas=(1,), Ks=(1, 2, 4, 8), offsets=(False, True),
                           lds=(False,), tprs=(), sell_lanes=(False, True), reps=10)
    L = solver_layout(best)
    lanes = best.get("lanes", 1) > 1
    if log:
        log(f"lockstep layout tuned on system {i} ({npts}-point {nx}^3): C={L['C']} K={L['K']} index={L['index']} "
            f"threads per row {best.get('lanes', 1)} ({best['ms'] * 1e3:.1f} us; slowest candidate "
            f"{table[-1]['ms'] * 1e3:.1f} us)")
    return {"C": L["C"], "K": L["K"], "index": L["index"], "lanes": lanes}


def run_lockstep(count=64, nrhs=4, seed=0, device=0, log=None, tune=False, use_graph=True, layout=None):
    """The same family solved in lockstep (f3r_batch): one block-diagonal
    layout of all systems x right-hand sides, every launch covering all of
    them; per-system SDAINV factorised on the device.  tune: the layout
    (C, K, index format, threads per row) autotuned first (tune_lockstep);
    layout: Batch keywords given instead.  use_graph: each outer step
    replayed from a HIP graph."""
    from . import Batch
    ctx = Context(device)
    fam = family(seed, count)
    lay = tune_lockstep(ctx, fam, log) if tune else dict(layout or {})
    t0 = time.time()
    As, Ps, bs, meta = [], [], [], []
    for (i, npts, nx, delta) in fam:
        A = make_system(npts, nx, delta)
        P = SDAINV(A, 0.1, 1.0, "f16", ctx=ctx, device=True)
        for r in range(nrhs):
            As.append(A)
            Ps.append(P)
            bs.append(uniform(A.n, -1.0, 1.0, seed=(seed << 20) + 16 * i + r))
            meta.append((i, npts, nx, delta, r))
    B = Batch(ctx, As, Ps, use_graph=use_graph, **lay)
    rows, segs = B.info()
    setup_s = time.time() - t0
    if log:
        log(f"lockstep set-up {setup_s:.1f} s: {segs} segments, {rows} padded rows")
    bv = [vector.from_numpy(ctx, b) for b in bs]
    xv = [vector(ctx, "f64", A.n) for A in As]
    walls = []
    for rep in range(3):
        for x in xv:
            x.fill(0.0)
        ctx.sync()
        res = B.solve(bv, xv)
        walls.append(res[0].solve_s)
        if log:
            log(f"lockstep solve {rep}: {res[0].solve_s * 1e3:.1f} ms for {segs} solves")
    worst = 0.0
    S64 = {}
    r = {}
    for s, A in enumerate(As):
        key = id(A)
        if key not in S64:
            S64[key] = SELL(ctx, A, "f64")
            r[key] = vector(ctx, "f64", A.n)
        S64[key].compute_residual(bv[s], xv[s], r[key])
        worst = max(worst, nrm(r[key]) / nrm(bv[s]))
    out = {"mode": "lockstep", "systems": count, "rhs": nrhs, "solves": segs, "padded_rows": rows,
           "rows_total": int(sum(A.n for A in As)), "wall_s": float(min(walls)), "walls_s": walls,
           "setup_s": setup_s, "all_solved": all(x.is_solved for x in res),
           "outer_steps_max": int(max(x.res_iter for x in res)), "outer_steps_mean": float(np.mean([x.res_iter for x in res])),
           "outer_steps_per_segment": [int(x.res_iter) for x in res],
           "layout": lay or {"C": 64, "K": 0, "index": 1, "lanes": False},
           "worst_true_rel_res": worst}
    if log:
        log(f"lockstep: {segs} solves in {out['wall_s'] * 1e3:.1f} ms wall (best of 3), all solved "
            f"{out['all_solved']}, worst true residual {worst:.2e}")
    return out


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="batched F3R solves of the seeded 64-system family")
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--rhs", type=int, default=4)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="no layout autotuning")
    ap.add_argument("--mode", choices=["lockstep", "streams"], default="lockstep",
                    help="lockstep: all solves in shared launches (f3r_batch); streams: one solver per system")
    ap.add_argument("--no-graph", action="store_true", help="lockstep: every kernel launched eagerly (no graphs)")
    ap.add_argument("--layout", default=None, metavar="C,K,INDEX[,lanes]",
                    help="lockstep: this layout instead of the tuned one, e.g. 64,4,0,lanes (K threads per row)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.time()

    def log(m):
        print(f"[batch {time.time() - t0:6.1f}s] {m}", flush=True)

    if a.mode == "lockstep":
        lay = None
        if a.layout:
            f = a.layout.split(",")
            lay = {"C": int(f[0]), "K": int(f[1]), "index": int(f[2]), "lanes": len(f) > 3 and f[3] == "lanes"}
        out = run_lockstep(a.count, a.rhs, a.seed, log=log, tune=not a.no_tune and lay is None,
                           use_graph=not a.no_graph, layout=lay)
    else:
        out = run(a.count, a.rhs, a.streams, a.seed, not a.no_tune, log=log)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "results"}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
