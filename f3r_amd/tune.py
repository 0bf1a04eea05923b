"""Autotuning of the SpMV format for one matrix: SELL-C-sigma (chunk height
C, sorting window sigma, K-slot interleave, threads per row, chunk-offset
index format, LDS-staged x gathers) and CSR (lanes per row).

Every candidate is built on the device and its SpMV timed with HIP events on
the context's stream (time_spmv); the fastest wins.  The reference fixes C =
32 and sigma = 1 (SELL<32, T, device>, work/src/f3r-gpu.cpp:67, :214); the
layout is a free parameter of this build, passed to F3R(C=, sigma=, K=,
index=) and Batch(C=, K=, index=).  Timing with sigma > 1 includes the row
permutation of x and y, which the solver avoids by renumbering the system,
so the ranking leans slightly towards sigma = 1.  Every SELL layout with one
thread per row and the one-lane-per-row CSR sum each row in CSR order, so
they give the same bits; K threads per row (SELL) and several lanes per row
(CSR) reassociate the sums.  The solver runs either SELL form
(F3R(..., lanes=True) for K threads per row).
"""
import itertools

import numpy as np

from . import SELL, CSRDevice, vector

DEFAULT_C = (32, 64, 128, 256)
DEFAULT_SIGMA = (1, 256)
DEFAULT_K = (1, 2, 4, 8)
DEFAULT_OFFSETS = (False, True)
DEFAULT_LDS = (False, True)
DEFAULT_TPR = (1, 2, 4, 8, 16, 32)
DEFAULT_SELL_LANES = (False, True)


def candidates(Cs=DEFAULT_C, sigmas=DEFAULT_SIGMA, Ks=DEFAULT_K, offsets=DEFAULT_OFFSETS, lds=DEFAULT_LDS,
               tprs=DEFAULT_TPR, sell_lanes=DEFAULT_SELL_LANES):
    """every valid candidate as a dict: SELL (sigma = 1 or sigma >= C; the
    chunk-offset format needs sigma = 1 and C >= 64; LDS-staged gathers the
    int32-index format; `lanes` threads per row: 1, or K for K >= 2 in the
    int32-index format without staging) and CSR (tpr)"""
    out = []
    for C, sigma, K, off, st in itertools.product(Cs, sigmas, Ks, offsets, lds):
        if sigma > 1 and sigma < C:
            continue
        if off and (sigma != 1 or C < 64):
            continue
        if off and st:
            continue
        out.append({"format": "sell", "C": C, "sigma": sigma, "K": K, "offsets": off, "lds": st, "lanes": 1})
        if True in sell_lanes and K >= 2 and not off and not st:
            out.append({"format": "sell", "C": C, "sigma": sigma, "K": K, "offsets": off, "lds": st, "lanes": K})
    for t in tprs:
        out.append({"format": "csr", "tpr": t})
    return out


def build(ctx, A, dtype, cand):
    if cand["format"] == "csr":
        return CSRDevice(ctx, A, dtype, tpr=cand["tpr"])
    return SELL(ctx, A, dtype, cand["C"], cand["sigma"], cand["K"], offsets=cand["offsets"], lds=cand["lds"],
                lanes=cand.get("lanes", 1))


def autotune(ctx, A, dtype="f16", reps=20, check=False, **space):
    """returns (best, table): best = the fastest candidate's dict with "ms";
    table lists (ms, candidate) fastest first.  `space` narrows the search
    (Cs, sigmas, Ks, offsets, lds, tprs; tprs=() leaves CSR out).  check: every
    candidate's y is also kept (table entries get "y")"""
    x = vector.from_numpy(ctx, np.random.default_rng(0).uniform(-1, 1, A.n), dtype)
    y = vector(ctx, dtype, A.n)
    table = []
    for cand in candidates(**space):
        M = build(ctx, A, dtype, cand)
        ms = M.time_spmv(x, y, reps)
        entry = dict(cand, ms=ms)
        if check:
            entry["y"] = y.download()
        del M
        table.append(entry)
    table.sort(key=lambda e: e["ms"])
    best = {k: v for k, v in table[0].items() if k != "y"}
    return best, table


def solver_layout(best):
    """F3R keywords of a SELL winner (CSR winners map to the SELL layout of
    one column index per slot that sums alike: C 64, K 1; the solver's fused
    kernels gather from global memory, so LDS staging is dropped; a K
    threads per row winner keeps them: lanes=True)"""
    if best["format"] == "csr":
        return {"C": 64, "sigma": 1, "K": 1, "index": 0}
    kw = {"C": best["C"], "sigma": best["sigma"], "K": best["K"], "index": 1 if best["offsets"] else 0}
    if best.get("lanes", 1) > 1:
        kw["lanes"] = True
    return kw


def random_rows(n, maxlen, seed=0):
    """CSR of n rows with uniform lengths in [1, maxlen]: the diagonal and
    distinct random columns in ascending order, off-diagonals in [-1, 1],
    diagonal 1 + sum |off| (seeded; a matrix with long, uneven rows)"""
    from . import CSR
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, maxlen + 1, n)
    rptr = np.zeros(n + 1, np.int64)
    cols, vals = [], []
    for i in range(n):
        c = np.unique(np.concatenate(([i], rng.integers(0, n, lens[i] - 1))))
        v = rng.uniform(-1.0, 1.0, c.size)
        v[c == i] = 1.0 + np.abs(v[c != i]).sum()
        cols.append(c)
        vals.append(v)
        rptr[i + 1] = rptr[i] + c.size
    return CSR(n, rptr.astype(np.int32), np.concatenate(cols).astype(np.int32), np.concatenate(vals))


def _problem(spec):
    """stencil27:<nx>[:<beta>] | stencil7:<nx>,<ny>,<nz> | random:<n>:<maxlen>[:<seed>] | mtx:<path>"""
    from . import CSR, algorithm, io
    kind, _, arg = spec.partition(":")
    if kind == "stencil27":
        nx, _, beta = arg.partition(":")
        A = CSR.stencil27(int(nx), beta=float(beta or 0.0))
    elif kind == "stencil7":
        nx, ny, nz = (int(t) for t in arg.split(","))
        A = CSR.dominant(7, nx, ny, nz, margin=1.0)
    elif kind == "random":
        f = arg.split(":")
        A = random_rows(int(f[0]), int(f[1]), int(f[2]) if len(f) > 2 else 0)
    elif kind == "mtx":
        A = io.mm.read_matrix(arg)
    else:
        raise ValueError(f"unknown problem {spec}")
    algorithm.scaling(A)
    return A


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="autotune the SpMV format of one matrix and print the table")
    ap.add_argument("--problem", default="stencil27:128:0.5",
                    help="stencil27:<nx>[:<beta>] | stencil7:<nx>,<ny>,<nz> (2-D: nz = 1) | "
                         "random:<n>:<maxlen>[:<seed>] | mtx:<path>")
    ap.add_argument("--dtype", default="f16")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--no-csr", action="store_true")
    a = ap.parse_args(argv)
    from . import Context
    ctx = Context(0)
    A = _problem(a.problem)
    best, table = autotune(ctx, A, a.dtype, reps=a.reps, **({"tprs": ()} if a.no_csr else {}))
    def us(e):  # the table's times in microseconds
        return {("us" if k == "ms" else k): (round(v * 1e3, 2) if k == "ms" else v) for k, v in e.items()}
    for e in table:
        print(json.dumps(us(e)), flush=True)
    print(json.dumps({"problem": a.problem, "n": A.n, "nnz": A.nnz, "dtype": a.dtype, "best": us(best)}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
