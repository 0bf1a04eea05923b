"""Autotuning of the SELL-C-sigma layout (chunk height C, sorting window
sigma, K-slot interleave) for one matrix.

Every candidate layout is built on the device and its SpMV timed with HIP
events on the context's stream (SELL.time_spmv); the fastest wins.  The
reference fixes C = 32 and sigma = 1 (SELL<32, T, device>,
work/src/f3r-gpu.cpp:67, :214); the layout is a free parameter of this build,
passed to F3R(C=, sigma=, K=).  Timing with sigma > 1 includes the row
permutation of x and y, which the solver avoids by renumbering the system,
so the ranking leans slightly towards sigma = 1.
"""
import itertools

import numpy as np

from . import SELL, vector

DEFAULT_C = (32, 64, 128, 256)
DEFAULT_SIGMA = (1, 256)
DEFAULT_K = (1, 2, 4)


def autotune(ctx, A, dtype="f16", Cs=DEFAULT_C, sigmas=DEFAULT_SIGMA, Ks=DEFAULT_K, reps=20):
    """returns (best, table): best = {"C", "sigma", "K", "ms"}; table lists
    every candidate as (ms, C, sigma, K), fastest first"""
    x = vector.from_numpy(ctx, np.ones(A.n), dtype)
    y = vector(ctx, dtype, A.n)
    table = []
    for C, sigma, K in itertools.product(Cs, sigmas, Ks):
        if sigma > 1 and sigma < C:
            continue
        S = SELL(ctx, A, dtype, C, sigma, K)
        ms = S.time_spmv(x, y, reps)
        del S
        table.append((ms, C, sigma, K))
    table.sort()
    ms, C, sigma, K = table[0]
    return {"C": C, "sigma": sigma, "K": K, "ms": ms}, table
