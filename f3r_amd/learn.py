"""Learned correction of SDAINV for the innermost fp16 level.

A per-row MLP 12 -> 64 -> 64 -> 1 (ReLU, bf16 weights, fp32 accumulation on
MFMA) maps local features of the scaled matrix and of its SDAINV factors to
y_i; the corrected preconditioner is M = Z diag(s) W^T with
s_i = exp(clamp(y_i, -3, 3)) -- SDAINV (sparse/preconditioner/ainv.hpp:57-64)
with the columns of Z rescaled, on the same sparsity pattern, applied by the
unchanged solver.  The reference has no learned component; this module has no
reference counterpart.

Training minimises the mean over probe vectors z of
min_w ||(I - w A M) z||^2 / ||z||^2 (the residual after the best weight, which
is what the adaptive Richardson weights leave; LearnSystem.loss kind 1) or of
||(I - A M) z||^2 / ||z||^2 (kind 0); the probes are the vectors the solver
feeds to its innermost level (F3R.capture), or random vectors.  Forward,
backward, the loss and its gradient and Adam all run on the GPU.

Family of systems (7-point lognormal diffusion):
    A = convdiff7(nx, nx, nx, sigma, corr, peclet=0, seed),
    sigma ~ U[0.5, 2], corr ~ U[2, 16] drawn from the system seed.
"""
import ctypes
import json
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import CSR, F3R, SDAINV, Context, _Precond, _p, algorithm, uniform, vector
from ._capi import check, lib

NIN = 16
NFEAT = 12
HIDDEN = 64
_n = ctypes.c_int64()
check(lib.f3r_mlp_nparams(ctypes.byref(_n)), "mlp_nparams")
NPARAMS = _n.value

_SHAPES = [("w1", (HIDDEN, NIN)), ("b1", (HIDDEN,)), ("w2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
           ("w3", (HIDDEN,)), ("b3", (1,))]


def unpack(params):
    """flat fp32 parameter vector -> dict of arrays (views)"""
    out, off = {}, 0
    for name, shape in _SHAPES:
        size = int(np.prod(shape))
        out[name] = params[off:off + size].reshape(shape)
        off += size
    assert off == NPARAMS
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float32).reshape(-1) for name, _ in _SHAPES])


class MLP:
    """The per-row network on one context; fp32 master weights with Adam."""

    def __init__(self, ctx, seed=0, params=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib.f3r_mlp_create(ctx.handle, seed, ctypes.byref(h)), "mlp_create")
        self.handle = h
        if params is not None:
            self.set(params)

    def get(self):
        out = np.empty(NPARAMS, np.float32)
        check(lib.f3r_mlp_get(self.handle, _p(out)), "mlp_get")
        return out

    def set(self, params):
        a = np.ascontiguousarray(params, np.float32)
        assert a.size == NPARAMS
        check(lib.f3r_mlp_set(self.handle, _p(a)), "mlp_set")

    def forward(self, feat):
        f = np.ascontiguousarray(feat, np.float32)
        assert f.ndim == 2 and f.shape[1] == NIN
        y = np.empty(f.shape[0], np.float32)
        check(lib.f3r_mlp_forward(self.handle, f.shape[0], _p(f), _p(y)), "mlp_forward")
        return y

    def grad_zero(self):
        check(lib.f3r_mlp_grad_zero(self.handle), "mlp_grad_zero")

    def backward(self, feat, gy):
        f = np.ascontiguousarray(feat, np.float32)
        g = np.ascontiguousarray(gy, np.float32)
        assert f.ndim == 2 and f.shape[1] == NIN and g.shape == (f.shape[0],)
        check(lib.f3r_mlp_backward(self.handle, f.shape[0], _p(f), _p(g)), "mlp_backward")

    def grad(self):
        out = np.empty(NPARAMS, np.float32)
        check(lib.f3r_mlp_grad_get(self.handle, _p(out)), "mlp_grad_get")
        return out

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib.f3r_mlp_adam(self.handle, lr, beta1, beta2, eps), "mlp_adam")

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def features(A, P, diag=None):
    """(n, 16) float32 row features of the scaled matrix A and its SDAINV P;
    `diag` is the diagonal before scaling (feature 0 = log|a_ii|)"""
    out = np.zeros((A.n, NIN), np.float32)
    d = None if diag is None else np.ascontiguousarray(diag, np.float64)
    check(lib.f3r_learn_features(A.handle(None), P.handle, None if d is None else _p(d), _p(out)),
          "learn_features")
    return out


class Rescaled(_Precond):
    """SDAINV with Z's column i scaled by s_i (the corrected preconditioner)"""

    def __init__(self, P, s, ctx=None):
        self.n = P.n
        self.kind = 0
        self.dtype = P.dtype
        a = np.ascontiguousarray(s, np.float32)
        assert a.shape == (P.n,)
        h = ctypes.c_void_p()
        check(lib.f3r_sdainv_rescale(ctx.handle if ctx else None, P.handle, _p(a), ctypes.byref(h)),
              "sdainv_rescale")
        self.handle = h


class LearnSystem:
    """One training system: A, A^T, Z, Z^T, W^T (fp32 SELL) and features on
    the device, plus probe vectors."""

    def __init__(self, ctx, A, P, feat):
        self.ctx, self.n = ctx, A.n
        f = np.ascontiguousarray(feat, np.float32)
        assert f.shape == (A.n, NIN)
        h = ctypes.c_void_p()
        check(lib.f3r_lsys_create(ctx.handle, A.handle(ctx), P.handle, _p(f), ctypes.byref(h)), "lsys_create")
        self.handle = h
        self.nprobes = 0

    def set_probes(self, Z):
        z = np.ascontiguousarray(Z, np.float32)
        assert z.ndim == 2 and z.shape[1] == self.n
        check(lib.f3r_lsys_set_probes(self.handle, z.shape[0], _p(z)), "lsys_set_probes")
        self.nprobes = z.shape[0]

    def loss(self, mlp, grad=False, kind=1):
        """kind 1: mean of 1 - <z,q>^2/(|z|^2 |q|^2), q = A M z (the residual
        after the best weight); kind 0: mean of |z - q|^2 / |z|^2"""
        out = ctypes.c_double()
        check(lib.f3r_lsys_loss(self.handle, mlp.handle, kind, 1 if grad else 0, ctypes.byref(out)), "lsys_loss")
        return out.value

    def scales(self, mlp, normalize=True):
        """s = exp(clamp(MLP(features))), divided by its geometric mean unless
        normalize=False (the scale-blind loss leaves the global scale free)"""
        s = np.empty(self.n, np.float32)
        check(lib.f3r_lsys_scales(self.handle, mlp.handle, 1 if normalize else 0, _p(s)), "lsys_scales")
        return s

    def __del__(self):
        try:
            if self.handle:
                lib.f3r_lsys_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ------------------------------------------------------------ the family ---

def family_params(seed):
    """(sigma, corr) of system `seed`: sigma ~ U[0.5, 2], corr ~ U[2, 16]"""
    u = uniform(2, 0.0, 1.0, seed=0x5EED0000 + seed)
    return 0.5 + 1.5 * float(u[0]), 2.0 + 14.0 * float(u[1])


def host_system(nx, seed, tol=0.1, acc=1.0):
    """host part of system `seed`: scaled CSR, unscaled diagonal, host SDAINV,
    features.  Pure host work (thread-safe; the C calls release the GIL)."""
    sigma, corr = family_params(seed)
    A = CSR.convdiff7(nx, nx, nx, sigma=sigma, corr=corr, peclet=0.0, seed=seed)
    diag = A.diagonal()
    algorithm.scaling(A)
    P = SDAINV(A, tol, acc, "f16", ctx=None)
    feat = features(A, P, diag)
    return {"seed": seed, "sigma": sigma, "corr": corr, "A": A, "P": P, "feat": feat}


def host_systems(nx, seeds, threads=8):
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda s: host_system(nx, s), seeds))


def rhs(n, seed):
    """b ~ U[-1, 1] of system `seed`"""
    return uniform(n, -1.0, 1.0, seed=0xB0B00000 + seed)


def solver_probes(ctx, A, P, k=8, every=4, seed=0):
    """k vectors the solver feeds its innermost level during one level-2
    application to the unit right-hand side (the plain SDAINV
    preconditioner); level 2 expects a unit-norm input, as the outer Krylov
    vectors it is applied to are"""
    Pd = Rescaled(P, np.ones(A.n, np.float32), ctx)
    S = F3R(ctx, A, Pd, use_graph=False)
    S.capture(k, every)
    bh = rhs(A.n, seed)
    b = vector.from_numpy(ctx, bh / np.linalg.norm(bh))
    out = vector(ctx, "f64", A.n)
    S.level_apply(2, b, out)
    Z = S.captured()
    S.capture(0)
    return Z


def random_probes(n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(np.array([-1.0, 1.0], np.float32), size=(k, n))


def train(ctx, systems, epochs=20, lr=1e-2, seed=0, log=None, kind=0):
    """Adam over the systems (one step per system, shuffled each epoch).
    `systems` are LearnSystem objects with probes.  Returns the MLP and the
    mean loss of each epoch."""
    mlp = MLP(ctx, seed)
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(systems)):
            mlp.grad_zero()
            tot += systems[i].loss(mlp, grad=True, kind=kind)
            mlp.adam(lr)
        hist.append(tot / len(systems))
        if log:
            log(f"epoch {ep}: mean loss {hist[-1]:.6f}")
    return mlp, hist


def evaluate(ctx, mlp, hs, reps=1, normalize=True):
    """plain vs learned F3R on one host system `hs`: outer steps, solve
    times and true residuals"""
    A, P = hs["A"], hs["P"]
    n = A.n
    b = vector.from_numpy(ctx, rhs(n, hs["seed"]))
    out = {}
    L = LearnSystem(ctx, A, P, hs["feat"])
    s = L.scales(mlp, normalize)
    del L
    for name, prec in (("plain", Rescaled(P, np.ones(n, np.float32), ctx)), ("learned", Rescaled(P, s, ctx))):
        S = F3R(ctx, A, prec)
        x = vector(ctx, "f64", n)
        best = None
        for _ in range(reps):
            x.fill(0.0)
            res = S.solve(b, x)
            best = res.solve_s if best is None else min(best, res.solve_s)
        Sd = SELLcheck(ctx, A)
        out[name] = {"outer": res.res_iter, "solved": res.is_solved, "solve_s": best,
                     "true_rel_res": Sd.true_rel_res(b, x)}
    out["s_range"] = [float(s.min()), float(s.max())]
    return out


class SELLcheck:
    """fp64 SELL copy of A for the true residual ||b - A x|| / ||b||"""

    def __init__(self, ctx, A):
        from . import SELL, nrm
        self.ctx, self.nrm = ctx, nrm
        self.S = SELL(ctx, A, "f64")
        self.r = vector(ctx, "f64", A.n)

    def true_rel_res(self, b, x):
        self.S.compute_residual(b, x, self.r)
        return self.nrm(self.r) / self.nrm(b)


def main(argv=None):
    """python -m f3r_amd.learn --nx 96 --train 512 --test 64 --out params.npz"""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--nx", type=int, default=96)
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--probes", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--kind", type=int, default=0, help="loss: 0 = |(I - AM)z|^2/|z|^2, 1 = best-weight residual")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--out", default="learned_sdainv.npz")
    ap.add_argument("--report", default=None)
    ap.add_argument("--evaluate", default=None, metavar="PARAMS.npz",
                    help="no training: evaluate these parameters on the held-out systems")
    a = ap.parse_args(argv)
    ctx = Context(0)
    t0 = time.time()

    def log(m):
        print(f"[learn {time.time() - t0:7.1f}s] {m}", flush=True)

    chunk = max(a.threads * 4, 16)
    if a.evaluate:
        d = np.load(a.evaluate, allow_pickle=False)
        mlp = MLP(ctx, params=d["params"])
        a.kind = int(d["kind"]) if "kind" in d.files else 1
        a.train = int(d["train"]) if "train" in d.files else 0
        base, hist = float(d["base_loss"]) if "base_loss" in d.files else float("nan"), \
            [float(v) for v in d["loss"]] if "loss" in d.files else [float("nan")]
        log(f"evaluating {a.evaluate} (loss kind {a.kind}, trained on {a.train} systems)")
        return _held_out(ctx, mlp, a, base, hist, chunk, log)
    systems = []
    for c0 in range(0, a.train, chunk):
        for hs in host_systems(a.nx, range(c0, min(a.train, c0 + chunk)), a.threads):
            L = LearnSystem(ctx, hs["A"], hs["P"], hs["feat"])
            L.set_probes(solver_probes(ctx, hs["A"], hs["P"], k=a.probes, seed=hs["seed"]))
            systems.append(L)
        log(f"{len(systems)} training systems ready")
    m0 = MLP(ctx, 0)
    base = float(np.mean([L.loss(m0, kind=a.kind) for L in systems]))
    log(f"plain SDAINV loss {base:.6f}")
    mlp, hist = train(ctx, systems, a.epochs, a.lr, log=log, kind=a.kind)
    np.savez(a.out, params=mlp.get(), loss=np.array(hist), base_loss=base, nx=a.nx, train=a.train,
             probes=a.probes, epochs=a.epochs, kind=a.kind, lr=a.lr)
    log(f"parameters saved to {a.out}")
    del systems
    return _held_out(ctx, mlp, a, base, hist, chunk, log)


def _held_out(ctx, mlp, a, base, hist, chunk, log):
    """plain vs learned F3R on the held-out seeds 100000..; summary printed"""
    rep = []
    test_seeds = range(100000, 100000 + a.test)
    for c0 in range(0, a.test, chunk):
        for hs in host_systems(a.nx, list(test_seeds)[c0:c0 + chunk], a.threads):
            r = evaluate(ctx, mlp, hs, reps=3, normalize=a.kind == 1)
            r.update(seed=hs["seed"], sigma=hs["sigma"], corr=hs["corr"])
            rep.append(r)
            log(f"held-out {hs['seed']}: plain {r['plain']['outer']} steps {r['plain']['solve_s'] * 1e3:.1f} ms, "
                f"learned {r['learned']['outer']} steps {r['learned']['solve_s'] * 1e3:.1f} ms")
    ps = float(np.mean([r["plain"]["solve_s"] for r in rep]))
    ls = float(np.mean([r["learned"]["solve_s"] for r in rep]))
    po = float(np.mean([r["plain"]["outer"] for r in rep]))
    lo = float(np.mean([r["learned"]["outer"] for r in rep]))
    ok = all(r[k]["solved"] and r[k]["true_rel_res"] <= 1e-8 for r in rep for k in ("plain", "learned"))
    summary = {"train": a.train, "test": a.test, "nx": a.nx, "kind": a.kind, "base_loss": base,
               "final_loss": hist[-1],
               "plain_mean_s": ps, "learned_mean_s": ls, "speedup": ps / ls, "plain_mean_outer": po,
               "learned_mean_outer": lo, "all_solved": ok}
    log(f"held-out mean: plain {ps * 1e3:.2f} ms ({po:.2f} outer steps), learned {ls * 1e3:.2f} ms "
        f"({lo:.2f}); speedup {ps / ls:.4f}; all solved to 1e-8: {ok}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump({"summary": summary, "loss": hist, "held_out": rep}, f, indent=1)
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
