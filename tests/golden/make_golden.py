"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs the binaries that oracle/ref_harness/Makefile builds from the reference
checkout (oracle/_ref/: ref_driver over the reference's headers, the
reference's CPU driver work/src/f3r.cpp in three precisions, its study
drivers work/src/depth.cpp and work/src/static.cpp, and its matrix generator
matrix/hpcg.cpp) on small inputs, single-threaded so that every
reduction runs in sequential order, and stores inputs and outputs as .npz
data.  Run in the container that has the reference checkout:

    make -C oracle/ref_harness && python tests/golden/make_golden.py [units f3r steps mm cpu ablation]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import problems as PB  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.path.join(ROOT, "oracle", "_ref")
ENV = dict(os.environ, OMP_NUM_THREADS="1")


def run(exe, *args, cwd=None):
    return subprocess.run([os.path.join(REF, exe)] + [str(a) for a in args], check=True,
                          capture_output=True, text=True, env=ENV, cwd=cwd).stdout


def matrices():
    yield "s27b", PB.stencil27(6, beta=0.5)
    yield "s7", PB.stencil7(6)
    yield "rnd", PB.random_sparse(70, 9, 3)


def unit_fixtures(T):
    """Scaling, SELL-32 layout, SDAINV, ILU(0) and SpMV of the reference."""
    for name, (n, rp, ix, v) in matrices():
        out = {"n": n, "rptr": rp, "idx": ix, "val": v}
        PB.write_csr_bin(f"{T}/A.bin", n, rp, ix, v)
        run("ref_driver", "scale", f"{T}/A.bin", f"{T}/As.bin")
        _, _, _, vs = PB.read_csr_bin(f"{T}/As.bin")
        out["val_scaled"] = vs
        run("ref_driver", "sell32", f"{T}/As.bin", f"{T}/S")
        out["sell32_sptr"] = PB.read_vec_bin(f"{T}/S", np.int32)
        out["sell32_idx"] = PB.read_vec_bin(f"{T}/S.idx", np.int32)
        out["sell32_val"] = PB.read_vec_bin(f"{T}/S.val", np.float64)
        for k, (tol, acc) in enumerate([(0.1, 1.0), (0.0, 1.0), (0.05, 1.3)]):
            run("ref_driver", "sdainv", f"{T}/As.bin", tol, acc, f"{T}/Z.bin", f"{T}/W.bin")
            for tag, fn in (("Z", "Z.bin"), ("W", "W.bin")):
                _, r, i, vv = PB.read_csr_bin(f"{T}/{fn}")
                out[f"sdainv{k}_{tag}_rptr"] = r
                out[f"sdainv{k}_{tag}_idx"] = i
                out[f"sdainv{k}_{tag}_val"] = vv
            out[f"sdainv{k}_params"] = np.array([tol, acc])
        for k, (blocks, acc) in enumerate([(1, 1.0), (4, 1.1)]):
            run("ref_driver", "ilu0", f"{T}/As.bin", blocks, acc, f"{T}/L.bin", f"{T}/U.bin")
            for tag, fn in (("L", "L.bin"), ("U", "U.bin")):
                _, r, i, vv = PB.read_csr_bin(f"{T}/{fn}")
                out[f"ilu{k}_{tag}_rptr"] = r
                out[f"ilu{k}_{tag}_idx"] = i
                out[f"ilu{k}_{tag}_val"] = vv
            out[f"ilu{k}_params"] = np.array([blocks, acc])
        x = np.random.default_rng(11).uniform(-1, 1, n)
        out["x"] = x
        for mode, dt, odt in ((64, np.float64, np.float64), (32, np.float32, np.float32),
                              (1616, np.float16, np.float16), (1632, np.float32, np.float32)):
            PB.write_vec_bin(f"{T}/x.bin", x.astype(dt))
            run("ref_driver", "spmv", f"{T}/As.bin", f"{T}/x.bin", mode, f"{T}/y.bin")
            out[f"spmv{mode}"] = PB.read_vec_bin(f"{T}/y.bin", odt)
        np.savez_compressed(os.path.join(HERE, f"units_{name}.npz"), **out)
        print("wrote units", name)


def f3r_fixtures(T):
    """Nested solves of the GPU-path composition (SDAINV, SELL-32) on the host."""
    cases = [("s27b8", PB.stencil27(8, beta=0.5)), ("s7_10", PB.stencil7(10))]
    for name, (n, rp, ix, v) in cases:
        PB.write_csr_bin(f"{T}/A.bin", n, rp, ix, v)
        run("ref_driver", "scale", f"{T}/A.bin", f"{T}/As.bin")
        _, _, _, vs = PB.read_csr_bin(f"{T}/As.bin")
        b = np.random.default_rng(5).uniform(-1, 1, n)
        PB.write_vec_bin(f"{T}/b.bin", b)
        out = {"n": n, "rptr": rp, "idx": ix, "val_scaled": vs, "b": b}
        for prec in (64, 32, 16):
            line = run("ref_driver", "f3r", f"{T}/As.bin", f"{T}/b.bin", prec, 8, 4, 2, 64, 20, 1.0,
                       f"{T}/x.bin", f"{T}/h.bin").strip()
            s, it, impl, expl = line.split(",")
            out[f"p{prec}_x"] = PB.read_vec_bin(f"{T}/x.bin", np.float64)
            out[f"p{prec}_hist"] = PB.read_vec_bin(f"{T}/h.bin", np.float64)
            out[f"p{prec}_res"] = np.array([float(s), float(it), float(impl), float(expl)])
        np.savez_compressed(os.path.join(HERE, f"f3r_{name}.npz"), **out)
        print("wrote f3r", name)


def step_cases():
    """Problems on which the reference's nest needs many outer steps, so that
    a +-10% window on its outer-step count is at least one step: the inner
    levels are shortened (m2, m3, m4 are arguments of the reference's drivers,
    work/suite-gpu.py:49), and the 64^3 case is the CPU-path configuration
    (symmetric 27-point, b = A*1, restart 3).  Each entry: name, problem
    description (regenerated by the tests), (m2, m3, m4), restart, precisions."""
    yield ("s27b24", {"kind": "stencil27", "nx": 24, "beta": 0.5, "b": "uniform", "seed": 5}, (2, 2, 1), 20,
           (16, 32, 64))
    yield ("s27b24", {"kind": "stencil27", "nx": 24, "beta": 0.5, "b": "uniform", "seed": 5}, (2, 2, 2), 20, (16,))
    yield ("ln24", {"kind": "convdiff7", "nx": 24, "sigma": 1.5, "corr": 8.0, "peclet": 0.0, "gseed": 3,
                    "b": "uniform", "seed": 5}, (2, 2, 1), 20, (16,))
    yield ("ln24", {"kind": "convdiff7", "nx": 24, "sigma": 1.5, "corr": 8.0, "peclet": 0.0, "gseed": 3,
                    "b": "uniform", "seed": 5}, (3, 2, 1), 20, (16,))
    yield ("cd24", {"kind": "convdiff7", "nx": 24, "sigma": 2.0, "corr": 8.0, "peclet": 10.0, "gseed": 3,
                    "b": "uniform", "seed": 5}, (2, 2, 1), 20, (16,))
    yield ("c1s27_64", {"kind": "stencil27", "nx": 64, "beta": 0.0, "b": "rowsum"}, (8, 4, 2), 3, (16, 64))


def f3r_step_fixtures(T):
    """Outer-step counts and residual histories of the reference's nest on
    problems that need 3 to 14 outer steps (one .npz, keys per case)."""
    import json
    out = {}
    names = []
    for name, d, ms, restart, precs in step_cases():
        n, rp, ix, v, b = PB.from_desc(d)
        PB.write_csr_bin(f"{T}/A.bin", n, rp, ix, v)
        run("ref_driver", "scale", f"{T}/A.bin", f"{T}/As.bin")
        _, _, _, vs = PB.read_csr_bin(f"{T}/As.bin")
        PB.write_vec_bin(f"{T}/b.bin", b)
        key = f"{name}_m{ms[0]}{ms[1]}{ms[2]}"
        names.append(key)
        out[f"{key}_desc"] = np.array(json.dumps(d))
        out[f"{key}_m"] = np.array(ms + (64, restart), np.int32)
        out[f"{key}_sha_unscaled"] = PB.csr_sha256(rp, ix, v)
        out[f"{key}_sha_scaled"] = PB.csr_sha256(rp, ix, vs)
        out[f"{key}_sha_b"] = np.frombuffer(__import__("hashlib").sha256(b.tobytes()).digest(), np.uint8)
        for prec in precs:
            line = run("ref_driver", "f3r", f"{T}/As.bin", f"{T}/b.bin", prec, *ms, 64, restart, 1.0,
                       f"{T}/x.bin", f"{T}/h.bin").strip()
            s, it, impl, expl = line.split(",")
            out[f"{key}_p{prec}_res"] = np.array([float(s), float(it), float(impl), float(expl)])
            out[f"{key}_p{prec}_hist"] = PB.read_vec_bin(f"{T}/h.bin", np.float64)
            print("step case", key, prec, line, flush=True)
    out["cases"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "f3r_ref_steps.npz"), **out)
    print("wrote f3r_ref_steps")


MM_FILES = ("general", "symmetric", "zeros", "skew", "integer", "rect")


def mm_fixtures(T):
    """The reference's MatrixMarket reader (io::mm::read_system + io::CSR) on
    the small files of tests/golden/mm/: CSR arrays and right-hand sides."""
    out = {}
    for name in MM_FILES:
        path = os.path.join(HERE, "mm", name + ".mtx")
        run("ref_driver", "mm", path, f"{T}/A.bin", f"{T}/b.bin")
        with open(f"{T}/A.bin", "rb") as f:
            n, nc = np.frombuffer(f.read(8), np.int32)
            nnz = int(np.frombuffer(f.read(8), np.int64)[0])
            out[f"{name}_shape"] = np.array([n, nc], np.int32)
            out[f"{name}_rptr"] = np.frombuffer(f.read(4 * (n + 1)), np.int32).copy()
            out[f"{name}_idx"] = np.frombuffer(f.read(4 * nnz), np.int32).copy()
            out[f"{name}_val"] = np.frombuffer(f.read(8 * nnz), np.float64).copy()
        out[f"{name}_b"] = PB.read_vec_bin(f"{T}/b.bin", np.float64)
    np.savez_compressed(os.path.join(HERE, "mm_ref.npz"), **out)
    print("wrote mm_ref")


def cpu_driver_fixture(T):
    """The reference's CPU driver work/src/f3r.cpp (BJILU0, 112 blocks) on the
    reference generator's 32^3 HPCG matrix with a seeded right-hand side."""
    mdir = os.path.join(T, "matrix")
    wdir = os.path.join(T, "work")
    os.makedirs(mdir, exist_ok=True)
    os.makedirs(wdir, exist_ok=True)
    run("hpcg", 5, 5, 5, cwd=mdir)
    n = 32 ** 3
    b = np.random.default_rng(7).uniform(-1, 1, n)
    with open(os.path.join(mdir, "hpcg_5_5_5_b.mtx"), "w") as f:
        f.write("%%%%MatrixMarket matrix array real general\n%d 1\n" % n)
        f.write("\n".join("%.17e" % t for t in b) + "\n")
    # digest of the generator's matrix as CSR, compared with tests/problems.py
    import hashlib
    import scipy.io
    M = scipy.io.mmread(os.path.join(mdir, "hpcg_5_5_5.mtx")).tocsr()
    M.sort_indices()
    h = hashlib.sha256()
    for a in (M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)):
        h.update(np.ascontiguousarray(a).tobytes())
    out = {"b": b, "hpcg_sha256": np.frombuffer(h.digest(), np.uint8)}
    for prec in (64, 32, 16):
        line = run(f"f3r_{prec}", "hpcg_5_5_5.mtx", 1.0, 1, 8, 4, 2, 64, 3, cwd=wdir).strip()
        f = line.split(",")
        # Problem,Method,Prec,M2,M3,M4,W,Precond,ACC,Time,Iter,ImplRes,ExplRes
        out[f"p{prec}_csv"] = np.array([float(f[10]), float(f[11]), float(f[12])])
    np.savez_compressed(os.path.join(HERE, "cpu_driver_hpcg32.npz"), **out)
    print("wrote cpu driver fixture")


def ablation_fixture(T):
    """The reference's study drivers on the 32^3 HPCG matrix of the CPU-driver
    fixture (same seeded right-hand side, BJILU0 on 112 blocks, restart 3):
    the nesting-depth nests of work/src/depth.cpp (F4, F3, F3H, F2, F2H) and
    the static-weight Richardson of work/src/static.cpp (weights 0.8, 1.0,
    1.2; 1.2 does not converge within 3 x 100 outer steps)."""
    mdir = os.path.join(T, "matrix")
    wdir = os.path.join(T, "work")
    os.makedirs(mdir, exist_ok=True)
    os.makedirs(wdir, exist_ok=True)
    run("hpcg", 5, 5, 5, cwd=mdir)
    n = 32 ** 3
    b = np.random.default_rng(7).uniform(-1, 1, n)
    with open(os.path.join(mdir, "hpcg_5_5_5_b.mtx"), "w") as f:
        f.write("%%%%MatrixMarket matrix array real general\n%d 1\n" % n)
        f.write("\n".join("%.17e" % t for t in b) + "\n")
    out = {"b_seed": np.array(7)}
    for nest in ("f4", "f3", "f3h", "f2", "f2h"):
        line = run(nest, "hpcg_5_5_5.mtx", 1.0, 1, 8, 4, 2, 64, 3, cwd=wdir).strip()
        f = line.split(",")
        # Problem,Method,Prec,M2,M3,M4,W,Precond,ACC,Time,Iter,ImplRes,ExplRes (Iter = outer steps x 64)
        out[f"{nest}_csv"] = np.array([float(f[10]), float(f[11]), float(f[12])])
        print("depth", nest, line, flush=True)
    for w in (0.8, 1.0, 1.2):
        line = run("static", "hpcg_5_5_5.mtx", 1.0, 1, 8, 4, 2, 0, 3, w, cwd=wdir).strip()
        f = line.split(",")
        # Iter = outer steps x m2 m3 m4
        out[f"static{w:.1f}_csv"] = np.array([float(f[10]), float(f[11]), float(f[12])])
        print("static", w, line, flush=True)
    np.savez_compressed(os.path.join(HERE, "ablation_hpcg32.npz"), **out)
    print("wrote ablation fixture")


if __name__ == "__main__":
    which = sys.argv[1:] or ["units", "f3r", "steps", "mm", "cpu", "ablation"]
    with tempfile.TemporaryDirectory() as T:
        if "units" in which:
            unit_fixtures(T)
        if "f3r" in which:
            f3r_fixtures(T)
        if "steps" in which:
            f3r_step_fixtures(T)
        if "mm" in which:
            mm_fixtures(T)
        if "cpu" in which:
            cpu_driver_fixture(T)
        if "ablation" in which:
            ablation_fixture(T)
