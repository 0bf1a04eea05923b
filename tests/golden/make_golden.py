"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs the binaries that oracle/ref_harness/Makefile builds from the reference
checkout (oracle/_ref/: ref_driver over the reference's headers, the
reference's CPU driver work/src/f3r.cpp in three precisions, and its matrix
generator matrix/hpcg.cpp) on small inputs, single-threaded so that every
reduction runs in sequential order, and stores inputs and outputs as .npz
data.  Run in the container that has the reference checkout:

    make -C oracle/ref_harness && python tests/golden/make_golden.py
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


if __name__ == "__main__":
    with tempfile.TemporaryDirectory() as T:
        unit_fixtures(T)
        f3r_fixtures(T)
        cpu_driver_fixture(T)
