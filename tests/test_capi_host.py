"""CPU-side checks of libf3r_amd.so: the library loads and exports every
entry point that include/f3r_amd.h declares, and its host set-up code
(generators, scaling, SELL-C-sigma layout, SDAINV, ILU(0)) matches the golden
vectors written by the reference and the oracle.  No GPU compute is called."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import problems as PB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "f3r_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(f3r_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import f3r_amd._capi as C
    names = _declared()
    assert len(names) >= 45
    for nm in names:
        assert hasattr(C.lib, nm), nm
    bound = {s[0] for s in C.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_error_paths_without_gpu():
    import f3r_amd as F
    from f3r_amd._capi import lib
    assert lib.f3r_strerror(-1) == b"invalid argument"
    h = ctypes.c_void_p()
    # bad CSR: unsorted columns
    rp = np.array([0, 2], np.int32)
    ix = np.array([0, 0], np.int32)
    v = np.ones(2)
    st = lib.f3r_mat_create_csr(None, 1, 1, 2, rp.ctypes.data, ix.ctypes.data, v.ctypes.data, ctypes.byref(h))
    assert st == -1 and b"sorted" in lib.f3r_last_error()
    # ILU(0) with more blocks than rows
    A = F.CSR(*PB.stencil7(3))
    with pytest.raises(F.F3RError):
        F.ILU0(A, blocks=100)
    # a host-only preconditioner cannot be applied
    P = F.SDAINV(A, 0.1, 1.0)
    assert lib.f3r_prec_apply(P.handle, None, None) == -1


@pytest.mark.parametrize("rp,ix,msg", [
    ([0, 5, 2], [0, 1], b"monotone"),       # a row pointer beyond nnz
    ([0, -1, 2], [0, 1], b"monotone"),      # a negative row pointer
    ([0, 1, 2], [0, 2], b"out of range"),   # a column outside the matrix
    ([0, 1, 2], [0, -1], b"out of range"),
    ([0, 2, 2], [1, 1], b"sorted"),         # a repeated column
    ([0, 1, 3], [0, 1], b"rptr[n]"),        # rptr[n] != nnz
])
def test_malformed_csr_rejected(rp, ix, msg):
    # every malformed CSR is refused before any array is read out of bounds
    from f3r_amd._capi import lib
    rp = np.array(rp, np.int32)
    ix = np.array(ix, np.int32)
    v = np.ones(len(ix))
    h = ctypes.c_void_p()
    st = lib.f3r_mat_create_csr(None, 2, 2, len(ix), rp.ctypes.data, ix.ctypes.data, v.ctypes.data, ctypes.byref(h))
    assert st == -1 and msg in lib.f3r_last_error(), lib.f3r_last_error()


def test_empty_csr_on_the_host():
    # a 0 x 0 matrix is a valid CSR; its host set-up is empty as well
    import f3r_amd as F
    A = F.CSR(0, np.array([0], np.int32), np.array([], np.int32), np.array([]))
    assert A.n == 0 and A.nnz == 0
    F.algorithm.scaling(A)
    F.SDAINV(A, 0.1, 1.0)


def test_generator_matches_reference_hpcg(golden):
    import f3r_amd as F
    import hashlib
    A = F.CSR.stencil27(32)
    h = hashlib.sha256()
    for a in (A.rptr, A.idx, A.val):
        h.update(np.ascontiguousarray(a).tobytes())
    assert np.frombuffer(h.digest(), np.uint8).tolist() == golden("cpu_driver_hpcg32")["hpcg_sha256"].tolist()


def test_generators_match_numpy_forms():
    import f3r_amd as F
    A = F.CSR.stencil27(7, 5, 6, beta=0.5)
    n, rp, ix, v = PB.stencil27(7, 5, 6, beta=0.5)
    assert A.n == n and np.array_equal(A.rptr, rp) and np.array_equal(A.idx, ix) and np.array_equal(A.val, v)
    assert A.nnz == (3 * 7 - 2) * (3 * 5 - 2) * (3 * 6 - 2)
    B = F.CSR.dominant(7, 6, margin=0.5)
    n, rp, ix, v = PB.stencil7(6)
    assert np.array_equal(B.rptr, rp) and np.array_equal(B.idx, ix)
    offs = np.diff(rp) - 1
    diag = B.val[ix == np.repeat(np.arange(n), np.diff(rp))]
    np.testing.assert_allclose(diag, 1.5 * offs)
    # variable-coefficient convection-diffusion: M-matrix sign pattern,
    # diffusion part symmetric, convection only on the -x neighbour
    D = F.CSR.convdiff7(10, sigma=1.5, corr=3.0, peclet=0.0, seed=4)
    Dd = PB.to_dense(D.n, D.rptr, D.idx, D.val)
    np.testing.assert_allclose(Dd, Dd.T, rtol=1e-14)
    assert np.all(np.diag(Dd) > 0) and np.all(Dd - np.diag(np.diag(Dd)) <= 0)
    Cv = F.CSR.convdiff7(10, sigma=1.5, corr=3.0, peclet=10.0, seed=4)
    Cd = PB.to_dense(Cv.n, Cv.rptr, Cv.idx, Cv.val)
    Dif = Cd - Dd
    i = np.arange(D.n)
    assert np.allclose(np.diag(Dif), 10.0)
    west = (i % 10) > 0
    assert np.allclose(Dif[i[west], i[west] - 1], -10.0)
    Dif[i, i] = 0
    Dif[i[west], i[west] - 1] = 0
    assert np.all(Dif == 0)
    # the lognormal field has unit log-variance scaled by sigma
    Cs = F.CSR.convdiff7(24, sigma=1.0, corr=2.0, peclet=0.0, seed=1)
    dg = Cs.val[Cs.idx == np.repeat(np.arange(Cs.n), np.diff(Cs.rptr))]
    assert 0.2 < np.std(np.log(dg)) < 2.0
    u = F.uniform(1000, -1, 1, seed=3)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.1
    assert np.array_equal(u, F.uniform(1000, -1, 1, seed=3))


@pytest.mark.parametrize("name", ["s27b", "s7", "rnd"])
def test_scaling_matches_reference(golden, name):
    import f3r_amd as F
    g = golden("units_" + name)
    A = F.CSR(int(g["n"]), g["rptr"], g["idx"], g["val"].copy())
    F.algorithm.scaling(A)
    assert np.array_equal(A.val, g["val_scaled"])


@pytest.mark.parametrize("name", ["s27b", "s7", "rnd"])
def test_sell32_layout_matches_reference(golden, name):
    import f3r_amd as F
    g = golden("units_" + name)
    A = F.CSR(int(g["n"]), g["rptr"], g["idx"], g["val_scaled"])
    cptr, cw, sidx, sval, perm = F.sell_layout(A, 32, 1, 1)
    assert np.array_equal(cptr[:-1], g["sell32_sptr"][:-1] * 32)
    assert np.array_equal(sidx, g["sell32_idx"]) and np.array_equal(sval, g["sell32_val"])
    assert np.array_equal(perm, np.arange(A.n))


@pytest.mark.parametrize("C,sigma,K", [(32, 1, 1), (64, 64, 4), (16, 256, 2), (64, 1, 4)])
def test_sell_sigma_layout_matches_oracle(C, sigma, K):
    import f3r_amd as F
    n, rp, ix, v = PB.random_sparse(500, 14, 11)
    A = F.CSR(n, rp, ix, v)
    cptr, cw, sidx, sval, perm = F.sell_layout(A, C, sigma, K)
    oc, oi, ov, op = O.csr_to_sell(n, rp, ix, v, C, sigma)
    assert np.array_equal(perm, op)
    # same row content in slot order: decode both
    for s in range(len(cw)):
        tc = min(C, n - s * C)
        for j in range(tc):
            row = perm[s * C + j]
            L = rp[row + 1] - rp[row]
            ours = [sidx[cptr[s] + ((q // K) * tc + j) * K + q % K] for q in range(L)]
            theirs = [oi[oc[s] + q * tc + j] for q in range(L)]
            assert ours == theirs
            assert cw[s] * K >= L and cw[s] * K - L < K or cw[s] * K >= L
    # padding stays below K slots per row over the chunk maximum
    lens = np.diff(rp)[perm]
    for s in range(len(cw)):
        mx = lens[s * C:(s + 1) * C].max()
        assert cw[s] * K == ((mx + K - 1) // K) * K


@pytest.mark.parametrize("name", ["s27b", "s7", "rnd"])
def test_sdainv_host_matches_reference(golden, name):
    import f3r_amd as F
    g = golden("units_" + name)
    A = F.CSR(int(g["n"]), g["rptr"], g["idx"], g["val_scaled"])
    for k in range(3):
        tol, acc = g[f"sdainv{k}_params"]
        (zr, zi, zv), (wr, wi, wv) = F.SDAINV(A, tol, acc).factors()
        for tag, got in (("Z", (zr, zi, zv)), ("W", (wr, wi, wv))):
            for part, arr in zip(("rptr", "idx", "val"), got):
                assert np.array_equal(arr, g[f"sdainv{k}_{tag}_{part}"]), (k, tag, part)


@pytest.mark.parametrize("name", ["s27b", "s7", "rnd"])
def test_ilu0_host_matches_reference(golden, name):
    import f3r_amd as F
    g = golden("units_" + name)
    A = F.CSR(int(g["n"]), g["rptr"], g["idx"], g["val_scaled"])
    for k in range(2):
        blocks, acc = g[f"ilu{k}_params"]
        (lr, li, lv), (ur, ui, uv) = F.ILU0(A, int(blocks), acc).factors()
        for tag, got in (("L", (lr, li, lv)), ("U", (ur, ui, uv))):
            for part, arr in zip(("rptr", "idx", "val"), got):
                assert np.array_equal(arr, g[f"ilu{k}_{tag}_{part}"]), (k, tag, part)


def test_sdainv_host_matches_oracle_on_lognormal():
    import f3r_amd as F
    A = F.CSR.convdiff7(9, sigma=2.0, corr=2.0, peclet=0.0, seed=2)
    F.algorithm.scaling(A)
    (zr, zi, zv), (wr, wi, wv) = F.SDAINV(A, 0.1, 1.0).factors()
    (ozr, ozi, ozv), (owr, owi, owv) = O.sdainv(A.n, A.rptr, A.idx, A.val, 0.1, 1.0)
    assert np.array_equal(zr, ozr) and np.array_equal(zi, ozi) and np.array_equal(zv, ozv)
    assert np.array_equal(wr, owr) and np.array_equal(wi, owi) and np.array_equal(wv, owv)
    assert len(zi) > A.n  # non-trivial pattern on the 7-point lognormal family


def _rows_from_offset_layout(L, n, C, K):
    """(columns, values) of every row, in slot order, from a chunk-offset layout"""
    rows = []
    nch = (n + C - 1) // C
    for s in range(nch):
        r0 = s * C
        tc = min(C, n - r0)
        g, half = L["cwidth"][s] & ((1 << 30) - 1), bool(L["cwidth"][s] & (1 << 30))
        assert not half or L["iptr"][s] < 0  # half groups only in offset chunks
        w = g * K + (K // 2 if half else 0)
        assert L["cptr"][s + 1] - L["cptr"][s] == w * tc
        for j in range(tc):
            r = r0 + j
            cols, vals = [], []
            for t in range(w):
                if t < g * K:
                    pos = L["cptr"][s] + ((t // K) * tc + j) * K + t % K
                else:
                    pos = L["cptr"][s] + g * K * tc + j * (K // 2) + t - g * K
                if L["iptr"][s] < 0:
                    c = r + L["coff"][-L["iptr"][s] - 1 + t]
                else:
                    c = L["idx"][L["iptr"][s] + pos - L["cptr"][s]]
                cols.append(c)
                vals.append(L["val"][pos])
            rows.append((np.array(cols), np.array(vals), L["iptr"][s] < 0))
    return rows


@pytest.mark.parametrize("kind,C,K", [("27pt", 64, 4), ("27pt", 64, 8), ("27pt", 128, 2), ("7pt", 64, 1),
                                      ("ragged", 64, 4)])
def test_offset_layout_reproduces_csr_rows(kind, C, K):
    # every nonzero of every row appears once, in CSR (column) order; the other
    # slots hold 0 at an in-range column; stencil chunks use offsets
    import f3r_amd as F
    if kind == "27pt":
        A = F.CSR.stencil27(9, 8, 10, beta=0.5)
    elif kind == "7pt":
        A = F.CSR.convdiff7(11, sigma=1.0, corr=3.0, peclet=5.0, seed=2)
    else:
        n, rp, ix, v = PB.random_sparse(700, 30, 4)
        A = F.CSR(n, rp, ix, v)
    L = F.sell_offset_layout(A, C, K)
    rows = _rows_from_offset_layout(L, A.n, C, K)
    n_off = 0
    for r, (cols, vals, is_off) in enumerate(rows):
        lo, hi = A.rptr[r], A.rptr[r + 1]
        assert np.all((cols >= 0) & (cols < A.n))
        if is_off:
            n_off += 1
            nz = vals != 0
            assert np.array_equal(cols[nz], A.idx[lo:hi][A.val[lo:hi] != 0])
            assert np.array_equal(vals[nz], A.val[lo:hi][A.val[lo:hi] != 0])
            assert np.all(np.diff(cols[nz]) > 0)      # nonzeros summed in column (CSR) order
        else:
            k = hi - lo
            assert np.array_equal(cols[:k], A.idx[lo:hi]) and np.array_equal(vals[:k], A.val[lo:hi])
            assert np.all(vals[k:] == 0)
    if kind != "ragged":
        assert n_off >= A.n // 2
    if kind != "ragged":
        # consecutive chunks with the same offsets share one table
        offc = L["iptr"] < 0
        assert len(np.unique(L["iptr"][offc])) < offc.sum()
        for s in range(1, len(L["iptr"])):
            if offc[s] and offc[s - 1] and L["iptr"][s] != L["iptr"][s - 1]:
                w = lambda c: (L["cwidth"][c] & ((1 << 30) - 1)) * K + (K // 2 if L["cwidth"][c] & (1 << 30) else 0)
                t0, t1 = -L["iptr"][s - 1] - 1, -L["iptr"][s] - 1
                assert w(s) != w(s - 1) or not np.array_equal(L["coff"][t0:t0 + w(s)], L["coff"][t1:t1 + w(s)])
    if kind == "27pt" and K == 8:
        # 27 offsets: three groups of 8 and a half group of 4 (28 slots, not 32)
        half = (L["cwidth"] & (1 << 30)) != 0
        assert half.any() and np.all(L["iptr"][half] < 0)
        assert np.all(L["cwidth"][half] & ((1 << 30) - 1) == 3)
    # eligibility rule: an offset chunk has at most 5/4 of the int32 layout's
    # slots, so the layout never grows by more than that
    plain = F.sell_layout(A, C, 1, K)[3]
    assert L["val"].size <= 1.25 * plain.size + K * ((A.n + C - 1) // C) * C


def test_tuner_random_rows_problem():
    # the autotuner's long-row problem: seeded, rows of 1..maxlen entries with
    # the diagonal, ascending distinct columns, strictly dominant diagonal
    from f3r_amd.tune import random_rows, candidates
    A = random_rows(500, 40, seed=3)
    B = random_rows(500, 40, seed=3)
    assert np.array_equal(A.rptr, B.rptr) and np.array_equal(A.idx, B.idx) and np.array_equal(A.val, B.val)
    lens = np.diff(A.rptr)
    assert lens.min() >= 1 and lens.max() <= 40
    for i in range(A.n):
        c, v = A.idx[A.rptr[i]:A.rptr[i + 1]], A.val[A.rptr[i]:A.rptr[i + 1]]
        assert np.all(np.diff(c) > 0) and i in c
        assert v[c == i][0] > np.abs(v[c != i]).sum()
    # the SELL threads-per-row candidates: K lanes for K >= 2, int32 indices, no staging
    sells = [e for e in candidates(tprs=()) if e["lanes"] > 1]
    assert sells and all(e["lanes"] == e["K"] >= 2 and not e["offsets"] and not e["lds"] for e in sells)


def test_header_compiles_as_c99(tmp_path):
    # the boundary header is plain C: a C caller (cgo-, ctypes- or JNI-style
    # bindings) includes it without a C++ compiler
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "use.c"
    src.write_text('#include "f3r_amd.h"\n'
                   "int main(void) { f3r_params p; f3r_result r; (void)r;\n"
                   "  return f3r_params_default(&p) == F3R_OK && p.m1 > 0 ? 0 : 1; }\n")
    res = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                          os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "use.o")],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_setup_where_argument_checks():
    # the build-location queries refuse null handles and a CSR matrix (they
    # report SELL layouts), without touching a GPU
    from f3r_amd._capi import lib
    w = ctypes.c_int(-7)
    assert lib.f3r_mat_setup_where(None, ctypes.byref(w)) == -1 and w.value == -7
    assert lib.f3r_solver_setup_where(None, ctypes.byref(w)) == -1 and w.value == -7
    rp, ix, v = np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.ones(2)
    h = ctypes.c_void_p()
    assert lib.f3r_mat_create_csr(None, 2, 2, 2, rp.ctypes.data, ix.ctypes.data, v.ctypes.data, ctypes.byref(h)) == 0
    try:
        assert lib.f3r_mat_setup_where(h, None) == -1
        assert lib.f3r_mat_setup_where(h, ctypes.byref(w)) == -1 and b"SELL" in lib.f3r_last_error()
    finally:
        lib.f3r_mat_destroy(h)
