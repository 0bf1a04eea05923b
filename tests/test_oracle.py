"""Pins the CPU oracle (oracle/f3r_oracle.cpp) before anything is checked
against it: golden vectors written by the reference's own code
(tests/golden/make_golden.py), closed forms, brute force on tiny inputs and
invariants chosen so that a dropped term, a wrong index or a transposed
operand fails one of them."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import problems as PB

UNITS = ("s27b", "s7", "rnd")


# ------------------------------------------------------------------ fp16 ----
def test_h2f_exhaustive_against_numpy():
    bits = np.arange(1 << 16, dtype=np.uint16)
    out = np.zeros(1 << 16, np.float32)
    O.lib().orc_h2f_array(len(bits), O.P(bits), O.P(out))
    ref = bits.view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(out), nan)
    assert np.array_equal(out[~nan].view(np.uint32), ref[~nan].view(np.uint32))


@pytest.mark.parametrize("src", ["f32", "f64"])
def test_to_half_round_to_nearest_even(src):
    rng = np.random.default_rng(0)
    h = np.arange(1 << 16, dtype=np.uint16).view(np.float16)
    h = h[np.isfinite(h)].astype(np.float64)
    hs = np.sort(np.unique(h))
    mids = (hs[1:] + hs[:-1]) / 2          # exact ties: round to even
    vals = np.concatenate([hs, mids, np.nextafter(mids, np.inf), np.nextafter(mids, -np.inf),
                           rng.standard_normal(200000) * 10.0 ** rng.integers(-9, 6, 200000),
                           [65504.0, 65519.99, 65520.0, 1e6, -1e6, 2.0 ** -25, 2.0 ** -26,
                            3 * 2.0 ** -26, 0.0, -0.0]])
    out = np.zeros(vals.size, np.uint16)
    if src == "f32":
        v = vals.astype(np.float32)
        O.lib().orc_f2h_array(v.size, O.P(v), O.P(out))
        ref = v.astype(np.float16)
    else:
        v = vals.astype(np.float64)
        O.lib().orc_d2h_array(v.size, O.P(v), O.P(out))
        ref = v.astype(np.float16)      # numpy: direct, correctly rounded
    assert np.array_equal(out, ref.view(np.uint16))


# ------------------------------------------------------- golden vectors ----
@pytest.mark.parametrize("name", UNITS)
def test_scaling_matches_reference(golden, name):
    g = golden("units_" + name)
    vs = O.scaling(int(g["n"]), g["rptr"], g["idx"], g["val"])
    assert np.array_equal(vs, g["val_scaled"])


def test_scaling_closed_form():
    n, rp, ix, v = PB.stencil27(4)
    vs = O.scaling(n, rp, ix, v)
    # constant diagonal 26: every entry / 26 (up to the rounding of sqrt(26)^2)
    np.testing.assert_allclose(vs, v / 26.0, rtol=5e-16, atol=0)
    assert np.all(vs[ix == np.repeat(np.arange(n), np.diff(rp))] == 26.0 / (np.sqrt(26.0) ** 2))


@pytest.mark.parametrize("name", UNITS)
def test_sell32_layout_matches_reference(golden, name):
    g = golden("units_" + name)
    n = int(g["n"])
    cptr, sidx, sval, perm = O.csr_to_sell(n, g["rptr"], g["idx"], g["val_scaled"], 32, 1)
    # the reference stores sptr in units of 32 slots; the oracle stores element offsets
    assert np.array_equal(cptr[:-1], g["sell32_sptr"][:-1] * 32)
    assert np.array_equal(sidx, g["sell32_idx"])
    assert np.array_equal(sval, g["sell32_val"])
    assert np.array_equal(perm, np.arange(n))


@pytest.mark.parametrize("C,sigma", [(32, 1), (64, 1), (16, 64), (64, 256)])
def test_sell_sigma_round_trip(C, sigma):
    n, rp, ix, v = PB.random_sparse(300, 12, 7)
    cptr, sidx, sval, perm = O.csr_to_sell(n, rp, ix, v, C, sigma)
    assert sorted(perm) == list(range(n))
    D = PB.to_dense(n, rp, ix, v)
    R = np.zeros_like(D)
    for s in range(len(cptr) - 1):
        r0 = s * C
        tc = min(C, n - r0)
        width = (cptr[s + 1] - cptr[s]) // tc
        for j in range(tc):
            row = perm[r0 + j]
            for k in range(width):
                R[row, sidx[cptr[s] + k * tc + j]] += sval[cptr[s] + k * tc + j]
    assert np.array_equal(R, D)
    lens = np.diff(rp)[perm]
    for w in range(0, n, sigma):
        seg = lens[w:w + sigma]
        assert np.all(seg[:-1] >= seg[1:]) or sigma == 1


@pytest.mark.parametrize("name", UNITS)
def test_sdainv_matches_reference(golden, name):
    g = golden("units_" + name)
    n = int(g["n"])
    for k in range(3):
        tol, acc = g[f"sdainv{k}_params"]
        (zr, zi, zv), (wr, wi, wv) = O.sdainv(n, g["rptr"], g["idx"], g["val_scaled"], tol, acc)
        for tag, (r, i, v) in (("Z", (zr, zi, zv)), ("W", (wr, wi, wv))):
            assert np.array_equal(r, g[f"sdainv{k}_{tag}_rptr"]), (k, tag)
            assert np.array_equal(i, g[f"sdainv{k}_{tag}_idx"]), (k, tag)
            assert np.array_equal(v, g[f"sdainv{k}_{tag}_val"]), (k, tag)


def test_sdainv_exact_inverse_on_tridiagonal():
    # The biconjugation only runs over A's own pattern; on a tridiagonal matrix
    # that drops nothing, so with tol = 0 and acc = 1: Z D^-1 W^T = A^-1.
    n = 40
    rng = np.random.default_rng(5)
    A = (np.diag(rng.uniform(3, 4, n)) + np.diag(rng.uniform(-1, 1, n - 1), 1)
         + np.diag(rng.uniform(-1, 1, n - 1), -1))
    r, c = np.nonzero(A)
    rp = np.searchsorted(r, np.arange(n + 1)).astype(np.int32)
    ix, v = c.astype(np.int32), A[r, c]
    (zr, zi, zv), (wr, wi, wv) = O.sdainv(n, rp, ix, v, 0.0, 1.0)
    Z = PB.to_dense(n, zr, zi, zv)          # columns z_i / d_i
    Wt = PB.to_dense(n, wr, wi, wv)         # rows w_i
    A = PB.to_dense(n, rp, ix, v)
    np.testing.assert_allclose(Z @ Wt, np.linalg.inv(A), rtol=0, atol=1e-10)


def test_sdainv_identity_on_scaled_27pt():
    # scaled 27-point: every |alpha| < tol = 0.1, so W = I and Z = I / acc
    n, rp, ix, v = PB.stencil27(6, beta=0.5)
    v = O.scaling(n, rp, ix, v)
    for acc in (1.0, 1.5):
        (zr, zi, zv), (wr, wi, wv) = O.sdainv(n, rp, ix, v, 0.1, acc)
        assert np.array_equal(zi, np.arange(n)) and np.array_equal(wi, np.arange(n))
        assert np.array_equal(wv, np.ones(n))
        np.testing.assert_allclose(zv, 1.0 / (acc * np.ones(n)), rtol=1e-15)


@pytest.mark.parametrize("name", UNITS)
def test_ilu0_matches_reference(golden, name):
    g = golden("units_" + name)
    n = int(g["n"])
    for k in range(2):
        blocks, acc = g[f"ilu{k}_params"]
        (lr, li, lv), (ur, ui, uv) = O.ilu0(n, g["rptr"], g["idx"], g["val_scaled"], int(blocks), acc)
        assert np.array_equal(lr, g[f"ilu{k}_L_rptr"]) and np.array_equal(li, g[f"ilu{k}_L_idx"])
        assert np.array_equal(lv, g[f"ilu{k}_L_val"])
        assert np.array_equal(ur, g[f"ilu{k}_U_rptr"]) and np.array_equal(ui, g[f"ilu{k}_U_idx"])
        assert np.array_equal(uv, g[f"ilu{k}_U_val"])


def test_ilu0_reproduces_pattern_and_tridiagonal_lu():
    n, rp, ix, v = PB.random_sparse(50, 7, 9)
    (lr, li, lv), (ur, ui, uv) = O.ilu0(n, rp, ix, v, 1, 1.0)
    LU = PB.to_dense(n, lr, li, lv) @ PB.to_dense(n, ur, ui, uv)
    A = PB.to_dense(n, rp, ix, v)
    mask = A != 0
    np.testing.assert_allclose(LU[mask], A[mask], rtol=1e-12, atol=1e-12)
    # tridiagonal: ILU(0) is the exact LU
    n = 30
    A = np.diag(np.full(n, 4.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -2.0), -1)
    r, c = np.nonzero(A)
    rp = np.searchsorted(r, np.arange(n + 1)).astype(np.int32)
    (lr, li, lv), (ur, ui, uv) = O.ilu0(n, rp, c.astype(np.int32), A[r, c], 1, 1.0)
    np.testing.assert_allclose(PB.to_dense(n, lr, li, lv) @ PB.to_dense(n, ur, ui, uv), A, atol=1e-14)


def test_bj_solve_inverts_lu():
    n, rp, ix, v = PB.random_sparse(60, 7, 4)
    h = O.lib().orc_ilu0(n, O.P(rp), O.P(ix), O.P(v), 1, 1.0)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    out, tmp = np.zeros(n), np.zeros(n)
    O.lib().orc_bj_solve(h, O.P(x), O.P(out), O.P(tmp))
    nl = O.lib().orc_factor_nnz(h, 0)
    nu = O.lib().orc_factor_nnz(h, 1)
    L = [np.zeros(n + 1, np.int32), np.zeros(nl, np.int32), np.zeros(nl)]
    U = [np.zeros(n + 1, np.int32), np.zeros(nu, np.int32), np.zeros(nu)]
    O.lib().orc_factor_get(h, 0, *[O.P(a) for a in L])
    O.lib().orc_factor_get(h, 1, *[O.P(a) for a in U])
    O.lib().orc_factor_free(h)
    M = PB.to_dense(n, *L) @ PB.to_dense(n, *U)
    np.testing.assert_allclose(M @ out, x, atol=1e-12)


@pytest.mark.parametrize("name", UNITS)
def test_spmv_matches_reference(golden, name):
    g = golden("units_" + name)
    n = int(g["n"])
    x = g["x"]
    O.lib().orc_set_fma(0)   # the reference was built without contraction
    try:
        for mode, args in ((64, (0, 0, 0, 0)), (32, (1, 1, 1, 1)), (1616, (2, 2, 2, 2)),
                           (1632, (2, 1, 1, 1))):
            y = O.spmv(n, g["rptr"], g["idx"], g["val_scaled"], x, *args)
            ref = g[f"spmv{mode}"]
            assert np.array_equal(y.view(np.uint8), ref.view(np.uint8)), mode
    finally:
        O.lib().orc_set_fma(1)


def test_spmv_fp64_brute_force_and_row_sums():
    n, rp, ix, v = PB.random_sparse(80, 9, 2)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    y = O.spmv(n, rp, ix, v, x, 0, 0, 0, 0)
    D = PB.to_dense(n, rp, ix, v)
    np.testing.assert_allclose(y, D @ x, rtol=1e-14, atol=1e-14)
    n, rp, ix, v = PB.stencil27(6)
    y = O.spmv(n, rp, ix, v, np.ones(n), 0, 0, 0, 0)
    # interior rows of the 27-point stencil sum to 26 - 26 = 0
    interior = np.zeros((6, 6, 6), bool)
    interior[1:-1, 1:-1, 1:-1] = True
    assert np.all(y[interior.ravel()] == 0)
    assert np.all(y[~interior.ravel()] > 0)


# --------------------------------------------------------- nested solve ----
@pytest.mark.parametrize("case", ["s27b8", "s7_10"])
@pytest.mark.parametrize("prec,variant", [(64, 2), (32, 1), (16, 0)])
def test_f3r_bitexact_against_reference_composition(golden, case, prec, variant):
    g = golden("f3r_" + case)
    n = int(g["n"])
    x, r, hist = O.f3r_solve(n, g["rptr"], g["idx"], g["val_scaled"], g["b"], variant=variant,
                             ref_arith=1, nthreads=1)
    s, it, impl, _ = g[f"p{prec}_res"]
    assert r.is_solved == int(s) and r.res_iter == int(it)
    assert r.res_nrm2 / r.nrm_b == impl
    assert np.array_equal(x, g[f"p{prec}_x"])
    h = g[f"p{prec}_hist"]
    assert np.array_equal(hist[:len(h)], h)


def _step_cases():
    import json
    g = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "f3r_ref_steps.npz"),
                     allow_pickle=False))
    out = []
    for key in g["cases"]:
        for prec, variant in ((16, 0), (32, 1), (64, 2)):
            if f"{key}_p{prec}_res" in g:
                out.append((str(key), prec, variant))
    return out


@pytest.mark.parametrize("key,prec,variant", _step_cases())
def test_f3r_step_counts_bitexact_against_reference(golden, key, prec, variant):
    # problems on which the reference's nest needs 3 to 14 outer steps
    # (shortened inner levels, and the 64^3 CPU-path configuration with
    # b = A*1 and restart 3): outer count, implicit residual and the whole
    # residual history equal to the reference's own run
    import json
    g = golden("f3r_ref_steps")
    d = json.loads(str(g[f"{key}_desc"]))
    m2, m3, m4, c, restart = (int(t) for t in g[f"{key}_m"])
    n, rp, ix, v, b = PB.from_desc(d)
    assert PB.csr_sha256(rp, ix, v).tolist() == g[f"{key}_sha_unscaled"].tolist()
    vs = O.scaling(n, rp, ix, v)
    assert PB.csr_sha256(rp, ix, vs).tolist() == g[f"{key}_sha_scaled"].tolist()
    _, r, hist = O.f3r_solve(n, rp, ix, vs, b, variant=variant, ref_arith=1, nthreads=1, m2=m2, m3=m3, m4=m4,
                             c=c, restart=restart)
    s, it, impl, _ = g[f"{key}_p{prec}_res"]
    assert r.is_solved == int(s) and r.res_iter == int(it)
    assert r.res_nrm2 / r.nrm_b == impl
    h = g[f"{key}_p{prec}_hist"]
    assert np.array_equal(hist[:len(h)], h)


@pytest.mark.parametrize("prec,variant", [(64, 2), (32, 1), (16, 0)])
def test_f3r_bjilu0_matches_reference_cpu_driver(golden, prec, variant):
    g = golden("cpu_driver_hpcg32")
    n, rp, ix, v = PB.stencil27(32)
    import hashlib
    h = hashlib.sha256()
    for a in (rp, ix, v):
        h.update(np.ascontiguousarray(a).tobytes())
    assert np.frombuffer(h.digest(), np.uint8).tolist() == g["hpcg_sha256"].tolist()
    vs = O.scaling(n, rp, ix, v)
    x, r, _ = O.f3r_solve(n, rp, ix, vs, g["b"], variant=variant, ref_arith=1, precond=1,
                          bj_blocks=112, restart=3, nthreads=1)
    it, impl, expl = g[f"p{prec}_csv"]
    assert r.res_iter * 64 == int(it)
    # the CSV line prints 7 significant digits
    assert abs(r.res_nrm2 / r.nrm_b - impl) <= 5e-7 * impl
    # the explicit residual is recomputed in fp64 by each side: agreement to
    # the rounding floor of ||b - A x|| (about 1e-16 relative to ||b||)
    assert abs(r.expl_res - expl) <= 1e-15


def test_fgmres_exact_in_n_steps_with_identity_levels():
    # DOUBLE variant on a tiny system: the nest reaches the fp64 floor at once
    n, rp, ix, v = PB.random_sparse(30, 5, 1)
    v = O.scaling(n, rp, ix, v)
    b = np.random.default_rng(2).uniform(-1, 1, n)
    x, r, _ = O.f3r_solve(n, rp, ix, v, b, variant=2, eps=1e-13)
    assert r.is_solved and r.expl_res < 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(PB.to_dense(n, rp, ix, v), b), atol=1e-10)


def test_gmres_oracle_solves():
    n, rp, ix, v = PB.stencil7(10)
    v = O.scaling(n, rp, ix, v)
    b = np.random.default_rng(0).uniform(-1, 1, n)
    x, it, rr = O.gmres(n, rp, ix, v, b, m=40, tol=1e-11)
    assert rr < 1e-10
    np.testing.assert_allclose(x, np.linalg.solve(PB.to_dense(n, rp, ix, v), b), atol=1e-8)


def test_build_arithmetic_close_to_reference_arithmetic():
    # the build's fp16 level (fp32 sums, fp32 weights, range factor) versus the
    # reference's fp16 arithmetic: same outer iteration count on a 16^3 problem
    n, rp, ix, v = PB.stencil27(16, beta=0.5)
    v = O.scaling(n, rp, ix, v)
    b = np.random.default_rng(1).uniform(-1, 1, n)
    _, r_ref, _ = O.f3r_solve(n, rp, ix, v, b, variant=0, ref_arith=1)
    _, r_bld, _ = O.f3r_solve(n, rp, ix, v, b, variant=0, ref_arith=0, range_scale=16.0)
    assert r_ref.is_solved and r_bld.is_solved
    assert abs(r_ref.res_iter - r_bld.res_iter) <= max(1, 0.1 * r_ref.res_iter)
    assert r_bld.expl_res < 1e-8


def test_mlp_forward_against_numpy():
    # the oracle's forward pass: ReLU(W1 x + b1) -> ReLU(W2 . + b2) -> w3 . + b3,
    # against numpy's matrix products in fp64 (a dropped bias, a missing ReLU
    # or a transposed weight fails by O(1))
    rng = np.random.default_rng(3)
    x = rng.standard_normal((9, 12))
    w1, b1 = rng.standard_normal((6, 12)), rng.standard_normal(6)
    w2, b2 = rng.standard_normal((5, 6)), rng.standard_normal(5)
    w3, b3 = rng.standard_normal((1, 5)), rng.standard_normal(1)
    a1 = np.maximum(x @ w1.T + b1, 0.0)
    a2 = np.maximum(a1 @ w2.T + b2, 0.0)
    ref = (a2 @ w3.T + b3).reshape(-1)
    y = O.mlp_forward(x, w1, b1, w2, b2, w3, b3)
    assert np.max(np.abs(y - ref)) <= 1e-12 * (1 + np.max(np.abs(ref)))
    assert np.any(a1 == 0) and np.any(a2 == 0)  # both ReLUs clip something here


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((7, 12))
    w1, b1 = rng.standard_normal((5, 12)) * 0.3, rng.standard_normal(5) * 0.1
    w2, b2 = rng.standard_normal((4, 5)) * 0.3, rng.standard_normal(4) * 0.1
    w3, b3 = rng.standard_normal((1, 4)) * 0.3, rng.standard_normal(1) * 0.1
    g = rng.standard_normal(7)
    grads = O.mlp_backward(x, w1, b1, w2, b2, w3, b3, g)
    params = [w1, b1, w2, b2, w3, b3]
    eps = 1e-6
    for p, gp in zip(params, grads):
        flat = p.reshape(-1)
        for k in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[k]
            flat[k] = old + eps
            fp = g @ O.mlp_forward(x, *params)
            flat[k] = old - eps
            fm = g @ O.mlp_forward(x, *params)
            flat[k] = old
            assert abs((fp - fm) / (2 * eps) - gp.reshape(-1)[k]) < 1e-6


@pytest.mark.skipif(not os.path.exists(os.path.join(O.REF_DIR, "ref_driver")),
                    reason="oracle/_ref not built (needs the reference checkout)")
def test_f3r_against_live_reference_build(tmp_path):
    # 27-point nonsymmetric 12^3 through the reference's own code, live
    n, rp, ix, v = PB.stencil27(12, beta=0.5)
    v = O.scaling(n, rp, ix, v)
    b = np.random.default_rng(9).uniform(-1, 1, n)
    PB.write_csr_bin(tmp_path / "A.bin", n, rp, ix, v)
    PB.write_vec_bin(tmp_path / "b.bin", b)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([os.path.join(O.REF_DIR, "ref_driver"), "f3r", str(tmp_path / "A.bin"),
                    str(tmp_path / "b.bin"), "16", "8", "4", "2", "64", "20", "1.0",
                    str(tmp_path / "x.bin"), str(tmp_path / "h.bin")], check=True, env=env,
                   capture_output=True)
    xr = PB.read_vec_bin(tmp_path / "x.bin", np.float64)
    x, r, _ = O.f3r_solve(n, rp, ix, v, b, variant=0, ref_arith=1, nthreads=1)
    assert np.array_equal(x, xr)


# -------------------------------------------------------------- bfloat16 ----
def _bf16_value(bits):
    return np.frombuffer((np.asarray(bits, np.uint32) << 16).astype(np.uint32).tobytes(), np.float32).astype(np.float64)


def test_bf16_roundtrip_exhaustive():
    L = O.lib()
    for b in range(1 << 16):
        f = L.orc_b2f(b)
        assert f == _bf16_value(b)[0] or (np.isnan(f) and np.isnan(_bf16_value(b)[0]))
        if not np.isnan(f):
            assert L.orc_f2b(f) == b, hex(b)


def test_bf16_from_fp32_against_neighbours():
    # fp32 -> bf16 nearest even, from the two bf16 numbers around each value
    # (differences of fp32 numbers are exact in fp64)
    L = O.lib()
    u = np.random.default_rng(5).integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32)
    u = u[(u & 0x7f800000) != 0x7f800000]  # finite
    u = np.concatenate([u, (u & 0xffff0000) | 0x8000])  # exact ties
    f = np.frombuffer(u.tobytes(), np.float32)
    lo = u >> 16
    vlo, vhi = _bf16_value(lo), _bf16_value(lo + 1)
    vhi = np.where(((lo + 1) & 0x7fff) == 0x7f80, 2.0 ** 128, vhi)  # rounding up past the largest: inf
    x = np.abs(f.astype(np.float64))
    dl, dh = x - np.abs(vlo), np.abs(vhi) - x
    want = np.where((dh < dl) | ((dh == dl) & (lo % 2 == 1)), lo + 1, lo).astype(np.uint16)
    got = np.array([L.orc_f2b(float(t)) for t in f], np.uint16)
    assert np.array_equal(got, want)


def test_bf16_from_fp64_exact_rounding():
    # fp64 -> bf16 against exact rational arithmetic: normal and subnormal
    # ranges, exact ties and their fp64 neighbours, overflow
    from fractions import Fraction
    L = O.lib()
    rng = np.random.default_rng(6)
    xs = list(rng.uniform(-1, 1, 500) * 2.0 ** rng.integers(-140, 128, 500))
    for b in rng.integers(1, 0x7f7f, 300):  # ties between bf16 neighbours, and next to them
        m = (float(_bf16_value(b)[0]) + float(_bf16_value(b + 1)[0])) / 2
        xs += [m, np.nextafter(m, 0.0), np.nextafter(m, np.inf), -m]
    xs += [3.4e38, 3.3961e38, 2.0 ** -134, 2.0 ** -133 * 1.5, 2.0 ** -135, 0.0, -0.0]
    grid = {}  # every finite non-negative bf16 value, by value
    for b in range(0x7f80):
        grid[b] = Fraction(float(_bf16_value(b)[0]))
    vals = sorted((v, b) for b, v in grid.items())
    import bisect
    keys = [v for v, _ in vals]
    for x in xs:
        a = Fraction(abs(x))
        i = bisect.bisect_left(keys, a)
        if i >= len(keys):
            want = vals[-1][1]  # above the largest finite: it, or inf (below)
        elif keys[i] == a:
            want = vals[i][1]
        else:
            lo, hi = vals[i - 1], vals[i]
            dl, dh = a - lo[0], hi[0] - a
            want = hi[1] if dh < dl or (dh == dl and hi[1] % 2 == 0) else lo[1]
        if a >= Fraction(2) ** 128 - Fraction(2) ** 119:  # past the midpoint above the largest
            want = 0x7f80
        if np.signbit(x):
            want |= 0x8000
        assert L.orc_d2b(float(x)) == want, (x, hex(L.orc_d2b(float(x))), hex(want))
