"""Independent numpy generators of the stencil matrices used by the tests.

These are written separately from the product's C++ generators so that the
two can check each other; the 27-point form follows the closed form of the
reference's generator (matrix/hpcg.cpp:13-57): lexicographic order
row = (k*ny + j)*nx + i, columns sorted ascending, diagonal 26, neighbours -1.
"""
import numpy as np


def _assemble(n, R, C, V):
    o = np.lexsort((C, R))
    R, C, V = R[o], C[o], V[o]
    rptr = np.zeros(n + 1, np.int64)
    np.add.at(rptr, R + 1, 1)
    rptr = np.cumsum(rptr).astype(np.int32)
    return n, rptr, C.astype(np.int32), V.astype(np.float64)


def stencil27(nx, ny=None, nz=None, beta=0.0):
    """27-point stencil: diagonal 26, neighbours -1 + beta*sign(dx)."""
    ny = ny or nx
    nz = nz or nx
    n = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k, j, i = k.ravel(), j.ravel(), i.ravel()
    row = np.arange(n)
    R, C, V = [], [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ok = ((k + dz >= 0) & (k + dz < nz) & (j + dy >= 0) & (j + dy < ny)
                      & (i + dx >= 0) & (i + dx < nx))
                col = ((k + dz) * ny + (j + dy)) * nx + (i + dx)
                v = 26.0 if (dz, dy, dx) == (0, 0, 0) else -1.0 + beta * np.sign(dx)
                R.append(row[ok])
                C.append(col[ok])
                V.append(np.full(ok.sum(), v))
    return _assemble(n, np.concatenate(R), np.concatenate(C), np.concatenate(V))


def stencil7(nx, ny=None, nz=None, diag=6.0, off=-1.0):
    ny = ny or nx
    nz = nz or nx
    n = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k, j, i = k.ravel(), j.ravel(), i.ravel()
    row = np.arange(n)
    R, C, V = [row], [row], [np.full(n, diag)]
    for dz, dy, dx in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        ok = ((k + dz >= 0) & (k + dz < nz) & (j + dy >= 0) & (j + dy < ny)
              & (i + dx >= 0) & (i + dx < nx))
        col = ((k + dz) * ny + (j + dy)) * nx + (i + dx)
        R.append(row[ok])
        C.append(col[ok])
        V.append(np.full(ok.sum(), off))
    return _assemble(n, np.concatenate(R), np.concatenate(C), np.concatenate(V))


def random_sparse(n, per_row, seed, diag_dom=1.5, symmetric_pattern=False):
    """Random row lengths in [1, per_row], diagonally dominant, sorted columns."""
    rng = np.random.default_rng(seed)
    R, C, V = [], [], []
    for r in range(n):
        k = int(rng.integers(0, per_row))
        cols = rng.choice(n, size=min(k, n), replace=False)
        cols = cols[cols != r]
        vals = rng.uniform(-1, 1, cols.size)
        R += [r] * (cols.size + 1)
        C += list(cols) + [r]
        V += list(vals) + [diag_dom * (np.abs(vals).sum() + 1.0)]
    return _assemble(n, np.array(R), np.array(C), np.array(V))


def from_desc(d):
    """(n, rptr, idx, val, b) of a problem described by a dict (the step-count
    fixtures of tests/golden/f3r_ref_steps.npz): unscaled matrix, and b as the
    reference driver receives it (it scales the matrix, not b).  The 7-point
    lognormal matrices come from the library's generator."""
    if d["kind"] == "stencil27":
        n, rp, ix, v = stencil27(d["nx"], beta=d["beta"])
    else:
        import f3r_amd as F
        A = F.CSR.convdiff7(d["nx"], sigma=d["sigma"], corr=d["corr"], peclet=d["peclet"], seed=d["gseed"])
        n, rp, ix, v = A.n, A.rptr, A.idx, A.val
    if d["b"] == "rowsum":
        b = np.add.reduceat(v, rp[:-1])  # b = A * 1
    else:
        b = np.random.default_rng(d["seed"]).uniform(-1, 1, n)
    return n, rp, ix, v, b


def csr_sha256(rp, ix, v):
    import hashlib
    h = hashlib.sha256()
    for a in (np.asarray(rp, np.int32), np.asarray(ix, np.int32), np.asarray(v, np.float64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


def to_dense(n, rptr, idx, val):
    D = np.zeros((n, n))
    for r in range(n):
        for k in range(rptr[r], rptr[r + 1]):
            D[r, idx[k]] += val[k]
    return D


def write_csr_bin(path, n, rptr, idx, val):
    with open(path, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.int64(len(idx)).tobytes())
        f.write(np.asarray(rptr, np.int32).tobytes())
        f.write(np.asarray(idx, np.int32).tobytes())
        f.write(np.asarray(val, np.float64).tobytes())


def read_csr_bin(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.int32)[0])
        nnz = int(np.frombuffer(f.read(8), np.int64)[0])
        rptr = np.frombuffer(f.read(4 * (n + 1)), np.int32).copy()
        idx = np.frombuffer(f.read(4 * nnz), np.int32).copy()
        val = np.frombuffer(f.read(8 * nnz), np.float64).copy()
    return n, rptr, idx, val


def write_vec_bin(path, v):
    v = np.ascontiguousarray(v)
    with open(path, "wb") as f:
        f.write(np.int64(v.size).tobytes())
        f.write(v.tobytes())


def read_vec_bin(path, dtype):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        return np.frombuffer(f.read(), dtype)[:n].copy()
