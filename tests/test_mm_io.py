"""MatrixMarket reader of the library (f3r_mm_read, f3r_mm_read_vector)
against the reference's own reader (io::mm::read_system + io::CSR,
sparse/matrix/io.hpp:42-64,104-217) on the small files of tests/golden/mm/;
the expected arrays were written by the reference (tests/golden/make_golden.py mm)."""
import ctypes
import os

import numpy as np
import pytest

import f3r_amd as F
from f3r_amd._capi import lib

MM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mm")
FILES = ("general", "symmetric", "zeros", "skew", "integer", "rect")


@pytest.mark.parametrize("name", FILES)
def test_reader_matches_reference(golden, name):
    g = golden("mm_ref")
    A, b = F.io.mm.read_system(os.path.join(MM, name + ".mtx"))
    assert [A.n, A.ncols] == g[f"{name}_shape"].tolist()
    assert np.array_equal(A.rptr, g[f"{name}_rptr"])
    assert np.array_equal(A.idx, g[f"{name}_idx"])
    assert np.array_equal(A.val.view(np.uint64), g[f"{name}_val"].view(np.uint64))
    # right-hand side: general_b.mtx, else U[0,1) of std::mt19937_64(0)
    assert np.array_equal(b, g[f"{name}_b"])


def test_skew_symmetric_is_not_mirrored(golden):
    A = F.io.mm.read_matrix(os.path.join(MM, "skew.mtx"))
    # only the three stored entries, none mirrored and none negated
    assert A.nnz == 3
    D = np.zeros((4, 4))
    for r in range(4):
        D[r, A.idx[A.rptr[r]:A.rptr[r + 1]]] = A.val[A.rptr[r]:A.rptr[r + 1]]
    assert np.count_nonzero(np.triu(D)) == 0


def test_zero_entries_dropped_off_the_diagonal_only():
    A = F.io.mm.read_matrix(os.path.join(MM, "zeros.mtx"))
    rows = np.repeat(np.arange(A.n), np.diff(A.rptr))
    zero = A.val == 0.0
    assert np.all(rows[zero] == A.idx[zero]) and zero.sum() == 1


def _status(path, fn="read_matrix"):
    with pytest.raises(F.F3RError) as e:
        getattr(F.io.mm, fn)(os.path.join(MM, path))
    return e.value.status


def test_reader_errors():
    assert _status("missing.mtx") == -1
    assert _status("bad_index.mtx") == -1      # row 4 of a 3 x 3 matrix
    assert _status("pattern.mtx") == -4        # no values: unsupported
    assert _status("bad_b.mtx", "read_system") == -1  # two-column right-hand side
    assert _status("general.mtx", "read_vector") == -1  # a coordinate file is not a vector
    h = ctypes.c_void_p()
    assert lib.f3r_mm_read(None, None, ctypes.byref(h)) == -1


def test_reader_refuses_repeats_and_overlong_headers(tmp_path):
    # a repeated (row, column) entry, and headers that declare more entries or
    # rows than the file holds, are errors before anything large is allocated
    cases = {
        "repeat.mtx": ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 2.0\n1 1 3.0\n",
                       "read_matrix", -4),
        "long.mtx": ("%%MatrixMarket matrix coordinate real general\n2 2 1000000000000\n1 1 1.0\n",
                     "read_matrix", -1),
        "long_b.mtx": ("%%MatrixMarket matrix array real general\n2000000000 1\n1.0\n", "read_vector", -1),
    }
    for name, (text, fn, st) in cases.items():
        path = tmp_path / name
        path.write_text(text)
        with pytest.raises(F.F3RError) as e:
            getattr(F.io.mm, fn)(str(path))
        assert e.value.status == st, (name, e.value.status)


def test_rhs_path_and_fallback():
    buf = ctypes.create_string_buffer(64)
    assert lib.f3r_mm_rhs_path(b"../matrix/hpcg_5_5_5.mtx", buf, 64) == 0
    assert buf.value == b"../matrix/hpcg_5_5_5_b.mtx"
    assert lib.f3r_mm_rhs_path(b"../matrix/hpcg_5_5_5.mtx", buf, 8) == -1
    # the fallback generator is std::uniform_real_distribution over mt19937_64
    b = np.empty(5)
    assert lib.f3r_gen_uniform_mt(5, 0.0, 1.0, 0, b.ctypes.data_as(ctypes.c_void_p)) == 0
    assert np.all((b >= 0) & (b < 1)) and len(set(b.tolist())) == 5


def test_reader_round_trip_of_a_generated_stencil(tmp_path):
    # a 27-point matrix written as a symmetric file (lower triangle) reads back
    # as the full matrix
    A = F.CSR.stencil27(5)
    rows = np.repeat(np.arange(A.n), np.diff(A.rptr))
    low = A.idx <= rows
    p = tmp_path / "s27.mtx"
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{A.n} {A.n} {int(low.sum())}\n")
        for r, c, v in zip(rows[low], A.idx[low], A.val[low]):
            f.write(f"{r + 1} {c + 1} {float(v)!r}\n")
    B = F.io.mm.read_matrix(p)
    assert np.array_equal(A.rptr, B.rptr) and np.array_equal(A.idx, B.idx) and np.array_equal(A.val, B.val)
