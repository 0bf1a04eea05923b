"""Level sets of the device factorisations (host logic): every row comes
after each row it waits for, and at the smallest such level; the level
counts of the stencil orders follow their closed forms."""
import numpy as np
import pytest

import problems


def _check_levels(F, n, rptr, idx, val, both):
    A = F.CSR(n, rptr, idx, val)
    lev, nl = A.factor_levels(both)
    deps = [set() for _ in range(n)]
    for i in range(n):
        for j in idx[rptr[i]:rptr[i + 1]]:
            if j < i:
                deps[i].add(int(j))
            elif j > i and both:
                deps[int(j)].add(i)
    for i in range(n):
        want = 1 + max(lev[j] for j in deps[i]) if deps[i] else 0
        assert lev[i] == want, i
    assert nl == lev.max() + 1
    return nl


@pytest.mark.parametrize("both", [True, False])
def test_levels_random_nonsymmetric(both):
    import f3r_amd as F
    n, rptr, idx, val = problems.random_sparse(300, 6, seed=11)
    nl = _check_levels(F, n, rptr, idx, val, both)
    assert 1 < nl <= n


def test_levels_symmetric_pattern_agree():
    import f3r_amd as F
    n, rptr, idx, val = problems.stencil7(5, 4, 3)
    A = F.CSR(n, rptr, idx, val)
    a, na = A.factor_levels(True)
    b, nb = A.factor_levels(False)
    assert na == nb and np.array_equal(a, b)


@pytest.mark.parametrize("m", [4, 7])
def test_levels_stencil_closed_forms(m):
    # lexicographic order: level = x + y + z (7-point), x + 2y + 4z (27-point)
    import f3r_amd as F
    n, rptr, idx, val = problems.stencil7(m)
    assert _check_levels(F, n, rptr, idx, val, True) == 3 * (m - 1) + 1
    n, rptr, idx, val = problems.stencil27(m, beta=0.5)
    assert _check_levels(F, n, rptr, idx, val, True) == 7 * (m - 1) + 1
