"""Host logic of the scaling-and-squaring propagator against SciPy's implementation of the same
method (scipy.sparse.linalg._expm_multiply, Al-Mohy and Higham 2011): the theta table, p_max,
condition 3.13 and the choice of (m*, s) in code fragment 3.1, for given norms."""

import math

import numpy as np
import pytest

sp = pytest.importorskip("scipy.sparse.linalg._expm_multiply")


def _expm():
    from paces_amd.aux import expm
    return expm


class _Norms:
    """Stands in for SciPy's LazyOperatorNormInfo with fixed norms."""

    def __init__(self, norm1, d):
        self.norm1, self.dfun = norm1, d

    def onenorm(self):
        return self.norm1

    def d(self, p):
        return self.dfun(p)

    def alpha(self, p):
        return max(self.d(p), self.d(p + 1))


def test_theta_table_and_p_max_equal_scipy():
    e = _expm()
    assert e.THETA == sp._theta
    for m_max in (5, 10, 30, 55, 56, 100):
        assert e.p_max(m_max) == sp._compute_p_max(m_max)


@pytest.mark.parametrize("norm1", [1e-3, 0.05, 0.7, 3.0, 12.5, 63.0, 63.4, 64.0, 150.0, 2e3])
@pytest.mark.parametrize("decay", [1.0, 0.6, 0.25])
def test_fragment_3_1_equals_scipy(norm1, decay):
    """(m*, s) equal SciPy's for the same ||tA||_1 and estimates d_p = ||tA|| decay^(p-1)."""
    e = _expm()

    def d(p):
        return norm1 * decay ** (p - 1)

    assert e.condition_3_13(norm1, 1) == sp._condition_3_13(norm1, 1, 55, 2)
    got = e.fragment_3_1(norm1, d)
    want = sp._fragment_3_1(_Norms(norm1, d), 1, 2.0 ** -53)
    assert got == want, (norm1, decay, got, want)
    m, s = got
    assert s >= 1 and m in e.THETA and m * s > 0
    assert math.isfinite(m * s)
