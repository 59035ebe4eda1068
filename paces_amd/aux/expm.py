"""Scaling-and-squaring action of the matrix exponential (paces/aux/_scipy/_expm_multiply.py:
106-202, 395-450) and the block 1-norm estimator it calls (paces/aux/_scipy/_onenormest.py), for
A = -i H with H the real symmetric Hamiltonian held as a diagonal plus per-term CSR matrices.

The method (Al-Mohy and Higham 2011, algorithm 3.2): shift A by mu = trace(A)/n, choose the
number of Taylor terms m* and scaling steps s from ||t A||_1 and estimates of ||(t A)^p||_1
(code fragment 3.1 with condition 3.13), then apply s times a truncated Taylor series of
exp(t A / s) that stops when the last two terms are negligible, and multiply by exp(t mu / s)
after each. The block estimator is Higham and Tisseur's algorithm 2.4 with two columns.

Everything over vectors runs in libpaces_hip.so: the series on the Taylor-term kernel with the
shifted diagonal, the estimator's products on the H.x kernel with its two real columns held as
one complex vector, and its reductions, signs and index searches on the pair kernels. The host
handles only scalars and the estimator's few indices.

Divergences from the reference: the reference's step calls this with 1j*H (time_evolution.py:
580-581), which evolves with exp(+i dt H) while its Taylor and Lanczos paths use exp(-i dt H);
here A = -i H, the direction of the other paths. Its estimator draws random columns from CuPy's
generator; here they come from a hash of a seed and a draw counter (the estimate is a lower
bound of the same norm either way, but the draws differ).
"""

from __future__ import annotations

import cmath
import ctypes as C
import math

import torch

from .. import _hip
from .._hip import check, ptr, stream
from ..core import engine

_lib = _hip.load()

TOL = 2.0 ** -53

# theta_m: the largest ||t A||_1 for which m Taylor terms give a backward error below 2^-53 in
# double precision. Published values: Higham, Functions of Matrices (2008), Table A.3 (m <= 30),
# and Al-Mohy and Higham (2011), Table 3.1 (m = 35 .. 55).
THETA = {
    1: 2.29e-16, 2: 2.58e-8, 3: 1.39e-5, 4: 3.40e-4, 5: 2.40e-3, 6: 9.07e-3, 7: 2.38e-2,
    8: 5.00e-2, 9: 8.96e-2, 10: 1.44e-1, 11: 2.14e-1, 12: 3.00e-1, 13: 4.00e-1, 14: 5.14e-1,
    15: 6.41e-1, 16: 7.81e-1, 17: 9.31e-1, 18: 1.09, 19: 1.26, 20: 1.44, 21: 1.62, 22: 1.82,
    23: 2.01, 24: 2.22, 25: 2.43, 26: 2.64, 27: 2.86, 28: 3.08, 29: 3.31, 30: 3.54, 35: 4.7,
    40: 6.0, 45: 7.2, 50: 8.5, 55: 9.9,
}
M_MAX = 55
ELL = 2


class HamiltonianOperator:
    """A = scale * H with H = diag(diag) + the sum of the term CSRs (real, symmetric)."""

    def __init__(self, diag, mats, scale=-1j):
        self.diag = diag
        self.mats = list(mats)
        self.scale = complex(scale)
        n = int(diag.shape[0])
        self.shape = (n, n)
        self._ws = _hip.Workspace(diag.device)

    def dot(self, x):
        y = engine.spmv(self.diag, self.mats, x)
        return engine.axpby(self.scale, y, 0.0, y)

    def diag_mean(self):
        """trace(H) / n (the terms have no diagonal entries)."""
        n = self.shape[0]
        out = torch.empty(1, dtype=torch.float64, device=self.diag.device)
        wsp, wsn = self._ws.get(_lib.paces_pair_workspace_bytes())
        check(_lib.paces_sum_f64(ptr(self.diag), n, ptr(out), wsp, wsn, stream()), "sum_f64")
        return float(out.item()) / float(n)

    def shifted_onenorm(self, m):
        """||H - m I||_1 (exact; H is symmetric, so the largest row sum of magnitudes)."""
        out = torch.empty(1, dtype=torch.float64, device=self.diag.device)
        wsp, wsn = self._ws.get(_lib.paces_pair_workspace_bytes())
        s = engine.csr_set(self.mats)
        check(_lib.paces_csr_shifted_onenorm(self.shape[0], ptr(self.diag), m, C.byref(s),
                                             ptr(out), wsp, wsn, stream()),
              "csr_shifted_onenorm")
        return float(out.item())


# ---- fragment 3.1 and condition 3.13 (_expm_multiply.py:376-460) ---------------------------------

def p_max(m_max):
    """The largest p with p(p - 1) <= m_max + 1."""
    p = 1
    while (p + 1) * p <= m_max + 1:
        p += 1
    return p


def condition_3_13(norm1, n0, m_max=M_MAX, ell=ELL):
    """True when ||t A||_1 is small enough that the estimates of higher powers are not needed:
    ||t A||_1 <= 2 ell p_max (p_max + 3) theta_(m_max) / (n0 m_max)."""
    pm = p_max(m_max)
    return norm1 <= (2 * ell * pm * (pm + 3)) * (THETA[m_max] / float(n0 * m_max))


def fragment_3_1(norm1, d, n0=1, m_max=M_MAX, ell=ELL):
    """(m*, s) minimising m * s, the number of H.x products (code fragment 3.1). norm1 is
    ||t A||_1; d(p) estimates ||(t A)^p||_1^(1/p) and is called only when condition 3.13 fails.
    Ties keep the first candidate in the order the fragment lists them."""
    best = None
    if condition_3_13(norm1, n0, m_max, ell):
        for m, theta in THETA.items():
            s = int(math.ceil(norm1 / theta))
            if best is None or m * s < best[0] * best[1]:
                best = (m, s)
        return best
    for p in range(2, p_max(m_max) + 1):
        alpha = max(d(p), d(p + 1))
        for m in range(p * (p - 1) - 1, m_max + 1):
            if m in THETA:
                s = int(math.ceil(alpha / THETA[m]))
                if best is None or m * s < best[0] * best[1]:
                    best = (m, s)
    return best[0], max(best[1], 1)


# ---- block 1-norm estimator (_onenormest.py:202-345) ---------------------------------------------

class _Pairs:
    """Pair-vector operations of the estimator on the device."""

    def __init__(self, n, device):
        self.n = n
        self.device = device
        self.ws = _hip.Workspace(device)
        self.out = torch.empty(4, dtype=torch.float64, device=device)
        self.idx = torch.empty(1, dtype=torch.int64, device=device)

    def new(self):
        return torch.empty(self.n, dtype=torch.complex128, device=self.device)

    def fill(self, x, re, im, i0=-1, i1=-1):
        check(_lib.paces_pair_fill(ptr(torch.view_as_real(x)), self.n, re, im, i0, i1, stream()),
              "pair_fill")

    def random_column(self, x, col, seed, scale=1.0):
        check(_lib.paces_pair_random_column(ptr(torch.view_as_real(x)), self.n, col,
                                             seed & ((1 << 64) - 1), scale, stream()),
              "pair_random_column")

    def abs_sums(self, y):
        wsp, wsn = self.ws.get(_lib.paces_pair_workspace_bytes())
        check(_lib.paces_pair_abs_sums(ptr(torch.view_as_real(y)), self.n, ptr(self.out), wsp,
                                       wsn, stream()), "pair_abs_sums")
        return self.out[:2].tolist()

    def dots(self, a, b):
        wsp, wsn = self.ws.get(_lib.paces_pair_workspace_bytes())
        check(_lib.paces_pair_dots(ptr(torch.view_as_real(a)), ptr(torch.view_as_real(b)),
                                   self.n, ptr(self.out), wsp, wsn, stream()), "pair_dots")
        return self.out.tolist()

    def signs(self, y):
        s = self.new()
        check(_lib.paces_pair_signs(ptr(torch.view_as_real(y)), self.n,
                                    ptr(torch.view_as_real(s)), stream()), "pair_signs")
        return s

    def rowmax(self, z):
        h = torch.empty(self.n, dtype=torch.float64, device=self.device)
        check(_lib.paces_pair_rowmax(ptr(torch.view_as_real(z)), self.n, ptr(h), stream()),
              "pair_rowmax")
        return h

    def largest(self, h, k):
        """The first k indices of h by decreasing value, then decreasing index."""
        wsp, wsn = self.ws.get(_lib.paces_pair_workspace_bytes())
        out, v, i = [], 0.0, -1
        for _ in range(min(k, self.n)):
            check(_lib.paces_argmax_after(ptr(h), self.n, v, i, ptr(self.idx), wsp, wsn,
                                          stream()), "argmax_after")
            i = int(self.idx.item())
            if i < 0:
                break
            v = float(h[i].item())
            out.append(i)
        return out


def onenormest_symmetric(apply, n, device, itmax=5, seed=0):
    """A lower bound of ||M||_1 for a real symmetric M given as apply(x) = M x on a pair (two
    real columns in one complex vector): Higham and Tisseur's algorithm 2.4 with t = 2 columns
    (_onenormest.py:202-345), where M's transpose is M. For n <= 2 the norm is exact."""
    pv = _Pairs(n, device)
    t = 2
    if n <= t:
        x = pv.new()
        pv.fill(x, 0.0, 0.0, 0, 1 if n > 1 else -1)
        sums = pv.abs_sums(apply(x))
        return max(sums[:n])
    draws = [0]

    def resample(x, col):
        draws[0] += 1
        pv.random_column(x, col, hash((seed, draws[0])))

    def parallel(d):
        return d == float(n)  # columns of +-1 entries are parallel when their product is n

    # starting matrix: a column of ones and a random +-1 column not parallel to it, scaled to
    # unit 1-norm columns
    xu = pv.new()
    pv.fill(xu, 1.0, 0.0)
    resample(xu, 1)
    while parallel(pv.dots(xu, xu)[1]):
        resample(xu, 1)
    x = engine.axpby(1.0 / n, xu, 0.0, pv.new())
    hist = []
    est_old = 0.0
    s_cur = pv.new()
    pv.fill(s_cur, 0.0, 0.0)
    k = 1
    ind = None
    ind_best = None
    while True:
        y = apply(x)
        mags = pv.abs_sums(y)
        best_j = 0 if mags[0] >= mags[1] else 1
        est = mags[best_j]
        if (est > est_old or k == 2) and k >= 2:
            ind_best = ind[best_j]
        if k >= 2 and est <= est_old:
            return est_old
        est_old = est
        s_old = s_cur
        if k > itmax:
            return est
        s_cur = pv.signs(y)
        d = pv.dots(s_cur, s_old)  # s0.o0, s0.o1, s1.o0, s1.o1
        if (parallel(d[0]) or parallel(d[1])) and (parallel(d[2]) or parallel(d[3])):
            return est
        # no column of S parallel to an earlier one or to a column of S_old
        for col in range(t):
            while True:
                d = pv.dots(s_cur, s_old)
                need = parallel(d[2 * col]) or parallel(d[2 * col + 1])
                if col == 1:
                    need = need or parallel(pv.dots(s_cur, s_cur)[1])
                if not need:
                    break
                resample(s_cur, col)
        z = apply(s_cur)
        h = pv.rowmax(z)
        top = pv.largest(h, t + len(hist))
        if k >= 2 and float(h[top[0]].item()) == float(h[ind_best].item()):
            return est
        if all(i in hist for i in top[:t]):
            return est
        ind = [i for i in top if i not in hist] + [i for i in top if i in hist]
        pv.fill(x, 0.0, 0.0, ind[0], ind[1])
        hist += [i for i in ind[:t] if i not in hist]
        k += 1


class _NormInfo:
    """||t A'||_1 and lazily estimated d_p = ||(t A')^p||_1^(1/p) for A' = -i (H - m I)."""

    def __init__(self, op, m, t, norm1, seed):
        self.op = op
        self.t = t
        self.norm1 = norm1
        self.seed = seed
        self._d = {}
        self._dshift = torch.empty_like(op.diag)
        check(_lib.paces_shift_f64(ptr(op.diag), op.shape[0], m, ptr(self._dshift), stream()),
              "shift_f64")

    def d(self, p):
        if p not in self._d:
            n = self.op.shape[0]

            def apply(x):
                for _ in range(p):
                    x = engine.spmv(self._dshift, self.op.mats, x)
                return x

            est = onenormest_symmetric(apply, n, self.op.diag.device, seed=(self.seed, p))
            self._d[p] = self.t * est ** (1.0 / p)
        return self._d[p]


def expm_multiply_simple(A, B, t=1.0, balance=False, return_dbg=False, seed=0):
    """exp(t A) B for A = -i H (a HamiltonianOperator), B a complex128 vector
    (_expm_multiply.py:106-202). With return_dbg, also [converged, the total number of Taylor
    terms, the last two terms' inf-norm sum, its ratio to the result's inf-norm]."""
    if balance:
        raise NotImplementedError("balancing is not implemented (nor in the reference)")
    if not isinstance(A, HamiltonianOperator) or A.scale != -1j:
        raise NotImplementedError("A must be a HamiltonianOperator with scale -1j (A = -i H)")
    n = A.shape[0]
    if B.ndim != 1 or B.shape[0] != n:
        raise ValueError("B must be a vector of the operator's size")
    m = A.diag_mean()  # mu = trace(A) / n = -i m
    norm1 = A.shifted_onenorm(m)  # ||A - mu I||_1, |scale| = 1
    info = _NormInfo(A, m, t, t * norm1, seed)
    if t * norm1 == 0:
        m_star, s = 0, 1
    else:
        m_star, s = fragment_3_1(t * norm1, info.d)
    eta = cmath.exp(t * (-1j * m) / s)
    F = B.clone()
    dbg = [False, 0, 0.0, 0.0]
    for _ in range(s):
        if m_star > 0:
            F, d = engine.simple_taylor(info._dshift, A.mats, F, None, m_star, TOL,
                                        coeff=lambda j: t / float(s * (j + 1)))
            # converged once any scaling step's series met the tolerance (as the reference
            # keeps its flag); the terms are counted over all steps
            dbg = [dbg[0] or d[0], dbg[1] + d[1] + 1, d[2], d[3]]
        if eta != 1:
            engine.axpby(eta, F, 0.0, F)
    A.last_choice = (m_star, s)
    return (F, dbg) if return_dbg else F


def expm_multiply(A, B, start=None, stop=None, num=None, endpoint=None, return_dbg=False):
    """exp(A) B at a single time (_expm_multiply.py:58-103); time grids are not supported."""
    if any(a is not None for a in (start, stop, num, endpoint)):
        raise NotImplementedError("expm_multiply over a time grid is not implemented")
    return expm_multiply_simple(A, B, return_dbg=return_dbg)
