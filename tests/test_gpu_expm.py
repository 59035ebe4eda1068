"""The scaling-and-squaring propagator on the GPU (expm_multiply_simple and its block 1-norm
estimator, paces/aux/_scipy/_expm_multiply.py:106-202, 395-450, _onenormest.py) against SciPy:
the result against scipy.sparse.linalg.expm_multiply on the same sparse H, the estimate against
the exact 1-norms of powers of H, and a timeline with ExpmParams(use_scaling=True) against the
oracle's Taylor timeline."""

import numpy as np
import pytest

from holstein_cases import CASES, gpu_setup, gpu_step, oracle_run

pytestmark = pytest.mark.gpu


SMALL = dict(N=6, dim=3, geo=1, periodic=False, terms=CASES["chain16"]["terms"], pos0=3,
             ms=150, dt=0.05, steps=2, enlarge=1, fillfac=0.2)


def _matrices(steps=4, case="chain16"):
    """A case's step matrices after a few steps on the GPU: (te, diag, mats, H)."""
    import scipy.sparse
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _, te = gpu_setup(case)
    t = 0.05 * np.arange(steps + 1)
    for i in range(steps + 1):
        gpu_step(te, i, t)
    mats = [te.sparse_mats_dict[k] for k in te.hamobj.offdiag_terms]
    H = scipy.sparse.diags(te.diag_vals.cpu().numpy()).tocsr()
    for m in mats:
        H = H + m.to_scipy()
    return te, te.diag_vals, mats, H


@pytest.mark.parametrize("t", [0.05, 0.4, 12.0])
def test_expm_multiply_simple_equals_scipy(t):
    """exp(-i t H) psi equals SciPy's expm_multiply to 1e-12 of the norm; t = 12 takes the
    branch that estimates norms of powers of H (||t A||_1 above condition 3.13's bound)."""
    import scipy.sparse.linalg
    from paces_amd.aux import expm
    te, diag, mats, H = _matrices()
    op = expm.HamiltonianOperator(diag, mats, scale=-1j)
    psi = te.vector
    got, dbg = expm.expm_multiply_simple(op, psi, t=t, return_dbg=True)
    want = scipy.sparse.linalg.expm_multiply(-1j * t * H.astype(np.complex128),
                                             psi.cpu().numpy())
    err = np.max(np.abs(got.cpu().numpy() - want))
    m_star, s = op.last_choice
    norm1 = t * op.shifted_onenorm(op.diag_mean())
    print(f"t={t}: ||tA||_1={norm1:.3g} m*={m_star} s={s} terms={dbg[1]} err={err:.2e}")
    assert err <= 1e-12 * max(1.0, float(np.linalg.norm(want))), err
    assert dbg[0] and 1 <= dbg[1] <= m_star * s
    assert expm.condition_3_13(norm1, 1) == (t < 1.0)


def test_onenormest_bounds_exact_norms():
    """The estimate of ||(H - m I)^p||_1 is a lower bound within a factor 3 of the exact norm of
    the dense power (Higham and Tisseur: 'almost invariably ... correct to within a factor 3'),
    exact on the matrices here; two-row operators are exact."""
    import torch
    from paces_amd.aux import expm
    from paces_amd.core import engine
    te, diag, mats, H = _matrices(steps=2, case=SMALL)
    n = H.shape[0]
    op = expm.HamiltonianOperator(diag, mats)
    m = op.diag_mean()
    assert abs(m - float(np.mean(diag.cpu().numpy()))) <= 1e-12 * max(1.0, abs(m))
    assert n < 2000, n
    Hs = H.toarray() - m * np.eye(n)
    assert abs(op.shifted_onenorm(m) - np.abs(Hs).sum(axis=0).max()) <= 1e-12 * n
    dshift = diag - m
    for p in (1, 2, 3, 5):
        def apply(x, p=p):
            for _ in range(p):
                x = engine.spmv(dshift, mats, x)
            return x
        est = expm.onenormest_symmetric(apply, n, diag.device, seed=p)
        exact = np.abs(np.linalg.matrix_power(Hs, p)).sum(axis=0).max()
        print(f"p={p}: estimate {est:.6g} exact {exact:.6g}")
        assert est <= exact * (1 + 1e-12) and est >= exact / 3.0, (p, est, exact)
    # a 2 x 2 operator: the exact branch
    d2 = torch.tensor([0.5, -1.5], dtype=torch.float64, device=diag.device)
    est2 = expm.onenormest_symmetric(lambda x: engine.spmv(d2, [], x), 2, diag.device)
    assert est2 == 1.5


def test_scaling_timeline_matches_oracle():
    """ExpmParams(use_scaling=True) over the 16-site chain's timeline against the oracle's
    Taylor timeline: norms and site densities within 1e-8. The two propagators agree to about
    1e-16 per step (test above), which is enough to move exact ties at the maxstates cut, so
    the bases may differ by a few rows of negligible weight and are not compared."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.core.time_evolution import ExpmParams
    ham, recs, t = oracle_run("chain16")
    _, te = gpu_setup("chain16")
    te.expm_params = ExpmParams(use_scaling=True)
    worst = 0.0
    for i in range(len(t)):
        gpu_step(te, i, t)
        rec = recs[i]
        err = float(np.max(np.abs(te.obsobj.last_results["n_exc"] - rec.obs["n_exc"])))
        worst = max(worst, err)
        assert err <= 1e-8, (i, err)
        assert abs(te._reducer.norm(te.vector) - rec.obs["norm"]) <= 1e-8, i
        if i:
            assert te.last_expm_dbg[0] and te.last_expm_choice[1] >= 1
    print(f"max site-density difference from the Taylor timeline {worst:.2e}")
