"""Full adaptive-basis timelines on the GPU against the oracle on the same inputs.

Bar: basis rows, selected rows and every term's CSR (indptr, indices, data) bit-exact at every
step; amplitudes, norm, n_exc, n_pho and energy within 1e-10 (the FP64 parity order makes the
amplitudes bit-exact in practice); the Taylor stopping index identical.
"""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from holstein_cases import CASES, gpu_setup, gpu_step, oracle_run  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("case", ["chain16", "ring64_small", "torus6", "square4_hop2"])
def test_timeline_parity(torch_mod, case):
    oham, recs, t = oracle_run(case)
    ham, te = gpu_setup(case)
    for i, rec in enumerate(recs):
        print(f"[{case}] step {i}", flush=True)
        gpu_step(te, i, t)
        rows = te.whoami.cpu().numpy().view(np.uint64)
        assert np.array_equal(te.last_select.cpu().numpy().view(np.uint64), rec.select), i
        assert np.array_equal(rows, rec.whoami), i
        for name in oham.offdiag_terms:
            ip, ix, dv = rec.csrs[name]
            m = te.sparse_mats_dict[name]
            assert np.array_equal(m.indptr.cpu().numpy(), ip), (i, name)
            assert np.array_equal(m.indices.cpu().numpy(), ix), (i, name)
            assert np.array_equal(m.data.cpu().numpy(), dv), (i, name)
        vec = te.vector.cpu().numpy()
        assert np.max(np.abs(vec - rec.vec), initial=0) <= TOL, i
        res = te.obsobj.last_results
        assert np.max(np.abs(res["n_exc"] - rec.obs["n_exc"])) <= TOL, i
        assert np.max(np.abs(res["n_pho"] - rec.obs["n_pho"])) <= TOL, i
        if i > 0:
            assert te.last_expm_dbg[:2] == rec.expm_dbg[:2], i
        assert abs(res["total_energy"].real - np.vdot(rec.vec, _oracle_h(rec)).real) <= TOL


def _oracle_h(rec):
    from oracle import paces_oracle as po
    return po.total_ham(rec.diag, list(rec.csrs.values()), rec.vec)


def test_lanczos_matches_oracle(torch_mod):
    from oracle import paces_oracle as po
    case = dict(CASES["chain16"])
    case["steps"] = 3
    oham, recs, t = oracle_run(case)
    ham, te = gpu_setup(case)
    for i in range(len(recs)):
        gpu_step(te, i, t)
    te.expm_params.lanczos_order = 14
    # propagate the oracle's adapted vector at the last step with Lanczos on both sides
    rec = recs[-1]
    exp = po.lanczos(rec.diag, [rec.csrs[n] for n in oham.offdiag_terms], rec.vec_adapted,
                     float(t[-1] - t[-2]), 14)
    import torch
    te.vector = torch.from_numpy(rec.vec_adapted).to(te.device)
    got, _ = te._lanczos(float(t[-1] - t[-2]))
    assert np.max(np.abs(got.cpu().numpy() - exp)) <= TOL
    # and Lanczos agrees with the exact Taylor result to its truncation error
    assert np.max(np.abs(got.cpu().numpy() - rec.vec)) <= 1e-8
