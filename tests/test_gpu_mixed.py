"""Mixed precision (complex64 series terms, FP32 matrix and accumulation, complex128 state) against
the FP64 oracle: the maximum absolute site-density error over every 10th step and the last must
stay at or below 1e-4 (the accuracy bar of the batched/learned forms); the norm stays within 1e-6
of the FP64 value."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from holstein_cases import CASES, gpu_setup, gpu_step, oracle_run  # noqa: E402

pytestmark = pytest.mark.gpu

DENSITY_TOL = 1e-4


@pytest.mark.parametrize("case,steps", [("chain16", 60), ("ring64_small", 40)])
def test_mixed_precision_density_error(case, steps):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = dict(CASES[case])
    c["steps"] = steps
    oham, recs, t = oracle_run(c)
    ham, te = gpu_setup(c)
    te.expm_params.precision = "mixed"
    worst = 0.0
    for i, rec in enumerate(recs):
        gpu_step(te, i, t)
        if i % 10 == 0 or i == len(recs) - 1:
            err = np.max(np.abs(te.obsobj.last_results["n_exc"] - rec.obs["n_exc"]))
            worst = max(worst, err)
            assert err <= DENSITY_TOL, (i, err)
            nrm = te._reducer.norm(te.vector)
            assert abs(nrm - rec.obs["norm"]) <= 1e-6, (i, nrm, rec.obs["norm"])
    print(f"[{case}] max density error {worst:.3e} over {steps} steps")
    # the series stopped early: FP32 terms need fewer of them than the FP64 test
    assert te.last_expm_dbg[1] < recs[-1].expm_dbg[1]
