"""The batched step's incremental adaptation against the full construction.

After the first step, BatchedTimeEvolution finds each basis by marking the previous basis
through the previous slot matrix, with explicit images only for the marked rows outside the
previous expanded basis. It must give the full construction's rows, slot matrix (each term read
as a CSR), diagonal and carried-over amplitudes bit for bit, so the whole run is bit-identical:
checked at every step on draws of the 256-chain workload's distribution (3-word rows), with the
series replayed from graphs, in FP64 and with complex64 terms. With incremental_max = 1 the
incremental path also runs through the growth from one state, where most marked rows lie outside
the previous expanded basis and many rows are new: the explicit images, the new rows' merge and
the explicit pairs are exercised there.
"""

import numpy as np
import pytest

from test_gpu_workloads import DIM, _gpu, chain_draws

pytestmark = pytest.mark.gpu


def _engine(draws, precision, incremental, incremental_max=0.05):
    from paces_amd.models import holstein
    bham = holstein.BatchedHamiltonian([d[0] for d in draws], DIM, [-d[1] for d in draws],
                                       [d[2] for d in draws])
    bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                        precision=precision, graph=True,
                                        incremental=incremental, incremental_max=incremental_max)
    bte.create_single_exciton_states([d[0] // 2 for d in draws])
    return bte


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("incremental_max", [0.05, 1.0])
def test_incremental_adaptation_equals_full(precision, incremental_max):
    _gpu()
    import torch
    draws = chain_draws(16, seed=5, ms_hi=3000, ms_lo=300)
    full = _engine(draws, precision, False)
    inc = _engine(draws, precision, True, incremental_max)
    assert inc.incremental and not full.incremental
    kinds = []
    for i in range(16):
        print(f"[incremental {precision} {incremental_max}] step {i}", flush=True)
        full.step(i, 0.05 if i else 0.0)
        inc.step(i, 0.05 if i else 0.0)
        kinds.append(inc.last_adapt)
        assert torch.equal(full.whoami, inc.whoami), i
        assert torch.equal(full.seg_off, inc.seg_off), i
        assert torch.equal(full.diag_vals, inc.diag_vals), i
        assert torch.equal(full.vector, inc.vector), i
        if i > 0:
            assert full.last_expm.tobytes() == inc.last_expm.tobytes(), i
        for name in full.hamobj.offdiag_terms:
            a, b = full.sparse_mats_dict[name], inc.sparse_mats_dict[name]
            assert torch.equal(a.indptr, b.indptr) and torch.equal(a.indices, b.indices), (i, name)
            assert (a.code is None) == (b.code is None), (i, name)
            if a.code is not None:
                assert torch.equal(a.code, b.code), (i, name)
        na, pa, ra = full.observables()
        nb, pb, rb = inc.observables()
        assert np.array_equal(na, nb) and np.array_equal(pa, pb) and np.array_equal(ra, rb), i
    incremental = [k for k in kinds if isinstance(k, tuple)]
    assert incremental, kinds
    print(f"[incremental {precision} {incremental_max}] {kinds}")
    if incremental_max >= 1.0:
        # the growth steps: explicit rows and new rows were handled by the incremental path
        assert any(k[1] + k[2] > 0 for k in incremental) and any(k[3] > 0 for k in incremental)
    else:
        assert kinds[1] == "full"  # the first steps grow from one state: full construction
