"""Full-size properties of the plain step (the 64-site ring at maxstates=2e6, ~8e6 basis rows,
five uint64 words per row), where running the oracle for every step is too slow.

Checked after several steps: the basis is sorted and unique; every term's CSR is canonical and
exactly symmetric with only the values the model allows (J for hops, g*sqrt(n) for vib pairs);
<H> is real to rounding; the norm decreases only through truncation (never grows).
"""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))

pytestmark = pytest.mark.gpu


def _sorted_unique(rows):
    a, b = rows[:-1], rows[1:]
    less = np.zeros(a.shape[0], dtype=bool)
    eq = np.ones(a.shape[0], dtype=bool)
    for k in range(rows.shape[1]):
        less |= eq & (a[:, k] < b[:, k])
        eq &= a[:, k] == b[:, k]
    return bool(np.all(less))


def test_ring64_full_size_properties():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import scipy.sparse as sp
    from holstein_cases import terms
    from paces_amd.config import devices
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    J, g = 1.0, 1.5
    ham = holstein.Hamiltonian(64, [16] * 64, periodic=True, use_terms=terms(J, g))
    te = holstein.TimeEvolution(ham, holstein.Observables, ["n_exc", "n_pho", "total_energy"],
                                None, TimeEvoParams(maxstates=2_000_000, enlarge_steps=1,
                                                    shuffle_seed=None), ExpmParams())
    te.create_nonuniform_basis([1] * 64, minpos=32, maxpos=33)
    te.grow_optimal_basis(fillfac=0.1)
    raw = np.zeros((1, 65), dtype=np.int64)
    raw[0, 0] = 32
    te.create_initial_vector([1.0], raw)
    te._perform_first_step(0.0)
    norms = [1.0]
    for i in range(1, 4):
        print(f"[ring64] step {i}", flush=True)
        te._evolve_single_step(i, 0.02 * i, 0.02)
        norms.append(te._reducer.norm(te.vector))
    R = te.numstates
    assert R > 7_000_000
    rows = te.whoami.cpu().numpy().view(np.uint64)
    assert _sorted_unique(rows)
    allowed_vib = set((np.sqrt(np.arange(1, 16)) * g).tolist())
    for name, m in te.sparse_mats_dict.items():
        A = m.to_scipy()
        assert A.has_canonical_format
        D = (A - A.T).tocsr()
        D.eliminate_zeros()
        assert D.nnz == 0, (name, D.nnz)
        vals = set(np.unique(A.data).tolist())
        if name == "hopping":
            assert vals == {-J}
        else:
            assert vals <= allowed_vib
    assert all(norms[k + 1] <= norms[k] * (1 + 1e-12) for k in range(len(norms) - 1))
    e = te.obsobj.last_results["total_energy"]
    assert abs(e.imag) < 1e-12
