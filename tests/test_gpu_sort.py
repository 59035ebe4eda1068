"""General sort + dedup of packed rows (paces_sort_rows, LSD radix) against NumPy, through the
C ABI: bit-exact rows, the permutation maps every output row to its source, equal rows keep
their input order (stability), and the generic plugin-based expansion equals the Holstein
union-of-runs expansion bit for bit."""

import numpy as np
import pytest

from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from paces_amd.config import devices
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not devices.configure_called:
        devices.configure()
    from paces_amd import _hip
    from paces_amd.core import engine
    return dict(torch=torch, dev=devices.vector_dev, hip=_hip, engine=engine)


@pytest.mark.parametrize("W", [1, 2, 5, 13])
@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 100003])
def test_sort_unique_matches_numpy(env, W, n):
    hip, eng = env["hip"], env["engine"]
    rng = np.random.default_rng(W * 1000 + n)
    for hi in (5, None):
        if hi is None:
            a = rng.integers(0, 2 ** 63, size=(n, W), dtype=np.int64).astype(np.uint64)
            a[:, 0] |= np.uint64(1) << np.uint64(63) * (rng.integers(0, 2, n).astype(np.uint64))
            if n > 10:
                a[n // 2:n // 2 + 5] = a[:5]  # duplicates of full-width rows
        else:
            a = rng.integers(0, hi, size=(n, W)).astype(np.uint64)
        ws = hip.Workspace(env["dev"])
        d = hip.rows_from_numpy(a, env["dev"])
        u, pu = eng.sort_rows(d, ws, unique=True, want_perm=True)
        np.testing.assert_array_equal(hip.rows_to_numpy(u), po.unique_rows(a))
        if n:
            np.testing.assert_array_equal(a[pu.cpu().numpy()], hip.rows_to_numpy(u))
        s, ps = eng.sort_rows(d, ws, unique=False, want_perm=True)
        order = po.lexsort_rows(a)  # stable
        np.testing.assert_array_equal(hip.rows_to_numpy(s), a[order])
        np.testing.assert_array_equal(ps.cpu().numpy(), order)


@pytest.mark.parametrize("case", ["chain16", "ring64"])
def test_generic_expansion_equals_union_of_runs(env, case):
    from paces_amd.core.hamiltonian import HamiltonianFramework
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein
    N, dim, periodic = (16, 8, False) if case == "chain16" else (64, 16, True)
    ham = holstein.Hamiltonian(N, [dim] * N, periodic=periodic, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
        "hopping": {"J": -1.0}, "vib_coupling": {"g": 1.3}})
    te = holstein.TimeEvolution(ham, holstein.Observables, ["n_exc"], None,
                                TimeEvoParams(maxstates=5000, enlarge_steps=1, shuffle_seed=None,
                                              diagnostics=False), ExpmParams())
    te.create_nonuniform_basis([1] * N, minpos=N // 2, maxpos=N // 2 + 1)
    basis = te.whoami
    for _ in range(7):
        fused = ham.enlarge_basis_set(basis)
        generic = HamiltonianFramework.enlarge_basis_set(ham, basis)
        assert env["torch"].equal(fused, generic)
        basis = fused
    assert basis.shape[0] > 500
