"""Initial bases and states (paces/models/holstein.py:406-569; growth,
paces/core/time_evolution.py:290-319) against the oracle: the generators' packed rows equal the
oracle's tuples packed, growth equals the oracle's growth, and timelines started from these
states match the oracle's step by step."""

import numpy as np
import pytest

from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu

TERMS = {"diag": {"eps_sys": 0.1, "hbar_omega": 1.0, "delta_eps": 0.0},
         "hopping": {"J": -1.0}, "vib_coupling": {"g": 0.8}}


def _setup(N, dim, ms, enlarge=1):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    ham = holstein.Hamiltonian(N, [dim] * N, use_terms=TERMS)
    te = holstein.TimeEvolution(ham, holstein.Observables, ["n_exc", "n_pho"], None,
                                TimeEvoParams(maxstates=ms, enlarge_steps=enlarge,
                                              shuffle_seed=None), ExpmParams())
    oham = po.HolsteinOracle(N, [dim] * N, use_terms=TERMS)
    return ham, te, oham


def _rows_equal(te, oham, raw):
    mine = te.whoami.cpu().numpy().view(np.uint64)
    np.testing.assert_array_equal(mine, oham.layout.pack(raw))


def _timeline_matches(te, oham, ms, steps=5, dt=0.05):
    """Runs te and an oracle started from te's basis and vector over the same times."""
    ote = po.TimeEvolutionOracle(oham, ms, te.te_params.enlarge_steps)
    ote.whoami = te.whoami.cpu().numpy().view(np.uint64).copy()
    ote.vector = te.vector.cpu().numpy().copy()
    t = dt * np.arange(steps + 1)
    recs = []
    ote.generate_timeline(t, recs.append)
    for i in range(steps + 1):
        if i == 0:
            te._perform_first_step(t[0])
        else:
            te._evolve_single_step(i, t[i], t[i] - t[i - 1])
        rec = recs[i]
        np.testing.assert_array_equal(te.whoami.cpu().numpy().view(np.uint64), rec.whoami)
        assert np.array_equal(te.vector.cpu().numpy(), rec.vec), i
        assert np.max(np.abs(te.obsobj.last_results["n_exc"] - rec.obs["n_exc"])) <= 1e-10


def test_nonuniform_basis_with_truncation_lists():
    """Per-site truncations and lowest levels (a truncation of 1 keeps the site at 0 whatever
    its lowest level) and an exciton window, then a uniform truncated basis."""
    N, dim = 8, 4
    ham, te, oham = _setup(N, dim, 5000)
    trunc = [1, 2, 3, 1, 2, 1, 1, 2]
    low = [0, 1, 0, 2, 0, 0, 0, 1]
    te.create_nonuniform_basis(trunc, lowest_d_list=low, minpos=2, maxpos=5)
    _rows_equal(te, oham, po.nonuniform_basis_tuples(N, trunc, low, 2, 5))
    te.create_uniform_truncated_basis(2)
    _rows_equal(te, oham, po.nonuniform_basis_tuples(N, [2] * N))


def test_moving_gaussian_basis_grow_and_timeline():
    """The Gaussian basis, grown twice (n_max binding stops one expansion early), then a
    timeline from a state on it, each equal to the oracle's."""
    N, dim, ms = 8, 8, 40000
    ham, te, oham = _setup(N, dim, ms)
    te.create_moving_gaussian_obc_basis(super_mu=3.5, super_sigma=2.0, sigma=1.2, maxval=3)
    raw = po.moving_gaussian_tuples(N, 3.5, 2.0, 1.2, 3)
    _rows_equal(te, oham, raw)
    ote = po.TimeEvolutionOracle(oham, ms, 1)
    ote.whoami = oham.layout.pack(raw)
    n_gpu = te.grow_optimal_basis(n_max=3, fillfac=1.0)
    n_ora = ote.grow_optimal_basis(n_max=3, fillfac=1.0)
    assert n_gpu == n_ora
    np.testing.assert_array_equal(te.whoami.cpu().numpy().view(np.uint64), ote.whoami)
    coo = np.zeros((2, N + 1), dtype=np.int64)
    coo[0, 0], coo[1, 0], coo[1, 4] = 4, 3, 1
    te.create_initial_vector([0.6, 0.8j], coo)
    _timeline_matches(te, oham, ms)


def test_tensor_init_state_and_timeline():
    """A tensor-product phonon state on two sites over a uniform truncated basis: the vector
    equals the oracle's coefficients placed on their rows, and the timeline from it matches."""
    N, dim, ms = 4, 4, 3000
    ham, te, oham = _setup(N, dim, ms)
    te.create_uniform_truncated_basis(3)
    te.create_tensor_init_state([0, 2], [0.6, 0.8], excsite=2, sites=[0, 3])
    coef, raw = po.tensor_state(N, [0, 2], [0.6, 0.8], 2, sites=[0, 3])
    rows = oham.layout.pack(raw)
    basis = te.whoami.cpu().numpy().view(np.uint64)
    idx, hit = po.lookup(basis, rows)
    assert hit.all()
    want = np.zeros(basis.shape[0], dtype=np.complex128)
    want[idx] = coef / np.linalg.norm(coef)
    np.testing.assert_allclose(te.vector.cpu().numpy(), want, rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        te.create_tensor_init_state([0, 1], [1.0, 1.0], excsite=0)  # not normalised
    _timeline_matches(te, oham, ms)
