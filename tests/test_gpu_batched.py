"""Batched trajectories against the oracle run one trajectory at a time.

FP64 batched: each trajectory's basis (unpacked fields), amplitudes and observables equal the
oracle's (the per-trajectory stopping test keeps every trajectory's series identical), amplitudes
within 1e-10. Mixed precision batched: maximum site-density error <= 1e-4.
"""

import numpy as np
import pytest

from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu


def _draws(n, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.choice([6, 8, 11, 16]))
        J = float(rng.uniform(0.5, 1.5))
        g = float(rng.uniform(0.5, 2.0))
        ms = int(round(10 ** rng.uniform(2.3, 3.3)))
        out.append((N, J, g, ms))
    return out


def _oracle(N, J, g, ms, dim, steps, dt, periodic=False):
    ham = po.HolsteinOracle(N, [dim] * N, periodic=periodic, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
        "hopping": {"J": -J}, "vib_coupling": {"g": g}})
    te = po.TimeEvolutionOracle(ham, ms, 1)
    te.create_single_state_basis(N // 2)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = N // 2
    te.create_initial_vector([1.0], raw)
    recs = []
    te.generate_timeline(dt * np.arange(steps + 1), recs.append)
    return ham, recs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_trajectories_match_per_trajectory_oracle(precision):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    draws = _draws(7)
    dim, steps, dt = 8, 12, 0.05
    oracles = [_oracle(N, J, g, ms, dim, steps, dt) for (N, J, g, ms) in draws]
    bham = holstein.BatchedHamiltonian([d[0] for d in draws], dim, [-d[1] for d in draws],
                                       [d[2] for d in draws])
    bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                        precision=precision)
    bte.create_single_exciton_states([d[0] // 2 for d in draws])
    worst = 0.0
    t_arr = dt * np.arange(steps + 1)
    for i in range(steps + 1):
        print(f"[batched {precision}] step {i}", flush=True)
        bte.step(i, t_arr[i] - t_arr[i - 1] if i else 0.0)
        n_exc, n_pho, nrm2 = bte.observables()
        raw = bham.decompress_ind_arr(bte.whoami).cpu().numpy()
        vec = bte.vector.cpu().numpy()
        for t, ((N, J, g, ms), (oham, recs)) in enumerate(zip(draws, oracles)):
            rec = recs[i]
            err = np.max(np.abs(n_exc[t, :N] - rec.obs["n_exc"]))
            worst = max(worst, err)
            assert np.all(n_exc[t, N:] == 0)
            if precision == "fp64":
                sel = raw[:, 0] == t
                mine = raw[sel][:, 1:2 + N]
                np.testing.assert_array_equal(mine, oham.layout.unpack(rec.whoami))
                assert np.all(raw[sel][:, 2 + N:] == 0)
                # the same per-step dt (t[i] - t[i-1]) as the oracle: bit-identical amplitudes
                assert np.array_equal(vec[sel], rec.vec), (i, t)
                assert err <= 1e-10, (i, t, err)
                assert np.max(np.abs(n_pho[t, :N] - rec.obs["n_pho"])) <= 1e-10
                assert abs(np.sqrt(nrm2[t]) - rec.obs["norm"]) <= 1e-10
                if i > 0:
                    assert int(bte.last_expm["j"][t]) == rec.expm_dbg[1], (i, t)
            else:
                assert err <= 1e-4, (i, t, err)
    print(f"[batched {precision}] max density error {worst:.3e}")


def test_batched_periodic_rings_match_per_trajectory_oracle():
    """FP64 batched rings of different lengths (periodic hops, whose wrap moves differ by
    trajectory, so the CSR is built through atomic cursors and a per-row sort rather than fixed
    slots): basis, amplitudes and observables equal each ring's own oracle run."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    draws = [(N, J, g, ms) for (N, J, g, ms) in _draws(4, seed=5)]
    dim, steps, dt = 8, 6, 0.05
    oracles = [_oracle(N, J, g, ms, dim, steps, dt, periodic=True) for (N, J, g, ms) in draws]
    bham = holstein.BatchedHamiltonian([d[0] for d in draws], dim, [-d[1] for d in draws],
                                       [d[2] for d in draws], periodic=True)
    bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                        precision="fp64")
    bte.create_single_exciton_states([d[0] // 2 for d in draws])
    t_arr = dt * np.arange(steps + 1)
    for i in range(steps + 1):
        print(f"[batched periodic] step {i}", flush=True)
        bte.step(i, t_arr[i] - t_arr[i - 1] if i else 0.0)
        n_exc, n_pho, nrm2 = bte.observables()
        raw = bham.decompress_ind_arr(bte.whoami).cpu().numpy()
        vec = bte.vector.cpu().numpy()
        for t, ((N, J, g, ms), (oham, recs)) in enumerate(zip(draws, oracles)):
            rec = recs[i]
            sel = raw[:, 0] == t
            np.testing.assert_array_equal(raw[sel][:, 1:2 + N], oham.layout.unpack(rec.whoami))
            assert np.max(np.abs(vec[sel] - rec.vec), initial=0) <= 1e-10, (i, t)
            assert np.max(np.abs(n_exc[t, :N] - rec.obs["n_exc"])) <= 1e-10, (i, t)
            assert np.max(np.abs(n_pho[t, :N] - rec.obs["n_pho"])) <= 1e-10, (i, t)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_series_equals_eager(precision):
    """The series replayed from a HIP graph (captured once per row capacity and term count, the
    device descriptor rewritten every step) gives the eager launches' state bit for bit at every
    step, across growing bases (new captures) and repeated replays."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    draws = _draws(7, seed=5)
    runs = []
    for graph in (False, True):
        bham = holstein.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                           [d[2] for d in draws])
        bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                            precision=precision, graph=graph)
        bte.create_single_exciton_states([d[0] // 2 for d in draws])
        runs.append(bte)
    for i in range(14):
        for bte in runs:
            bte.step(i, 0.05)
        a, b = runs
        assert torch.equal(a.whoami, b.whoami), i
        assert torch.equal(a.vector, b.vector), i
        if i > 0:
            assert a.last_expm.tobytes() == b.last_expm.tobytes(), i
    assert len(runs[1].graph_pool.graphs) >= 2  # several capacities were captured, then replayed


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_series_continued_past_the_graph(precision):
    """With at most 2 (then 3) terms in the graph every series continues eagerly past it; in the
    mixed form the step's observables pass is launched right after the graph (before the host
    reads whether every series stopped) and a second pass adds the remaining terms. The state,
    the stopping states and the observables equal the eager run's bit for bit at every step."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    draws = _draws(7, seed=5)
    runs = []
    for graph in (False, True):
        bham = holstein.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                           [d[2] for d in draws])
        bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                            precision=precision, graph=graph)
        bte.create_single_exciton_states([d[0] // 2 for d in draws])
        runs.append(bte)
    runs[1].graph_terms_max = 2
    for i in range(12):
        if i == 6:
            runs[1].graph_terms_max = 3
        print(f"[continued {precision}] step {i}", flush=True)
        for bte in runs:
            bte.step(i, 0.05)
        a, b = runs
        assert torch.equal(a.whoami, b.whoami), i
        assert torch.equal(a.vector, b.vector), i
        if i > 0:
            assert a.last_expm.tobytes() == b.last_expm.tobytes(), i
            assert int(b.last_expm["j"].max()) >= 3, i  # every step continued past the graph
        x, y = a.observables(), b.observables()
        assert all(np.array_equal(p, q) for p, q in zip(x, y)), i


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("graph", [False, True])
def test_coded_values_equal_explicit_values(precision, graph):
    """The series over coded CSR values (a code byte per vib entry, each trajectory's values
    formed in LDS) and over the slot matrix (a uint32 per row and op: column and code) gives the
    explicit-value series' basis, amplitudes and stopping states bit for bit at every step; the
    coded matrices' values, formed on demand, equal the FP64 fill's (and, in FP32, the FP32
    fill's) bit for bit, and each term of the slot matrix read as a CSR is the coded CSR."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    draws = _draws(6, seed=9)
    runs = []
    for csr in ("explicit", "coded", "slots"):
        bham = holstein.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                           [d[2] for d in draws])
        bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                            precision=precision, graph=graph, csr_values=csr)
        bte.create_single_exciton_states([d[0] // 2 for d in draws])
        runs.append(bte)
    assert [r.csr_values for r in runs] == ["explicit", "coded", "slots"]
    for i in range(10):
        for bte in runs:
            bte.step(i, 0.05)
        a = runs[0]
        for b in runs[1:]:
            assert torch.equal(a.whoami, b.whoami), (b.csr_values, i)
            assert torch.equal(a.vector, b.vector), (b.csr_values, i)
            if i > 0:
                assert a.last_expm.tobytes() == b.last_expm.tobytes(), (b.csr_values, i)
    a, b, c = runs
    for name in a.hamobj.offdiag_terms:
        ma, mb, mc = a.sparse_mats_dict[name], b.sparse_mats_dict[name], c.sparse_mats_dict[name]
        assert torch.equal(ma.indptr, mb.indptr) and torch.equal(ma.indices, mb.indices), name
        assert (mb.code is not None) == (name == "vib_coupling")
        assert mb.data.dtype == torch.float64
        assert torch.equal(mb.data.to(ma.data.dtype), ma.data), name
        assert torch.equal(mc.indptr, mb.indptr) and torch.equal(mc.indices, mb.indices), name
        assert (mc.code is None) == (mb.code is None), name
        if mb.code is not None:
            assert torch.equal(mc.code, mb.code), name
        assert torch.equal(mc.data, mb.data), name


@pytest.mark.parametrize("csr_values", ["explicit", "coded", "slots"])
def test_single_trajectory_batch_matches_oracle(csr_values):
    """A batch of one trajectory (a zero-width trajectory field): basis, amplitudes and
    observables equal the oracle's, with explicit and coded CSR values and the slot matrix."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    N, J, g, ms = 12, 0.9, 1.3, 800
    dim, steps, dt = 8, 8, 0.05
    oham, recs = _oracle(N, J, g, ms, dim, steps, dt)
    bham = holstein.BatchedHamiltonian([N], dim, [-J], [g])
    bte = holstein.BatchedTimeEvolution(bham, [ms], enlarge_steps=1, precision="fp64",
                                        csr_values=csr_values)
    bte.create_single_exciton_states([N // 2])
    t_arr = dt * np.arange(steps + 1)
    for i in range(steps + 1):
        print(f"[single trajectory {csr_values}] step {i}", flush=True)
        bte.step(i, t_arr[i] - t_arr[i - 1] if i else 0.0)
        n_exc, n_pho, _ = bte.observables()
        raw = bham.decompress_ind_arr(bte.whoami).cpu().numpy()
        rec = recs[i]
        assert np.all(raw[:, 0] == 0)
        np.testing.assert_array_equal(raw[:, 1:2 + N], oham.layout.unpack(rec.whoami))
        assert np.array_equal(bte.vector.cpu().numpy(), rec.vec), i
        assert np.max(np.abs(n_exc[0, :N] - rec.obs["n_exc"])) <= 1e-10, i
        assert np.max(np.abs(n_pho[0, :N] - rec.obs["n_pho"])) <= 1e-10, i


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_finish_equals_separate_finish(precision):
    """The step's observables pass also finishes the series (adds the last even term of the
    trajectories that stopped on one) and forms the next select's keys. Every amplitude must be
    finished exactly once: the step equals the one with the separate finish and key passes bit
    for bit, at every step, twice over (a workgroup's waves share each row's amplitude, so a
    wave re-reading a finished amplitude would show here)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    from paces_amd.models import holstein
    if not devices.configure_called:
        devices.configure()
    draws = _draws(7, seed=5)
    for rep in range(2):
        runs = []
        for graph in (False, True):
            bham = holstein.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                               [d[2] for d in draws])
            bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                                precision=precision, graph=graph)
            bte.create_single_exciton_states([d[0] // 2 for d in draws])
            runs.append(bte)
        fused, separate = runs
        for i in range(12):
            print(f"[fused finish {precision}] rep {rep} step {i}", flush=True)
            fused.step(i, 0.05)
            separate.adapt()
            separate._resolve_observables()
            if i > 0:
                separate.propagate(0.05, fuse_finish=False)
            separate._launch_observables(i)
            separate._keys = None  # the separate form's next select forms its own keys
            assert torch.equal(fused.whoami, separate.whoami), (rep, i)
            assert torch.equal(fused.vector, separate.vector), (rep, i)
            a, b = fused.observables(), separate.observables()
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), (rep, i)
