"""The batched workloads at their own row layouts, against the oracle run one trajectory at a time.

* 256-trajectory chains (N in {16, 24, 32, 48}, phonon cap 7): rows of 1 to 3 uint64 words with
  a trajectory field and phonon fields straddling word boundaries. Sixteen draws from the
  workload's distribution (maxstates scaled down so the oracle finishes in seconds), run the
  way the benchmark runs them (series replayed from a HIP graph, tuned launch shape, coded CSR
  values): FP64 rows and amplitudes bit-identical to the oracle at every step, complex64 terms
  within 1e-4 site density.
* 16x16 periodic torus, 16 trajectories with g in linspace(0.8, 1.6), enlarge_steps=0:
  13-word rows; FP64 parity with the oracle at reduced maxstates.
* Full sizes (the benchmark's 256 seeded draws; the torus at a total maxstates of 5e7) through
  size-independent properties: every trajectory's rows sorted and unique, the CSR symmetric and
  canonical with only the trajectory's J and g*sqrt(m) values, norms that do not grow, and the
  complex64-term run within 1e-4 site density of the FP64 run.
"""

import numpy as np
import pytest

from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu

DIM = 8  # phonon cap 7


def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paces_amd.config import devices
    if not devices.configure_called:
        devices.configure()


def _oracle(N, J, g, ms, steps, dt, geo=1, periodic=False, enlarge=1, pos=None):
    ham = po.HolsteinOracle(N, [DIM] * N, geo_dim=geo, periodic=periodic, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
        "hopping": {"J": -J}, "vib_coupling": {"g": g}})
    te = po.TimeEvolutionOracle(ham, ms, enlarge)
    p = N // 2 if pos is None else pos
    te.create_single_state_basis(p)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = p
    te.create_initial_vector([1.0], raw)
    recs = []
    te.generate_timeline(dt * np.arange(steps + 1), recs.append)
    return ham, recs


def chain_draws(n, seed, ms_hi=2e5, ms_lo=1e3):
    """The 256-chain workload's distribution: N in {16,24,32,48}, J ~ U[0.5,1.5],
    g ~ U[0.5,2], maxstates = round(10^U[log10 ms_lo, log10 ms_hi])."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.choice([16, 24, 32, 48]))
        J = float(rng.uniform(0.5, 1.5))
        g = float(rng.uniform(0.5, 2.0))
        ms = int(round(10 ** rng.uniform(np.log10(ms_lo), np.log10(ms_hi))))
        out.append((N, J, g, ms))
    return out


def _run_against_oracle(bham, bte, draws, oracles, t_arr, fp64, label, every=1):
    """Steps the batched run over t_arr and compares each trajectory with its oracle run:
    FP64 rows and amplitudes exactly, observables within 1e-10; otherwise the site density at
    every `every`-th step and the last within 1e-4. Returns the worst density error."""
    worst = 0.0
    steps = len(t_arr) - 1
    for i in range(steps + 1):
        print(f"[{label}] step {i}", flush=True)
        bte.step(i, t_arr[i] - t_arr[i - 1] if i else 0.0)
        if not fp64 and i % every and i != steps:
            continue
        n_exc, n_pho, nrm2 = bte.observables()
        if fp64:
            raw = bham.decompress_ind_arr(bte.whoami).cpu().numpy()
            vec = bte.vector.cpu().numpy()
        for t, ((N, J, g, ms), (oham, recs)) in enumerate(zip(draws, oracles)):
            rec = recs[i]
            err = float(np.max(np.abs(n_exc[t, :N] - rec.obs["n_exc"])))
            worst = max(worst, err)
            if fp64:
                sel = raw[:, 0] == t
                np.testing.assert_array_equal(raw[sel][:, 1:2 + N], oham.layout.unpack(rec.whoami))
                assert np.all(raw[sel][:, 2 + N:] == 0)
                assert np.array_equal(vec[sel], rec.vec), (i, t)
                assert err <= 1e-10, (i, t, err)
                assert np.max(np.abs(n_pho[t, :N] - rec.obs["n_pho"])) <= 1e-10, (i, t)
                assert abs(np.sqrt(nrm2[t]) - rec.obs["norm"]) <= 1e-10, (i, t)
                if i > 0:
                    assert int(bte.last_expm["j"][t]) == rec.expm_dbg[1], (i, t)
            else:
                assert err <= 1e-4, (i, t, err)
    return worst


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chain_workload_layout_matches_oracle(precision):
    """Sixteen draws of the 256-chain workload's distribution (at least two 48-site chains:
    3-word rows) in the benchmark's configuration, against each chain's oracle run."""
    _gpu()
    from paces_amd.models import holstein
    draws = chain_draws(16, seed=11, ms_hi=3000, ms_lo=300)
    assert sum(d[0] == 48 for d in draws) >= 2 and {d[0] for d in draws} == {16, 24, 32, 48}
    steps, dt = 14, 0.05
    t_arr = dt * np.arange(steps + 1)
    oracles = [_oracle(N, J, g, ms, steps, dt) for (N, J, g, ms) in draws]
    bham = holstein.BatchedHamiltonian([d[0] for d in draws], DIM, [-d[1] for d in draws],
                                       [d[2] for d in draws])
    assert bham.totwordwidth == 3
    bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                        precision=precision, graph=True, autotune=True,
                                        csr_values="slots")
    assert bte.csr_values == "slots"
    bte.create_single_exciton_states([d[0] // 2 for d in draws])
    worst = _run_against_oracle(bham, bte, draws, oracles, t_arr, precision == "fp64",
                                f"chains {precision}", every=10 if precision == "mixed" else 1)
    print(f"[chains {precision}] max site-density error vs oracle {worst:.3e}")


TORUS_G = np.linspace(0.8, 1.6, 16)
TORUS_POS = 8 * 16 + 8  # exciton at (8, 8)


def test_torus_workload_layout_matches_oracle():
    """Sixteen 16x16 periodic tori (g = linspace(0.8, 1.6, 16), J = 1, cap 7, 13-word rows,
    enlarge_steps = 0) at reduced maxstates: FP64 rows and amplitudes equal each torus's oracle
    run at every step."""
    _gpu()
    from paces_amd.models import holstein
    ms = [600 + 50 * k for k in range(16)]
    draws = [(256, 1.0, float(g), m) for g, m in zip(TORUS_G, ms)]
    steps, dt = 7, 0.02
    t_arr = dt * np.arange(steps + 1)
    oracles = [_oracle(N, J, g, m, steps, dt, geo=2, periodic=True, enlarge=0, pos=TORUS_POS)
               for (N, J, g, m) in draws]
    bham = holstein.BatchedHamiltonian([256] * 16, DIM, [-1.0] * 16, list(TORUS_G), geo_dim=2,
                                       periodic=True)
    assert bham.totwordwidth == 13
    bte = holstein.BatchedTimeEvolution(bham, ms, enlarge_steps=0, precision="fp64")
    bte.create_single_exciton_states([TORUS_POS] * 16)
    _run_against_oracle(bham, bte, draws, oracles, t_arr, True, "torus fp64")


# ---- full sizes: size-independent properties -----------------------------------------------------

def _as_unsigned_order(rows):
    """int64 rows whose signed order is the rows' unsigned order (flip the sign bit)."""
    import torch
    return rows ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=rows.device)


def check_rows_sorted_unique(rows):
    """Rows strictly increasing in lexicographic (unsigned) order, hence sorted and unique."""
    import torch
    if rows.shape[0] < 2:
        return
    u = _as_unsigned_order(rows)
    a, b = u[:-1], u[1:]
    less = torch.zeros(a.shape[0], dtype=torch.bool, device=rows.device)
    equal = torch.ones_like(less)
    for k in range(rows.shape[1]):
        less |= equal & (a[:, k] < b[:, k])
        equal &= a[:, k] == b[:, k]
    assert bool(less.all()), "rows are not strictly increasing"


def check_csr(bham, rows, mat, seg_off, name):
    """Canonical (columns strictly increasing per row), symmetric, block diagonal by trajectory,
    and only the trajectory's value: J for hopping, g*sqrt(m) with 1 <= m < cap for vib."""
    import torch
    R = rows.shape[0]
    ip = mat.indptr.long()
    ix = mat.indices.long()
    assert int(ip[0]) == 0 and int(ip[-1]) == ix.numel() and bool((ip[1:] >= ip[:-1]).all())
    counts = ip[1:] - ip[:-1]
    r = torch.repeat_interleave(torch.arange(R, device=rows.device), counts)
    assert bool(((ix >= 0) & (ix < R)).all())
    same_row = r[1:] == r[:-1]
    assert bool((ix[1:][same_row] > ix[:-1][same_row]).all()), f"{name}: columns not ascending"
    key = r * R + ix
    tkey = ix * R + r
    assert torch.equal(torch.sort(key).values, torch.sort(tkey).values), f"{name}: not symmetric"
    traj = bham.get_traj(rows).long()
    assert torch.equal(traj[r], traj[ix]), f"{name}: entries across trajectories"
    vals = mat.data
    term = list(bham.offdiag_terms).index(name)
    tv = bham._d_traj_val[traj[r], term]
    if name == "vib_coupling":
        m2 = (vals / tv) ** 2
        m = torch.round(m2)
        assert bool(((m >= 1) & (m <= DIM - 1)).all()), f"{name}: phonon count out of range"
        assert torch.equal(vals, tv * torch.sqrt(m)), f"{name}: values are not g*sqrt(m)"
    else:
        assert torch.equal(vals, tv), f"{name}: values are not the trajectory's J"


def test_chain_workload_full_size_properties():
    """The benchmark's 256 seeded draws (maxstates up to 2e5, about 4e7 rows once grown),
    complex64 terms and FP64 side by side for 8 steps: rows per trajectory sorted and unique,
    CSR canonical, symmetric and with the allowed values, norms that do not grow (FP64), and
    site density of the complex64-term run within 1e-4 of the FP64 run at every step."""
    _gpu()
    import torch
    from paces_amd.models import holstein
    draws = chain_draws(256, seed=0)  # the benchmark's draws
    runs = {}
    for prec in ("fp64", "mixed"):
        bham = holstein.BatchedHamiltonian([d[0] for d in draws], DIM, [-d[1] for d in draws],
                                           [d[2] for d in draws])
        bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                            precision=prec, graph=True, autotune=True)
        bte.create_single_exciton_states([d[0] // 2 for d in draws])
        runs[prec] = (bham, bte)
    steps, dt = 8, 0.05
    t_arr = dt * np.arange(steps + 1)
    prev_norm = None
    worst = 0.0
    for i in range(steps + 1):
        obs = {}
        for prec, (bham, bte) in runs.items():
            print(f"[full chains {prec}] step {i}", flush=True)
            bte.step(i, t_arr[i] - t_arr[i - 1] if i else 0.0)
            obs[prec] = bte.observables()
        bham, bte = runs["fp64"]
        n_exc64, _, nrm2 = obs["fp64"]
        assert np.allclose(n_exc64.sum(axis=1), nrm2, rtol=1e-12, atol=0)
        if prev_norm is not None:
            assert np.all(nrm2 <= prev_norm * (1 + 1e-12)), i
        prev_norm = nrm2
        worst = max(worst, float(np.max(np.abs(obs["mixed"][0] - n_exc64))))
        assert worst <= 1e-4, (i, worst)
    bham, bte = runs["fp64"]
    print(f"[full chains] {bte.numstates} rows; checking order and CSR", flush=True)
    seg = bte.seg_off.cpu().numpy()
    check_rows_sorted_unique(bte.whoami)
    traj = bham.get_traj(bte.whoami).long().cpu().numpy()
    assert np.all(np.diff(traj) >= 0) and np.array_equal(
        np.searchsorted(traj, np.arange(len(draws) + 1)), seg)
    for name in bham.offdiag_terms:
        check_csr(bham, bte.whoami, bte.sparse_mats_dict[name], bte.seg_off, name)
    print(f"[full chains] max mixed-vs-FP64 site-density error {worst:.3e}")
    del runs
    torch.cuda.empty_cache()


def test_torus_workload_full_size_properties():
    """Sixteen 16x16 tori sharing a total maxstates of 5e7 (3.125e6 each, enlarge_steps = 0,
    complex64 terms), grown from one state each until every selection is cut at maxstates, then
    two more steps: rows per trajectory sorted and unique, the CSR canonical and symmetric with
    the allowed values, the site densities summing to the norms, norms below 1 + 1e-6."""
    _gpu()
    import torch
    from paces_amd.models import holstein
    ms = [50_000_000 // 16] * 16
    bham = holstein.BatchedHamiltonian([256] * 16, DIM, [-1.0] * 16, list(TORUS_G), geo_dim=2,
                                       periodic=True)
    bte = holstein.BatchedTimeEvolution(bham, ms, enlarge_steps=0, precision="mixed",
                                        graph=True)
    bte.create_single_exciton_states([TORUS_POS] * 16)
    dt = 0.02
    i, extra = 0, 0
    while True:
        print(f"[full torus] step {i}: basis {bte.numstates}", flush=True)
        seg = bte.seg_off.cpu().numpy() if bte.seg_off is not None else None
        bte.step(i, dt)
        n_exc, _, nrm2 = bte.observables()
        assert np.allclose(n_exc.sum(axis=1), nrm2, rtol=1e-9, atol=0), i
        assert np.all(np.sqrt(nrm2) <= 1 + 1e-6), i
        if seg is not None and np.all(np.diff(seg) > np.asarray(ms)):
            extra += 1
            if extra == 2:
                break
        i += 1
        assert i < 20, "the tori did not reach their maxstates"
    print(f"[full torus] {bte.numstates} rows; checking order and CSR", flush=True)
    check_rows_sorted_unique(bte.whoami)
    for name in bham.offdiag_terms:
        check_csr(bham, bte.whoami, bte.sparse_mats_dict[name], bte.seg_off, name)
    del bte
    torch.cuda.empty_cache()
