"""Learned proposal (candidate scorer in the step) through the C ABI.

* Features (paces_holstein_features) are restated in NumPy from the unpacked union rows, the
  carried amplitudes and the per-term CSR pattern. Integer-derived columns are bit-exact; the
  log-amplitude columns are within one bf16 ulp, since device and NumPy log10 may differ in the
  last double ulp. The added-rows form (what the step scores) equals them at the added rows.
* The plan from the added rows' scores equals the union plan with the basis rows forced.
* Compaction (paces_prune_plan, paces_compact_kept) is checked against NumPy boolean indexing.
* With no scorer the learned step keeps every row and equals the plain step bit for bit.
* Ablation on the 16-site chain (cap 7, maxstates 2e4):
  * labels are logged from plain runs at three couplings (the rows each first expansion adds,
    labelled by the next selection), and the scorer is trained on the device;
  * at a fourth coupling the trained scorer drops a share of the added rows before the second
    expansion, so the final basis is smaller than the plain step's;
  * the maximum site-density error against the FP64 oracle stays within 1e-4, with no more
    than one fallback;
  * an untrained network keeps everything, so it gains nothing over the method without it.
"""

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu

DT = 0.05


@pytest.fixture(scope="module")
def env():
    import torch
    from paces_amd.config import devices
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not devices.configure_called:
        devices.configure()
    from paces_amd import _hip
    from paces_amd.core import engine, scorer
    from paces_amd.models import holstein
    return dict(torch=torch, dev=devices.vector_dev, hip=_hip, engine=engine, scorer=scorer,
                holstein=holstein)


def _terms(J, g):
    return {"diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
            "hopping": {"J": -J}, "vib_coupling": {"g": g}}


def _make(env, cls, N, dim, g, ms, periodic=False, **kw):
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    h = env["holstein"]
    ham = h.Hamiltonian(N, [dim] * N, periodic=periodic, use_terms=_terms(1.0, g))
    te = cls(ham, h.Observables, ["n_exc", "n_pho"], None,
             TimeEvoParams(maxstates=ms, enlarge_steps=1, shuffle_seed=None, diagnostics=False),
             ExpmParams(), **kw)
    te.create_nonuniform_basis([1] * N, minpos=N // 2, maxpos=N // 2 + 1)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = N // 2
    te.create_initial_vector([1.0], raw)
    te._perform_first_step(0.0)
    return te


def _union(env, te):
    """The learned step's union and everything the features need, for the current state."""
    eng = env["engine"]
    ham = te.hamobj
    sel = te._determine_select_whoami(te.vector)
    basis = ham.enlarge_basis_set(sel)
    imgs, offs, src, nup = ham._images(basis, want_src=True)
    runs = [(basis, 0, basis.shape[0])]
    runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
    U, inb, cpos = eng.merge_unique(runs, ham.totwordwidth, ham._ws, want_from0=True,
                                    want_cand_pos=True)
    psi = eng.reassign(te.whoami, te.vector, U, ham._ws)
    diag = ham.generate_mel_diag(U)
    mats = ham._pairs_csr(U.shape[0], basis.shape[0], offs, cpos, inb, src, nup, basis)
    return ham, U, psi, diag, basis, offs, cpos, src, mats


@pytest.mark.parametrize("case", ["chain16", "ring24"])
def test_features_match_numpy(env, case):
    torch, hip = env["torch"], env["hip"]
    N, dim, periodic, g = (16, 8, False, 1.3) if case == "chain16" else (24, 16, True, 0.9)
    te = _make(env, env["holstein"].TimeEvolution, N, dim, g, 3000, periodic=periodic)
    for i in range(1, 6):
        te._evolve_single_step(i, DT * i, DT)
    ham, U, psi, diag, basis, offs, cpos, src, mats = _union(env, te)
    X = ham.features(U, psi, diag, basis.shape[0], offs, cpos, src)
    got = X.cpu().view(torch.bfloat16).to(torch.float64).numpy()
    R = U.shape[0]
    lay = po.Layout([N] + [dim] * N)
    raw = lay.unpack(hip.rows_to_numpy(U)).astype(np.int64)
    pos = raw[:, 0]
    ph = raw[:, 1:]
    w2 = np.abs(psi.cpu().numpy()) ** 2
    par = np.zeros(R)
    for m in mats.values():
        A = m.to_scipy().tocsr()
        A.data = np.ones_like(A.data)
        P = A.multiply(w2[None, :]).tocsr()
        par = np.maximum(par, P.max(axis=1).toarray().ravel())
    exp = np.zeros((R, 64), dtype=np.float64)
    for c in range(33):
        s = pos + c - 16
        if periodic:
            s = np.mod(s, N)
            inr = np.ones(R, dtype=bool)
        else:
            inr = (s >= 0) & (s < N)
        v = np.full(R, -1.0, dtype=np.float32)
        sc = np.clip(s, 0, N - 1)
        v[inr] = (ph[np.arange(R), sc].astype(np.float32) / np.float32(dim - 1))[inr]
        exp[:, c] = mo.bf16(v)

    def logamp(x):
        return np.where(x > 1e-40, 0.5 * np.log10(np.maximum(x, 1e-300)), -20.0) / 10.0

    exp[:, 33] = mo.bf16(logamp(par).astype(np.float32))
    exp[:, 34] = mo.bf16(((diag.cpu().numpy() - 0.0) * (1.0 / (1.0 * N))).astype(np.float32))
    exp[:, 35] = mo.bf16(logamp(w2).astype(np.float32))
    exp[:, 36] = mo.bf16(np.float32(g))
    exp[:, 37] = mo.bf16(np.float32(-1.0))
    exp[:, 38] = mo.bf16((pos.astype(np.float32) / np.float32(N)))
    exact = [c for c in range(64) if c not in (33, 35)]
    np.testing.assert_array_equal(got[:, exact], exp[:, exact])
    for c in (33, 35):
        tol = np.abs(exp[:, c]) * 2.0 ** -7
        assert np.all(np.abs(got[:, c] - exp[:, c]) <= tol), c
    assert np.any(par > 0) and np.any(exp[:, 33] > -2.0)


@pytest.mark.parametrize("case", ["chain16", "ring24"])
def test_features_of_added_rows_equal_full_features(env, case):
    """paces_holstein_features_added = the full-union features at the rows outside the basis,
    bit for bit (an added row is never the basis end of an image pair, and the inline diagonal is
    k_diag's), with their union positions in order."""
    torch = env["torch"]
    N, dim, periodic, g = (16, 8, False, 1.3) if case == "chain16" else (24, 16, True, 0.9)
    te = _make(env, env["holstein"].TimeEvolution, N, dim, g, 3000, periodic=periodic)
    for i in range(1, 6):
        te._evolve_single_step(i, DT * i, DT)
    ham, U, psi, diag, basis, offs, cpos, src, mats = _union(env, te)
    full = ham.features(U, psi, diag, basis.shape[0], offs, cpos, src)
    # the union's basis flags as the learned step has them
    inb = torch.zeros(U.shape[0], dtype=torch.uint8, device=U.device)
    inb[cpos[:basis.shape[0]].long()] = 1
    nadd = U.shape[0] - basis.shape[0]
    X, idx = ham.features_added(U, psi, basis.shape[0], offs, cpos, src, inb, nadd)
    want_idx = torch.nonzero(inb == 0).flatten()
    assert nadd > 1000
    np.testing.assert_array_equal(idx.cpu().numpy(), want_idx.cpu().numpy())
    assert torch.equal(X, full.index_select(0, want_idx))


def test_prune_plan_of_added_rows_equals_union_plan(env):
    """paces_prune_plan_rows (scores of the added rows at their union positions) equals
    paces_prune_plan on the scores scattered to the union with the basis rows forced;
    paces_scatter_u8 writes out[idx[j]] = src[j]."""
    torch, hip = env["torch"], env["hip"]
    lib = hip.load()
    rng = np.random.default_rng(9)
    d = env["dev"]
    n = 40_011
    inb = (rng.uniform(size=n) < 0.4).astype(np.uint8)
    added = np.nonzero(inb == 0)[0].astype(np.int32)
    m = added.shape[0]
    prob_a = rng.uniform(size=m).astype(np.float32)
    prob_u = np.ones(n, dtype=np.float32)
    prob_u[added] = prob_a
    forced = np.nonzero(inb)[0].astype(np.int32)
    k1 = torch.empty(n, dtype=torch.uint8, device=d)
    k2 = torch.empty(n, dtype=torch.uint8, device=d)
    t = {k: torch.from_numpy(v).to(d) for k, v in
         dict(pa=prob_a, pu=prob_u, idx=added, inb=inb, forced=forced).items()}
    for bits in (0, 3, 6):
        hip.check(lib.paces_prune_plan(hip.ptr(t["pu"]), n, 0.35, hip.ptr(t["forced"]),
                                       forced.shape[0], bits, 1234567, hip.ptr(k1), hip.stream()))
        hip.check(lib.paces_prune_plan_rows(hip.ptr(t["pa"]), hip.ptr(t["idx"]), m,
                                            hip.ptr(t["inb"]), n, 0.35, bits, 1234567,
                                            hip.ptr(k2), hip.stream()))
        assert torch.equal(k1, k2), bits
        k = k2.cpu().numpy()
        assert np.all(k[inb == 1] == 1)
        low = k[added[prob_a < 0.35]]
        assert np.all(low == 2) if bits == 0 else (np.any(low == 0) and np.any(low == 2))
    src = torch.from_numpy(rng.integers(0, 3, m).astype(np.uint8)).to(d)
    out = torch.full((n,), 7, dtype=torch.uint8, device=d)
    hip.check(lib.paces_scatter_u8(hip.ptr(t["idx"]), m, hip.ptr(src), hip.ptr(out),
                                   hip.stream()))
    want = np.full(n, 7, dtype=np.uint8)
    want[added] = src.cpu().numpy()
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_compaction_matches_numpy(env):
    torch, hip = env["torch"], env["hip"]
    lib = hip.load()
    rng = np.random.default_rng(5)
    n, W, Cn = 50_003, 2, 120_000
    rows = rng.integers(0, 2 ** 62, (n, W), dtype=np.int64).astype(np.uint64)
    prob = rng.uniform(size=n).astype(np.float32)
    forced = rng.choice(n, 3000, replace=False).astype(np.int32)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    inb = rng.integers(0, 2, n).astype(np.uint8)
    cpos = rng.integers(0, n, Cn).astype(np.int32)
    d = env["dev"]
    keep = torch.empty(n, dtype=torch.uint8, device=d)
    d_prob = torch.from_numpy(prob).to(d)
    d_forced = torch.from_numpy(forced).to(d)
    hip.check(lib.paces_prune_plan(hip.ptr(d_prob), n, 0.3, hip.ptr(d_forced), 3000, 4, 77,
                                   hip.ptr(keep), hip.stream()))
    k = keep.cpu().numpy()
    assert set(np.unique(k)) <= {0, 1, 2}
    assert np.all(k[prob >= 0.3] == 1) and np.all(k[forced] == 1)
    dropped = (prob < 0.3) & ~np.isin(np.arange(n), forced)
    frac = np.mean(k[dropped] == 2)
    assert 1 / 32 < frac < 1 / 8  # about 1/16 of the dropped rows are sampled
    d_rows = hip.rows_from_numpy(rows, d)
    d_psi = torch.from_numpy(psi).to(d)
    d_inb = torch.from_numpy(inb).to(d)
    d_cpos = torch.from_numpy(cpos).to(d)
    o_rows = torch.empty((n, W), dtype=torch.int64, device=d)
    o_psi = torch.empty(n, dtype=torch.complex128, device=d)
    o_inb = torch.empty(n, dtype=torch.uint8, device=d)
    o_smp = torch.empty(n, dtype=torch.uint8, device=d)
    o_cpos = torch.empty(Cn, dtype=torch.int32, device=d)
    cnt = torch.empty(1, dtype=torch.int32, device=d)
    ws = hip.Workspace(d)
    wsp, wsn = ws.get(lib.paces_compact_workspace_bytes(n))
    hip.check(lib.paces_compact_kept(hip.ptr(keep), n, hip.ptr(d_rows), W,
                                     hip.ptr(torch.view_as_real(d_psi)), hip.ptr(d_inb),
                                     hip.ptr(d_cpos), Cn, hip.ptr(o_rows),
                                     hip.ptr(torch.view_as_real(o_psi)), hip.ptr(o_inb),
                                     hip.ptr(o_smp), hip.ptr(o_cpos), hip.ptr(cnt), wsp, wsn,
                                     hip.stream()))
    m = k != 0
    nk = int(m.sum())
    assert int(cnt.item()) == nk
    np.testing.assert_array_equal(hip.rows_to_numpy(o_rows[:nk]), rows[m])
    np.testing.assert_array_equal(o_psi[:nk].cpu().numpy(), psi[m])
    np.testing.assert_array_equal(o_inb[:nk].cpu().numpy(), inb[m])
    np.testing.assert_array_equal(o_smp[:nk].cpu().numpy(), (k[m] == 2).astype(np.uint8))
    newpos = np.cumsum(m) - 1
    np.testing.assert_array_equal(o_cpos.cpu().numpy(), np.where(m[cpos], newpos[cpos], -1))


def test_keep_all_path_equals_plain(env):
    torch, h = env["torch"], env["holstein"]
    a = _make(env, h.TimeEvolution, 16, 8, 1.0, 20000)
    b = _make(env, h.LearnedTimeEvolution, 16, 8, 1.0, 20000)
    for i in range(1, 11):
        a._evolve_single_step(i, DT * i, DT)
        b._evolve_single_step(i, DT * i, DT)
        assert torch.equal(a.whoami, b.whoami), i
        assert torch.equal(a.vector, b.vector), i
        np.testing.assert_array_equal(a.obsobj.last_results["n_exc"],
                                      b.obsobj.last_results["n_exc"])


def test_trained_scorer_gains_over_untrained_at_1e4_density_error(env):
    torch, h, sc = env["torch"], env["holstein"], env["scorer"]
    rng = np.random.default_rng(7)
    gs = rng.uniform(0.5, 2.0, 4)
    log = h.LabelLog(env["dev"], 6_000_000)
    for g in gs[:3]:
        te = _make(env, h.LearnedTimeEvolution, 16, 8, g, 20000, label_log=log)
        for i in range(1, 25):
            te._evolve_single_step(i, DT * i, DT)
    X, y = log.data()
    print(f"[learned] {log.n} labelled candidates, retained fraction "
          f"{float(y.float().mean()):.4f}", flush=True)
    trained = sc.CandidateScorer(env["dev"], seed=1)
    hist = trained.train(X, y, epochs=6, batch=8192, lr=1e-3)
    print(f"[learned] training loss by epoch {['%.4f' % v for v in hist]}", flush=True)
    assert hist[-1] < hist[0]
    g = gs[3]
    # the FP64 oracle over the same times (the plain GPU step equals it bit for bit)
    oham = po.HolsteinOracle(16, [8] * 16, use_terms=_terms(1.0, g))
    ote = po.TimeEvolutionOracle(oham, 20000, 1)
    ote.create_single_state_basis(8)
    raw = np.zeros((1, 17), dtype=np.int64)
    raw[0, 0] = 8
    ote.create_initial_vector([1.0], raw)
    recs = []
    ote.generate_timeline(DT * np.arange(41), recs.append)
    plain = _make(env, h.TimeEvolution, 16, 8, g, 20000)
    runs = {"trained": _make(env, h.LearnedTimeEvolution, 16, 8, g, 20000, scorer=trained),
            "untrained": _make(env, h.LearnedTimeEvolution, 16, 8, g, 20000,
                               scorer=sc.CandidateScorer(env["dev"], seed=2))}
    err = {k: 0.0 for k in runs}
    rows_plain = []
    t = DT * np.arange(41)
    for i in range(1, 41):
        plain._evolve_single_step(i, t[i], t[i] - t[i - 1])
        rows_plain.append(plain.numstates)
        ref = recs[i].obs["n_exc"]
        for k, te in runs.items():
            te._evolve_single_step(i, t[i], t[i] - t[i - 1])
            err[k] = max(err[k], float(np.max(np.abs(te.obsobj.last_results["n_exc"] - ref))))
    summary = {}
    for k, te in runs.items():
        st = te.stats[10:]  # the basis has saturated maxstates
        kept = float(np.mean([s["kept"] / max(s["scored"], 1) for s in st]))
        rows = float(np.mean([s["rows"] / r for s, r in zip(st, rows_plain[10:])]))
        fb = sum(s["fallback"] for s in te.stats)
        summary[k] = (kept, fb, err[k], rows)
        print(f"[learned] {k}: mean kept fraction of the added rows {kept:.3f}, final rows "
              f"{rows:.3f} of the plain step's, fallbacks {fb}, max density error vs the oracle "
              f"{err[k]:.2e}", flush=True)
    kept_t, fb_t, err_t, rows_t = summary["trained"]
    kept_u, _, _, rows_u = summary["untrained"]
    assert err_t <= 1e-4
    assert fb_t <= 1
    assert kept_t <= 0.8 and rows_t < 0.95
    # an untrained network passes (nearly) every row: no smaller basis than the plain step's
    assert kept_u >= 0.99 and abs(rows_u - 1.0) < 0.01 and kept_t < kept_u - 0.1
