"""GPU parity of each device primitive against the oracle, through the C ABI.

Integer/row results must be bit-exact; FP64 products in the parity order must be bit-exact;
reductions (norms, observables) must agree to 1e-12 relative.
"""

import ctypes as C

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
    return dict(torch=torch, dev=devices.vector_dev, hip=_hip, engine=engine, lib=_hip.load())


def _rows(env, a):
    return env["hip"].rows_from_numpy(a, env["dev"])


def _np(env, t):
    return env["hip"].rows_to_numpy(t)


def _rand_unique_rows(rng, n, W, hi=None):
    if hi is None:
        a = rng.integers(0, 2 ** 63, size=(n, W), dtype=np.int64).astype(np.uint64)
        a[:, 0] >>= np.uint64(rng.integers(0, 8))
    else:
        a = rng.integers(0, hi, size=(n, W)).astype(np.uint64)
    return po.unique_rows(a)


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 70000, 5_000_003])
def test_scan_u8_and_i32(env, n):
    torch, lib, hip = env["torch"], env["lib"], env["hip"]
    rng = np.random.default_rng(n)
    a8 = rng.integers(0, 2, n).astype(np.uint8)
    a32 = rng.integers(0, 9, n).astype(np.int32)
    ws = hip.Workspace(env["dev"])
    for arr, fn in ((a8, lib.paces_exclusive_scan_u8), (a32, lib.paces_exclusive_scan_i32)):
        d_in = torch.from_numpy(arr).to(env["dev"])
        d_out = torch.empty(n + 1, dtype=torch.int32, device=env["dev"])
        wsp, wsn = ws.get(lib.paces_scan_workspace_bytes(n))
        hip.check(fn(hip.ptr(d_in) if n else None, n, hip.ptr(d_out), wsp, wsn, hip.stream()))
        exp = np.concatenate([[0], np.cumsum(arr.astype(np.int64))])
        np.testing.assert_array_equal(d_out.cpu().numpy(), exp)


@pytest.mark.parametrize("W", [1, 2, 5, 13, 16])
def test_lower_bound_bit_exact(env, W):
    torch, eng = env["torch"], env["engine"]
    rng = np.random.default_rng(W)
    hi = {1: 50000, 2: 300}.get(W, 7)
    book = _rand_unique_rows(rng, 20000, W, hi=hi)
    present = book[rng.choice(book.shape[0], min(3000, book.shape[0]), replace=False)]
    absent = rng.integers(0, hi, size=(3000, W)).astype(np.uint64)
    q = np.concatenate([present, absent, book[:1], book[-1:], np.full((1, W), 2 ** 64 - 1,
                                                                      dtype=np.uint64)])
    for qq, srt in ((q, False), (po.unique_rows(q), False), (po.unique_rows(q), True),
                    (np.sort(q.view(f"V{8 * W}").ravel()).view(np.uint64).reshape(-1, W), True)):
        if srt:  # the sorted path needs ascending queries (duplicates allowed)
            qq = qq[po.lexsort_rows(qq)]
        idx, hit, miss = eng.lower_bound(_rows(env, book), _rows(env, qq), True, True, srt)
        exp_idx, exp_hit = po.lookup(book, qq)
        np.testing.assert_array_equal(idx.cpu().numpy(), exp_idx)
        np.testing.assert_array_equal(hit.cpu().numpy().astype(bool), exp_hit)
        assert int(miss.item()) == int((~exp_hit).sum())
    # empty phonebook and single-row phonebook
    idx, hit, _ = eng.lower_bound(_rows(env, book[:0]), _rows(env, q[:5]), True)
    np.testing.assert_array_equal(idx.cpu().numpy(), 0)
    idx, hit, _ = eng.lower_bound(_rows(env, book[3:4]), _rows(env, q), True)
    np.testing.assert_array_equal(idx.cpu().numpy(), po.lower_bound(book[3:4], q))


@pytest.mark.parametrize("dims", [[16] + [8] * 16, [64] + [16] * 64, [256] + [8] * 256,
                                  [5, 3, 7, 1, 2, 9]])
def test_pack_get_add_fields(env, dims):
    import paces_amd._hip as hip
    torch, lib = env["torch"], env["lib"]
    lay = po.Layout(dims)
    rng = np.random.default_rng(len(dims))
    raw = np.stack([rng.integers(0, d, 500) for d in lay.max_dims], axis=1)
    dev = env["dev"]
    d_off = torch.tensor(lay.offsets, dtype=torch.int32, device=dev)
    d_w = torch.tensor(lay.bitwidths, dtype=torch.int32, device=dev)
    d_m = torch.tensor(lay.max_dims, dtype=torch.int32, device=dev)
    L = hip.Layout(lay.W, lay.nfields, d_off.data_ptr(), d_w.data_ptr(), d_m.data_ptr(), None, 0)
    d_raw = torch.tensor(raw, dtype=torch.int32, device=dev)
    rows = torch.empty((500, lay.W), dtype=torch.int64, device=dev)
    hip.check(lib.paces_pack_rows(hip.ptr(d_raw), 500, C.byref(L), hip.ptr(rows), hip.stream()))
    np.testing.assert_array_equal(_np(env, rows), lay.pack(raw))
    out = torch.empty(500, dtype=torch.int32, device=dev)
    for f in range(lay.nfields):
        hip.check(lib.paces_get_field(hip.ptr(rows), 500, C.byref(L), f, hip.ptr(out),
                                      hip.stream()))
        np.testing.assert_array_equal(out.cpu().numpy(), raw[:, f])
    for f in range(lay.nfields):
        if lay.bitwidths[f] == 0:
            continue
        for delta, sel in ((+1, raw[:, f] + 1 < lay.max_dims[f]), (-1, raw[:, f] > 0)):
            sub = raw[sel]
            d = _rows(env, lay.pack(sub))
            hip.check(lib.paces_add_field(hip.ptr(d), sub.shape[0], C.byref(L), f, delta,
                                          hip.stream()))
            exp = sub.copy()
            exp[:, f] += delta
            np.testing.assert_array_equal(_np(env, d), lay.pack(exp))


@pytest.mark.parametrize("dims", [[256, 64] + [8] * 64, [64] + [16] * 64, [2, 4] + [128] * 40])
def test_fields_at_word_straddles_repeatable(env, dims):
    """Fields that straddle two uint64 words (e.g. site 16 of the 64-site batched layout, bits
    62-64) read and step identically on every call: 200k rows, values at the field limits so the
    carry/borrow crosses the word boundary, four repetitions, each bit-exact against the oracle."""
    import paces_amd._hip as hip
    torch, lib = env["torch"], env["lib"]
    lay = po.Layout(dims)
    straddle = [f for f in range(lay.nfields)
                if lay.bitwidths[f] and lay.offsets[f] // 64
                != (lay.offsets[f] + lay.bitwidths[f] - 1) // 64]
    assert straddle
    n = 200_000
    rng = np.random.default_rng(7)
    dev = env["dev"]
    d_off = torch.tensor(lay.offsets, dtype=torch.int32, device=dev)
    d_w = torch.tensor(lay.bitwidths, dtype=torch.int32, device=dev)
    d_m = torch.tensor(lay.max_dims, dtype=torch.int32, device=dev)
    L = hip.Layout(lay.W, lay.nfields, d_off.data_ptr(), d_w.data_ptr(), d_m.data_ptr(), None, 0)
    out = torch.empty(n, dtype=torch.int32, device=dev)
    for f in straddle:
        md = int(lay.max_dims[f])
        raw = np.stack([rng.integers(0, d, n) for d in lay.max_dims], axis=1)
        # half the rows sit at a boundary value whose +-1 moves bits across the two words
        edge = rng.integers(0, 2, n).astype(bool)
        raw[edge, f] = rng.choice([0, md // 2 - 1, md // 2, md - 1], int(edge.sum()))
        packed = lay.pack(raw)
        for rep in range(4):
            rows = _rows(env, packed)
            hip.check(lib.paces_get_field(hip.ptr(rows), n, C.byref(L), f, hip.ptr(out),
                                          hip.stream()))
            np.testing.assert_array_equal(out.cpu().numpy(), raw[:, f], err_msg=f"{f} {rep}")
            for delta, sel in ((+1, raw[:, f] + 1 < md), (-1, raw[:, f] > 0)):
                sub = raw[sel]
                d = _rows(env, lay.pack(sub))
                hip.check(lib.paces_add_field(hip.ptr(d), sub.shape[0], C.byref(L), f, delta,
                                              hip.stream()))
                exp = sub.copy()
                exp[:, f] += delta
                np.testing.assert_array_equal(_np(env, d), lay.pack(exp),
                                              err_msg=f"{f} {delta} {rep}")


@pytest.mark.parametrize("W,K", [(1, 5), (3, 7), (5, 11), (13, 3)])
def test_merge_unique_matches_numpy_unique(env, W, K):
    eng, hip = env["engine"], env["hip"]
    rng = np.random.default_rng(W * 100 + K)
    pool = _rand_unique_rows(rng, 30000, W, hi=9)
    runs = []
    for k in range(K):
        n = int(rng.integers(0, 6000)) if k else 5000
        if k == 2:
            n = 0
        runs.append(pool[np.sort(rng.choice(pool.shape[0], min(n, pool.shape[0]),
                                            replace=False))])
    d_runs = [(_rows(env, r) if r.shape[0] else _rows(env, pool[:1]), 0, r.shape[0])
              for r in runs]
    ws = hip.Workspace(env["dev"])
    out, from0 = eng.merge_unique(d_runs, W, ws, want_from0=True)
    exp = po.unique_rows(np.concatenate(runs))
    np.testing.assert_array_equal(_np(env, out), exp)
    _, in0 = po.lookup(runs[0], exp)
    np.testing.assert_array_equal(from0.cpu().numpy().astype(bool), in0)
    # all runs empty
    out, _ = eng.merge_unique([(d_runs[0][0], 0, 0)], W, ws)
    assert out.shape[0] == 0


@pytest.mark.parametrize("W,K", [(1, 5), (3, 5), (3, 9), (5, 3), (13, 6)])
def test_merge_unique_large(env, W, K):
    """Unions of 2^16 and more rows: rows, run-0 flags and the output index of every input row
    against NumPy. One run is a dense block of consecutive rows (a narrow key range where the
    partition table's ranges are uneven), one is empty."""
    eng, hip = env["engine"], env["hip"]
    rng = np.random.default_rng(W * 1000 + K)
    pool = _rand_unique_rows(rng, 400000, W)
    runs = []
    for k in range(K):
        if k == 1:
            a = int(rng.integers(0, pool.shape[0] - 60000))
            r = pool[a:a + 60000]  # a dense block of consecutive rows
        elif k == K - 1:
            r = pool[:0]
        else:
            n = int(rng.integers(20000, 120000))
            r = pool[np.sort(rng.choice(pool.shape[0], n, replace=False))]
        runs.append(r)
    d_runs = [(_rows(env, r) if r.shape[0] else _rows(env, pool[:1]), 0, r.shape[0])
              for r in runs]
    ws = hip.Workspace(env["dev"])
    out, from0, cpos = eng.merge_unique(d_runs, W, ws, want_from0=True, want_cand_pos=True)
    cat = np.concatenate(runs)
    exp = po.unique_rows(cat)
    assert cat.shape[0] >= 1 << 16
    np.testing.assert_array_equal(_np(env, out), exp)
    _, in0 = po.lookup(runs[0], exp)
    np.testing.assert_array_equal(from0.cpu().numpy().astype(bool), in0)
    np.testing.assert_array_equal(cpos.cpu().numpy(), po.lower_bound(exp, cat))


@pytest.mark.parametrize("R,k", [(1, 1), (10, 3), (5000, 1), (5000, 4999), (5000, 5000),
                                 (200000, 20000), (200000, 300000)])
def test_select_topk_stable_ties(env, R, k):
    torch, eng, hip = env["torch"], env["engine"], env["hip"]
    rng = np.random.default_rng(R + k)
    vec = (rng.integers(0, 5, R) * 0.25 + 1j * rng.integers(0, 3, R) * 0.5).astype(np.complex128)
    vec[rng.integers(0, R, max(1, R // 10))] *= 1e-300  # subnormal-ish keys
    rows = _rand_unique_rows(rng, R * 2, 2, hi=2 ** 40)[:R]
    assert rows.shape[0] == R
    out = eng.select_topk(torch.from_numpy(vec).to(env["dev"]), _rows(env, rows), k,
                          hip.Workspace(env["dev"]))
    np.testing.assert_array_equal(_np(env, out), rows[po.select_indices(vec, k)])


def test_reassign_and_reductions(env):
    torch, eng = env["torch"], env["engine"]
    rng = np.random.default_rng(5)
    old = _rand_unique_rows(rng, 50000, 3, hi=6)
    new = _rand_unique_rows(rng, 60000, 3, hi=6)
    vec = rng.standard_normal(old.shape[0]) + 1j * rng.standard_normal(old.shape[0])
    d_old_vec = torch.from_numpy(vec).to(env["dev"])
    out = eng.reassign(_rows(env, old), d_old_vec, _rows(env, new),
                       env["hip"].Workspace(env["dev"]))
    np.testing.assert_array_equal(out.cpu().numpy(), po.reassign(old, vec, new))
    red = eng.Reducer(env["dev"])
    assert abs(red.norm2(d_old_vec) - np.sum(np.abs(vec) ** 2)) <= 1e-12 * np.sum(np.abs(vec) ** 2)
    y = rng.standard_normal(old.shape[0]) + 1j * rng.standard_normal(old.shape[0])
    v = red.vdot(d_old_vec, torch.from_numpy(y).to(env["dev"]))
    assert abs(v - np.vdot(vec, y)) <= 1e-12 * np.sum(np.abs(vec) * np.abs(y))
    z = torch.from_numpy(y).to(env["dev"])
    eng.axpby(0.5 - 2j, d_old_vec, 1.5 + 0.25j, z)
    np.testing.assert_allclose(z.cpu().numpy(), (0.5 - 2j) * vec + (1.5 + 0.25j) * y,
                               rtol=0, atol=1e-13)


GEN_CASES = ["chain16", "ring64_small", "torus6", "square4_hop2"]


@pytest.mark.parametrize("case", GEN_CASES)
def test_expand_csr_diag_taylor_observables(env, case):
    """generate_mel on the oracle's own step-1 basis: rows, per-term CSR (FP64 values, and FP32
    values equal to their rounding), diagonal bit-exact (and its FP32 rounding);
    H.x and the Taylor series bit-exact; observables to 1e-12."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from holstein_cases import CASES, oracle_run
    torch, eng, hip = env["torch"], env["engine"], env["hip"]
    c = dict(CASES[case])
    c["steps"] = min(c["steps"], 2)
    oham, recs, t = oracle_run(c)
    from paces_amd.models import holstein
    ham = holstein.Hamiltonian(c["N"], [c["dim"]] * c["N"], geo_dim=c["geo"],
                               periodic=c["periodic"], use_terms=c["terms"])
    rec = recs[-1]
    sel = _rows(env, rec.select)
    (diag, new_inds), mats, dbg = ham.generate_mel(sel, enlarge_steps=c["enlarge"])
    np.testing.assert_array_equal(_np(env, ham.last_basis_e), rec.basis_e)
    np.testing.assert_array_equal(_np(env, new_inds), rec.whoami)
    np.testing.assert_array_equal(diag.cpu().numpy(), rec.diag)
    d32 = ham.generate_mel_diag(new_inds, torch.float32)
    np.testing.assert_array_equal(d32.cpu().numpy(), rec.diag.astype(np.float32))
    for name in oham.offdiag_terms:
        ip, ix, dv = rec.csrs[name]
        m = mats[name]
        np.testing.assert_array_equal(m.indptr.cpu().numpy(), ip)
        np.testing.assert_array_equal(m.indices.cpu().numpy(), ix)
        np.testing.assert_array_equal(m.data.cpu().numpy(), dv)
    if "vib_coupling" in oham.offdiag_terms:
        assert dbg["vib_coupling"] == rec.dbg["vib_coupling"]
    # FP32 values (the mixed-precision engine's): the same entries, each value rounded to nearest
    _, mats32, _ = ham.generate_mel(sel, enlarge_steps=c["enlarge"], value_dtype=torch.float32)
    for name in oham.offdiag_terms:
        ip, ix, dv = rec.csrs[name]
        m = mats32[name]
        assert m.data.dtype == torch.float32
        np.testing.assert_array_equal(m.indptr.cpu().numpy(), ip)
        np.testing.assert_array_equal(m.indices.cpu().numpy(), ix)
        np.testing.assert_array_equal(m.data.cpu().numpy(), dv.astype(np.float32))
    # the lookup-based CSR entry point agrees with the union-based one
    _, in_b, _ = eng.lower_bound(ham.last_basis_e, new_inds, True, False, True)
    mats2 = ham._csr(new_inds, in_b)
    for name in oham.offdiag_terms:
        np.testing.assert_array_equal(mats2[name].indptr.cpu().numpy(), rec.csrs[name][0])
        np.testing.assert_array_equal(mats2[name].indices.cpu().numpy(), rec.csrs[name][1])
        np.testing.assert_array_equal(mats2[name].data.cpu().numpy(), rec.csrs[name][2])
    csr_list = [rec.csrs[n] for n in oham.offdiag_terms]
    mat_list = [mats[n] for n in oham.offdiag_terms]
    v = rec.vec_adapted
    d_v = torch.from_numpy(v).to(env["dev"])
    hx = eng.spmv(diag, mat_list, d_v).cpu().numpy()
    np.testing.assert_array_equal(hx, po.total_ham(rec.diag, csr_list, v))
    res, dbg_g = eng.simple_taylor(diag, mat_list, d_v, float(t[-1] - t[-2]))
    exp, dbg_o = po.simple_taylor(rec.diag, csr_list, v, float(t[-1] - t[-2]))
    np.testing.assert_array_equal(res.cpu().numpy(), exp)
    assert dbg_g[:2] == dbg_o[:2] and dbg_g[2] == dbg_o[2] and dbg_g[3] == dbg_o[3]
    n_exc, n_pho, nrm2 = ham.site_observables(new_inds, torch.from_numpy(rec.vec).to(env["dev"]))
    np.testing.assert_allclose(n_exc, rec.obs["n_exc"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(n_pho, rec.obs["n_pho"], rtol=0, atol=1e-12)
    assert abs(np.sqrt(nrm2) - rec.obs["norm"]) < 1e-12


@pytest.mark.parametrize("case", ["chain16", "square4_hop2"])
def test_generate_mel_two_expansions_match_oracle(env, case):
    """Two expansions before the pairs: the second union takes the images of the rows the last
    expansion added, and the pairs take the earlier images as a separate part
    (paces_holstein_pairs_*_split). Rows, expanded basis, diagonal (FP64 and FP32, written by
    the union's pass), CSR entries and values (FP64 and FP32) and the vib ceiling count must
    equal the oracle's full-image-list result."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from holstein_cases import CASES, oracle_run
    torch = env["torch"]
    c = dict(CASES[case])
    c["steps"] = 1
    oham, recs, _ = oracle_run(c)
    from paces_amd.models import holstein
    ham = holstein.Hamiltonian(c["N"], [c["dim"]] * c["N"], geo_dim=c["geo"],
                               periodic=c["periodic"], use_terms=c["terms"])
    sel = recs[-1].select
    diag, new_inds, coo, dbg, basis_e = oham.generate_mel(sel, 2)
    R = new_inds.shape[0]
    (gdiag, gnew), mats, gdbg = ham.generate_mel(_rows(env, sel), enlarge_steps=2)
    np.testing.assert_array_equal(_np(env, ham.last_basis_e), basis_e)
    np.testing.assert_array_equal(_np(env, gnew), new_inds)
    np.testing.assert_array_equal(gdiag.cpu().numpy(), diag)
    (gd32, _), mats32, _ = ham.generate_mel(_rows(env, sel), enlarge_steps=2,
                                            value_dtype=torch.float32, diag_dtype=torch.float32)
    assert gd32.dtype == torch.float32
    np.testing.assert_array_equal(gd32.cpu().numpy(), diag.astype(np.float32))
    for name in oham.offdiag_terms:
        ip, ix, dv = po.coo_to_csr(*coo[name], R)
        for m, vals in ((mats[name], dv), (mats32[name], dv.astype(np.float32))):
            np.testing.assert_array_equal(m.indptr.cpu().numpy(), ip)
            np.testing.assert_array_equal(m.indices.cpu().numpy(), ix)
            np.testing.assert_array_equal(m.data.cpu().numpy(), vals)
    if "vib_coupling" in oham.offdiag_terms:
        assert gdbg["vib_coupling"] == dbg["vib_coupling"]
    # with no room for bounded outputs the host waits for the exact sizes: the same result
    holstein._BOUND_BYTES, keep = 0, holstein._BOUND_BYTES
    try:
        (gd2, gn2), mats2, _ = ham.generate_mel(_rows(env, sel), enlarge_steps=2)
    finally:
        holstein._BOUND_BYTES = keep
    np.testing.assert_array_equal(_np(env, gn2), new_inds)
    np.testing.assert_array_equal(gd2.cpu().numpy(), diag)
    for name in oham.offdiag_terms:
        ip, ix, dv = po.coo_to_csr(*coo[name], R)
        np.testing.assert_array_equal(mats2[name].indptr.cpu().numpy(), ip)
        np.testing.assert_array_equal(mats2[name].indices.cpu().numpy(), ix)
        np.testing.assert_array_equal(mats2[name].data.cpu().numpy(), dv)


@pytest.mark.parametrize("layout_case", ["ring64_cap15", "torus16_cap7"])
def test_generate_mel_at_word_straddling_fields(env, layout_case):
    """Bases concentrated on sites whose phonon field straddles a uint64 word boundary, with
    every occupation 0..cap-1: rows, CSR and values bit-exact against the oracle."""
    from paces_amd.models import holstein
    if layout_case == "ring64_cap15":
        N, dim, geo = 64, 16, 1
    else:
        N, dim, geo = 256, 8, 2
    terms = {"diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
             "hopping": {"J": -1.0}, "vib_coupling": {"g": 1.5}}
    oham = po.HolsteinOracle(N, [dim] * N, geo_dim=geo, use_terms=terms, periodic=True)
    lay = oham.layout
    straddle = [s for s in range(N)
                if (lay.offsets[s + 1] // 64) != ((lay.offsets[s + 1] + lay.bitwidths[s + 1] - 1)
                                                  // 64)]
    assert straddle
    rng = np.random.default_rng(11)
    raws = []
    for s in straddle:
        for n in range(dim):
            r = np.zeros(N + 1, dtype=np.int64)
            r[0] = s
            r[s + 1] = n
            raws.append(r.copy())
            r2 = r.copy()
            r2[1 + (s + 1) % N] = rng.integers(0, dim)
            r2[1 + (s - 1) % N] = rng.integers(0, dim)
            raws.append(r2)
    basis = po.unique_rows(lay.pack(np.array(raws)))
    # keep every other row so that some partners lie outside the basis
    basis = basis[::2]
    diag, new_inds, coo, dbg, _ = oham.generate_mel(basis, 0)
    R = new_inds.shape[0]
    ham = holstein.Hamiltonian(N, [dim] * N, geo_dim=geo, periodic=True, use_terms=terms)
    (gdiag, gnew), mats, gdbg = ham.generate_mel(_rows(env, basis), 0)
    np.testing.assert_array_equal(_np(env, gnew), new_inds)
    np.testing.assert_array_equal(gdiag.cpu().numpy(), diag)
    for name in oham.offdiag_terms:
        ip, ix, dv = po.coo_to_csr(*coo[name], R)
        np.testing.assert_array_equal(mats[name].indptr.cpu().numpy(), ip)
        np.testing.assert_array_equal(mats[name].indices.cpu().numpy(), ix)
        np.testing.assert_array_equal(mats[name].data.cpu().numpy(), dv)
    # phonon-at-exciton read and +-1 through the C ABI on the same rows
    occ = ham.get_phonon_at_exc(_rows(env, basis)).cpu().numpy()
    np.testing.assert_array_equal(occ, oham.get_phonon_at_exc(basis))
    dn = basis[oham.get_phonon_at_exc(basis) > 0]
    d = _rows(env, dn)
    ham.rem_phonon_at_exc(d)
    exp = dn.copy()
    oham.add_phonon_at_exc(exp, -1)
    np.testing.assert_array_equal(_np(env, d), exp)
