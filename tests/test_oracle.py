"""Pins for the CPU oracle: every check compares against something the reference's algorithm is
fixed by (tuple arithmetic, brute force on tiny lattices, closed forms, SciPy/NumPy routines),
never against a retyped formula."""

import bisect
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg

from oracle import paces_oracle as po


def _rand_raw(rng, lay, n):
    return np.stack([rng.integers(0, d, size=n) for d in lay.max_dims], axis=1)


LAYOUTS = [
    [16] + [8] * 16,        # 16-site chain, cap 7: 4 + 48 bits, one word
    [64] + [16] * 64,       # 64-site ring, cap 15: 6 + 256 bits, five words, fields straddle words
    [256] + [8] * 256,      # 16x16 torus, cap 7: 8 + 768 bits, 13 words
    [5] + [3, 7, 1, 2, 9],  # ragged widths incl. a zero-bit field
    [24] + [8] * 24,        # 24-site chain, cap 7: 77 bits
]


@pytest.mark.parametrize("dims", LAYOUTS)
def test_pack_roundtrip_and_order(dims):
    rng = np.random.default_rng(1)
    lay = po.Layout(dims)
    raw = _rand_raw(rng, lay, 300)
    rows = lay.pack(raw)
    assert rows.shape == (300, lay.W)
    np.testing.assert_array_equal(lay.unpack(rows), raw)
    # lexicographic row order == tuple order (brute force)
    order_rows = po.lexsort_rows(rows)
    tuples = [tuple(r) for r in raw.tolist()]
    order_tup = sorted(range(len(tuples)), key=lambda k: (tuples[k], k))
    assert [tuples[k] for k in order_rows] == [tuples[k] for k in order_tup]


@pytest.mark.parametrize("dims", LAYOUTS)
def test_identity_with_reference_uint32_layout(dims):
    rng = np.random.default_rng(2)
    lay = po.Layout(dims)
    raw = _rand_raw(rng, lay, 40)
    w32 = lay.pack_ref32(raw)
    np.testing.assert_array_equal(po.Layout.ref32_to_build64(w32), lay.pack(raw))


@pytest.mark.parametrize("dims", LAYOUTS)
def test_field_add_matches_tuple_arithmetic(dims):
    rng = np.random.default_rng(3)
    lay = po.Layout(dims)
    raw = _rand_raw(rng, lay, 200)
    for i in range(lay.nfields):
        if lay.bitwidths[i] == 0:
            continue
        up = raw[:, i] + 1 < lay.max_dims[i]
        rows = lay.pack(raw[up])
        lay.add_field(rows, 1, i)
        exp = raw[up].copy()
        exp[:, i] += 1
        np.testing.assert_array_equal(lay.unpack(rows), exp)
        dn = raw[:, i] > 0
        rows = lay.pack(raw[dn])
        lay.add_field(rows, -1, i)
        exp = raw[dn].copy()
        exp[:, i] -= 1
        np.testing.assert_array_equal(lay.unpack(rows), exp)


def test_unique_rows_matches_numpy_unique():
    rng = np.random.default_rng(4)
    for W in (1, 3, 5):
        rows = rng.integers(0, 4, size=(500, W)).astype(np.uint64)
        np.testing.assert_array_equal(po.unique_rows(rows), np.unique(rows, axis=0))
    assert po.unique_rows(np.zeros((0, 2), np.uint64)).shape == (0, 2)


def test_lower_bound_matches_bisect():
    rng = np.random.default_rng(5)
    for W in (1, 2, 5):
        book = po.unique_rows(rng.integers(0, 5, size=(300, W)).astype(np.uint64))
        q = rng.integers(0, 6, size=(200, W)).astype(np.uint64)
        bt = [tuple(r) for r in book.tolist()]
        exp = [bisect.bisect_left(bt, tuple(r)) for r in q.tolist()]
        np.testing.assert_array_equal(po.lower_bound(book, q), exp)
    # single-row and empty phonebooks
    book = np.array([[3, 4]], dtype=np.uint64)
    q = np.array([[3, 4], [3, 5], [2, 9], [4, 0]], dtype=np.uint64)
    np.testing.assert_array_equal(po.lower_bound(book, q), [0, 1, 0, 1])
    assert po.lower_bound(book, np.zeros((0, 2), np.uint64)).shape == (0,)
    with pytest.raises(ValueError):
        po.searchsorted_noesc(book, q)


def _tiny_ham(N, dim, geo=1, periodic=False, J=0.7, g=0.9, w=1.3, eps=0.2, de=0.0, J2=None):
    terms = {"diag": {"eps_sys": eps, "hbar_omega": w, "delta_eps": de}, "hopping": {"J": J}}
    if J2 is not None:
        terms["hopping2"] = {"J2": J2}
    terms["vib_coupling"] = {"g": g}
    return po.HolsteinOracle(N, [dim] * N, geo_dim=geo, use_terms=terms, periodic=periodic)


CASES = [
    dict(N=4, dim=3),
    dict(N=4, dim=3, de=0.3),
    dict(N=5, dim=2, J2=0.4),
    dict(N=4, dim=3, periodic=False, geo=2),
    dict(N=3, dim=3, periodic=True),
    dict(N=4, dim=2, periodic=True),
    dict(N=9, dim=2, periodic=True, geo=2),
]


@pytest.mark.parametrize("case", CASES)
def test_generate_mel_against_dense_hamiltonian(case):
    """CSR = entries of the dense H with at least one endpoint in the basis; no duplicates."""
    ham = _tiny_ham(**case)
    raw_all = po.all_states(ham)
    H = po.dense_hamiltonian(ham, raw_all)
    rows_all = ham.layout.pack(raw_all)
    rng = np.random.default_rng(6)
    for trial in range(4):
        k = max(1, int(rng.integers(1, min(60, rows_all.shape[0]))))
        pick = np.sort(rng.choice(rows_all.shape[0], size=k, replace=False))
        basis = rows_all[pick]
        diag, new_inds, coo, dbg, basis_e = ham.generate_mel(basis, 0)
        # new basis = basis plus every state reachable by one off-diagonal element
        reach = set(pick.tolist())
        for c in pick:
            reach |= set(np.nonzero(H[:, c])[0].tolist())
        reach = sorted(reach)
        np.testing.assert_array_equal(new_inds, rows_all[reach])
        R = new_inds.shape[0]
        Hs = np.diag(diag)
        for t in ham.offdiag_terms:
            ip, ix, dv = po.coo_to_csr(*coo[t], R)
            assert np.all(np.diff(ip) >= 0)
            for r in range(R):  # canonical: strictly ascending columns
                assert np.all(np.diff(ix[ip[r]:ip[r + 1]]) > 0)
            Hs = Hs + scipy.sparse.csr_matrix((dv, ix, ip), shape=(R, R)).toarray()
        sub = H[np.ix_(reach, reach)]
        inB = np.isin(reach, pick)
        touch = inB[:, None] | inB[None, :]
        exp = np.where(touch | np.eye(R, dtype=bool), sub, 0.0)
        np.testing.assert_allclose(Hs, exp, rtol=0, atol=1e-15)


def test_vib_values_and_ceiling_hits():
    ham = _tiny_ham(N=3, dim=4)
    raw = np.array([[0, 3, 0, 0], [1, 0, 2, 0], [2, 0, 0, 0]])
    basis = ham.layout.pack(raw)
    (plus, minus), dbg = ham.generate_mel_vib_coupling(basis)
    assert dbg == [1]  # the n=3 row hits the ceiling of dim 4
    np.testing.assert_array_equal(plus.vals, 0.9 * np.sqrt([3.0, 1.0]))
    np.testing.assert_array_equal(ham.layout.unpack(plus.inds), [[1, 0, 3, 0], [2, 0, 0, 1]])


def test_enlarge_equals_brute_force_neighbourhood():
    ham = _tiny_ham(N=4, dim=3)
    raw_all = po.all_states(ham)
    H = po.dense_hamiltonian(ham, raw_all)
    rows_all = ham.layout.pack(raw_all)
    pick = np.array([10, 50, 200])
    exp = set(pick.tolist())
    for c in pick:
        exp |= set(np.nonzero(H[:, c] - np.diag(np.diag(H))[:, c])[0].tolist())
    np.testing.assert_array_equal(ham.enlarge_basis_set(rows_all[pick]), rows_all[sorted(exp)])


def test_coo_to_csr_matches_scipy():
    rng = np.random.default_rng(7)
    n = 50
    r = rng.integers(0, n, 200)
    c = rng.integers(0, n, 200)
    keys = np.unique(r * n + c)
    r, c = keys // n, keys % n
    v = rng.standard_normal(r.shape[0])
    ip, ix, dv = po.coo_to_csr(v, r, c, n)
    m = scipy.sparse.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    m.sort_indices()
    np.testing.assert_array_equal(ip, m.indptr)
    np.testing.assert_array_equal(ix, m.indices)
    np.testing.assert_array_equal(dv, m.data)


def test_total_ham_matches_dense_and_taylor_matches_expm():
    ham = _tiny_ham(N=4, dim=3)
    raw_all = po.all_states(ham)
    rows = ham.layout.pack(raw_all)
    diag, new_inds, coo, _, _ = ham.generate_mel(rows, 0)
    np.testing.assert_array_equal(new_inds, rows)
    R = rows.shape[0]
    csrs = [po.coo_to_csr(*coo[t], R) for t in ham.offdiag_terms]
    H = po.dense_hamiltonian(ham, raw_all)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(R) + 1j * rng.standard_normal(R)
    np.testing.assert_allclose(po.total_ham(diag, csrs, x), H @ x, rtol=0, atol=1e-12)
    x /= np.linalg.norm(x)
    res, dbg = po.simple_taylor(diag, csrs, x, 0.05)
    exact = scipy.linalg.expm(-1j * 0.05 * H) @ x
    assert dbg[0] and dbg[3] <= po.TAYLOR_TOL
    np.testing.assert_allclose(res, exact, rtol=0, atol=1e-13)
    Hs = scipy.sparse.csr_matrix(H)
    np.testing.assert_allclose(res, scipy.sparse.linalg.expm_multiply(-1j * 0.05 * Hs, x),
                               rtol=0, atol=1e-13)
    lz = po.lanczos(diag, csrs, x, 0.05, 12)
    np.testing.assert_allclose(lz, exact, rtol=0, atol=1e-11)


def test_taylor_coeff_is_correctly_rounded_quotient():
    for dt in (0.05, 0.02, 0.1 + 0.2):
        for j in range(30):
            assert po.taylor_coeff(dt, j) == dt / (j + 1)


def test_select_matches_stable_rank():
    rng = np.random.default_rng(9)
    v = (rng.integers(0, 4, 300) + 1j * rng.integers(0, 3, 300)).astype(np.complex128)
    for ms in (1, 17, 150, 299, 300, 500):
        kept = po.select_indices(v, ms)
        key = (v.real ** 2 + v.imag ** 2)
        # brute force: element i is kept iff fewer than ms elements rank above it, where
        # "above" means larger key, or equal key and larger index
        exp = [i for i in range(300)
               if sum(1 for j in range(300) if key[j] > key[i] or (key[j] == key[i] and j > i)) < ms]
        np.testing.assert_array_equal(kept, exp)


def test_reassign_definition():
    rng = np.random.default_rng(10)
    old = po.unique_rows(rng.integers(0, 6, size=(80, 2)).astype(np.uint64))
    new = po.unique_rows(rng.integers(0, 6, size=(90, 2)).astype(np.uint64))
    vec = rng.standard_normal(old.shape[0]) + 1j * rng.standard_normal(old.shape[0])
    out = po.reassign(old, vec, new)
    od = {tuple(r): vec[k] for k, r in enumerate(old.tolist())}
    exp = [od.get(tuple(r), 0) for r in new.tolist()]
    np.testing.assert_array_equal(out, exp)


def _exact_timeline(ham, pos0, t_array):
    raw_all = po.all_states(ham)
    H = po.dense_hamiltonian(ham, raw_all)
    psi = np.zeros(raw_all.shape[0], complex)
    start = np.zeros(ham.nchain + 1, dtype=np.int64)
    start[0] = pos0
    psi[np.nonzero(np.all(raw_all == start, axis=1))[0][0]] = 1
    out = []
    for t in t_array:
        phi = scipy.linalg.expm(-1j * t * H) @ psi
        w = np.abs(phi) ** 2
        out.append((np.bincount(raw_all[:, 0], weights=w, minlength=ham.nchain),
                    np.array([np.sum(w * raw_all[:, i + 1]) for i in range(ham.nchain)])))
    return out


@pytest.mark.parametrize("periodic", [False, True])
def test_untruncated_timeline_equals_exact_evolution(periodic):
    ham = _tiny_ham(N=3, dim=3, periodic=periodic, J=-1.0, g=0.8, w=1.0, eps=0.0)
    full = po.all_states(ham).shape[0]
    te = po.TimeEvolutionOracle(ham, maxstates=full, enlarge_steps=1)
    te.whoami = ham.layout.pack(po.all_states(ham))  # the whole space: the step is exact
    raw = np.zeros((1, 4), dtype=np.int64)
    raw[0, 0] = 1
    te.create_initial_vector([1.0], raw)
    t = 0.05 * np.arange(21)
    recs = []
    te.generate_timeline(t, recs.append)
    exact = _exact_timeline(ham, 1, t)
    e0 = recs[1].post_energy
    for rec, (nx, npho) in zip(recs, exact):
        np.testing.assert_allclose(rec.obs["n_exc"], nx, atol=1e-12)
        np.testing.assert_allclose(rec.obs["n_pho"], npho, atol=1e-12)
        assert abs(rec.obs["n_exc"].sum() - rec.obs["norm"] ** 2) < 1e-13
        assert abs(rec.post_energy - e0) < 1e-12


def test_free_exciton_closed_form():
    """g=0: OBC chain modes sin(k pi (j+1)/(N+1)), energies 2J cos(k pi/(N+1))."""
    N, J, j0 = 9, -1.0, 4
    ham = po.HolsteinOracle(N, [2] * N, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
        "hopping": {"J": J}, "vib_coupling": {"g": 0.0}})
    te = po.TimeEvolutionOracle(ham, maxstates=10 ** 6, enlarge_steps=1)
    te.create_single_state_basis(j0)
    te.grow_optimal_basis(n_max=3, fillfac=1.0)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = j0
    te.create_initial_vector([1.0], raw)
    t = 0.1 * np.arange(16)
    recs = []
    te.generate_timeline(t, recs.append)
    k = np.arange(1, N + 1)
    phi = np.sqrt(2.0 / (N + 1)) * np.sin(np.outer(np.arange(1, N + 1), k) * np.pi / (N + 1))
    E = 2 * J * np.cos(k * np.pi / (N + 1))
    for rec, tt in zip(recs, t):
        amp = phi @ (phi[j0] * np.exp(-1j * E * tt))
        np.testing.assert_allclose(rec.obs["n_exc"], np.abs(amp) ** 2, atol=1e-12)
        # mirror symmetry of the centred odd chain
        np.testing.assert_allclose(rec.obs["n_exc"], rec.obs["n_exc"][::-1], atol=1e-12)


def test_independent_boson_closed_form():
    """J=0: <n_pos>(t) = 2 (g/w)^2 (1 - cos w t) for the displaced oscillator."""
    N, g, w = 3, 0.4, 1.0
    ham = po.HolsteinOracle(N, [16] * N, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": w, "delta_eps": 0.0},
        "hopping": {"J": 0.0}, "vib_coupling": {"g": g}})
    te = po.TimeEvolutionOracle(ham, maxstates=10 ** 6, enlarge_steps=1)
    te.create_single_state_basis(1)
    te.grow_optimal_basis(fillfac=1.0, n_max=20)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = 1
    te.create_initial_vector([1.0], raw)
    t = 0.25 * np.arange(26)
    recs = []
    te.generate_timeline(t, recs.append)
    for rec, tt in zip(recs, t):
        assert abs(rec.obs["n_pho"][1] - 2 * (g / w) ** 2 * (1 - math.cos(w * tt))) < 1e-11
        assert abs(rec.obs["n_exc"][1] - 1.0) < 1e-13


def test_grow_optimal_basis_bfs_ball_and_quirks():
    ham = _tiny_ham(N=6, dim=3)
    te = po.TimeEvolutionOracle(ham, maxstates=100, enlarge_steps=0)
    te.create_single_state_basis(2)
    n = te.grow_optimal_basis(fillfac=1.0)
    # BFS ball sizes by brute force over the dense graph
    raw_all = po.all_states(ham)
    H = po.dense_hamiltonian(ham, raw_all)
    A = (H - np.diag(np.diag(H))) != 0
    start = np.zeros(7, dtype=np.int64)
    start[0] = 2
    ball = {int(np.nonzero(np.all(raw_all == start, axis=1))[0][0])}
    sizes = [1]
    while len(ball) <= 100 * 10:
        nb = set(ball)
        for c in ball:
            nb |= set(np.nonzero(A[:, c])[0].tolist())
        ball = nb
        sizes.append(len(ball))
    # the kept basis is the last ball not above maxstates*fillfac; n enlargements were made
    assert sizes[n - 1] == te.whoami.shape[0] <= 100 < sizes[n]
    te2 = po.TimeEvolutionOracle(ham, maxstates=100, enlarge_steps=0)
    te2.create_single_state_basis(2)
    te2.grow_optimal_basis(fillfac=1.0, n_max=2)  # n_max binds: keeps the ball after 1 step
    assert te2.whoami.shape[0] == sizes[1]
    te3 = po.TimeEvolutionOracle(ham, maxstates=0, enlarge_steps=0)
    te3.create_single_state_basis(2)
    with pytest.raises(UnboundLocalError):
        te3.grow_optimal_basis(fillfac=1.0)


def test_explosion_cutoff_raises():
    ham = _tiny_ham(N=3, dim=2)
    te = po.TimeEvolutionOracle(ham, maxstates=100, enlarge_steps=0, explosion_cutoff=0.5)
    te.create_single_state_basis(1)
    te.grow_optimal_basis(fillfac=1.0, n_max=2)
    raw = np.zeros((1, 4), dtype=np.int64)
    raw[0, 0] = 1
    te.create_initial_vector([1.0], raw)
    with pytest.raises(RuntimeError):
        te.generate_timeline(np.array([0.0, 0.05]))


def _full_space_csrs(ham):
    """Every state of a tiny lattice, its packed rows and the per-term CSR of the whole space
    (the full space expanded once is itself)."""
    raw = po.all_states(ham)
    rows = ham.layout.pack(raw)
    diag, new_inds, coo, _, _ = ham.generate_mel(rows, 0)
    np.testing.assert_array_equal(new_inds, rows)
    R = rows.shape[0]
    return raw, rows, {t: po.coo_to_csr(*coo[t], R) for t in ham.offdiag_terms}


def _bond_matrix(ham, raw, a, b):
    """Dense hopping restricted to the bond {a, b}: H entries between states that differ only
    by the exciton moving between sites a and b."""
    H = po.dense_hamiltonian(ham, raw)
    n = raw.shape[0]
    M = np.zeros((n, n))
    for k in range(n):
        for j in np.nonzero(H[:, k])[0]:
            if j != k and {raw[j, 0], raw[k, 0]} == {a, b} and np.array_equal(raw[j, 1:],
                                                                               raw[k, 1:]):
                M[j, k] = H[j, k]
    return M


@pytest.mark.parametrize("N,periodic", [(5, False), (4, False), (5, True), (3, True)])
def test_hopping_bonds_equal_brute_force_bond_energies(N, periodic):
    """calculate_hopping's board formula gives <psi| T_bond |psi> for each bond (brute force
    from the dense hopping matrix)."""
    ham = po.HolsteinOracle(N, [2] * N, periodic=periodic, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
        "hopping": {"J": -0.7}, "vib_coupling": {"g": 0.9}})
    raw, rows, csrs = _full_space_csrs(ham)
    hop = po.HolsteinOracle(N, [2] * N, periodic=periodic, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 0.0, "delta_eps": 0.0}, "hopping": {"J": -0.7}})
    rng = np.random.default_rng(N)
    psi = rng.normal(size=rows.shape[0]) + 1j * rng.normal(size=rows.shape[0])
    bonds = po.hopping_bonds(ham, rows, psi, csrs["hopping"])
    nb = N if periodic else N - 1
    want = np.array([np.vdot(psi, _bond_matrix(hop, raw, b, (b + 1) % N) @ psi)
                     for b in range(nb)])
    np.testing.assert_allclose(bonds, want, rtol=1e-12, atol=1e-12)
    v = po.per_position_products(ham, rows, psi, csrs["hopping"])
    assert abs(np.sum(bonds) - np.sum(v)) < 1e-9  # the bonds add up to <psi|T|psi>


def test_per_position_coupling_and_vacuum_sum_brute_force():
    """Per-position vib products equal <psi| P_p V |psi> from the dense matrices; the vacuum
    amplitude sum equals a sum over the unpacked all-zero phonon tuples."""
    ham = _tiny_ham(N=3, dim=3)
    raw, rows, csrs = _full_space_csrs(ham)
    vib = po.HolsteinOracle(3, [3] * 3, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 0.0, "delta_eps": 0.0}, "vib_coupling": {"g": 0.9}})
    V = po.dense_hamiltonian(vib, raw)
    rng = np.random.default_rng(4)
    psi = rng.normal(size=rows.shape[0]) + 1j * rng.normal(size=rows.shape[0])
    got = po.per_position_products(ham, rows, psi, csrs["vib_coupling"])
    for p in range(3):
        P = np.diag((raw[:, 0] == p).astype(float))
        assert abs(got[p] - np.vdot(psi, P @ V @ psi)) < 1e-12
    want = sum(psi[k] for k in range(raw.shape[0]) if not raw[k, 1:].any())
    assert abs(po.vacuum_amplitude_sum(ham, rows, psi) - want) < 1e-14


def test_initial_basis_generators_hand_examples():
    """The initial-basis generators on examples worked by hand."""
    raw = po.nonuniform_basis_tuples(3, [2, 1, 3], lowest=[1, 4, 0], minpos=1, maxpos=3)
    want = [(p, a, 0, b) for p in (1, 2) for a in (1, 2) for b in (0, 1, 2)]
    assert [tuple(r) for r in raw.tolist()] == want  # truncate 1 leaves the site at 0
    # maxval 2, super_sigma huge (height 3 everywhere), sigma 1 on 3 sites:
    # position 0: ceil(3 * [1, e^-.5, e^-2]) = [3, 2, 1]; position 1: [2, 3, 2]; 2: [1, 2, 3]
    raw = po.moving_gaussian_tuples(3, 1.0, 1e9, 1.0, 2)
    counts = [int(np.sum(raw[:, 0] == p)) for p in range(3)]
    assert counts == [6, 12, 6]
    assert raw[:, 1:].max(axis=0).tolist() == [2, 2, 2]
    blk = raw[raw[:, 0] == 1][:, 1:]
    assert sorted(map(tuple, blk.tolist())) == sorted(
        (a, b, c) for a in range(2) for b in range(3) for c in range(2))
    assert [tuple(r) for r in raw.tolist()] == sorted(map(tuple, raw.tolist()))


def test_tensor_state_equals_kronecker_product():
    """The tensor-product state's coefficients on its tuples equal the Kronecker product of the
    one-site states |b> = sum_j c_j |b_j> (vacuum off the chosen sites) in the full space."""
    N, dim = 3, 3
    ham = po.HolsteinOracle(N, [dim] * N, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0}, "hopping": {"J": 1.0}})
    bst, cf = [0, 2], np.array([0.6, 0.8j])
    coef, raw = po.tensor_state(N, bst, cf, excsite=1, sites=[0, 2])
    local = np.zeros(dim, dtype=complex)
    local[bst] = cf
    vac = np.zeros(dim)
    vac[0] = 1.0
    exc = np.zeros(N)
    exc[1] = 1.0
    full = np.kron(exc, np.kron(local, np.kron(vac, local)))
    allraw = po.all_states(ham)
    index = {tuple(r): k for k, r in enumerate(allraw.tolist())}
    got = np.zeros(allraw.shape[0], dtype=complex)
    for c, r in zip(coef, raw.tolist()):
        got[index[tuple(r)]] += c
    np.testing.assert_allclose(got, full, rtol=0, atol=1e-15)
    assert abs(np.linalg.norm(coef) - 1.0) < 1e-15


def test_pool_chain_worker_equals_timeline():
    """oracle.pool.run_chain (the CPU baseline's worker) runs the oracle's own timeline: the
    same site densities at the recorded steps and min(maxstates, R_prev) state-steps per step."""
    from oracle import pool
    N, J, g, ms, dt, nsteps = 8, 1.0, 0.8, 150, 0.05, 6
    out, units, secs = pool.run_chain((N, J, g, ms, dt, nsteps, [3, 6]))
    ham = po.HolsteinOracle(N, [8] * N, use_terms=pool.chain_terms(J, g))
    te = po.TimeEvolutionOracle(ham, ms, 1)
    te.create_single_state_basis(N // 2)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = N // 2
    te.create_initial_vector([1.0], raw)
    recs = []
    te.generate_timeline(dt * np.arange(nsteps + 1), recs.append)
    assert sorted(out) == [3, 6]
    for i in (3, 6):
        np.testing.assert_array_equal(out[i], recs[i].obs["n_exc"])
    assert units == sum(min(recs[i - 1].whoami.shape[0], ms) for i in range(1, nsteps + 1))
    assert secs > 0.0
