"""CPU oracle for the paces adaptive-basis time step -- TEST INFRASTRUCTURE ONLY.

This module is a NumPy restatement of the reference algorithm (rkevk/paces v1.2.0). It is the
checker for the HIP path and the CPU baseline of the benchmark; the product package
(``paces_amd``) never imports it. Each function cites the reference file:line it restates.

Conventions shared with the HIP path (the "build layout"):

* A basis state is a tuple ``(pos, n_0, ..., n_{N-1})``. Field ``i`` has width
  ``b_i = ceil(log2(max_dims[i]))`` and occupies bits ``[sum(b_<i), sum(b_<=i))`` counted from
  the most significant bit of word 0 of a big-endian concatenation
  (``paces/core/hamiltonian.py:218-248``). The build stores that bit string in uint64 words
  (the reference uses uint32 words); build word ``k`` equals ``(w32[2k] << 32) | w32[2k+1]``,
  so the lexicographic row order is the same.
* The ranking key of the top-``maxstates`` select is ``re*re + im*im`` evaluated without FMA.
  The reference ranks by ``cupy.abs`` (hypot); both orders agree except where rounding of
  the square root merges two keys into a tie.
* Sparse products sum each row sequentially in ascending column order without FMA, term by term
  in ``use_terms`` order after the diagonal product (``paces/core/time_evolution.py:397-400``).
* ``lower_bound`` is lexicographic for every phonebook size; the reference's one-row shortcut
  (``paces/aux/cupy_search.py:105-110``) is not lexicographic, but every caller on the step
  verifies the hit by row equality, so results are unaffected.

Parity status of each function is recorded in its docstring ("pinned by ...").
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

U64 = np.uint64
TAYLOR_TOL = 1.1102230246251565e-16  # 2**-53, paces/core/time_evolution.py:484


# ------------------------------------------------------------------------------------------------
# A1/A2/A3/A4: bit layout, pack/unpack, field read/add
# ------------------------------------------------------------------------------------------------

class Layout:
    """Packed-row layout (paces/core/hamiltonian.py:134-160, paces/models/holstein.py:42-66).

    Pinned by: pack/unpack round trips, row order == tuple order (brute force), and identity
    with the reference's uint32 layout restated in ``pack_ref32``.
    """

    def __init__(self, max_dims):
        self.max_dims = np.asarray(max_dims, dtype=np.int64)
        if np.any(self.max_dims < 1):
            raise ValueError("every dimension must be >= 1")
        # ceil(log2(d)), computed exactly on integers (hamiltonian.py:155)
        self.bitwidths = np.array([(int(d) - 1).bit_length() for d in self.max_dims],
                                  dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.bitwidths)[:-1]]).astype(np.int64)
        self.totbits = int(self.bitwidths.sum())
        self.nfields = len(self.max_dims)
        self.W = max(1, -(-self.totbits // 64))            # uint64 words per row
        self.W32 = max(1, -(-self.totbits // 32))          # reference uint32 words per row

    # -- pack / unpack --------------------------------------------------------------------------
    def pack(self, raw):
        """(n, nfields) small ints -> (n, W) uint64 rows (hamiltonian.py:218-248)."""
        raw = np.asarray(raw)
        if raw.ndim != 2 or raw.shape[1] != self.nfields:
            raise ValueError("raw states must have shape (n, nfields)")
        if np.any(raw < 0) or np.any(raw >= self.max_dims[None, :]):
            raise ValueError("field value out of range")
        n = raw.shape[0]
        out = np.zeros((n, self.W), dtype=U64)
        for i in range(self.nfields):
            b = int(self.bitwidths[i])
            if b == 0:
                continue
            self._or_field(out, raw[:, i].astype(U64), int(self.offsets[i]), b)
        return out

    @staticmethod
    def _or_field(out, vals, off, b):
        k, o = off // 64, off % 64
        end = o + b
        if end <= 64:
            out[:, k] |= vals << U64(64 - end)
        else:
            spill = end - 64
            out[:, k] |= vals >> U64(spill)
            out[:, k + 1] |= (vals << U64(64 - spill)) & U64(0xFFFFFFFFFFFFFFFF)

    def get_field(self, rows, i):
        """Value of field i for every row (hamiltonian.py:291-318)."""
        rows = np.asarray(rows, dtype=U64)
        b = int(self.bitwidths[i])
        if b == 0:
            return np.zeros(rows.shape[0], dtype=np.int64)
        off = int(self.offsets[i])
        k, o = off // 64, off % 64
        end = o + b
        mask = U64((1 << b) - 1)
        if end <= 64:
            v = (rows[:, k] >> U64(64 - end)) & mask
        else:
            spill = end - 64
            v = ((rows[:, k] << U64(spill)) | (rows[:, k + 1] >> U64(64 - spill))) & mask
        return v.astype(np.int64)

    def unpack(self, rows):
        """(n, W) rows -> (n, nfields) int64 (hamiltonian.py:210-215)."""
        rows = np.asarray(rows, dtype=U64)
        out = np.zeros((rows.shape[0], self.nfields), dtype=np.int64)
        for i in range(self.nfields):
            out[:, i] = self.get_field(rows, i)
        return out

    def add_field(self, rows, delta, i):
        """Add ``delta`` (int or per-row array) to field i in place (hamiltonian.py:321-360).

        Carries and borrows cross word boundaries (the reference implements the straddling case
        only for 8-bit words and in the phonon kernels, paces/models/_kernels/_holstein_kernels.py
        :96-149). Callers guarantee the result stays inside the field.
        """
        if rows.shape[0] == 0:
            return
        vals = self.get_field(rows, i) + np.asarray(delta, dtype=np.int64)
        if np.any(vals < 0) or np.any(vals >= (1 << int(self.bitwidths[i]))):
            raise ValueError("field over/underflow")
        self.set_field(rows, i, vals)

    def set_field(self, rows, i, vals):
        b = int(self.bitwidths[i])
        if b == 0:
            return
        off = int(self.offsets[i])
        k, o = off // 64, off % 64
        end = o + b
        vals = np.asarray(vals).astype(U64)
        full = U64(0xFFFFFFFFFFFFFFFF)
        if end <= 64:
            m = U64(((1 << b) - 1) << (64 - end))
            rows[:, k] = (rows[:, k] & (m ^ full)) | (vals << U64(64 - end))
        else:
            spill = end - 64
            m0 = U64((1 << (b - spill)) - 1)
            m1 = U64(((1 << spill) - 1) << (64 - spill))
            rows[:, k] = (rows[:, k] & (m0 ^ full)) | (vals >> U64(spill))
            rows[:, k + 1] = (rows[:, k + 1] & (m1 ^ full)) | ((vals << U64(64 - spill)) & full)

    # -- reference uint32 layout (for the identity check only) ----------------------------------
    def pack_ref32(self, raw):
        """Reference uint32 packing (hamiltonian.py:226-241), restated with Python integers."""
        raw = np.asarray(raw)
        out = np.zeros((raw.shape[0], self.W32), dtype=np.uint32)
        for r in range(raw.shape[0]):
            big = 0
            for i in range(self.nfields):
                big = (big << int(self.bitwidths[i])) | int(raw[r, i])
            big <<= 32 * self.W32 - self.totbits
            for k in range(self.W32):
                out[r, k] = (big >> (32 * (self.W32 - 1 - k))) & 0xFFFFFFFF
        return out

    @staticmethod
    def ref32_to_build64(w32):
        """Build word k = (w32[2k] << 32) | w32[2k+1], missing words = 0."""
        w32 = np.asarray(w32, dtype=np.uint64)
        n, m = w32.shape
        W = -(-m // 2)
        pad = np.zeros((n, 2 * W), dtype=np.uint64)
        pad[:, :m] = w32
        return (pad[:, 0::2] << U64(32)) | pad[:, 1::2]


# ------------------------------------------------------------------------------------------------
# A7 / B1: row sort + dedup, lexicographic lower_bound
# ------------------------------------------------------------------------------------------------

def _void_view(rows):
    rows = np.ascontiguousarray(np.asarray(rows, dtype=U64))
    W = rows.shape[1]
    return np.ascontiguousarray(rows.astype(">u8")).view(f"V{8 * W}").ravel()


def lexsort_rows(rows):
    """Permutation sorting rows lexicographically (the cupy.lexsort calls, helpers.py:71)."""
    rows = np.asarray(rows, dtype=U64)
    if rows.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    return np.lexsort(rows.T[::-1])


def unique_rows(rows):
    """Sorted unique rows (cupy_unique, paces/aux/helpers.py:63-75).

    Pinned by: numpy.unique(axis=0) on the same rows (library routine).
    """
    rows = np.asarray(rows, dtype=U64)
    if rows.shape[0] == 0:
        return rows.copy()
    s = rows[lexsort_rows(rows)]
    keep = np.empty(s.shape[0], dtype=bool)
    keep[0] = True
    keep[1:] = np.any(s[1:] != s[:-1], axis=1)
    return s[keep]


def lower_bound(book, q):
    """Lexicographic lower_bound of each query row in the sorted phonebook.

    Restates searchsorted_full_binary (paces/aux/cupy_search.py:78-93, kernel 149-202): result in
    [0, M], left = m+1 when book[m] < q. Pinned by: bisect.bisect_left on tuples (textbook).
    """
    book = np.asarray(book, dtype=U64)
    q = np.asarray(q, dtype=U64)
    if q.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    if book.shape[0] == 0:
        return np.zeros(q.shape[0], dtype=np.int64)
    return np.searchsorted(_void_view(book), _void_view(q), side="left").astype(np.int64)


def lookup(book, q):
    """lower_bound plus the row-equality hit test (time_evolution.py:783-785)."""
    idx = lower_bound(book, q)
    M = book.shape[0]
    hit = idx < M
    if np.any(hit):
        hit[hit] = np.all(book[idx[hit]] == q[hit], axis=1)
    return idx, hit


def searchsorted_noesc(book, q):
    """lower_bound that raises when a row is absent (full_binary_noesc, cupy_search.py:205-228)."""
    idx, hit = lookup(book, q)
    if not np.all(hit):
        raise ValueError("row not found in phonebook")
    return idx


# ------------------------------------------------------------------------------------------------
# Holstein model: terms, diagonal, H.c. + COO, CSR (C1-C3, D1, D2, A6, A8)
# ------------------------------------------------------------------------------------------------

@dataclass
class MelTriple:
    """(inds, vals, mask) as in paces/core/hamiltonian.py:27-42."""
    inds: np.ndarray
    vals: object
    mask: np.ndarray


class HolsteinOracle:
    """Holstein Hamiltonian (paces/models/holstein.py:22-287) on the build layout.

    ``periodic=True`` extends the reference (which raises at holstein.py:183-184, 193-194) with
    wrap-around hopping for rings/tori with lattice length >= 3; pinned by brute force against
    the dense Hamiltonian.
    """

    def __init__(self, nchain, max_ho_dims, geo_dim=1, use_terms=None, periodic=False):
        if nchain != len(max_ho_dims):
            raise ValueError("nchain should match the number of QHOs.")
        self.nchain = int(nchain)
        self.max_ho_dims = np.asarray(max_ho_dims, dtype=np.int64)
        self.geo_dim = int(geo_dim)
        self.lattice_l = int(round(nchain ** (1.0 / geo_dim)))
        if self.lattice_l ** geo_dim != nchain:
            raise ValueError(f"Number of sites {nchain} is incompatible with dimension {geo_dim}.")
        self.periodic = bool(periodic)
        if self.periodic and self.lattice_l < 3:
            raise ValueError("periodic hopping needs a lattice length >= 3")
        self.layout = Layout(np.concatenate([[nchain], self.max_ho_dims]))
        self.use_terms = use_terms if use_terms is not None else {
            "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
            "hopping": {"J": 1.0}, "vib_coupling": {"g": 1.0}}
        self.offdiag_terms = [k for k in self.use_terms if k != "diag"]
        self.W = self.layout.W

    # -- field helpers ----------------------------------------------------------------------------
    def get_pos(self, rows):
        return self.layout.get_field(rows, 0)

    def get_phonon_at_exc(self, rows):
        """n at the exciton site (_holstein_kernels.py:153-179)."""
        pos = self.get_pos(rows)
        out = np.zeros(rows.shape[0], dtype=np.int64)
        for p in np.unique(pos):
            sel = pos == p
            out[sel] = self.layout.get_field(rows[sel], int(p) + 1)
        return out

    def add_phonon_at_exc(self, rows, delta):
        """+-1 phonon at the exciton site in place (_holstein_kernels.py:96-149)."""
        pos = self.get_pos(rows)
        for p in np.unique(pos):
            sel = np.nonzero(pos == p)[0]
            sub = rows[sel]
            self.layout.add_field(sub, delta, int(p) + 1)
            rows[sel] = sub

    def sum_all_phonons(self, rows):
        """Sum of occupations per row (sum_occ_pte_kernel, _holstein_kernels.py:183-210)."""
        s = np.zeros(rows.shape[0], dtype=np.float64)
        for i in range(self.nchain):
            s += self.layout.get_field(rows, i + 1).astype(np.float64)
        return s

    # -- term images --------------------------------------------------------------------------------
    def _hop_images(self, rows, order_base, cur_dim):
        """Plus/minus hop images of the rows (holstein.py:177-199) + periodic extension.

        Returns (plus_mask, plus_pos, minus_mask, minus_pos) with target positions.
        """
        L = self.lattice_l
        if order_base < 1 or order_base > L - 1:
            raise ValueError("Invalid value for base_order of hopping operator.")
        pos = self.get_pos(rows)
        span = L ** (cur_dim + 1)
        order = order_base * L ** cur_dim
        mod = pos % span
        base = pos - mod
        if self.periodic:
            plus_mask = np.ones(pos.shape[0], dtype=bool)
            minus_mask = np.ones(pos.shape[0], dtype=bool)
            plus_pos = base + (mod + order) % span
            minus_pos = base + (mod - order) % span
        else:
            plus_mask = mod < (span - order)
            minus_mask = mod > order - 1
            plus_pos = pos + order
            minus_pos = pos - order
        return plus_mask, plus_pos, minus_mask, minus_pos

    def _with_pos(self, rows, newpos):
        out = rows.copy()
        self.layout.set_field(out, 0, newpos)
        return out

    def o_star_mask(self, basis, plus_inds):
        """mask[i] = basis[i] not in plus_inds (hamiltonian.py:363-374)."""
        if plus_inds.shape[0] == 0:
            return np.ones(basis.shape[0], dtype=bool)
        sortplus = plus_inds[lexsort_rows(plus_inds)]
        _, hit = lookup(sortplus, basis)
        return ~hit

    def _generate_mel_hopping(self, basis, t_sys, base_order=1, cur_dim=0, raw_map_to=False):
        """holstein.py:151-207."""
        pm, ppos, mm, mpos = self._hop_images(basis, base_order, cur_dim)
        plus_inds = self._with_pos(basis[pm], ppos[pm])
        if raw_map_to:
            return [plus_inds, self._with_pos(basis[mm], mpos[mm])]
        mm = mm & self.o_star_mask(basis, plus_inds)
        minus_inds = self._with_pos(basis[mm], mpos[mm])
        return [MelTriple(plus_inds, t_sys, pm), MelTriple(minus_inds, np.conj(t_sys), mm)]

    def _hopping_dim_wrapper(self, basis, t_sys, order, raw_map_to):
        """holstein.py:210-222."""
        res = self._generate_mel_hopping(basis, t_sys, order, 0, raw_map_to)
        for d in range(1, self.geo_dim):
            res += self._generate_mel_hopping(basis, t_sys, order, d, raw_map_to)
        if raw_map_to:
            return res
        return res, None

    def generate_mel_hopping(self, basis, raw_map_to=False):
        return self._hopping_dim_wrapper(basis, self.use_terms["hopping"]["J"], 1, raw_map_to)

    def generate_mel_hopping2(self, basis, raw_map_to=False):
        return self._hopping_dim_wrapper(basis, self.use_terms["hopping2"]["J2"], 2, raw_map_to)

    def generate_mel_vib_coupling(self, basis, raw_map_to=False):
        """holstein.py:242-277."""
        occ = self.get_phonon_at_exc(basis)
        pos = self.get_pos(basis)
        plus_mask = occ < (self.max_ho_dims[pos] - 1)
        plus_inds = basis[plus_mask].copy()
        self.add_phonon_at_exc(plus_inds, +1)
        minus_mask = occ > 0
        if raw_map_to:
            minus_inds = basis[minus_mask].copy()
            self.add_phonon_at_exc(minus_inds, -1)
            return [plus_inds, minus_inds]
        ceiling_hits = int(basis.shape[0] - plus_mask.sum())
        minus_mask = minus_mask & self.o_star_mask(basis, plus_inds)
        minus_inds = basis[minus_mask].copy()
        self.add_phonon_at_exc(minus_inds, -1)
        g = self.use_terms["vib_coupling"]["g"]
        plus_vals = np.sqrt((occ[plus_mask] + 1).astype(np.float64)) * g
        minus_vals = np.sqrt(occ[minus_mask].astype(np.float64)) * np.conj(g)
        return ([MelTriple(plus_inds, plus_vals, plus_mask),
                 MelTriple(minus_inds, minus_vals, minus_mask)], [ceiling_hits])

    def generate_mel_diag(self, rows):
        """holstein.py:280-287: eps + hbar_omega * sum(n) (+ delta_eps * pos)."""
        p = self.use_terms["diag"]
        d = p["eps_sys"] + p["hbar_omega"] * self.sum_all_phonons(rows)
        if p.get("delta_eps", 0.0) != 0:
            d = d + p["delta_eps"] * self.get_pos(rows).astype(np.float64)
        return np.asarray(d, dtype=np.float64)

    # -- expansion and matrix elements --------------------------------------------------------------
    def raw_images(self, basis):
        out = []
        for term in self.offdiag_terms:
            out += list(getattr(self, "generate_mel_" + term)(basis, raw_map_to=True))
        return out

    def enlarge_basis_set(self, basis):
        """hamiltonian.py:550-577: unique(basis + raw images)."""
        return unique_rows(np.concatenate([basis] + self.raw_images(basis), axis=0))

    def generate_mel(self, basis, enlarge_steps=0):
        """hamiltonian.py:383-492.

        Returns (diag, new_inds, coo_dict, dbg, basis_e) where coo_dict[term] = (vals, rows, cols).
        """
        if basis.shape[1] != self.W:
            raise ValueError("Shape of basis states does not match the specified number of bits.")
        for _ in range(enlarge_steps):
            basis = self.enlarge_basis_set(basis)
        melpacks, dbg = {}, {}
        parts = [basis]
        for term in self.offdiag_terms:
            melpack, dbg_terms = getattr(self, "generate_mel_" + term)(basis)
            if dbg_terms is not None:
                dbg[term] = dbg_terms
            melpacks[term] = melpack
            parts += [m.inds for m in melpack]
        new_inds = unique_rows(np.concatenate(parts, axis=0))
        if new_inds.shape[0] >= 2 ** 31:
            raise ValueError("The number of states is too large for signed 32-bit-integer indices.")
        diag = self.generate_mel_diag(new_inds)
        basis_lookup = searchsorted_noesc(new_inds, basis)
        coo = {term: self._add_hc_and_generate_coo(new_inds, basis_lookup, melpack)
               for term, melpack in melpacks.items()}
        return diag, new_inds, coo, dbg, basis

    def _add_hc_and_generate_coo(self, new_inds, basis_lookup, melpack):
        """hamiltonian.py:495-547: pairs (dst, src, v) and their H.c. (src, dst, conj v)."""
        rows, cols, vals = [], [], []
        for m in melpack:
            k = int(m.mask.sum())
            src = basis_lookup[m.mask]
            dst = searchsorted_noesc(new_inds, m.inds) if k else np.zeros(0, dtype=np.int64)
            v = np.broadcast_to(np.asarray(m.vals, dtype=np.float64), (k,)).copy()
            rows += [dst, src]
            cols += [src, dst]
            vals += [v, np.conj(v).real.astype(np.float64)]
        return (np.concatenate(vals) if vals else np.zeros(0),
                np.concatenate(rows) if rows else np.zeros(0, np.int64),
                np.concatenate(cols) if cols else np.zeros(0, np.int64))


def coo_to_csr(vals, rows, cols, n):
    """Canonical CSR (coo_matrix(...).tocsr(), time_evolution.py:369-371).

    Rows ascending, columns ascending within a row, int32 indptr/indices, float64 data. The
    reference's o_star masks make duplicates impossible; they are rejected here.
    """
    order = np.lexsort((cols, rows))
    r, c, v = rows[order], cols[order], vals[order]
    if r.shape[0] > 1:
        dup = (r[1:] == r[:-1]) & (c[1:] == c[:-1])
        if np.any(dup):
            raise ValueError("duplicate COO entries")
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=indptr[1:])
    return indptr.astype(np.int32), c.astype(np.int32), v.astype(np.float64)


# ------------------------------------------------------------------------------------------------
# E1 / E2 / E3: H.x, Taylor, Lanczos
# ------------------------------------------------------------------------------------------------

def csr_matvec_seq(indptr, indices, data, xr, xi):
    """Row-sequential CSR product, ascending columns, no FMA (the build's parity order)."""
    n = indptr.shape[0] - 1
    cnt = np.diff(indptr.astype(np.int64))
    tr = np.zeros(n)
    ti = np.zeros(n)
    if n == 0 or cnt.max(initial=0) == 0:
        return tr, ti
    for k in range(int(cnt.max())):
        rws = np.nonzero(cnt > k)[0]
        e = indptr[rws].astype(np.int64) + k
        a = data[e]
        j = indices[e]
        tr[rws] = tr[rws] + a * xr[j]
        ti[rws] = ti[rws] + a * xi[j]
    return tr, ti


def total_ham(diag, csrs, x):
    """time_evolution.py:382-400: y = d*x; y += M_t.x for each term in order."""
    xr = np.ascontiguousarray(x.real)
    xi = np.ascontiguousarray(x.imag)
    yr = diag * xr
    yi = diag * xi
    for (ip, ix, dv) in csrs:
        tr, ti = csr_matvec_seq(ip, ix, dv, xr, xi)
        yr = yr + tr
        yi = yi + ti
    return _cplx(yr, yi)


def _cplx(re, im):
    out = np.empty(re.shape[0], dtype=np.complex128)
    out.real = re
    out.imag = im
    return out


def absmax(x):
    """inf-norm as sqrt(max(re*re + im*im)) -- see the module docstring."""
    if x.shape[0] == 0:
        return 0.0
    xr, xi = x.real, x.imag
    return float(np.sqrt(np.max(xr * xr + xi * xi)))


def taylor_coeff(delta_t, j):
    """The imaginary part c of -1j*dt/(j+1) (time_evolution.py:492), as Python computes it."""
    return -((-1j * delta_t) / float(j + 1)).imag


def simple_taylor(diag, csrs, v, delta_t, m_star=100, tol=TAYLOR_TOL):
    """time_evolution.py:482-512. Returns (res, [converged, j_ind, c1+c2, ratio])."""
    term = v
    res = v.copy()
    c1 = absmax(res)
    converged = False
    c1pc2 = 0.0
    ratio = 0.0
    j = 0
    for j in range(m_star):
        c = taylor_coeff(delta_t, j)
        h = total_ham(diag, csrs, term)
        # (a + ib) * (-0 - ic) == (b*c, -(a*c)) exactly
        term = _cplx(h.imag * c, -(h.real * c))
        c2 = absmax(term)
        res = _cplx(res.real + term.real, res.imag + term.imag)
        c1pc2 = c1 + c2
        ratio = c1pc2 / absmax(res)
        if ratio <= tol:
            converged = True
            break
        c1 = c2
    return res, [converged, j, c1pc2, ratio]


def lanczos(diag, csrs, v, delta_t, m):
    """time_evolution.py:515-549 (no reorthogonalisation, break at beta < 1e-15)."""
    import scipy.linalg
    small = np.zeros((m + 1, m + 1), dtype=complex)
    n = v.shape[0]
    vmat = np.zeros((n, m + 1), dtype=complex)
    orig = np.linalg.norm(v)
    q = v / orig
    hv = total_ham(diag, csrs, q)
    a = np.vdot(hv, q)
    hv = hv - a * q
    small[0, 0] = a
    vmat[:, 0] = q
    for i in range(1, m + 1):
        b = np.linalg.norm(hv)
        if b < 1e-15:
            break
        nq = hv / b
        hv = total_ham(diag, csrs, nq) - b * q
        a = np.vdot(hv, nq)
        hv = hv - a * nq
        q = nq
        small[i, i] = a
        small[i - 1, i] = b
        small[i, i - 1] = b
        vmat[:, i] = q
    left = scipy.linalg.expm(-1j * delta_t * small)[:, 0]
    return orig * (vmat @ left)


# ------------------------------------------------------------------------------------------------
# F1 / G1 / H1: select, reassignment, observables
# ------------------------------------------------------------------------------------------------

def select_keys(vec):
    return vec.real * vec.real + vec.imag * vec.imag


def select_indices(vec, maxstates):
    """Indices (ascending) kept by _determine_select_whoami (time_evolution.py:833-910).

    Stable ascending argsort of the key, keep the last ``maxstates`` (ties at the cut keep the
    highest indices); the kept rows of a sorted basis are re-sorted by index. shuffle_seed=None.
    Pinned by: numpy.argsort(kind='stable') (library) and a brute-force rank count.
    """
    key = select_keys(vec)
    order = np.argsort(key, kind="stable")
    kept = order[-maxstates:] if maxstates > 0 else order[:0]
    return np.sort(kept)


def reassign(old_rows, old_vec, new_rows):
    """time_evolution.py:781-797: copy amplitudes of rows present in both bases, else 0."""
    idx, hit = lookup(old_rows, new_rows)
    out = np.zeros(new_rows.shape[0], dtype=np.complex128)
    out[hit] = old_vec[idx[hit]]
    return out


def observables(ham: HolsteinOracle, rows, vec):
    """n_exc, n_pho (holstein.py:311-322) and norm (time_evolution.py:602)."""
    w = select_keys(vec)
    pos = ham.get_pos(rows)
    n_exc = np.bincount(pos, weights=w, minlength=ham.nchain)[:ham.nchain]
    n_pho = np.zeros(ham.nchain)
    for i in range(ham.nchain):
        n_pho[i] = np.sum(w * ham.layout.get_field(rows, i + 1))
    return {"n_exc": n_exc, "n_pho": n_pho, "norm": float(np.sqrt(np.sum(w)))}


def per_position_products(ham: HolsteinOracle, rows, vec, csr):
    """_calculate_partitioned_sum (holstein.py:380-396) of conj(psi) * (M psi) for one term's
    CSR: per exciton position p, the sum over its rows.
    Pinned by: dense <psi| P_p M |psi> on tiny lattices (test_oracle.py)."""
    ip, ix, dv = csr
    tr, ti = csr_matvec_seq(ip, ix, dv, np.ascontiguousarray(vec.real),
                            np.ascontiguousarray(vec.imag))
    prod = np.conj(vec) * _cplx(tr, ti)
    pos = ham.get_pos(rows)
    out = np.zeros(ham.nchain, dtype=np.complex128)
    np.add.at(out, pos, prod)
    return out


def hopping_bonds(ham: HolsteinOracle, rows, vec, hop_csr):
    """calculate_hopping (holstein.py:325-359): bond energies from the per-position products V
    of the total hopping, bond i = sum_j (1 - B_ij) V_j - B_ij conj(V_j), B the board of
    holstein.py:348-352 (1 at j > i with i + j even and at j < i with i + j odd). Open chains
    return the first N - 1 entries, odd rings all N.
    Pinned by: brute-force bond energies <psi| T_(i,i+1) |psi> from the dense hopping matrix."""
    N = ham.nchain
    if ham.periodic and N % 2 == 0:
        raise NotImplementedError("even ring")
    v = per_position_products(ham, rows, vec, hop_csr)
    out = np.zeros(N, dtype=np.complex128)
    for i in range(N):
        for j in range(N):
            on_board = (j > i and (i + j) % 2 == 0) or (j < i and (i + j) % 2 == 1)
            out[i] += -np.conj(v[j]) if on_board else v[j]
    return out if ham.periodic else out[:-1]


def vacuum_amplitude_sum(ham: HolsteinOracle, rows, vec):
    """calculate_mu_to_0 (holstein.py:366-377): sum of psi over rows with every phonon zero.
    Pinned by: a per-row check of the unpacked fields."""
    raw = ham.layout.unpack(rows)
    mask = np.all(raw[:, 1:] == 0, axis=1)
    return complex(np.sum(vec[mask]))


def energy(diag, csrs, vec):
    """calculate_total_energy (observables.py:95-99): vdot(psi, H psi)."""
    return complex(np.vdot(vec, total_ham(diag, csrs, vec)))


# ------------------------------------------------------------------------------------------------
# G2: step drivers and timeline
# ------------------------------------------------------------------------------------------------

@dataclass
class StepRecord:
    i: int
    t: float
    select: np.ndarray
    basis_e: np.ndarray
    whoami: np.ndarray
    diag: np.ndarray
    csrs: dict
    vec_adapted: np.ndarray
    vec: np.ndarray
    obs: dict
    post_norm: float
    post_energy: float
    expm_dbg: list
    dbg: dict = field(default_factory=dict)


class TimeEvolutionOracle:
    """Plain adaptive-basis evolution (time_evolution.py:290-359, 426-458, 552-650, 735-910)."""

    def __init__(self, ham: HolsteinOracle, maxstates, enlarge_steps, m_star=100,
                 lanczos_order=0, renormalize=False, explosion_cutoff=2.0):
        self.ham = ham
        self.maxstates = int(maxstates)
        self.enlarge_steps = int(enlarge_steps)
        self.m_star = m_star
        self.lanczos_order = lanczos_order
        self.renormalize = renormalize
        self.explosion_cutoff = explosion_cutoff
        self.whoami = None
        self.vector = None
        self.csrs = None
        self.diag = None

    def create_single_state_basis(self, pos):
        """create_nonuniform_basis([1]*N, minpos=pos, maxpos=pos+1) (holstein.py:427-476)."""
        raw = np.zeros((1, self.ham.nchain + 1), dtype=np.int64)
        raw[0, 0] = pos
        self.whoami = self.ham.layout.pack(raw)

    def grow_optimal_basis(self, n_max=np.inf, fillfac=1.0):
        """time_evolution.py:290-319, including the keep-previous rule and its n_max quirk."""
        whoami = self.whoami
        n = 0
        previous = None
        while len(whoami) <= self.maxstates * fillfac:
            if n >= n_max:
                break
            n += 1
            previous = whoami
            whoami = self.ham.enlarge_basis_set(previous)
        if previous is None:
            raise UnboundLocalError("initial basis already exceeds maxstates * fillfac")
        self.whoami = previous
        return n

    def create_initial_vector(self, coeffs, raw_coos):
        """time_evolution.py:325-359 (auto_normalize=True)."""
        coeffs = np.asarray(coeffs, dtype=np.complex128)
        coeffs = coeffs / np.linalg.norm(coeffs)
        rows = self.ham.layout.pack(np.asarray(raw_coos))
        idx = searchsorted_noesc(self.whoami, rows)
        self.vector = np.zeros(self.whoami.shape[0], dtype=np.complex128)
        self.vector[idx] = coeffs

    def adapt(self):
        """_generate_new_hilbert_space_cupy (time_evolution.py:735-807)."""
        kept = select_indices(self.vector, self.maxstates)
        select = self.whoami[kept]
        diag, new_inds, coo, dbg, basis_e = self.ham.generate_mel(select, self.enlarge_steps)
        R = new_inds.shape[0]
        csrs = {t: coo_to_csr(v, r, c, R) for t, (v, r, c) in coo.items()}
        new_vec = reassign(self.whoami, self.vector, new_inds)
        self.whoami, self.vector, self.csrs, self.diag = new_inds, new_vec, csrs, diag
        csr_list = [csrs[t] for t in self.ham.offdiag_terms]
        post_norm = float(np.linalg.norm(new_vec))
        post_energy = energy(diag, csr_list, new_vec).real
        return select, basis_e, dbg, post_norm, post_energy

    def propagate(self, delta_t):
        csr_list = [self.csrs[t] for t in self.ham.offdiag_terms]
        if self.lanczos_order > 0:
            res = lanczos(self.diag, csr_list, self.vector, delta_t, self.lanczos_order)
            dbg = [False, 0, 0.0, 0.0]
        else:
            res, dbg = simple_taylor(self.diag, csr_list, self.vector, delta_t, self.m_star)
        if self.renormalize:
            res = res / np.linalg.norm(res)
        elif np.linalg.norm(res) > self.explosion_cutoff:
            raise RuntimeError(f"norm {np.linalg.norm(res)} above explosion_cutoff "
                               f"{self.explosion_cutoff}")
        self.vector = res
        return dbg

    def step(self, i, t_curr, delta_t):
        select, basis_e, dbg, pn, pe = self.adapt()
        vec_adapted = self.vector.copy()
        expm_dbg = [0, 0, 0.0, 0.0] if i == 0 else self.propagate(delta_t)
        obs = observables(self.ham, self.whoami, self.vector)
        return StepRecord(i, t_curr, select, basis_e, self.whoami, self.diag, self.csrs,
                          vec_adapted, self.vector, obs, pn, pe, expm_dbg, dbg)

    def generate_timeline(self, t_array, record: Optional[Callable[[StepRecord], None]] = None):
        """time_evolution.py:426-458: step 0 adapts only; step i>0 adapts then propagates."""
        for i, t in enumerate(t_array):
            dt = 0.0 if i == 0 else float(t - t_array[i - 1])
            rec = self.step(i, float(t), dt)
            if record is not None:
                record(rec)


# ------------------------------------------------------------------------------------------------
# Initial bases and states (holstein.py:406-569)
# ------------------------------------------------------------------------------------------------

def nonuniform_basis_tuples(nchain, truncate, lowest=None, minpos=0, maxpos=None):
    """create_nonuniform_basis (holstein.py:427-476): exciton at minpos..maxpos-1; a site with
    truncate[j] > 1 takes lowest[j] .. lowest[j] + truncate[j] - 1 phonons, every other site 0
    (its lowest value is not used). Lexicographic order.
    Pinned by: itertools enumeration of hand-written examples (test_oracle.py)."""
    import itertools
    maxpos = nchain if maxpos is None else maxpos
    lowest = [0] * nchain if lowest is None else list(lowest)
    ranges = [range(lowest[j], lowest[j] + truncate[j]) if truncate[j] > 1 else range(0, 1)
              for j in range(nchain)]
    out = [(p, *occ) for p in range(minpos, maxpos) for occ in itertools.product(*ranges)]
    return np.array(out, dtype=np.int64).reshape(-1, nchain + 1)


def moving_gaussian_tuples(nchain, super_mu, super_sigma, sigma, maxval):
    """create_moving_gaussian_obc_basis (holstein.py:479-526) without its bounds checks:
    position i, site j takes 0 .. ceil((1 + maxval g(i; super_mu, super_sigma)) g(j; i, sigma))
    - 1 phonons with g(x; m, s) = exp(-(x - m)^2 / (2 s^2)); positions ascending.
    Pinned by: hand-computed ranges and counts (test_oracle.py)."""
    import itertools
    rows = []
    for i in range(nchain):
        height = 1.0 + maxval * math.exp(-0.5 * ((i - super_mu) / super_sigma) ** 2)
        tops = [math.ceil(height * math.exp(-0.5 * ((j - i) / sigma) ** 2)) for j in range(nchain)]
        rows += [(i, *occ) for occ in itertools.product(*[range(t) for t in tops])]
    return np.array(rows, dtype=np.int64).reshape(-1, nchain + 1)


def tensor_state(nchain, bstates, coeffs, excsite, sites="all"):
    """create_tensor_init_state (holstein.py:528-569): (coefficients, tuples) of
    |excsite> (x) prod over the chosen sites of sum_j c_j |b_j>, the other sites in vacuum.
    Pinned by: the Kronecker product of the one-site states in the full space (test_oracle.py)."""
    import itertools
    chosen = list(range(nchain)) if isinstance(sites, str) else sorted(set(sites))
    coef, rows = [], []
    for combo in itertools.product(range(len(bstates)), repeat=len(chosen)):
        occ = [0] * nchain
        c = 1.0 + 0j
        for site, k in zip(chosen, combo):
            occ[site] = int(bstates[k])
            c *= complex(coeffs[k])
        coef.append(c)
        rows.append((excsite, *occ))
    return np.array(coef, dtype=np.complex128), np.array(rows, dtype=np.int64)


# ------------------------------------------------------------------------------------------------
# Brute-force helpers for tiny lattices (test pins)
# ------------------------------------------------------------------------------------------------

def all_states(ham: HolsteinOracle):
    """Every (pos, n_0..n_{N-1}) tuple of the untruncated space, lexicographic order."""
    grids = [np.arange(ham.nchain)] + [np.arange(int(d)) for d in ham.max_ho_dims]
    mesh = np.meshgrid(*grids, indexing="ij")
    return np.stack([m.ravel() for m in mesh], axis=1).astype(np.int64)


def dense_hamiltonian(ham: HolsteinOracle, raw):
    """Dense H over the given tuples, built entry by entry from the textbook operators."""
    index = {tuple(r): k for k, r in enumerate(raw.tolist())}
    n = raw.shape[0]
    H = np.zeros((n, n))
    L = ham.lattice_l
    p = ham.use_terms["diag"]
    for k, r in enumerate(raw.tolist()):
        pos = r[0]
        occ = r[1:]
        H[k, k] = p["eps_sys"] + p["hbar_omega"] * sum(occ) + p.get("delta_eps", 0.0) * pos
        coords = [(pos // L ** d) % L for d in range(ham.geo_dim)]
        for term, base in (("hopping", 1), ("hopping2", 2)):
            if term not in ham.use_terms:
                continue
            J = ham.use_terms[term]["J" if term == "hopping" else "J2"]
            for d in range(ham.geo_dim):
                for sgn in (+1, -1):
                    c = coords[d] + sgn * base
                    if ham.periodic:
                        c %= L
                    elif c < 0 or c >= L:
                        continue
                    npos = pos + (c - coords[d]) * L ** d
                    j = index.get((npos, *occ))
                    if j is not None:
                        H[j, k] += J
        if "vib_coupling" in ham.use_terms:
            g = ham.use_terms["vib_coupling"]["g"]
            n0 = occ[pos]
            if n0 + 1 < ham.max_ho_dims[pos]:
                up = list(occ)
                up[pos] += 1
                j = index.get((pos, *up))
                if j is not None:
                    H[j, k] += g * math.sqrt(n0 + 1)
                    H[k, j] += g * math.sqrt(n0 + 1)
    return H
