"""Single-exciton Holstein model on the GPU (paces/models/holstein.py:22-569).

The Hamiltonian's Hilbert-space work (images, union, per-term CSR, diagonal) and the observables
run as fused HIP kernels of libpaces_hip.so. The ``generate_mel_<term>`` plugin methods keep the
reference's contract for callers that use them directly.
"""

from __future__ import annotations

import ctypes as C
import inspect
import math

import numpy
import torch

from .. import _hip
from .._hip import check, ptr, stream
from ..aux.helpers import cartesian_product, debug_lister, obs_attrs
from ..config import INIT_VERBOSITY_LEVEL, devices
from ..core import engine
from ..core.engine import CodedCsrMatrix, CsrMatrix
from ..core.hamiltonian import HamiltonianFramework, MelTriple
from ..core.observables import ObservablesFramework
from ..core.time_evolution import TimeEvolutionFramework

_lib = _hip.load()


# Outputs sized by an upper bound (so that their kernels are launched before the exact size
# reaches the host) are only so sized up to this many bytes; beyond it the host waits for the size.
_BOUND_BYTES = 8 << 30


def _to_host_async(t):
    """A pinned host copy of a small device tensor, queued on the current stream without a
    sync: (host tensor, event); read the host tensor after event.synchronize()."""
    host = torch.empty(t.shape, dtype=t.dtype, device="cpu", pin_memory=True)
    host.copy_(t, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream())
    return host, ev


class Hamiltonian(HamiltonianFramework):
    """Holstein Hamiltonian H = sum_terms + diag (holstein.py:22-287).

    ``periodic`` (keyword, default False) wraps the exciton hops around a ring/torus of lattice
    length >= 3; the reference has the flag fixed to False (holstein.py:66) and raises for True.
    """

    def __init__(self, nchain, max_ho_dims, geo_dim=1, periodic=False, **kwargs):
        self.nchain = int(nchain)
        if nchain != len(max_ho_dims):
            raise ValueError("nchain should match the number of QHOs.")
        if nchain > _hip.MAX_SITES:
            raise NotImplementedError(f"At most {_hip.MAX_SITES} sites are supported.")
        max_dims = numpy.empty(len(max_ho_dims) + 1, dtype=numpy.int64)
        max_dims[0] = self.nchain
        max_dims[1:] = max_ho_dims
        self.geo_dim = int(geo_dim)
        self.lattice_l = int(numpy.round(nchain ** (1 / geo_dim)))
        if self.lattice_l ** geo_dim != nchain:
            raise ValueError(f"Number of sites {nchain} is incompatible with dimension {geo_dim}.")
        self.periodic = bool(periodic)
        if self.periodic and self.lattice_l < 3:
            raise ValueError("Periodic hopping needs a lattice length of at least 3.")
        super().__init__(max_dims=max_dims, **kwargs)
        self.input_args["nchain"] = nchain
        self.input_args["max_ho_dims"] = max_ho_dims
        self.input_args["geo_dim"] = geo_dim
        self.input_args["periodic"] = periodic
        self.posbitwidth = int(self.bitwidths[0])
        self.qhobitwidth = self.bitwidths[1:]
        self.max_ho_dims = numpy.asarray(max_ho_dims, dtype=numpy.int64)
        self._hs = self._make_holstein_struct(self.offdiag_terms)

    def _make_holstein_struct(self, terms):
        h = _hip.Holstein()
        h.nchain = self.nchain
        h.lattice_l = self.lattice_l
        h.geo_dim = self.geo_dim
        h.periodic = 1 if self.periodic else 0
        if len(terms) > _hip.MAX_TERMS:
            raise NotImplementedError("too many Hamiltonian terms")
        h.nterms = len(terms)
        for k, (name, params) in enumerate(terms.items()):
            if name == "hopping":
                h.terms[k] = _hip.Term(_hip.TERM_HOP, 1, float(numpy.real(params["J"])))
            elif name == "hopping2":
                if self.lattice_l < 3:
                    raise ValueError("Invalid value for base_order of hopping operator.")
                h.terms[k] = _hip.Term(_hip.TERM_HOP, 2, float(numpy.real(params["J2"])))
            elif name == "vib_coupling":
                h.terms[k] = _hip.Term(_hip.TERM_VIB, 0, float(numpy.real(params["g"])))
            else:
                raise NotImplementedError(f"Unknown Holstein term {name}.")
        d = self.use_terms["diag"]
        h.eps_sys = float(d["eps_sys"])
        h.hbar_omega = float(d["hbar_omega"])
        h.delta_eps = float(d.get("delta_eps", 0.0))
        # sqrt(m) for the vibronic elements; NumPy's sqrt is correctly rounded. One table per
        # Hamiltonian: every struct built from it points at the same device buffer.
        if getattr(self, "_d_sqrt", None) is None:
            tab = numpy.sqrt(numpy.arange(int(self.max_ho_dims.max()) + 1, dtype=numpy.float64))
            self._d_sqrt = torch.tensor(tab, dtype=torch.float64, device=self.device)
        h.d_sqrt_table = self._d_sqrt.data_ptr()
        h.sqrt_table_len = int(self._d_sqrt.numel())
        return h

    # ---- field helpers (holstein.py:68-147) --------------------------------------------------------
    def get_phonon_occ(self, arr, site):
        return self.get_n_at_site(arr, site + 1)

    def get_pos(self, arr):
        return self.get_n_at_site(arr, 0)

    def get_phonon_at_exc(self, arr):
        return self.get_n_at_site(arr, -1)

    def add_n_to_qho(self, arr, n, site):
        self.add_n_at_site(arr, n, site + 1)

    def add_n_to_pos(self, arr, n):
        self.add_n_at_site(arr, n, 0)

    def add_phonon_at_exc(self, arr):
        self.add_n_at_site(arr, 1, -1)

    def rem_phonon_at_exc(self, arr):
        self.add_n_at_site(arr, -1, -1)

    # ---- device expansion ---------------------------------------------------------------------------
    def _pooled(self, key, n, dtype):
        """A device tensor of n elements from a grow-only buffer kept under key across calls
        (for outputs whose contents are dead by the next call with the same key): the output of
        a size that changes every step then needs no new device allocation."""
        pool = self.__dict__.setdefault("_pool", {})
        t = pool.get(key)
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(int(n * 1.25) + 1, dtype=dtype, device=self.device)
            pool[key] = t
        return t[:n]

    def _images(self, basis, hstruct=None, want_src=False, skip=None, slot=None):
        """Image runs of a sorted unique basis: (images tensor, run offsets list) and, with
        want_src, the source basis index and vib upper phonon count of every image. With skip
        (uint8 per row), only the rows with skip == 0 give images. With slot, the outputs live in
        buffers kept under that name (_pooled): the next call with the same slot overwrites
        them."""
        h = self._hs if hstruct is None else hstruct
        n = basis.shape[0]
        nruns = _lib.paces_holstein_num_runs(C.byref(h))
        check(min(nruns, 0), "holstein_num_runs")
        offs = torch.empty(nruns + 1, dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_holstein_images_workspace_bytes(n, C.byref(h)))
        # The outputs are sized by a bound (a row gives at most one image per run): the count is
        # launched first, the outputs are allocated while it runs, then the offsets' copy to the
        # host and the emit are queued back to back, so the host reads the offsets while the
        # emit runs (the emit places images by the scanned tile counts in the workspace, not by
        # the host's copy). Only a bound beyond _BOUND_BYTES waits for the exact size first.
        W = self.totwordwidth
        cap = max(n * nruns, 1)
        bounded = cap * (8 * W + 5) <= _BOUND_BYTES

        def outputs(cap):
            if slot is None:
                imgs = torch.empty((cap, W), dtype=self.dtype, device=self.device)
            else:
                imgs = self._pooled(slot + ".rows", cap * W, self.dtype).view(cap, W)
            src = nup = None
            if want_src:
                if slot is None:
                    src = torch.empty(cap, dtype=torch.int32, device=self.device)
                    nup = torch.empty(cap, dtype=torch.uint8, device=self.device)
                else:
                    src = self._pooled(slot + ".src", cap, torch.int32)
                    nup = self._pooled(slot + ".nup", cap, torch.uint8)
            return imgs, src, nup

        if skip is not None:
            check(_lib.paces_holstein_images_count_skip(
                ptr(basis) if n else None, n, ptr(skip) if n else None, C.byref(self._layout),
                C.byref(h), ptr(offs), wsp, wsn, stream()), "holstein_images_count_skip")
        else:
            check(_lib.paces_holstein_images_count(ptr(basis) if n else None, n,
                                                   C.byref(self._layout), C.byref(h), ptr(offs),
                                                   wsp, wsn, stream()), "holstein_images_count")
        if bounded:
            imgs, src, nup = outputs(cap)
        offs_host, ev = _to_host_async(offs)
        if not bounded:
            ev.synchronize()
            imgs, src, nup = outputs(max(int(offs_host[-1]), 1))
        if skip is not None:
            check(_lib.paces_holstein_images_emit_skip(
                ptr(basis) if n else None, n, ptr(skip) if n else None, C.byref(self._layout),
                C.byref(h), ptr(imgs), ptr(src), ptr(nup), wsp, wsn, stream()),
                "holstein_images_emit_skip")
        else:
            check(_lib.paces_holstein_images_emit(ptr(basis) if n else None, n,
                                                  C.byref(self._layout), C.byref(h), ptr(imgs),
                                                  ptr(src), ptr(nup), wsp, wsn, stream()),
                  "holstein_images_emit")
        ev.synchronize()
        offs_h = offs_host.tolist()
        m = max(offs_h[-1], 1)
        if want_src:
            return imgs[:m], offs_h, src[:m], nup[:m]
        return imgs[:m], offs_h

    def _expand(self, basis, want_from0=False):
        imgs, offs = self._images(basis)
        runs = [(basis, 0, basis.shape[0])]
        runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
        rows, from0 = engine.merge_unique(runs, self.totwordwidth, self._ws, want_from0)
        return rows, from0, offs

    def enlarge_basis_set(self, basis_states):
        """Basis plus all first-degree neighbours, sorted and unique (hamiltonian.py:550-577)."""
        rows, _, _ = self._expand(basis_states)
        return rows

    def _vib_plus_count(self, offs):
        """Rows with a vib+ image = basis rows below the phonon ceiling (ceiling_hits, 266)."""
        k = 0
        for name, _ in self.offdiag_terms.items():
            if name in ("hopping", "hopping2"):
                k += 2 * self.geo_dim * (2 if self.periodic else 1)
            else:
                return offs[k + 1] - offs[k]
        return None

    def generate_mel(self, basis_states, enlarge_steps=0, log_ind=1.2,
                     value_dtype=torch.float64, diag_dtype=torch.float64, first=None,
                     want_start_pos=False):
        """((diag, new_inds), {term: CsrMatrix}, dbg) for the basis (hamiltonian.py:383-492).

        new_inds = the basis (after ``enlarge_steps`` expansions) plus its neighbours; each term's
        CSR holds the entries with at least one endpoint in that basis. value_dtype=torch.float32
        gives the CSR values rounded to FP32 (what the mixed-precision series reads);
        value_dtype="coded" (batched trajectories) gives coded CSRs (engine.CodedCsrMatrix);
        diag_dtype=torch.float32 the diagonal rounded to FP32.

        first=(expanded basis, n_prev, run offsets, src, nup, union positions, from0): an
        expansion already made by the caller (the learned step's, with the rows it pruned
        removed: their union positions are -1), continued here with enlarge_steps more.
        want_start_pos: set self.last_start_pos, the position in new_inds of each row of the basis
        the expansions started from (int32, device).
        """
        if basis_states.shape[1] != self.totwordwidth:
            raise ValueError("Shape of basis states does not match the specified number of bits.")
        basis = basis_states
        last = None  # the last expansion's images, with their union positions
        prior = None
        if first is not None:
            basis, n_prev, offs_f, src_f, nup_f, cpos_f, from0_f = first
            last = (n_prev, offs_f, src_f, nup_f, cpos_f, from0_f)
        start_pos = None  # positions of the starting rows in the current basis (None: identity)

        def follow(cpos, nb):
            # run 0 of a merge is the current basis: its rows' union positions lead the cand_pos
            return cpos[:nb] if start_pos is None else cpos.index_select(0, start_pos.long())

        for _ in range(enlarge_steps):
            imgs, offs, src, nup = self._images(basis, want_src=True, slot="expand")
            runs = [(basis, 0, basis.shape[0])]
            runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
            # each row's skip code for the next expansion: the basis rows give no images (theirs
            # are in already), and a row added by a positive-direction op skips the inverse op,
            # whose image is the row it came from
            rows, _, cpos, skipc = engine.merge_unique(runs, self.totwordwidth, self._ws,
                                                       want_cand_pos=True,
                                                       run_codes=self._inverse_skip_codes())
            last = (basis.shape[0], offs, src, nup, cpos, skipc)
            if want_start_pos:
                start_pos = follow(cpos, basis.shape[0])
            basis = rows
        if last is None:
            imgs, offs, src, nup = self._images(basis, want_src=True)
            runs = [(basis, 0, basis.shape[0])]
            runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
            new_inds, in_basis, cand_pos, diag = engine.merge_unique(
                runs, self.totwordwidth, self._ws, want_from0=True, want_cand_pos=True,
                diag=(self._layout, self._hs, diag_dtype))
        else:
            # The expanded basis already holds the images of the rows it was expanded from, so
            # the new rows are its union with the images of the rows the expansion added only;
            # the pairs still take every image of the expanded basis.
            n_prev, offs_a, src_a, nup_a, cpos_a, skip_a = last
            prior = (n_prev, offs_a, cpos_a, src_a, nup_a)
            imgs, offs, src, nup = self._images(basis, want_src=True, skip=skip_a,
                                                slot="expand2")
            runs = [(basis, 0, basis.shape[0])]
            runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
            new_inds, in_basis, cand_pos, diag = engine.merge_unique(
                runs, self.totwordwidth, self._ws, want_from0=True, want_cand_pos=True,
                diag=(self._layout, self._hs, diag_dtype))
        self._check_basis_size(new_inds.shape[0])
        if want_start_pos:
            self.last_start_pos = follow(cand_pos, basis.shape[0])
        mats = self._pairs_csr(new_inds.shape[0], basis.shape[0], offs, cand_pos, in_basis, src,
                               nup, basis, value_dtype, rows=new_inds, prior=prior)
        dbg = {}
        if "vib_coupling" in self.offdiag_terms:
            offs_all = offs if prior is None else [a + b for a, b in zip(prior[1], offs)]
            dbg["vib_coupling"] = [basis.shape[0] - self._vib_plus_count(offs_all)]
        self.last_basis_e = basis
        self.last_in_basis = in_basis  # the rows of new_inds that are rows of basis
        return (diag, new_inds), mats, dbg

    def _inverse_skip_codes(self):
        """paces_holstein_inverse_skip_codes: the union-run codes that make each expanded row's
        skip code for the next expansion."""
        codes = (C.c_uint8 * (_hip.MAX_RUNS + 1))()
        k = _lib.paces_holstein_inverse_skip_codes(C.byref(self._hs), codes, _hip.MAX_RUNS + 1)
        check(min(k, 0), "holstein_inverse_skip_codes")
        return list(codes[:k])

    @staticmethod
    def _prior_arg(prior):
        """paces_prior_images_t from (n_prev, run offsets, union positions, source, phonon count)
        of the previous expansion's images, or None."""
        if prior is None:
            return None
        n_prev, offs_a, cpos_a, src_a, nup_a = prior
        return _hip.PriorImages(n_prev, (C.c_int64 * len(offs_a))(*offs_a), ptr(cpos_a),
                                ptr(src_a), ptr(nup_a))

    def _pairs_count(self, R, nbasis, run_off, cand_pos, in_basis, src, pr):
        """Per-term nnz on its way to the host ((pinned tensor, event): the fill is launched
        before it arrives), an upper bound of every term's nnz (two entries per image) and the
        workspace the fill call reuses."""
        h = self._hs
        nnz = torch.empty(max(h.nterms, 1), dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_holstein_pairs_workspace_bytes(R, C.byref(h)))
        if pr is None:
            check(_lib.paces_holstein_pairs_count(R, nbasis, run_off, ptr(cand_pos),
                                                  ptr(in_basis), ptr(src), C.byref(h), ptr(nnz),
                                                  wsp, wsn, stream()), "holstein_pairs_count")
        else:
            check(_lib.paces_holstein_pairs_count_split(R, nbasis, run_off, ptr(cand_pos),
                                                        ptr(in_basis), ptr(src), C.byref(pr),
                                                        C.byref(h), ptr(nnz), wsp, wsn,
                                                        stream()), "holstein_pairs_count_split")
        # a term's entries come from its own ops' images, two per image (the pair and its
        # transpose): each term's arrays are bounded by its runs' images, earlier part included
        caps = []
        runs = self._term_runs()
        if runs[-1][1] != len(run_off) - 1:
            raise RuntimeError("image runs do not match the terms' ops")
        for b, e in runs:
            nimg = run_off[e] - run_off[b]
            if pr:
                nimg += pr.run_offsets[e] - pr.run_offsets[b]
            caps.append(max(2 * nimg, 1))
        nnz_host, ev = _to_host_async(nnz)
        if sum(caps) * 12 > _BOUND_BYTES:
            ev.synchronize()
            caps = [max(int(c), 1) for c in nnz_host.tolist()[:h.nterms]]
        return (nnz_host, ev), caps, wsp, wsn

    def _term_runs(self):
        """(first run, one past the last run) of each off-diagonal term, in term order: the
        image runs are the ops of the terms in turn, a hopping term's two directions per
        lattice dimension (and the wrap-around ones when periodic), a vib term's two."""
        out, k = [], 0
        for name in self.offdiag_terms:
            n = 2 * self.geo_dim * (2 if self.periodic else 1) if name in ("hopping",
                                                                           "hopping2") else 2
            out.append((k, k + n))
            k += n
        return out

    def _pairs_csr(self, R, nbasis, offs, cand_pos, in_basis, src, nup, basis=None,
                   value_dtype=torch.float64, rows=None, prior=None):
        """Per-term CSR from the union's candidate positions (paces_holstein_pairs_*), with FP64
        or FP32 values, or coded values ("coded": batched trajectories; rows = the new basis,
        for the values formed on demand). With prior (the previous expansion's images, see
        _prior_arg), the image list comes in two parts (paces_holstein_pairs_*_split)."""
        if value_dtype == "coded":
            return self._pairs_csr_coded(R, nbasis, offs, cand_pos, in_basis, src, nup, rows,
                                         prior)
        if value_dtype == "slots":
            return self._pairs_slots(R, nbasis, offs, cand_pos, in_basis, src, nup, rows, prior)
        if value_dtype not in (torch.float64, torch.float32):
            raise ValueError("CSR values are float64, float32 or coded.")
        h = self._hs
        run_off = (C.c_int64 * len(offs))(*offs)
        pr = self._prior_arg(prior)
        (nnz_host, ev), caps, wsp, wsn = self._pairs_count(R, nbasis, run_off, cand_pos, in_basis,
                                                           src, pr)
        parts = []
        arrs = [(C.c_void_p * _hip.MAX_TERMS)() for _ in range(3)]
        for t, name in enumerate(self.offdiag_terms):
            ip = torch.empty(R + 1, dtype=torch.int32, device=self.device)
            ix = torch.empty(caps[t], dtype=torch.int32, device=self.device)
            dv = torch.empty(caps[t], dtype=value_dtype, device=self.device)
            arrs[0][t], arrs[1][t], arrs[2][t] = ip.data_ptr(), ix.data_ptr(), dv.data_ptr()
            parts.append((name, ip, ix, dv))
        if pr is not None:
            f64 = value_dtype == torch.float64
            check(_lib.paces_holstein_pairs_fill_split(
                R, nbasis, run_off, ptr(cand_pos), ptr(in_basis), ptr(src), ptr(nup),
                C.byref(pr), ptr(basis), C.byref(self._layout), C.byref(h), arrs[0], arrs[1],
                arrs[2] if f64 else None, None if f64 else arrs[2], None, wsp, wsn, stream()),
                "holstein_pairs_fill_split")
        else:
            fill = _lib.paces_holstein_pairs_fill if value_dtype == torch.float64 else \
                _lib.paces_holstein_pairs_fill32
            check(fill(R, nbasis, run_off, ptr(cand_pos), ptr(in_basis), ptr(src), ptr(nup),
                       ptr(basis), C.byref(self._layout), C.byref(h), arrs[0], arrs[1], arrs[2],
                       wsp, wsn, stream()), "holstein_pairs_fill")
        ev.synchronize()
        nnz_h = nnz_host.tolist()
        return {name: CsrMatrix(ip, ix[:nnz_h[t]], dv[:nnz_h[t]], (R, R))
                for t, (name, ip, ix, dv) in enumerate(parts)}

    def _pairs_csr_coded(self, R, nbasis, offs, cand_pos, in_basis, src, nup, rows, prior=None):
        """Coded per-term CSR (paces_holstein_pairs_fill_coded): a code byte per vib entry."""
        h = self._hs
        if not h.batched:
            raise NotImplementedError("Coded CSR values are for batched trajectories.")
        run_off = (C.c_int64 * len(offs))(*offs)
        pr = self._prior_arg(prior)
        (nnz_host, ev), caps, wsp, wsn = self._pairs_count(R, nbasis, run_off, cand_pos, in_basis,
                                                           src, pr)
        parts = []
        arrs = [(C.c_void_p * _hip.MAX_TERMS)() for _ in range(3)]
        for t, name in enumerate(self.offdiag_terms):
            ip = torch.empty(R + 1, dtype=torch.int32, device=self.device)
            ix = torch.empty(caps[t], dtype=torch.int32, device=self.device)
            cd = None
            if name == "vib_coupling":
                cd = torch.empty(caps[t], dtype=torch.uint8, device=self.device)
                arrs[2][t] = cd.data_ptr()
            arrs[0][t], arrs[1][t] = ip.data_ptr(), ix.data_ptr()
            parts.append((name, ip, ix, cd))
        if pr is not None:
            check(_lib.paces_holstein_pairs_fill_split(
                R, nbasis, run_off, ptr(cand_pos), ptr(in_basis), ptr(src), ptr(nup),
                C.byref(pr), None, None, C.byref(h), arrs[0], arrs[1], None, None, arrs[2], wsp,
                wsn, stream()), "holstein_pairs_fill_split")
        else:
            check(_lib.paces_holstein_pairs_fill_coded(R, nbasis, run_off, ptr(cand_pos),
                                                       ptr(in_basis), ptr(src), ptr(nup),
                                                       C.byref(h), arrs[0], arrs[1], arrs[2], wsp,
                                                       wsn, stream()), "holstein_pairs_fill_coded")
        ev.synchronize()
        nnz_h = nnz_host.tolist()
        return {name: CodedCsrMatrix(ip, ix[:nnz_h[t]], cd[:nnz_h[t]] if cd is not None else None,
                                     (R, R), t, lambda m, rows=rows: self._coded_values(m, rows))
                for t, (name, ip, ix, cd) in enumerate(parts)}

    def slot_layout(self, R):
        """paces_slots_t host fields for R rows (paces_holstein_slot_layout), or None when this
        model has no slot form (not batched, periodic, too many ops or rows)."""
        lay = _hip.Slots()
        rc = _lib.paces_holstein_slot_layout(C.byref(self._hs), R, C.byref(lay))
        if rc == _hip.E_UNSUPPORTED:
            return None
        check(rc, "holstein_slot_layout")
        lay.term_value = self._hs.d_traj_value
        lay.sqrt_table = self._hs.d_sqrt_table
        return lay

    def _pairs_slots(self, R, nbasis, offs, cand_pos, in_basis, src, nup, rows, prior=None):
        """Every term's entries in one slot matrix (paces_holstein_pairs_slots, batched open
        lattices): no count pass and no host wait; each term is a SlotTermMatrix view."""
        lay = self.slot_layout(R)
        if lay is None:
            raise NotImplementedError("This model has no slot form (batched open lattices only).")
        slots = torch.empty((lay.nslots, R), dtype=torch.int32, device=self.device)
        run_off = (C.c_int64 * len(offs))(*offs)
        pr = self._prior_arg(prior)
        check(_lib.paces_holstein_pairs_slots(R, nbasis, run_off, ptr(cand_pos), ptr(in_basis),
                                              ptr(src), ptr(nup),
                                              C.byref(pr) if pr is not None else None,
                                              C.byref(self._hs), ptr(slots), stream()),
              "holstein_pairs_slots")
        sm = engine.SlotMatrix(slots, lay, (R, R))
        return {name: sm.term(t, lambda m, rows=rows: self._coded_values(m, rows))
                for t, name in enumerate(self.offdiag_terms)}

    def _coded_values(self, m, rows):
        """float64 values of a coded CSR: the term value of each entry's row's trajectory, times
        sqrt(code) for a vib term (the products paces_holstein_pairs_fill forms)."""
        R = m.shape[0]
        counts = (m.indptr[1:] - m.indptr[:-1]).long()
        traj = self.get_traj(rows).long()
        te = torch.repeat_interleave(traj, counts, output_size=m.nnz)
        v = self._d_traj_val[te, m.term]
        if m.code is not None:
            v = self._d_sqrt[m.code.long()] * v
        return v.contiguous()

    def _csr(self, rows, in_basis):
        """Per-term CSR of the rows by neighbour lookups (paces_holstein_csr_*); equal to
        _pairs_csr when in_basis marks the basis whose expansion gave the rows."""
        R = rows.shape[0]
        h = self._hs
        nterms = h.nterms
        nnz = torch.empty(max(nterms, 1), dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_holstein_csr_workspace_bytes(R, C.byref(h)))
        check(_lib.paces_holstein_csr_count(ptr(rows) if R else None, R,
                                            ptr(in_basis) if R else None, C.byref(self._layout),
                                            C.byref(h), ptr(nnz), wsp, wsn, stream()),
              "holstein_csr_count")
        nnz_h = nnz.tolist()
        mats = {}
        ip_arr, ix_arr, dv_arr = (C.c_void_p * _hip.MAX_TERMS)(), (C.c_void_p * _hip.MAX_TERMS)(), \
            (C.c_void_p * _hip.MAX_TERMS)()
        for t, name in enumerate(self.offdiag_terms):
            ip = torch.empty(R + 1, dtype=torch.int32, device=self.device)
            ix = torch.empty(max(nnz_h[t], 1), dtype=torch.int32, device=self.device)
            dv = torch.empty(max(nnz_h[t], 1), dtype=torch.float64, device=self.device)
            ip_arr[t], ix_arr[t], dv_arr[t] = ip.data_ptr(), ix.data_ptr(), dv.data_ptr()
            mats[name] = CsrMatrix(ip, ix[:nnz_h[t]], dv[:nnz_h[t]], (R, R))
        check(_lib.paces_holstein_csr_fill(ptr(rows) if R else None, R, C.byref(self._layout),
                                           C.byref(h), ip_arr, ix_arr, dv_arr, wsp, wsn,
                                           stream()), "holstein_csr_fill")
        return mats

    def generate_mel_diag(self, basis_states, dtype=torch.float64):
        """eps_sys + hbar_omega * sum(n) (+ delta_eps * pos) (holstein.py:280-287); float32
        gives each value rounded to nearest (what the mixed-precision series reads)."""
        R = basis_states.shape[0]
        out = torch.empty(R, dtype=dtype, device=self.device)
        fn = _lib.paces_holstein_diag if dtype == torch.float64 else _lib.paces_holstein_diag32
        check(fn(ptr(basis_states) if R else None, R, C.byref(self._layout), C.byref(self._hs),
                 ptr(out) if R else None, stream()), "holstein_diag")
        return out

    def features(self, rows, psi, diag, nbasis, offs, cand_pos, src):
        """Scorer inputs of the union rows (paces_holstein_features): (R, 64) bf16 bits in int16.
        nbasis, offs, cand_pos, src: the expanded basis size, image run offsets (host), the
        union's candidate positions and the images' source rows, as for the pairs CSR."""
        R = rows.shape[0]
        X = torch.empty((max(R, 1), 64), dtype=torch.int16, device=self.device)
        run_off = (C.c_int64 * len(offs))(*offs)
        wsp, wsn = self._ws.get(_lib.paces_holstein_features_workspace_bytes(R))
        check(_lib.paces_holstein_features(ptr(rows) if R else None, R,
                                           ptr(torch.view_as_real(psi)) if R else None,
                                           ptr(diag) if R else None, nbasis, run_off,
                                           ptr(cand_pos), ptr(src), C.byref(self._layout),
                                           C.byref(self._hs), ptr(X), wsp, wsn, stream()),
              "holstein_features")
        return X[:R]

    def features_added(self, rows, psi, nbasis, offs, cand_pos, src, in_basis, nadd):
        """Scorer inputs of the union rows outside the expanded basis only
        (paces_holstein_features_added): ((nadd, 64) bf16 bits in int16, int32 union row of each).
        in_basis: the union's 0/1 flags of the basis rows, nadd = rows - sum(in_basis) (known to
        the caller); the diagonal is computed from the rows."""
        R = rows.shape[0]
        X = torch.empty((max(nadd, 1), 64), dtype=torch.int16, device=self.device)
        idx = torch.empty(max(nadd, 1), dtype=torch.int32, device=self.device)
        run_off = (C.c_int64 * len(offs))(*offs)
        wsp, wsn = self._ws.get(_lib.paces_holstein_features_added_workspace_bytes(R))
        check(_lib.paces_holstein_features_added(
            ptr(rows) if R else None, R, ptr(torch.view_as_real(psi)) if R else None, nbasis,
            run_off, ptr(cand_pos), ptr(src), ptr(in_basis) if R else None,
            C.byref(self._layout), C.byref(self._hs), ptr(X), ptr(idx), wsp, wsn, stream()),
            "holstein_features_added")
        return X[:nadd], idx[:nadd]

    # ---- plugin methods with the reference's contract (holstein.py:151-277) ----------------------
    def _term_images(self, basis, name):
        h = self._make_holstein_struct({name: self.offdiag_terms[name]})
        return self._images(basis, h)

    def _plugin(self, basis_states, name, raw_map_to):
        imgs, offs = self._term_images(basis_states, name)
        runs = [imgs[offs[k]:offs[k + 1]] for k in range(len(offs) - 1)]
        if raw_map_to:
            return runs
        raise NotImplementedError("MelTriples are produced inside the fused generate_mel; call "
                                  "generate_mel(basis) for matrix elements.")

    def generate_mel_hopping(self, basis_states, raw_map_to=False):
        return self._plugin(basis_states, "hopping", raw_map_to)

    def generate_mel_hopping2(self, basis_states, raw_map_to=False):
        return self._plugin(basis_states, "hopping2", raw_map_to)

    @debug_lister(["ceiling_hits"])
    def generate_mel_vib_coupling(self, basis_states, raw_map_to=False):
        return self._plugin(basis_states, "vib_coupling", raw_map_to)

    # ---- observables kernel ------------------------------------------------------------------------
    def site_observables(self, rows, vector):
        """(n_exc[N], n_pho[N], sum|psi|^2) in one fused pass."""
        R = rows.shape[0]
        N = self.nchain
        out = torch.empty(2 * N + 1, dtype=torch.float64, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_holstein_observables_workspace_bytes(R, C.byref(self._hs)))
        check(_lib.paces_holstein_observables(ptr(rows) if R else None,
                                              ptr(torch.view_as_real(vector)) if R else None, R,
                                              C.byref(self._layout), C.byref(self._hs), ptr(out),
                                              wsp, wsn, stream()), "holstein_observables")
        h = out.cpu().numpy()
        return h[:N], h[N:2 * N], float(h[2 * N])


    def position_segments(self, rows):
        """seg_off (device int32, nchain + 1): the rows of exciton position p are
        [seg_off[p], seg_off[p + 1]) (the position is the leading field of a sorted basis)."""
        seg = torch.empty(self.nchain + 1, dtype=torch.int32, device=self.device)
        R = rows.shape[0]
        check(_lib.paces_batch_segments(ptr(rows) if R else None, R, self.totwordwidth,
                                        self.nchain, ptr(seg), stream()), "batch_segments")
        return seg

    def position_vdot(self, seg, x, y):
        """Per exciton position p: sum of conj(x) * y over its rows (complex128[N], host)."""
        out = torch.empty(2 * self.nchain, dtype=torch.float64, device=self.device)
        if x.shape[0] == 0:
            return numpy.zeros(self.nchain, dtype=numpy.complex128)
        check(_lib.paces_segment_vdot(ptr(torch.view_as_real(x)), ptr(torch.view_as_real(y)),
                                      ptr(seg), self.nchain, ptr(out), stream()),
              "segment_vdot")
        h = out.cpu().numpy()
        return h[0::2] + 1j * h[1::2]

    def vacuum_sum(self, rows, seg, x):
        """Sum of x over the rows with every phonon field zero (complex)."""
        if x.shape[0] == 0:
            return 0j
        out = torch.empty(2, dtype=torch.float64, device=self.device)
        check(_lib.paces_vacuum_sum(ptr(rows), self.totwordwidth, self.posbitwidth,
                                    ptr(torch.view_as_real(x)), ptr(seg), self.nchain, ptr(out),
                                    stream()), "vacuum_sum")
        h = out.tolist()
        return complex(h[0], h[1])


def bond_board(n):
    """The n x n sign board of calculate_hopping (holstein.py:348-352): 1 where the position
    sum of row i's bond runs through the conjugate of V_j, i.e. at (i, j) with j > i and i + j
    even, or j < i and i + j odd."""
    i, j = numpy.meshgrid(numpy.arange(n), numpy.arange(n), indexing="ij")
    return (((j > i) & ((i + j) % 2 == 0)) | ((j < i) & ((i + j) % 2 == 1))).astype(numpy.int64)


class Observables(ObservablesFramework):
    """Holstein observables (holstein.py:291-396)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._cache_key = None
        self._cache = None

    def primer_hook(self):
        self._cache_key = None
        self._seg_key = None

    def _site_obs(self, vector):
        key = (vector.data_ptr(), vector._version, self.teobj.whoami.data_ptr())
        if self._cache_key != key:
            self._cache = self.hamobj.site_observables(self.teobj.whoami, vector)
            self._cache_key = key
        return self._cache

    @obs_attrs(fname="n_exc", header=" ".join(f"n^{{exc}}_{i}" for i in range(3)) + " etc.")
    def calculate_n_exc(self, vector):
        """Exciton density per site: sum of |psi|^2 over rows with pos == site."""
        return self._site_obs(vector)[0]

    @obs_attrs(fname="n_pho", header=" ".join(f"n^{{pho}}_{i}" for i in range(3)) + " etc.")
    def calculate_n_pho(self, vector):
        """Phonon number per site: sum of |psi|^2 * n_site."""
        return self._site_obs(vector)[1]

    def _position_segments(self):
        whoami = self.teobj.whoami
        key = (whoami.data_ptr(), whoami.shape[0])
        if getattr(self, "_seg_key", None) != key:
            self._seg = self.hamobj.position_segments(whoami)
            self._seg_key = key
        return self._seg

    def _term_per_position(self, term, vector):
        """V_p = sum over rows at exciton position p of conj(psi) * (M_term psi)
        (_calculate_partitioned_sum of holstein.py:380-396 applied to those products)."""
        mat = self.teobj.sparse_mats_dict[term]
        return self.hamobj.position_vdot(self._position_segments(), vector, mat.dot(vector))

    @obs_attrs(fname="V_hop", header=" ".join(f"V_{i}" for i in range(3)) + " etc.")
    def calculate_hopping(self, vector):
        """Hopping energy of each bond (holstein.py:325-359): from the per-position products
        V_p = <psi| P_p T |psi> of the total hopping T, bond i's energy is
        sum_j (1 - B_ij) V_j - B_ij conj(V_j) with B = bond_board(N). An open chain gives its
        N - 1 bonds (the formula's last entry is zero), a ring of odd length its N bonds; rings
        of even length are not covered by the formula."""
        h = self.hamobj
        if h.periodic and h.nchain % 2 == 0:
            raise NotImplementedError("the bond formula needs an odd ring length (or open "
                                      "boundaries)")
        v = self._term_per_position("hopping", vector)
        board = bond_board(h.nchain)
        bonds = ((1 - board) * v[None, :] - board * numpy.conj(v)[None, :]).sum(axis=1)
        if h.periodic:
            return bonds
        if abs(bonds[-1]) >= 1e-12:
            raise RuntimeError(f"open chain: the bond past the last site has energy "
                               f"{bonds[-1]} (zero expected)")
        return bonds[:-1]

    @obs_attrs(fname="V_vib_coupl", header=" ".join(f"V_{i}" for i in range(3)) + " etc.")
    def calculate_coupling(self, vector):
        """Exciton-phonon coupling energy per site: sum over rows at exciton position p of
        conj(psi) * (M_vib psi) (holstein.py:360-365). The reference reads the matrix of a
        term named "coupling", which does not exist (its term is "vib_coupling", so it raises
        KeyError); here the vib_coupling matrix is used."""
        return self._term_per_position("vib_coupling", vector)

    @obs_attrs(fname="mu_to_0", header=r"<mu(t) mu(0)>")
    def calculate_mu_to_0(self, vector):
        """<mu(t) mu(0)> for a start in the global vacuum (holstein.py:366-377): the sum of the
        amplitudes of the rows with no phonons, at any exciton position."""
        return self.hamobj.vacuum_sum(self.teobj.whoami, self._position_segments(), vector)


class TimeEvolution(TimeEvolutionFramework):
    """Holstein initial bases and states (holstein.py:401-569)."""

    def create_nonuniform_basis(self, truncate_d_list, lowest_d_list=None, minpos=0, maxpos=None):
        """Product basis with per-site phonon ranges and exciton in [minpos, maxpos)
        (holstein.py:427-476)."""
        N = self.hamobj.nchain
        if maxpos is None:
            maxpos = N
        if lowest_d_list is None:
            lowest_d_list = numpy.zeros_like(truncate_d_list)
        if len(truncate_d_list) != N or len(lowest_d_list) != N:
            raise ValueError("Incorrect chain length.")
        mh = self.hamobj.max_ho_dims
        if any(lowest_d_list[i] + truncate_d_list[i] > mh[i] for i in range(N)):
            raise ValueError("Specified truncation value exceeds at least one of the max_ho_dims.")
        if numpy.prod(truncate_d_list) * (maxpos - minpos) > self.te_params.maxstates:
            raise ValueError("Number of states that would result from this value of truncate_d"
                             " exceeds maxstates!")
        if minpos < 0 or maxpos > N or minpos >= maxpos:
            raise ValueError("Invalid minpos or maxpos.")
        self.write_current_params_to_file(inspect.currentframe())
        nontriv = numpy.array(truncate_d_list) > 1
        dims = [numpy.arange(lowest_d_list[i], lowest_d_list[i] + truncate_d_list[i])
                for i in range(N) if nontriv[i]]
        nopad = cartesian_product(numpy.arange(minpos, maxpos), *dims)
        raw = numpy.zeros((nopad.shape[0], N + 1), dtype=numpy.int64)
        raw[:, numpy.nonzero(nontriv)[0] + 1] = nopad[:, 1:]
        raw[:, 0] = nopad[:, 0]
        self.whoami = self.hamobj.compress_ind_arr(raw)
        self.numstates = self.whoami.shape[0]

    def create_uniform_truncated_basis(self, truncate_d):
        """All states with at most truncate_d-1 phonons per site (holstein.py:406-425)."""
        N = self.hamobj.nchain
        if truncate_d > self.hamobj.max_ho_dims.min():
            raise ValueError("Specified basis truncation value exceeds at least one HO dimension.")
        if N * truncate_d ** N > self.te_params.maxstates:
            raise ValueError("Number of states that would result"
                             " from this value of truncate_d exceeds maxstates!")
        self.write_current_params_to_file(inspect.currentframe())
        raw = cartesian_product(numpy.arange(N), *([numpy.arange(truncate_d)] * N))
        self.whoami = self.hamobj.compress_ind_arr(raw)
        self.numstates = self.whoami.shape[0]

    @staticmethod
    def moving_gaussian_tuples(nchain, super_mu, super_sigma, sigma, maxval, maxstates):
        """The field tuples of create_moving_gaussian_obc_basis (holstein.py:479-526), in
        position order. With w(x, m, s) = exp(-(x - m)^2 / (2 s^2)): for exciton position i,
        site j holds 0 .. ceil((1 + maxval w(i, super_mu, super_sigma)) w(j, i, sigma)) - 1
        phonons (an empty range leaves position i without states), every combination of the
        sites' ranges once, site 0 varying slowest. Raises ValueError once the count passes
        maxstates."""
        sites = numpy.arange(nchain, dtype=numpy.float64)

        def w(x, m, s):
            return numpy.exp(-0.5 * ((x - m) / s) ** 2)

        height = 1.0 + maxval * w(sites, super_mu, super_sigma)
        blocks, total = [], 0
        for i in range(nchain):
            tops = numpy.ceil(height[i] * w(sites, i, sigma)).astype(numpy.int64)
            count = int(numpy.prod(numpy.maximum(tops, 0).astype(object)))
            total += count
            if total > maxstates:
                raise ValueError(f"the Gaussian basis has more than maxstates={maxstates} "
                                 f"states (position {i} brings it to {total})")
            if count == 0:
                continue
            block = numpy.zeros((count, nchain + 1), dtype=numpy.int64)
            block[:, 0] = i
            block[:, 1:] = cartesian_product(*[numpy.arange(t, dtype=numpy.int64) for t in tops])
            blocks.append(block)
        if not blocks:
            return numpy.zeros((0, nchain + 1), dtype=numpy.int64)
        return numpy.concatenate(blocks)

    def create_moving_gaussian_obc_basis(self, super_mu, super_sigma, sigma, maxval):
        """Initial basis of Gaussian phonon ranges around each exciton position, the heights
        peaked at super_mu (holstein.py:479-526); see moving_gaussian_tuples.

        The reference ends with a bounds check whose operands have N + 1 and N entries, which
        NumPy and CuPy refuse to broadcast, so it raises on every call; here the phonon
        columns are checked against max_ho_dims, the check it intends."""
        h = self.hamobj
        if maxval >= int(h.max_ho_dims.max()):
            raise ValueError(f"maxval={maxval} reaches the largest phonon dimension "
                             f"{int(h.max_ho_dims.max())}")
        self.write_current_params_to_file(inspect.currentframe())
        raw = self.moving_gaussian_tuples(h.nchain, super_mu, super_sigma, sigma, maxval,
                                          self.te_params.maxstates)
        if raw.shape[0] and numpy.any(raw[:, 1:].max(axis=0) >= h.max_ho_dims):
            raise ValueError("a Gaussian phonon range exceeds max_ho_dims")
        self.whoami = self.hamobj.compress_ind_arr(raw)
        self.numstates = int(self.whoami.shape[0])

    @staticmethod
    def tensor_state_tuples(nchain, bstates, coeffs, excsite, sites="all"):
        """(coefficients, field tuples) of create_tensor_init_state (holstein.py:528-569): the
        exciton at excsite times, on each chosen site, sum_j coeffs[j] |bstates[j]>, other sites
        in their vacuum; one entry per combination of the chosen sites' Fock states (first
        chosen site slowest), with the product of their coefficients."""
        bstates = numpy.asarray(bstates)
        coeffs = numpy.asarray(coeffs, dtype=numpy.complex128)
        if not (isinstance(sites, str) and sites == "all") and len(sites) > nchain:
            raise ValueError(f"{len(sites)} sites chosen on a lattice of {nchain}")
        if len(bstates) != len(coeffs):
            raise ValueError(f"{len(coeffs)} coefficients for {len(bstates)} Fock states")
        if isinstance(sites, str) and sites == "all":
            chosen = numpy.arange(nchain)
        else:
            chosen = numpy.array([i for i in range(nchain) if i in sites], dtype=numpy.int64)
        k = len(chosen)
        vec = cartesian_product(*([coeffs] * k)).prod(axis=1) if k else numpy.ones(1, complex)
        raw = numpy.zeros((vec.shape[0], nchain + 1), dtype=numpy.int64)
        raw[:, 0] = excsite
        if k:
            raw[:, 1 + chosen] = cartesian_product(*([bstates.astype(numpy.int64)] * k))
        return vec, raw

    def create_tensor_init_state(self, bstates, coeffs, excsite, sites="all"):
        """Initial vector |excsite> (x) prod over the chosen sites of sum_j c_j |j>, other sites
        in their vacuum (holstein.py:528-569), on the current basis, which must hold every
        product state; see tensor_state_tuples. The state must be normalised to 1e-12."""
        vec, raw = self.tensor_state_tuples(self.hamobj.nchain, bstates, coeffs, excsite, sites)
        norm = float(numpy.linalg.norm(vec))
        if abs(1.0 - norm) > 1e-12:
            raise ValueError(f"the tensor-product state has norm {norm}, not 1")
        self.create_initial_vector(vec, raw, auto_normalize=True)


# ------------------------------------------------------------------------------------------------
# Batched trajectories (no reference counterpart): many independent Holstein lattices advanced
# together. Rows carry the trajectory id as field 0 and the exciton position as field 1, phonon
# fields are padded to the largest lattice; the per-trajectory lattice length and term values
# live in device tables.
# ------------------------------------------------------------------------------------------------

class LabelLog:
    """Retained/discarded labels of candidate rows, logged from steps that keep every row.

    At step i the union rows outside the selection (sorted) are staged with their scorer
    features; at step i+1 each staged row is labelled 1 when the new selection kept it, else 0.
    Labels and features accumulate on the device up to ``capacity`` rows."""

    def __init__(self, device, capacity):
        self.device = device
        self.capacity = int(capacity)
        self.X = torch.empty((self.capacity, 64), dtype=torch.int16, device=device)
        self.y = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.n = 0
        self._pending = None

    @property
    def full(self):
        return self.n >= self.capacity

    def stage(self, rows, X):
        """Stage candidate rows (sorted, outside the selection) with their features."""
        self._pending = None if self.full else (rows, X)

    def resolve(self, select):
        if self._pending is None:
            return
        rows, X = self._pending
        self._pending = None
        if rows.shape[0] == 0:
            return
        _, hit, _ = engine.lower_bound(select, rows, want_hit=True, q_sorted=True)
        m = min(rows.shape[0], self.capacity - self.n)
        self.X[self.n:self.n + m].copy_(X[:m])
        self.y[self.n:self.n + m].copy_(hit[:m])
        self.n += m

    def data(self):
        return self.X[:self.n], self.y[:self.n]


class LearnedTimeEvolution(TimeEvolution):
    """The adaptive-basis step with the candidate scorer (no counterpart in the reference).

    Per step (enlarge_steps >= 1):
    * select, then expand the selection once: the basis plus its images (the first shell);
    * carry the amplitudes over onto that basis and score every row the expansion added;
    * drop the added rows whose predicted retain probability is below ``threshold``, keeping a
      hashed sample of 2^-sample_bits of them, so rows unlikely to survive the next selection,
      and the rows they would have reached, are never materialised;
    * run the remaining expansions, the union, the CSR and the propagation on the rest;
    * the weight that all dropped rows would have kept is estimated from the sample: its weight
      in the next selection times (dropped / sampled); when that exceeds ``fallback_tol`` the
      step is redone from the whole first expansion, which is the plain step.

    With ``scorer=None`` nothing is dropped (the plain step through the same code). A
    ``label_log`` then records training labels: the added rows of each first expansion with
    their features, labelled by the next selection. ``stats`` holds one dict per step: rows of
    the first expansion, rows scored and kept, sampled rows, final rows, the estimated dropped
    weight and whether the step fell back.
    """

    def __init__(self, *args, scorer=None, threshold=0.01, sample_bits=6, fallback_tol=1e-6,
                 label_log=None, **kwargs):
        super().__init__(*args, **kwargs)
        if self.te_params.enlarge_steps < 1:
            raise ValueError("the learned step prunes the first of enlarge_steps + 1 expansions: "
                             "it needs enlarge_steps >= 1")
        self.scorer = scorer
        self.threshold = float(threshold)
        self.sample_bits = int(sample_bits)
        self.fallback_tol = float(fallback_tol)
        self.label_log = label_log
        self.stats = []
        self._full = None
        self._sample = None
        self._next_select = None

    def _determine_select_whoami(self, vector, use_weight_function=False, tau=0.0,
                                 garbage_tol=-1):
        # the selection computed right after the last propagation (same rows and amplitudes)
        if self._next_select is not None:
            sel, self._next_select = self._next_select, None
            return sel
        return super()._determine_select_whoami(vector, use_weight_function, tau, garbage_tol)

    def _build(self, first, enlarge_steps, want_start_pos=False):
        """The rows, CSR and carried amplitudes of the step from its first expansion (see
        Hamiltonian.generate_mel's first); returns the debug columns."""
        ham = self.hamobj
        (diag, new_inds), mats, dbg = ham.generate_mel(first[0], enlarge_steps, first=first,
                                                       want_start_pos=want_start_pos)
        self.numstates = new_inds.shape[0]
        self.create_matrices(diag, mats)
        self.vector = engine.reassign(self.whoami, self.vector, new_inds, self._ws)
        self.whoami = new_inds
        return dbg

    def _generate_new_hilbert_space(self, enlarge_steps):
        ham = self.hamobj
        W = ham.totwordwidth
        select = self._determine_select_whoami(self.vector)
        self.last_select = select
        if self.label_log is not None:
            self.label_log.resolve(select)
        # first expansion: the selection and its images
        imgs, offs, src, nup = ham._images(select, want_src=True, slot="learned")
        runs = [(select, 0, select.shape[0])]
        runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
        B, in_sel, cand_pos = engine.merge_unique(runs, W, self._ws, want_from0=True,
                                                  want_cand_pos=True)
        nB = B.shape[0]
        nsel = select.shape[0]
        nadd = nB - nsel  # every selected row is a union row
        full = (B, nsel, offs, src, nup, cand_pos, in_sel)
        want_x = self.scorer is not None or (self.label_log is not None and
                                             not self.label_log.full)
        X = idx = None
        if want_x and nadd:
            # features of the added rows only, in union order, with their union positions
            psiB = engine.reassign(self.whoami, self.vector, B, self._ws)
            X, idx = ham.features_added(B, psiB, nsel, offs, cand_pos, src, in_sel, nadd)
            del psiB
        if self.label_log is not None and X is not None:
            self.label_log.stage(B.index_select(0, idx), X)
        st = {"rows_first": nB, "scored": 0, "kept": 0, "sampled": 0, "rows": 0,
              "dropped_weight": 0.0, "dropped_retained_weight": 0.0, "fallback": False}
        self._sample = None
        self._full = None
        if self.scorer is None:
            dbg = self._build(full, enlarge_steps - 1)
        else:
            prob = None
            if nadd:
                tuner = self._tuner()
                if tuner is not None:
                    from ..core import autotune
                    self.scorer.tile_rows = tuner.pick(
                        tuner.key("mlp", nadd), autotune.MLP_TILES,
                        lambda t: (lambda: self.scorer.infer(X, tile_rows=t)))
                prob = self.scorer.infer(X)
            del X
            keep = torch.empty(max(nB, 1), dtype=torch.uint8, device=self.device)
            # sample 2^-bits of the dropped rows, bits lowered on small expansions so the sample
            # holds a few hundred rows
            bits = max(0, min(self.sample_bits, int(math.log2(max(nB, 1) / 4096.0))
                              if nB >= 4096 else 0))
            check(_lib.paces_prune_plan_rows(ptr(prob) if nadd else None,
                                             ptr(idx) if nadd else None, nadd, ptr(in_sel), nB,
                                             self.threshold, bits,
                                             (self._step_index + 1) * 0x9E3779B97F4A7C15
                                             % (1 << 64), ptr(keep), stream()), "prune_plan_rows")
            del prob, idx
            # the kept rows in order, the selection's flags and the union positions renumbered
            # (-1 for a pruned row: its pairs are not formed)
            nC = cand_pos.shape[0]
            rows_k = torch.empty((max(nB, 1), W), dtype=B.dtype, device=self.device)
            inb_k = torch.empty(max(nB, 1), dtype=torch.uint8, device=self.device)
            smp_k = torch.empty(max(nB, 1), dtype=torch.uint8, device=self.device)
            pos_k = torch.empty(max(nC, 1), dtype=torch.int32, device=self.device)
            cnt = torch.empty(1, dtype=torch.int32, device=self.device)
            wsp, wsn = self._ws.get(_lib.paces_compact_workspace_bytes(nB))
            check(_lib.paces_compact_kept(ptr(keep), nB, ptr(B), W, None, ptr(in_sel),
                                          ptr(cand_pos), nC, ptr(rows_k), None, ptr(inb_k),
                                          ptr(smp_k), ptr(pos_k), ptr(cnt), wsp, wsn, stream()),
                  "compact_kept")
            nK = int(cnt.item())
            rows_k, inb_k, smp_k = rows_k[:nK], inb_k[:nK], smp_k[:nK]
            # the number of sampled rows reaches the host while the step runs on
            nsample = _to_host_async(smp_k.sum()) if nK else None
            st.update(scored=nadd, kept=nK - nsel)
            self._full = (full, enlarge_steps - 1, self.whoami, self.vector)
            dbg = self._build((rows_k, nsel, offs, src, nup, pos_k, inb_k), enlarge_steps - 1,
                              want_start_pos=True)
            # the sampled rows' flags at their positions in the new basis (every kept row is a
            # row of it)
            mark = torch.zeros(max(self.numstates, 1), dtype=torch.uint8, device=self.device)
            if nK:
                check(_lib.paces_scatter_u8(ptr(ham.last_start_pos), nK, ptr(smp_k), ptr(mark),
                                            stream()), "scatter_u8")
            self._sample = (mark, nsample, nB - nK)
        st["rows"] = self.numstates
        self.stats.append(st)
        return dbg, self._post_assignment_diagnostics()

    def _check_dropped(self, vec, expm_dbg, delta_t):
        """Estimate, from the sampled dropped rows, the weight of dropped rows that the next
        selection would keep; fall back to the whole first expansion when it exceeds
        fallback_tol."""
        smp, nsample_h, nd_unsampled = self._sample
        nsample = 0
        if nsample_h is not None:
            host, ev = nsample_h
            ev.synchronize()
            nsample = int(host.item())
        ndropped = nd_unsampled + nsample  # the pruned rows and the sampled ones
        st = self.stats[-1]
        st["sampled"] = nsample
        if ndropped == 0:
            return vec, expm_dbg
        R = self.whoami.shape[0]
        ms = self.te_params.maxstates
        sel, sidx = engine.select_topk(vec, self.whoami, ms, self._ws, want_index=True)
        if nsample:
            scale = ndropped / nsample
            mark = torch.zeros(R, dtype=torch.uint8, device=self.device)
            check(_lib.paces_mark_indices(ptr(sidx), sidx.shape[0], ptr(smp), ptr(mark),
                                          stream()), "mark_indices")
            w_smp, w_kept = self._reducer.norm2_masked_pair(vec, smp, mark)
            st["dropped_weight"] = w_smp * scale
            est = w_kept * scale
        else:
            est = float("inf")  # rows were dropped but none sampled: no estimate
        st["dropped_retained_weight"] = est
        if est > self.fallback_tol:
            st["fallback"] = True
            full, steps, old_rows, old_vec = self._full
            self.whoami, self.vector = old_rows, old_vec
            self._build(full, steps)
            st["rows"] = self.numstates
            return self._propagate(delta_t)
        if not self.expm_params.renormalize:
            self._next_select = sel
        return vec, expm_dbg

    def _propagate(self, delta_t):
        if self.expm_params.lanczos_order > 0:
            return self._lanczos(delta_t)
        return self._simple_taylor(delta_t)

    def _evolve_single_step(self, i, t_curr, delta_t, diag_coo_debug=None):
        dbg_dict, post = self._generate_new_hilbert_space(self.te_params.enlarge_steps)
        vec, expm_dbg = self._propagate(delta_t)
        if self._sample is not None:
            vec, expm_dbg = self._check_dropped(vec, expm_dbg, delta_t)
        self._full = None
        self._sample = None
        self.vector = vec
        norm = self._norm_after_propagation()
        self.last_expm_dbg = expm_dbg
        self._postprocess(t_curr, norm, i)
        self._save_diagnostic_data(t_curr, norm, post, expm_dbg, dbg_dict)
        self._step_index = i
        return None


class BatchedHamiltonian(Hamiltonian):
    """Holstein Hamiltonians of ``len(nchains)`` independent trajectories in one packed basis.

    ``hopping_J`` and ``vib_g`` are per-trajectory values in the reference's convention (the
    hopping element and the g of g*sqrt(n+1)); the diagonal (``eps_sys``, ``hbar_omega``), the
    phonon dimension, ``geo_dim`` and ``periodic`` are shared.
    """

    def __init__(self, nchains, max_ho_dim, hopping_J, vib_g, hbar_omega=1.0, eps_sys=0.0,
                 geo_dim=1, periodic=False, debug_verb=0):
        nchains = [int(n) for n in nchains]
        ntraj = len(nchains)
        if ntraj < 1 or len(hopping_J) != ntraj or len(vib_g) != ntraj:
            raise ValueError("One lattice size, J and g per trajectory are required.")
        self.ntraj = ntraj
        self.nchains = nchains
        self.nchain = max(nchains)
        if self.nchain > _hip.MAX_SITES:
            raise NotImplementedError(f"At most {_hip.MAX_SITES} sites are supported.")
        self.geo_dim = int(geo_dim)
        self.lattice_ls = [int(round(n ** (1 / geo_dim))) for n in nchains]
        for n, L in zip(nchains, self.lattice_ls):
            if L ** geo_dim != n:
                raise ValueError(f"Number of sites {n} is incompatible with dimension {geo_dim}.")
            if periodic and L < 3:
                raise ValueError("Periodic hopping needs a lattice length of at least 3.")
        self.lattice_l = max(self.lattice_ls)
        self.periodic = bool(periodic)
        self.max_ho_dims = numpy.full(self.nchain, int(max_ho_dim), dtype=numpy.int64)
        use_terms = {"diag": {"eps_sys": float(eps_sys), "hbar_omega": float(hbar_omega),
                              "delta_eps": 0.0},
                     "hopping": {"J": 1.0}, "vib_coupling": {"g": 1.0}}
        max_dims = numpy.array([ntraj, self.nchain] + [int(max_ho_dim)] * self.nchain)
        HamiltonianFramework.__init__(self, max_dims=max_dims, use_terms=use_terms,
                                      debug_verb=debug_verb)
        # the diagonal sums the phonon fields only (not the trajectory and position fields)
        masks, kb = _hip.sum_masks(self.bit_offsets, self.bitwidths, self.totwordwidth, first=2)
        self._d_sum_mask = torch.from_numpy(masks.view(numpy.int64).copy()).to(self.device)
        self._layout = _hip.Layout(self.totwordwidth, self.n_sites, self._d_off.data_ptr(),
                                   self._d_width.data_ptr(), self._d_maxdim.data_ptr(),
                                   self._d_sum_mask.data_ptr(), kb)
        self.posbitwidth = int(self.bitwidths[1])
        self.hopping_J = numpy.asarray(hopping_J, dtype=numpy.float64)
        self.vib_g = numpy.asarray(vib_g, dtype=numpy.float64)
        self._d_traj_L = torch.tensor(self.lattice_ls, dtype=torch.int32, device=self.device)
        vals = numpy.stack([self.hopping_J, self.vib_g], axis=1)
        self._d_traj_val = torch.tensor(vals, dtype=torch.float64, device=self.device)
        self._hs = self._make_holstein_struct(self.offdiag_terms)
        self._hs.batched = 1
        self._hs.ntraj = ntraj
        self._hs.d_traj_L = self._d_traj_L.data_ptr()
        self._hs.d_traj_value = self._d_traj_val.data_ptr()

    def get_pos(self, arr):
        return self.get_n_at_site(arr, 1)

    def get_traj(self, arr):
        return self.get_n_at_site(arr, 0)

    def segments(self, rows):
        """seg_off (device int32, ntraj+1): trajectory t owns rows [seg_off[t], seg_off[t+1])."""
        seg = torch.empty(self.ntraj + 1, dtype=torch.int32, device=self.device)
        check(_lib.paces_batch_segments(ptr(rows) if rows.shape[0] else None, rows.shape[0],
                                        self.totwordwidth, self.ntraj, ptr(seg), stream()),
              "batch_segments")
        return seg

    def site_observables_batched_async(self, rows, vector, seg_off, out=None, finish=None,
                                       keys=None, vmax=None):
        """Launches the per-trajectory observables into out (device float64, ntraj * (2N + 1):
        per trajectory n_exc[N], n_pho[N], sum|psi|^2) and returns out; no host sync. finish =
        (term buffers, series states, mixed, k0): the series' finish is first done on vector in
        place (paces_holstein_observables_batched_finish: mixed, terms k0 .. j added to v; FP64,
        the last even terms added); keys (int64, R) receives the rows' select keys and vmax
        (int64, ntraj) each trajectory's largest one."""
        R = rows.shape[0]
        N = self.nchain
        if out is None:
            out = torch.empty(self.ntraj * (2 * N + 1), dtype=torch.float64, device=self.device)
        wsp, wsn = self._ws.get(
            _lib.paces_holstein_observables_batched_workspace_bytes(R, self.ntraj, N))
        terms, ts, mixed, k0 = finish if finish is not None else ([], None, 0, 0)
        tp = (C.c_void_p * max(len(terms), 1))(*[ptr(torch.view_as_real(t)) for t in terms])
        check(_lib.paces_holstein_observables_batched_finish(
            ptr(rows) if R else None, ptr(torch.view_as_real(vector)) if R else None, R,
            ptr(seg_off), C.byref(self._layout), C.byref(self._hs), tp, len(terms), k0,
            1 if mixed else 0, ptr(ts) if ts is not None else None,
            ptr(keys) if keys is not None and R else None,
            ptr(vmax) if vmax is not None else None, ptr(out), wsp, wsn, stream()),
            "holstein_observables_batched_finish")
        return out

    def split_observables(self, flat):
        """(n_exc [ntraj, N], n_pho [ntraj, N], sum|psi|^2 [ntraj]) from the flat host array."""
        N = self.nchain
        h = numpy.asarray(flat).reshape(self.ntraj, 2 * N + 1)
        return h[:, :N], h[:, N:2 * N], h[:, 2 * N]

    def site_observables_batched(self, rows, vector, seg_off):
        """(n_exc [ntraj, N], n_pho [ntraj, N], sum|psi|^2 [ntraj]) per trajectory."""
        out = self.site_observables_batched_async(rows, vector, seg_off)
        return self.split_observables(out.cpu().numpy())


class SeriesGraphPool:
    """Captured series graphs and the device descriptor their kernels read (N1).

    A series graph depends only on the row capacity, the trajectory count, the term count, the
    precision and the launch shape: everything else comes from the descriptor, rewritten before
    each replay. Engines that pass the same pool (on the same device and stream) reuse each
    other's captures, so a run prepared once replays without capturing again."""

    def __init__(self, device):
        self.device = device
        self.desc = torch.empty(C.sizeof(_hip.BatchSeries), dtype=torch.uint8, device=device)
        self.graphs = {}


class BatchedTimeEvolution:
    """Adaptive-basis evolution of every trajectory of a BatchedHamiltonian in one set of
    launches per step: per-trajectory select, one shared expansion/union/CSR/reassignment, and a
    Taylor series whose stopping test runs per trajectory (the series continues until all stop).

    precision: "mixed" (complex64 terms, FP32 matrix, complex128 state) or "fp64".
    """

    def __init__(self, hamobj: BatchedHamiltonian, maxstates, enlarge_steps=1,
                 precision="mixed", m_star=100, taylor_tol=None, autotune=False, graph=False,
                 launch_shape=None, csr_values="slots", explosion_cutoff=2.0,
                 on_observables=None, graph_pool=None, incremental=True, incremental_max=0.05):
        self.hamobj = hamobj
        # every step ends with the per-trajectory norm and site observables (as the reference's
        # step does, time_evolution.py:585-605); they travel to pinned host memory without a
        # stream sync and are read once the next step's adaptation has waited on the device
        # (or by observables()): a norm above explosion_cutoff then raises, and
        # on_observables(i, n_exc, n_pho, norm2) receives each step's values
        self.explosion_cutoff = float(explosion_cutoff)
        self.on_observables = on_observables
        self._obs_dev = None
        self._obs_host = None
        self._obs_pending = None
        self._obs_last = None
        self.device = hamobj.device
        ms = numpy.asarray(maxstates, dtype=numpy.int64)
        if ms.shape != (hamobj.ntraj,) or numpy.any(ms < 1):
            raise ValueError("maxstates needs one positive value per trajectory.")
        self.maxstates = ms
        self._d_ms = torch.tensor(ms, dtype=torch.int64, device=self.device)
        # the largest |psi|^2 of each trajectory (bits of a double), from the observables pass;
        # it holds across an adaptation: the row attaining it is selected (every maxstates is at
        # least 1) and keeps its amplitude. With the complex64 copy of the amplitudes that the
        # incremental emit writes (_x0), the graph's series skips its begin pass over them.
        self._vmax = torch.zeros(len(ms), dtype=torch.int64, device=self.device)
        self._vmax_ok = False
        self._x0 = None
        self._prebegin = bool(numpy.min(ms) >= 1)
        self.enlarge_steps = int(enlarge_steps)
        if precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'.")
        self.precision = precision
        self.m_star = m_star
        self.tol = taylor_tol if taylor_tol is not None else (
            engine.TAYLOR_TOL_MIXED if precision == "mixed" else engine.TAYLOR_TOL)
        self._ws = hamobj._ws
        self._tws = _hip.Workspace(self.device)
        self.autotune = bool(autotune)
        # launch_shape=(threads, rows per thread): a fixed Taylor-term shape instead of the
        # default or the tuned one (every shape gives identical results)
        if launch_shape is not None:
            launch_shape = (int(launch_shape[0]), int(launch_shape[1]))
            from ..core import autotune as _at
            if launch_shape not in _at.SHAPES:
                raise ValueError(f"launch_shape must be one of {_at.SHAPES}")
        self.launch_shape = launch_shape
        # csr_values="slots": one slot matrix for every term (paces_slots_t: a uint32 per
        # (row, op) holding the column and a vib entry's code; open lattices), the values formed
        # per trajectory in LDS; "coded": a CSR per term with a code byte per vib entry
        # (paces_csr_coded_t); "explicit": a CSR per term with FP32 (mixed) or FP64 values.
        # All give identical results.
        if csr_values not in ("slots", "coded", "explicit"):
            raise ValueError("csr_values must be 'slots', 'coded' or 'explicit'.")
        if csr_values == "slots" and hamobj.slot_layout(0) is None:
            csr_values = "coded"  # periodic lattices: wrap-around hops order columns per row
        if csr_values == "coded" and hamobj._hs.sqrt_table_len > _hip.CODE_TABLE:
            csr_values = "explicit"  # codes index a table of at most CODE_TABLE entries
        self.csr_values = csr_values
        # graph=True: the Taylor series runs as one HIP graph per step (captured once per row
        # capacity and term count, replayed with a rewritten device descriptor); it runs at least
        # two terms, and a series with a tolerance of 1 or more could stop at term 0
        self.graph = bool(graph)
        if self.graph and (self.m_star < 2 or not self.tol < 1.0):
            raise ValueError("the graph-replayed series needs m_star >= 2 and taylor_tol < 1")
        self.graph_pool = graph_pool if graph_pool is not None else SeriesGraphPool(self.device)
        # incremental=True: with the slot matrix and one expansion, a step after the first finds
        # its basis by marking the previous basis through the previous slots (_adapt_incremental;
        # the same rows, matrix and amplitudes as the full construction). A step whose marked
        # rows outside the previous expanded basis exceed incremental_max of the previous rows
        # (the growth from a single state) takes the full construction instead.
        self.incremental = bool(incremental) and self.csr_values == "slots" and \
            self.enlarge_steps == 1
        self.incremental_max = float(incremental_max)
        self._prev_be = None  # the expanded basis's rows among whoami (uint8), for the next step
        self.last_adapt = None
        # the last series' finish (x_even, states, mixed), done by the observables pass, and the
        # select keys that pass formed for the current amplitudes
        self._finish = None
        self._keys = None
        self._graph_terms = 8 if precision == "mixed" else 16
        # at most this many terms in the graph (None: no limit); a smaller limit makes more
        # series continue past the graph
        self.graph_terms_max = None
        self.last_launch_shape = None
        self.whoami = None
        self.vector = None
        self.seg_off = None
        self.diag_vals = None
        self.sparse_mats_dict = None
        self._expm_pending = None
        self._expm_info = None

    @property
    def numstates(self):
        return int(self.whoami.shape[0])

    def create_single_exciton_states(self, positions):
        """One basis state per trajectory: exciton at positions[t], phonon vacuum, amplitude 1."""
        h = self.hamobj
        pos = numpy.asarray(positions, dtype=numpy.int64)
        if pos.shape != (h.ntraj,) or numpy.any(pos < 0) or numpy.any(pos >= h.nchains):
            raise ValueError("One exciton position inside each lattice is required.")
        raw = numpy.zeros((h.ntraj, 2 + h.nchain), dtype=numpy.int64)
        raw[:, 0] = numpy.arange(h.ntraj)
        raw[:, 1] = pos
        self.whoami = h.compress_ind_arr(raw)
        self.vector = torch.ones(h.ntraj, dtype=torch.complex128, device=self.device)
        self._keys = None
        self._finish = None
        self._vmax_ok = False
        self._x0 = None
        self._set_segments(h.segments(self.whoami))

    def _select(self):
        R, W = self.whoami.shape
        out = torch.empty((max(R, 1), W), dtype=torch.int64, device=self.device)
        cnt = torch.empty(1, dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_batch_select_workspace_bytes(R, self.hamobj.ntraj))
        check(_lib.paces_batch_select_topk_keys(ptr(torch.view_as_real(self.vector)),
                                                ptr(self._keys) if self._keys is not None else
                                                None, ptr(self.whoami), R, W, ptr(self.seg_off),
                                                self.hamobj.ntraj, ptr(self._d_ms), ptr(out),
                                                ptr(cnt), wsp, wsn, stream()),
              "batch_select_topk_keys")
        host, ev = self._seg_host
        ev.synchronize()  # recorded a step ago: long complete
        n = int(numpy.minimum(numpy.diff(host.numpy()), self.maxstates).sum())
        return out[:n]

    def adapt(self):
        self._x0 = None
        if self.incremental and self._prev_be is not None and self.numstates > 0:
            done = self._adapt_incremental()
            if done is not None:
                return done
            select = self._selected_rows
        else:
            select = self._select()
        # the mixed-precision series reads FP32 values only: the CSR is built with them
        vdt = torch.float32 if self.precision == "mixed" else torch.float64
        ddt = vdt  # the diagonal too (diag_vals is FP32 in the mixed form)
        if self.csr_values in ("coded", "slots"):
            vdt = self.csr_values
        (diag, new_inds), mats, dbg = self.hamobj.generate_mel(select, self.enlarge_steps,
                                                               value_dtype=vdt, diag_dtype=ddt)
        self.vector = engine.reassign(self.whoami, self.vector, new_inds, self._ws)
        self.whoami = new_inds
        self.diag_vals = diag
        self.sparse_mats_dict = mats
        self._prev_be = self.hamobj.last_in_basis if self.incremental else None
        self._keys = None
        self._obs_last = None
        self._set_segments(self.hamobj.segments(new_inds))
        self.last_adapt = "full"
        return dbg

    def _flag_indices(self, flags, n, mask):
        idx = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        cnt = torch.empty(1, dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_flag_indices_workspace_bytes(n))
        check(_lib.paces_flag_indices(ptr(flags) if n else None, n, mask, ptr(idx), ptr(cnt),
                                      wsp, wsn, stream()), "flag_indices")
        return idx, cnt

    def _gather(self, rows, idx, m):
        W = rows.shape[1]
        out = torch.empty((max(m, 1), W), dtype=rows.dtype, device=self.device)
        check(_lib.paces_gather_rows(ptr(rows), W, ptr(idx), m, ptr(out), stream()),
              "gather_rows")
        return out[:m]

    def _union(self, runs):
        """Sorted union of sorted unique row runs [(tensor, offset, length)], with the union
        index of every input row (paces_merge_unique)."""
        rows, _, cpos = engine.merge_unique(runs, self.hamobj.totwordwidth, self._ws,
                                            want_cand_pos=True)
        return rows, cpos

    def _explicit(self, X, flags, bit):
        """Images of the sorted unique rows X (rows outside the previous expanded basis): their
        union U, with every image's index in it; the rows of U found among the previous rows set
        bit in their flags, the others (sorted) are returned as new rows. The second element
        keeps what the pairs and positions need: (X, run offsets, sources, phonon counts, U, the
        images' indices in U, each U row's lower bound among the previous rows, the indices in
        U of the new rows), or None without images."""
        rows0 = self.whoami
        imgs, offs, src, nup = self.hamobj._images(X, want_src=True)
        nimg = offs[-1]
        if not nimg:
            return rows0[:0], None
        U, cpos = self._union([(imgs, offs[k], offs[k + 1] - offs[k])
                               for k in range(len(offs) - 1)])
        nU = U.shape[0]
        idx, hit, _ = engine.lower_bound(rows0, U, want_hit=True, q_sorted=False, ws=self._ws)
        miss = torch.empty(nU, dtype=torch.uint8, device=self.device)
        check(_lib.paces_mark_hits(ptr(idx), ptr(hit), nU, ptr(flags), bit, ptr(miss),
                                   stream()), "mark_hits")
        midx, mcnt = self._flag_indices(miss, nU, 1)
        m = int(mcnt.item())
        new = self._gather(U, midx, m)
        return new, (X, offs, src[:nimg], nup[:nimg], U, cpos, idx, midx[:m])

    @staticmethod
    def _move(dst, src, smap, dmap, m):
        """dst[dmap[j]] = src[smap[j]] for j < m (a None map is the identity)."""
        if m:
            check(_lib.paces_move_i32(ptr(src), ptr(smap) if smap is not None else None,
                                      ptr(dmap) if dmap is not None else None, m, ptr(dst),
                                      stream()), "move_i32")

    def _adapt_incremental(self):
        """The new basis from the previous one (paces_slots_mark / _warm_*): the selection as
        flags over the previous rows; the first and second expansions as marks through the
        previous slot matrix (every row of the previous expanded basis lists all its neighbours
        there); explicit images for the marked rows outside it, whose new neighbours join as new
        rows; then positions, the rows, amplitudes, diagonal and slots moved in one pass, and the
        explicit pairs placed. Returns None after leaving the selection's rows in
        _selected_rows when the explicit rows exceed incremental_max of the previous rows."""
        h = self.hamobj
        rows0, psi0, diag0 = self.whoami, self.vector, self.diag_vals
        R0, W = rows0.shape
        mats0 = self.sparse_mats_dict
        sm0 = next(iter(mats0.values())).slotmat
        lay = sm0.layout
        dev = self.device
        flags = torch.empty(R0, dtype=torch.uint8, device=dev)
        cnt = torch.empty(1, dtype=torch.int32, device=dev)
        wsp, wsn = self._ws.get(_lib.paces_batch_select_workspace_bytes(R0, h.ntraj))
        check(_lib.paces_batch_select_flags(ptr(torch.view_as_real(psi0)),
                                            ptr(self._keys) if self._keys is not None else None,
                                            R0, ptr(self.seg_off), h.ntraj, ptr(self._d_ms),
                                            ptr(self._prev_be), ptr(flags), ptr(cnt), wsp, wsn,
                                            stream()), "batch_select_flags")
        ecnt = torch.zeros(2, dtype=torch.int32, device=dev)
        e1 = torch.empty(R0, dtype=torch.uint8, device=dev)
        e2 = torch.empty(R0, dtype=torch.uint8, device=dev)
        slots0 = ptr(sm0.slots)

        def mark(src, skip, bit, expl, counter):
            check(_lib.paces_slots_mark(slots0, lay.stride, lay.nslots, lay.col_bits, R0,
                                        ptr(flags), src,
                                        skip, bit, ptr(expl), counter, stream()), "slots_mark")

        # bit 0 selected; bit 1 the first expansion (neighbours of selected rows); bit 2 the
        # second (neighbours of the rows the first one added, bit 1 without bit 0); bit 7 (from
        # the select) rows of the previous expanded basis
        mark(1, 0, 2, e1, ptr(ecnt))
        mark(2, 1, 4, e2, ecnt.data_ptr() + 4)
        n1, n2 = ecnt.tolist()
        if n1 + n2 > self.incremental_max * R0:
            sidx, scnt = self._flag_indices(flags, R0, 1)
            self._selected_rows = self._gather(rows0, sidx, int(scnt.item()))
            return None
        # Each explicit pass p leaves (its sources X_p, where their positions in the new basis
        # come from, its image record); the new rows' insertion points, their expanded-basis
        # bytes and every position the explicit pairs need are composed from the lookups of the
        # passes' image unions (no further searches).
        passes = []
        y1 = rows0[:0]
        if n1:
            idx1, _ = self._flag_indices(e1, R0, 1)
            y1, pr = self._explicit(self._gather(rows0, idx1, n1), flags, 2)
            if pr is not None:
                passes.append((pr, [(idx1, None, n1)]))
            # the rows found by the first explicit images are first-expansion rows too
            e2c = torch.zeros(1, dtype=torch.int32, device=dev)
            mark(2, 1, 4, e2, ptr(e2c))
            n2 = int(e2c.item())
        ny1 = y1.shape[0]
        x2 = None
        x2_src = []  # (indices into the previous rows or None for the rows of y1, their places)
        if n2:
            idx2, _ = self._flag_indices(e2, R0, 1)
            x2 = self._gather(rows0, idx2, n2)
            x2_src.append((idx2, None, n2))
        if ny1:
            # the new first-expansion rows are not previous rows: a disjoint sorted union
            if x2 is None:
                x2 = y1
                x2_src.append((None, None, ny1))
            else:
                x2, cx2 = self._union([(x2, 0, n2), (y1, 0, ny1)])
                x2_src = [(idx2, cx2[:n2], n2), (None, cx2[n2:n2 + ny1], ny1)]
        Y, cY = y1, None
        y2pr = None
        if x2 is not None:
            y2, pr = self._explicit(x2, flags, 4)
            if pr is not None:
                passes.append((pr, x2_src))
                y2pr = pr
            if y2.shape[0]:
                if not ny1:
                    Y = y2
                else:
                    Y, cY = self._union([(y1, 0, ny1), (y2, 0, y2.shape[0])])
        nY = Y.shape[0]
        ybe = None
        ins = None
        yidx = {}  # each pass's new rows: their indices in Y (None: the identity)
        if nY:
            ins = torch.empty(nY, dtype=torch.int32, device=dev)
            ybe = torch.zeros(nY, dtype=torch.uint8, device=dev)
            for pr, _ in passes:
                ym = pr[7].shape[0]
                if not ym:
                    continue
                if cY is None:
                    yi = None
                elif pr is y2pr:
                    yi = cY[ny1:ny1 + ym]
                else:
                    yi = cY[:ny1]
                yidx[id(pr)] = yi
                # a new row's insertion point is its lower bound found by its pass's lookup
                self._move(ins, pr[6], pr[7], yi, ym)
            if ny1:
                if cY is None:
                    ybe.fill_(1)
                else:
                    check(_lib.paces_fill_u8_at(ptr(ybe), ptr(cY), ny1, 1, stream()),
                          "fill_u8_at")
        NP = torch.empty(R0 + 1, dtype=torch.int32, device=dev)
        ypos = torch.empty(max(nY, 1), dtype=torch.int32, device=dev)
        tot = torch.empty(1, dtype=torch.int32, device=dev)
        wsp, wsn = self._ws.get(_lib.paces_warm_positions_workspace_bytes(R0))
        check(_lib.paces_warm_positions(ptr(flags), R0, ptr(ins) if nY else None, nY, ptr(NP),
                                        ptr(ypos), ptr(tot), wsp, wsn, stream()),
              "warm_positions")
        # The new basis has at most Rb rows: the outputs are allocated for Rb, the emit and the
        # explicit pairs are launched, and the new size is read from pinned memory while they
        # run (the slot matrix keeps Rb as its stride).
        Rb = R0 + nY
        tot_h = torch.empty(1, dtype=torch.int32, device="cpu", pin_memory=True)
        tot_h.copy_(tot, non_blocking=True)
        tot_ev = torch.cuda.Event()
        tot_ev.record()
        f32 = 1 if diag0.dtype == torch.float32 else 0
        rows1 = torch.empty((max(Rb, 1), W), dtype=rows0.dtype, device=dev)
        psi1 = torch.empty(max(Rb, 1), dtype=torch.complex128, device=dev)
        diag1 = torch.empty(max(Rb, 1), dtype=diag0.dtype, device=dev)
        be1 = torch.empty(max(Rb, 1), dtype=torch.uint8, device=dev)
        slots1 = torch.empty((lay.nslots, max(Rb, 1)), dtype=torch.int32, device=dev)
        # the mixed series' complex64 term input, written with the amplitudes
        x0 = torch.empty(max(Rb, 1), dtype=torch.complex64, device=dev) if f32 else None
        x0p = ptr(torch.view_as_real(x0)) if x0 is not None else None
        check(_lib.paces_warm_emit(ptr(rows0), W, ptr(torch.view_as_real(psi0)), ptr(diag0), f32,
                                   R0, ptr(flags), 3, ptr(NP), slots0, lay.stride, lay.nslots,
                                   lay.col_bits, ptr(rows1), ptr(torch.view_as_real(psi1)),
                                   ptr(diag1), ptr(be1), ptr(slots1), Rb, x0p, stream()),
              "warm_emit")
        if nY:
            ydiag = h.generate_mel_diag(Y, dtype=diag0.dtype)
            check(_lib.paces_warm_emit_new(ptr(Y), W, nY, ptr(ypos), ptr(ydiag), f32, ptr(ybe),
                                           lay.nslots, ptr(rows1),
                                           ptr(torch.view_as_real(psi1)), ptr(diag1), ptr(be1),
                                           ptr(slots1), Rb, x0p, stream()), "warm_emit_new")
        y1pos = None
        if ny1:
            # the rows of y1 in the new basis (they are the first pass's new rows)
            y1pos = torch.empty(ny1, dtype=torch.int32, device=dev)
            self._move(y1pos, ypos, cY[:ny1] if cY is not None else None, None, ny1)
        for pr, xsrc in passes:
            X, offs, src, nup, U, cpos, uidx, umiss = pr
            # the sources' positions: a previous row's through NP, a row of y1's through ypos
            sp = torch.empty(X.shape[0], dtype=torch.int32, device=dev)
            for idx, place, m in xsrc:
                if idx is not None:
                    self._move(sp, NP, idx, place, m)
                else:
                    self._move(sp, y1pos, None, place, m)
            # the images' rows: a previous row's through NP at its lower bound, a new row's
            # through ypos
            nU = U.shape[0]
            up = torch.empty(nU, dtype=torch.int32, device=dev)
            self._move(up, NP, uidx, None, nU)
            ym = umiss.shape[0]
            if ym:
                yi = yidx[id(pr)]
                self._move(up, ypos, yi, umiss, ym)
            run_off = (C.c_int64 * len(offs))(*offs)
            check(_lib.paces_holstein_slots_place_explicit(Rb, run_off, ptr(sp), ptr(up),
                                                           ptr(cpos), ptr(src), ptr(nup),
                                                           C.byref(h._hs), ptr(slots1), stream()),
                  "holstein_slots_place_explicit")
        tot_ev.synchronize()
        R1 = int(tot_h[0])
        lay1 = h.slot_layout(R1)
        if lay1 is None or lay1.col_bits != lay.col_bits or lay1.nslots != lay.nslots:
            raise RuntimeError("the slot layout changed between steps")
        lay1.stride = Rb
        rows1, psi1, diag1, be1 = rows1[:R1], psi1[:R1], diag1[:R1], be1[:R1]
        x0 = x0[:R1] if x0 is not None else None
        sm1 = engine.SlotMatrix(slots1[:, :R1], lay1, (R1, R1))
        self.whoami, self.vector, self.diag_vals = rows1, psi1, diag1
        self.sparse_mats_dict = {name: sm1.term(t, lambda m, rows=rows1: h._coded_values(m, rows))
                                 for t, name in enumerate(h.offdiag_terms)}
        self._prev_be = be1
        self._keys = None
        self._obs_last = None
        self._x0 = x0
        self._set_segments(h.segments(rows1))
        self.last_adapt = ("incremental", n1, n2, nY)
        return {}

    def _set_segments(self, seg_off):
        """The segment offsets, with a pinned host copy made without a stream sync (the next
        select sizes its output from it: sum over trajectories of min(maxstates, rows))."""
        self.seg_off = seg_off
        host = torch.empty(seg_off.shape, dtype=seg_off.dtype, device="cpu", pin_memory=True)
        host.copy_(seg_off, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        self._seg_host = (host, ev)

    def _term_operator(self, mats, mixed):
        """What one batched Taylor term needs, for the CSR form in use: (diag tensor, keep-alive
        objects, call(R, x_in, x_out, res, c, j, tol, ts, hdr, wsp, wsn, cfgref, st) launching one
        term, fill(desc) writing the Hamiltonian into a series descriptor)."""
        h = self.hamobj
        if self.csr_values == "slots":
            diag = engine.to_f32(self.diag_vals) if mixed else self.diag_vals
            sm = mats[0].slotmat
            lay = sm.layout

            def call(R, x_in, x_out, res, c, j, tol, ts, hdr, wsp, wsn, cref, st):
                check(_lib.paces_batch_taylor_term_slots(
                    R, ptr(diag), 1 if mixed else 0, C.byref(lay), x_in, x_out, res, c, j, tol,
                    ptr(self.seg_off), h.ntraj, ts, hdr, wsp, wsn, cref, st),
                    "batch_taylor_term_slots")

            def fill(d):
                d.nterms_h = lay.nterms
                d.coded = 1
                d.sqrt_len = lay.sqrt_len
                d.term_value = lay.term_value
                d.sqrt_table = lay.sqrt_table
                d.slots = lay.slots
                d.slot_stride = lay.stride
                d.nslots = lay.nslots
                d.col_bits = lay.col_bits
                for t in range(_hip.MAX_TERMS + 1):
                    d.slot_off[t] = lay.slot_off[t]
                for t in range(_hip.MAX_TERMS):
                    d.slot_vib[t] = lay.vib[t]
            return diag, (sm,), call, fill
        if self.csr_values == "coded":
            diag = engine.to_f32(self.diag_vals) if mixed else self.diag_vals
            hc = engine.csr_coded(mats, h._hs.sqrt_table_len, h._hs.d_traj_value,
                                  h._hs.d_sqrt_table)

            def call(R, x_in, x_out, res, c, j, tol, ts, hdr, wsp, wsn, cref, st):
                check(_lib.paces_batch_taylor_term_coded(
                    R, ptr(diag), 1 if mixed else 0, C.byref(hc), x_in, x_out, res, c, j, tol,
                    ptr(self.seg_off), h.ntraj, ts, hdr, wsp, wsn, cref, st),
                    "batch_taylor_term_coded")

            def fill(d):
                d.nterms_h = hc.nterms
                for t in range(hc.nterms):
                    d.indptr[t] = hc.indptr[t]
                    d.indices[t] = hc.indices[t]
                    d.data[t] = None
                    d.code[t] = hc.code[t]
                d.coded = 1
                d.sqrt_len = hc.sqrt_len
                d.term_value = hc.term_value
                d.sqrt_table = hc.sqrt_table
            return diag, (hc,), call, fill
        if mixed:
            mm = engine.MixedMats(self.diag_vals, mats)
            diag, sset, h64, h32 = mm.diag, mm.cset, None, C.byref(mm.cset)
            keep = (mm,)
        else:
            sset = engine.csr_set(mats)
            diag, h64, h32 = self.diag_vals, C.byref(sset), None
            keep = (sset,)

        def call(R, x_in, x_out, res, c, j, tol, ts, hdr, wsp, wsn, cref, st):
            check(_lib.paces_batch_taylor_term_tuned(
                R, ptr(diag), h64, h32, x_in, x_out, res, c, j, tol, ptr(self.seg_off), h.ntraj,
                ts, hdr, wsp, wsn, cref, st), "batch_taylor_term")

        def fill(d):
            d.nterms_h = sset.nterms
            for t in range(sset.nterms):
                d.indptr[t] = sset.indptr[t]
                d.indices[t] = sset.indices[t]
                d.data[t] = sset.data[t]
            d.coded = 0
        return diag, keep, call, fill

    def _propagate_graph(self, delta_t, fill, mixed, res, bufs, tbuf, ts, hdr, diag, cfg, st,
                         prebegun=False):
        """Terms 0 .. n-1 by replaying a captured graph; returns n (the eager loop continues
        from there if some trajectory has not stopped). bufs: x[0..2] of the descriptor; tbuf(k):
        the mixed form's buffer of term k."""
        h = self.hamobj
        R = self.numstates
        # capacity: R rounded up in sixteenths of its octave (at most 6.25% spare rows)
        q = 1 << max(8, R.bit_length() - 4)
        cap = ((R + q - 1) // q) * q
        info = self.last_expm  # the previous step's series (arrived long ago: no stall)
        if info is not None and info.size:
            # terms 0 .. j of the slowest trajectory; a longer series continues eagerly
            self._graph_terms = int(info["j"].max()) + 1
        # at least two terms: the graph's series reads v as the partial sum until term 1 stores it
        nterms = max(2, min(self.m_star, self._graph_terms, _hip.SERIES_MAX_TERMS,
                            self.graph_terms_max or _hip.SERIES_MAX_TERMS))
        wsp, wsn = self._tws.get(_lib.paces_batch_taylor_workspace_bytes(cap, h.ntraj))
        d = _hip.BatchSeries()
        d.R = R
        d.diag = ptr(diag)
        fill(d)
        d.v = ptr(torch.view_as_real(self.vector))
        d.res = ptr(torch.view_as_real(res))
        for k in range(3):
            if bufs[k] is not None:
                d.x[k] = ptr(torch.view_as_real(bufs[k]))
        if mixed:
            for k in range(nterms):
                d.xt[k] = ptr(torch.view_as_real(tbuf(k)))
        if prebegun:
            d.vmax = ptr(self._vmax)
        d.seg_off = ptr(self.seg_off)
        d.ts = ptr(ts)
        d.hdr = ptr(hdr)
        d.ws = wsp.value
        d.tol = self.tol
        for j in range(_hip.SERIES_MAX_TERMS):
            d.c[j] = engine.taylor_coeff(delta_t, j)
        pool = self.graph_pool
        # from pinned memory without a host wait (a pageable copy would wait for the stream to
        # drain); the caching host allocator keeps the block until the copy has run
        pool.desc.copy_(torch.frombuffer(bytearray(bytes(d)), dtype=torch.uint8).pin_memory(),
                        non_blocking=True)
        shape = (cfg.block, cfg.rows_per_thread) if cfg is not None else None
        nslots = d.nslots if d.slots else 0  # the term kernel instance: CSR (0) or slots
        key = (cap, h.ntraj, nterms, mixed, shape, nslots, prebegun)
        g = pool.graphs.get(key)
        if g is None:
            g = torch.cuda.CUDAGraph()
            cref = C.byref(cfg) if cfg is not None else None
            with torch.cuda.graph(g):
                check(_lib.paces_batch_series_launch(ptr(pool.desc), cap, h.ntraj,
                                                     1 if mixed else 0, nslots, nterms,
                                                     1 if prebegun else 0, cref,
                                                     stream()), "batch_series_launch")
            pool.graphs[key] = g
        g.replay()
        return nterms

    def propagate(self, delta_t, fuse_finish=False, obs_step=None):
        """The Taylor series of every trajectory. FP64: complex128 terms in two alternating
        buffers, the partial sum stored on odd terms. Mixed: complex64 terms, each in a buffer of
        its own; no term forms the partial sum, and the finish adds t_0 .. t_j to v in place (at
        most SERIES_MAX_TERMS terms). With fuse_finish the step's observables pass (for step
        obs_step) does the finish; in the mixed graph form it is launched right after the graph,
        before the host learns whether every series has stopped, and a series that continues
        past the graph gets its remaining terms added by a second pass (terms are added in
        order, so the sum is the same)."""
        h = self.hamobj
        R = self.numstates
        mats = [self.sparse_mats_dict[t] for t in h.offdiag_terms]
        mixed = self.precision == "mixed"
        vdt = torch.complex64 if mixed else torch.complex128
        tb = []

        def tbuf(k):
            while len(tb) <= k:
                tb.append(torch.empty(R, dtype=vdt, device=self.device))
            return tb[k]

        # the graph's series starts from the emit's complex64 copy of v and the observables'
        # per-trajectory maxima when both describe the current v (no begin pass)
        prebegun = (self.graph and mixed and self._prebegin and self._vmax_ok and
                    self._x0 is not None and self._x0.shape[0] == R)
        # the series changes v: the maxima and the copy describe it again only after the
        # observables pass (launched below with fuse_finish) and the next incremental emit
        self._vmax_ok = False
        x0_keep = self._x0
        self._x0 = None
        if mixed:
            res = self.vector  # the finish forms v + t_0 + ... + t_j in place
            bufs = [None, None, x0_keep if prebegun else
                    torch.empty(R, dtype=vdt, device=self.device)]
            m_star = min(self.m_star, _hip.SERIES_MAX_TERMS)
        else:
            res = torch.empty_like(self.vector)
            bufs = [torch.empty(R, dtype=vdt, device=self.device) for _ in range(3)]
            m_star = self.m_star
        ts = torch.zeros(h.ntraj * C.sizeof(_hip.BatchTaylor), dtype=torch.uint8,
                         device=self.device)
        hdr = torch.zeros(16, dtype=torch.uint8, device=self.device)
        wsp, wsn = self._tws.get(_lib.paces_batch_taylor_workspace_bytes(R, h.ntraj))
        st = stream()
        diag, keep, term, fill = self._term_operator(mats, mixed)
        cfg = self._launch_shape(R, mats, mixed, term, wsp, wsn, st)
        self.last_launch_shape = (cfg.block, cfg.rows_per_thread) if cfg is not None else None
        cref = C.byref(cfg) if cfg is not None else None
        j = 0
        spec = 0  # terms the speculative observables pass finished
        if self.graph:
            j = self._propagate_graph(delta_t, fill, mixed, res, bufs, tbuf, ts, hdr, diag, cfg,
                                      st, prebegun)
            wsp, wsn = self._tws.get(0)  # the graph's workspace (same buffer, grown if needed)
            x_in = tb[j - 1] if mixed else bufs[(j - 1) % 2]
            if fuse_finish and mixed:
                # res is v itself: the observables pass finishes it in place
                self._keys = None
                self._obs_last = None
                self._finish = (tb[:j], ts, mixed, 0)
                self._launch_observables(obs_step)
                spec = j
        else:
            check(_lib.paces_batch_taylor_begin_tuned(R, ptr(torch.view_as_real(self.vector)),
                                                      None if mixed else
                                                      ptr(torch.view_as_real(res)),
                                                      ptr(torch.view_as_real(bufs[2])),
                                                      1 if mixed else 0, ptr(self.seg_off),
                                                      h.ntraj, ptr(ts), wsp, wsn, cref, st),
                  "batch_taylor_begin")
            x_in = bufs[2]
        while j < m_star:
            if self.graph and int(hdr[:4].view(torch.int32).item()) == 0:
                break
            for _ in range(4):
                if j >= m_star:
                    break
                x_out = tbuf(j) if mixed else bufs[j % 2]
                term(R, ptr(torch.view_as_real(x_in)), ptr(torch.view_as_real(x_out)),
                     None if mixed else ptr(torch.view_as_real(res)),
                     engine.taylor_coeff(delta_t, j), j, self.tol, ptr(ts), ptr(hdr), wsp, wsn,
                     cref, st)
                x_in = x_out
                j += 1
            if int(hdr[:4].view(torch.int32).item()) == 0:
                break
        # the finish: mixed, v + t_0 + ... + t_j (the j terms run); FP64, the last term of the
        # trajectories whose series ended on an even term (even terms wrote bufs[0])
        terms = tb[:j] if mixed else [bufs[0]]
        if not fuse_finish:
            tp = (C.c_void_p * len(terms))(*[ptr(torch.view_as_real(t)) for t in terms])
            check(_lib.paces_batch_taylor_finish(R, tp, len(terms), 1 if mixed else 0,
                                                 ptr(torch.view_as_real(res)),
                                                 ptr(self.seg_off), h.ntraj, ptr(ts), wsp, wsn,
                                                 cref, st), "batch_taylor_finish")
        del keep
        self.vector = res
        if j > spec:  # (no second pass when the speculative one saw every term)
            self._keys = None
            self._obs_last = None
            if fuse_finish:
                self._finish = (terms, ts, mixed, spec)
                self._launch_observables(obs_step)
        # the per-trajectory series states go to pinned host memory without a stream sync; they
        # are parsed when first read (last_expm, or the next step's graph size)
        host = torch.empty(ts.shape, dtype=ts.dtype, device="cpu", pin_memory=True)
        host.copy_(ts, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        self._expm_pending = (host, ev)
        self._expm_info = None

    _EXPM_DTYPE = numpy.dtype([("c1", "f8"), ("c1pc2", "f8"), ("ratio", "f8"), ("done", "i4"),
                               ("j", "i4"), ("rn", "f8")])

    @property
    def last_expm(self):
        """The last series' per-trajectory states: c1, c1 + c2, ratio, done and the last term
        index j (a numpy structured array), or None before the first propagation."""
        if self._expm_info is None and self._expm_pending is not None:
            host, ev = self._expm_pending
            ev.synchronize()
            self._expm_info = numpy.frombuffer(host.numpy().tobytes(), dtype=self._EXPM_DTYPE)
            self._expm_pending = None
        return self._expm_info

    def _launch_shape(self, R, mats, mixed, term, wsp, wsn, st):
        """The tuned launch shape of the batched term (None: the default), timed on scratch
        buffers with the stopping test off the first time a size bucket is seen."""
        if self.launch_shape is not None:
            from ..core import autotune
            return autotune.launch_cfg(self.launch_shape)
        if not self.autotune or R == 0:
            return None
        from ..core import autotune
        h = self.hamobj
        vdt = torch.complex64 if mixed else torch.complex128
        tuner = autotune.default_tuner()
        key = tuner.key("btay32" if mixed else "btay64", R, len(mats), h.ntraj, self.csr_values)

        def run(shape):
            cfg = _hip.LaunchCfg(*shape)
            res = torch.empty_like(self.vector)
            xa = torch.empty(R, dtype=vdt, device=self.device)
            xb = torch.empty(R, dtype=vdt, device=self.device)
            ts = torch.zeros(h.ntraj * C.sizeof(_hip.BatchTaylor), dtype=torch.uint8,
                             device=self.device)
            hdr = torch.zeros(16, dtype=torch.uint8, device=self.device)
            check(_lib.paces_batch_taylor_begin_tuned(
                R, ptr(torch.view_as_real(self.vector)), ptr(torch.view_as_real(res)),
                ptr(torch.view_as_real(xa)), 1 if mixed else 0, ptr(self.seg_off), h.ntraj,
                ptr(ts), wsp, wsn, C.byref(cfg), st), "batch_taylor_begin_tuned")
            return lambda: term(R, ptr(torch.view_as_real(xa)), ptr(torch.view_as_real(xb)),
                                ptr(torch.view_as_real(res)), 1e-3, 0, -1.0, ptr(ts), ptr(hdr),
                                wsp, wsn, C.byref(cfg), st)

        return autotune.launch_cfg(tuner.pick(key, autotune.SHAPES, run))

    def step(self, i, delta_t):
        """Step 0 adapts only; later steps adapt then propagate (time_evolution.py:426-458).
        Every step ends with each trajectory's norm, n_exc and n_pho (time_evolution.py:585-605;
        holstein.py:311-322)."""
        dbg = self.adapt()
        self._resolve_observables()  # the last step's: the adaptation has waited on the device
        if i > 0:
            self.propagate(delta_t, fuse_finish=True, obs_step=i)  # launches the observables
        else:
            self._launch_observables(i)
        return dbg

    def _launch_observables(self, i):
        h = self.hamobj
        n = h.ntraj * (2 * h.nchain + 1)
        if self._obs_dev is None or self._obs_dev.numel() != n:
            self._obs_dev = torch.empty(n, dtype=torch.float64, device=self.device)
            self._obs_host = torch.empty(n, dtype=torch.float64, device="cpu", pin_memory=True)
        keys = torch.empty(self.numstates, dtype=torch.int64, device=self.device)
        h.site_observables_batched_async(self.whoami, self.vector, self.seg_off, self._obs_dev,
                                         finish=self._finish, keys=keys, vmax=self._vmax)
        self._vmax_ok = True
        self._finish = None
        self._keys = keys
        self._obs_host.copy_(self._obs_dev, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        self._obs_pending = (i, ev)

    def _resolve_observables(self):
        """Reads the pending step's observables: checks the norms against explosion_cutoff and
        hands the values to on_observables."""
        if self._obs_pending is None:
            return
        i, ev = self._obs_pending
        self._obs_pending = None
        ev.synchronize()
        n_exc, n_pho, nrm2 = self.hamobj.split_observables(self._obs_host.numpy().copy())
        self._obs_last = (n_exc, n_pho, nrm2)
        bad = numpy.nonzero(numpy.sqrt(nrm2) > self.explosion_cutoff)[0]
        if bad.size:
            raise RuntimeError(f"step {i}: the norm of trajectories {bad.tolist()} is above "
                               f"explosion_cutoff {self.explosion_cutoff}: the series diverged")
        if self.on_observables is not None:
            self.on_observables(i, n_exc, n_pho, nrm2)

    def observables(self):
        """(n_exc [ntraj, N], n_pho [ntraj, N], norm^2 [ntraj]) of the current state."""
        if self._obs_pending is not None:
            self._resolve_observables()
            return self._obs_last
        if self._obs_last is not None:
            return self._obs_last
        return self.hamobj.site_observables_batched(self.whoami, self.vector, self.seg_off)
