"""Single-exciton Holstein model on the GPU (paces/models/holstein.py:22-569).

The Hamiltonian's Hilbert-space work (images, union, per-term CSR, diagonal) and the observables
run as fused HIP kernels of libpaces_hip.so. The ``generate_mel_<term>`` plugin methods keep the
reference's contract for callers that use them directly.
"""

from __future__ import annotations

import ctypes as C
import inspect

import numpy
import torch

from .. import _hip
from .._hip import check, ptr, stream
from ..aux.helpers import cartesian_product, debug_lister, obs_attrs
from ..config import INIT_VERBOSITY_LEVEL, devices
from ..core import engine
from ..core.engine import CsrMatrix
from ..core.hamiltonian import HamiltonianFramework, MelTriple
from ..core.observables import ObservablesFramework
from ..core.time_evolution import TimeEvolutionFramework

_lib = _hip.load()


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
    def _images(self, basis, hstruct=None, want_src=False):
        """Image runs of a sorted unique basis: (images tensor, run offsets list) and, with
        want_src, the source basis index and vib upper phonon count of every image."""
        h = self._hs if hstruct is None else hstruct
        n = basis.shape[0]
        nruns = _lib.paces_holstein_num_runs(C.byref(h))
        check(min(nruns, 0), "holstein_num_runs")
        offs = torch.empty(nruns + 1, dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_holstein_images_workspace_bytes(n, C.byref(h)))
        check(_lib.paces_holstein_images_count(ptr(basis) if n else None, n,
                                               C.byref(self._layout), C.byref(h), ptr(offs), wsp,
                                               wsn, stream()), "holstein_images_count")
        offs_h = offs.tolist()
        imgs = torch.empty((max(offs_h[-1], 1), self.totwordwidth), dtype=self.dtype,
                           device=self.device)
        src = nup = None
        if want_src:
            src = torch.empty(max(offs_h[-1], 1), dtype=torch.int32, device=self.device)
            nup = torch.empty(max(offs_h[-1], 1), dtype=torch.uint8, device=self.device)
        check(_lib.paces_holstein_images_emit(ptr(basis) if n else None, n,
                                              C.byref(self._layout), C.byref(h), ptr(imgs),
                                              ptr(src), ptr(nup), wsp, wsn, stream()),
              "holstein_images_emit")
        if want_src:
            return imgs, offs_h, src, nup
        return imgs, offs_h

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

    def generate_mel(self, basis_states, enlarge_steps=0, log_ind=1.2):
        """((diag, new_inds), {term: CsrMatrix}, dbg) for the basis (hamiltonian.py:383-492).

        new_inds = the basis (after ``enlarge_steps`` expansions) plus its neighbours; each term's
        CSR holds the entries with at least one endpoint in that basis.
        """
        if basis_states.shape[1] != self.totwordwidth:
            raise ValueError("Shape of basis states does not match the specified number of bits.")
        basis = basis_states
        for _ in range(enlarge_steps):
            basis = self.enlarge_basis_set(basis)
        imgs, offs, src, nup = self._images(basis, want_src=True)
        runs = [(basis, 0, basis.shape[0])]
        runs += [(imgs, offs[k], offs[k + 1] - offs[k]) for k in range(len(offs) - 1)]
        new_inds, in_basis, cand_pos = engine.merge_unique(runs, self.totwordwidth, self._ws,
                                                           want_from0=True, want_cand_pos=True)
        self._check_basis_size(new_inds.shape[0])
        diag = self.generate_mel_diag(new_inds)
        mats = self._pairs_csr(new_inds.shape[0], basis.shape[0], offs, cand_pos, in_basis, src,
                               nup)
        dbg = {}
        if "vib_coupling" in self.offdiag_terms:
            dbg["vib_coupling"] = [basis.shape[0] - self._vib_plus_count(offs)]
        self.last_basis_e = basis
        return (diag, new_inds), mats, dbg

    def _pairs_csr(self, R, nbasis, offs, cand_pos, in_basis, src, nup):
        """Per-term CSR from the union's candidate positions (paces_holstein_pairs_*)."""
        h = self._hs
        nterms = h.nterms
        run_off = (C.c_int64 * len(offs))(*offs)
        nnz = torch.empty(max(nterms, 1), dtype=torch.int32, device=self.device)
        wsp, wsn = self._ws.get(_lib.paces_holstein_pairs_workspace_bytes(R, C.byref(h)))
        check(_lib.paces_holstein_pairs_count(R, nbasis, run_off, ptr(cand_pos), ptr(in_basis),
                                              ptr(src), C.byref(h), ptr(nnz), wsp, wsn, stream()),
              "holstein_pairs_count")
        nnz_h = nnz.tolist()
        mats = {}
        arrs = [(C.c_void_p * _hip.MAX_TERMS)() for _ in range(3)]
        for t, name in enumerate(self.offdiag_terms):
            ip = torch.empty(R + 1, dtype=torch.int32, device=self.device)
            ix = torch.empty(max(nnz_h[t], 1), dtype=torch.int32, device=self.device)
            dv = torch.empty(max(nnz_h[t], 1), dtype=torch.float64, device=self.device)
            arrs[0][t], arrs[1][t], arrs[2][t] = ip.data_ptr(), ix.data_ptr(), dv.data_ptr()
            mats[name] = CsrMatrix(ip, ix[:nnz_h[t]], dv[:nnz_h[t]], (R, R))
        check(_lib.paces_holstein_pairs_fill(R, nbasis, run_off, ptr(cand_pos), ptr(in_basis),
                                             ptr(src), ptr(nup), C.byref(h), arrs[0], arrs[1],
                                             arrs[2], wsp, wsn, stream()), "holstein_pairs_fill")
        return mats

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

    def generate_mel_diag(self, basis_states):
        """eps_sys + hbar_omega * sum(n) (+ delta_eps * pos) (holstein.py:280-287)."""
        R = basis_states.shape[0]
        out = torch.empty(R, dtype=torch.float64, device=self.device)
        check(_lib.paces_holstein_diag(ptr(basis_states) if R else None, R,
                                       C.byref(self._layout), C.byref(self._hs),
                                       ptr(out) if R else None, stream()), "holstein_diag")
        return out

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


class Observables(ObservablesFramework):
    """Holstein observables (holstein.py:291-396)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._cache_key = None
        self._cache = None

    def primer_hook(self):
        self._cache_key = None

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
