"""HamiltonianFramework: packed-basis layout and Hilbert-space primitives on the GPU
(paces/core/hamiltonian.py:46-577).

Rows are stored as (n, W) torch.int64 device tensors whose bits are uint64 words; field 0 (the
exciton position for the Holstein model) sits in the most significant bits of word 0, so the
lexicographic order of rows is the reference's (its uint32 word k pairs become our word k//2).
"""

from __future__ import annotations

import abc
import ctypes as C
import sys
from typing import NamedTuple

import numpy
import torch

from .. import _hip
from .._hip import check, ptr, stream
from ..config import HILBERT_SPACE_DETAILED_LEVEL, devices
from . import engine

if sys.version_info < (3, 10):
    raise RuntimeError("Python >= 3.10 is required.")


class MelTriple(NamedTuple):
    """(inds, vals, mask): images of the masked input rows and their matrix elements
    (paces/core/hamiltonian.py:27-42)."""
    inds: torch.Tensor
    vals: object
    mask: torch.Tensor


def _bitwidth(d):
    return (int(d) - 1).bit_length() if int(d) > 1 else 0


class HamiltonianFramework:
    """Base class of a model Hamiltonian on a packed Fock basis (hamiltonian.py:46-172).

    ``wordsize`` is accepted for signature compatibility and validated as in the reference, but
    rows are always packed into uint64 words on the device (``totwordwidth`` counts those).
    """
    __metaclass__ = abc.ABCMeta

    def __init__(self, max_dims, use_terms, use_complex_type=numpy.complex128, use_module=None,
                 wordsize=32, debug_verb=0, search_mindiff=32):
        if not devices.configure_called:
            raise NameError("The ROCm device must be configured before starting the calculation"
                            " (by calling paces_amd.config.devices.configure).")
        self.input_args = {"max_dims": max_dims, "use_complex_type": use_complex_type,
                           "use_module": use_module, "wordsize": wordsize,
                           "debug_verb": debug_verb, "search_mindiff": search_mindiff}
        self.use_module = use_module
        self.complex_type = use_complex_type
        if numpy.dtype(use_complex_type) != numpy.complex128:
            raise NotImplementedError("The state vector is complex128.")
        self.debug_verb = debug_verb
        self.search_mindiff = search_mindiff
        self.device = devices.vector_dev

        self.max_dims = numpy.asarray(max_dims).astype(numpy.int64)
        self.n_sites = len(self.max_dims)
        self.wordsize = wordsize
        if self.wordsize not in (8, 16, 32, 64):
            raise ValueError(f"The specified wordsize ({self.wordsize}) is illegal.")
        if self.max_dims.max() > 2 ** 32:
            raise ValueError("At least one of the max dimensions exceeds 2**32 states.")
        self.bitwidths = numpy.array([_bitwidth(d) for d in self.max_dims], dtype=numpy.int64)
        self.bit_offsets = numpy.concatenate([[0], numpy.cumsum(self.bitwidths)[:-1]])
        self.totbits = int(self.bitwidths.sum())
        self.totwordwidth = max(1, -(-self.totbits // 64))
        if self.totwordwidth > _hip.MAX_ROW_WORDS:
            raise NotImplementedError(f"Rows of {self.totwordwidth} uint64 words exceed the "
                                      f"compiled limit of {_hip.MAX_ROW_WORDS}.")
        self.dtype = torch.int64  # bit patterns of uint64 words
        self.max_word_dim = 2 ** 31  # int32 row indices (cupy's sparse index limit, 456-457)

        dev = self.device
        self._d_off = torch.tensor(self.bit_offsets, dtype=torch.int32, device=dev)
        self._d_width = torch.tensor(self.bitwidths, dtype=torch.int32, device=dev)
        self._d_maxdim = torch.tensor(self.max_dims, dtype=torch.int32, device=dev)
        masks, kb = _hip.sum_masks(self.bit_offsets, self.bitwidths, self.totwordwidth)
        self._d_sum_mask = torch.from_numpy(masks.view(numpy.int64).copy()).to(dev)
        self._layout = _hip.Layout(self.totwordwidth, self.n_sites, self._d_off.data_ptr(),
                                   self._d_width.data_ptr(), self._d_maxdim.data_ptr(),
                                   self._d_sum_mask.data_ptr(), kb)
        self._ws = _hip.Workspace(dev)

        self.use_terms = use_terms
        self.offdiag_terms = {k: v for k, v in use_terms.items() if k != "diag"}
        self.extra_dbg_list = []
        for term in self.offdiag_terms:
            func = getattr(self, "generate_mel_" + term)
            if hasattr(func, "dbg_list"):
                self.extra_dbg_list += func.dbg_list

    # ---- lookups ----------------------------------------------------------------------------------
    def searchsorted(self, phonebook, findme, allow_escapes=False):
        """Index of each row of findme in the sorted phonebook (hamiltonian.py:175-202).

        With allow_escapes=False an absent row raises ValueError (the reference's device assert).
        """
        idx, hit, miss = engine.lower_bound(phonebook, findme, want_hit=False,
                                            count_misses=not allow_escapes)
        if not allow_escapes and findme.shape[0] and int(miss.item()) != 0:
            raise ValueError(f"{int(miss.item())} rows were not found in the phonebook.")
        return idx

    # ---- packing (hamiltonian.py:210-360) ----------------------------------------------------------
    def compress_ind_arr(self, raw_states):
        """(n, n_sites) small integers -> packed (n, W) rows."""
        raw = torch.as_tensor(numpy.asarray(raw_states.cpu() if torch.is_tensor(raw_states)
                                            else raw_states), device="cpu")
        if raw.ndim != 2 or raw.shape[1] != self.n_sites:
            raise ValueError("Raw states must have one column per site.")
        rawn = raw.numpy().astype(numpy.int64)
        if numpy.any(rawn < 0) or numpy.any(rawn >= self.max_dims[None, :]):
            raise ValueError("A raw state exceeds its site's maximum dimension.")
        d_raw = torch.tensor(rawn, dtype=torch.int32, device=self.device)
        out = torch.empty((raw.shape[0], self.totwordwidth), dtype=self.dtype, device=self.device)
        check(_hip.load().paces_pack_rows(ptr(d_raw), raw.shape[0], C.byref(self._layout),
                                          ptr(out), stream()), "pack_rows")
        return out

    def get_n_at_site(self, arr, index):
        out = torch.empty(arr.shape[0], dtype=torch.int32, device=self.device)
        if arr.shape[0]:
            check(_hip.load().paces_get_field(ptr(arr), arr.shape[0], C.byref(self._layout),
                                              int(index), ptr(out), stream()), "get_field")
        return out

    def decompress_ind_arr(self, arr):
        return torch.stack([self.get_n_at_site(arr, i) for i in range(self.n_sites)], dim=1)

    def add_n_at_site(self, arr, n, index):
        """Add n to field ``index`` of every row, in place (carries cross words)."""
        if n == 0 or arr.shape[0] == 0:
            return
        check(_hip.load().paces_add_field(ptr(arr), arr.shape[0], C.byref(self._layout),
                                          int(index), int(n), stream()), "add_field")

    @abc.abstractmethod
    def generate_mel_diag(self, basis_states):
        return

    def enlarge_basis_set(self, basis_states):
        """Basis plus its first-degree neighbours (hamiltonian.py:550-577): every off-diagonal
        term's images (generate_mel_<term>(basis, raw_map_to=True)) are concatenated after the
        basis and sorted with duplicates dropped on the device (paces_sort_rows). Models whose
        image runs are sorted override this with the union of runs."""
        parts = [basis_states]
        for term in self.offdiag_terms:
            imgs = getattr(self, "generate_mel_" + term)(basis_states, raw_map_to=True)
            parts.extend(x for x in imgs if x.shape[0])
        allrows = torch.cat(parts, dim=0) if len(parts) > 1 else basis_states
        return engine.sort_rows(allrows.contiguous(), self._ws, unique=True)

    def generate_mel(self, basis_states, enlarge_steps=0, log_ind=1.2):
        """((diag_vals, new_inds), {term: CsrMatrix}, dbg_dict) for the given basis
        (hamiltonian.py:383-492). Models implement it with their fused device kernels."""
        raise NotImplementedError("This model does not provide device matrix elements.")

    def _check_basis_size(self, statenum):
        if statenum >= 2147483648:
            raise ValueError("The number of states is too large for signed 32-bit-integer indices.")
        if self.debug_verb > HILBERT_SPACE_DETAILED_LEVEL:
            print(f"      New basis holds {statenum} states.")
