"""TimeEvolutionFramework: the adaptive-basis time step on the GPU
(paces/core/time_evolution.py:27-910).

One step i > 0 (``_evolve_single_step``, 552-613): keep the ``maxstates`` largest |psi| rows,
expand them ``enlarge_steps + 1`` times through the Hamiltonian, build the per-term CSR on the new
basis, carry the amplitudes over, propagate with the Taylor series (or Lanczos), then compute the
norm and the observables. Step 0 adapts without propagating (616-650).
"""

from __future__ import annotations

import abc
import inspect
import math
import os
import pprint
import sys
import time
from dataclasses import dataclass
from typing import Optional

import numpy
import torch

from .. import __version__, _hip
from ..aux.helpers import flatten_dbg_dict
from ..config import (BASE_STEP_LEVEL, EXPM_EVO_LEVEL, HILBERT_SPACE_LEVEL,
                      INIT_VERBOSITY_LEVEL, MEM_INFO_LEVEL, devices)
from . import engine
from .engine import CsrMatrix


@dataclass
class CoeffSaveParams:
    """When to save wavefunction coefficients (time_evolution.py:27-53)."""
    save_every: int = 0
    save_first: bool = False
    save_last: bool = False
    max_i: int = 0

    def save_necessary(self, i: int):
        if self.max_i == 0:
            raise ValueError("max_i should be set before calling this method.")
        necessary = (self.save_first and i == 0)
        necessary |= (self.save_last and i == self.max_i)
        if self.save_every != 0:
            necessary |= (self.save_every is not None and i % self.save_every == 0)
        return necessary


@dataclass
class ExpmParams:
    """Matrix-exponential parameters (time_evolution.py:57-82).

    Three fields extend the reference: ``precision`` ("fp64": the reference's complex128 series;
    "mixed": complex64 terms with FP32 matrix values and accumulation, summed into the complex128
    state), ``taylor_tol`` (None: 2**-53 for fp64, 2**-24 for mixed) and ``autotune`` (pick the
    Taylor-term launch shape by timing the candidates; results are identical).
    """
    m_star: int = 100
    explosion_cutoff: float = 2.0
    use_scaling: bool = False
    renormalize: bool = False
    lanczos_order: int = 0
    precision: str = "fp64"
    taylor_tol: Optional[float] = None
    autotune: bool = False


@dataclass
class TimeEvoParams:
    """Generic time-evolution parameters (time_evolution.py:85-119)."""
    maxstates: int
    enlarge_steps: int
    diagnostics: bool = True
    shuffle_seed: Optional[int] = 0
    weighting_name: str = "coherence"
    tau: float = 0.0
    garbage_tol: float = -1
    use_two_streams: bool = False


class TimeEvolutionFramework:
    """Runs the adaptive-basis evolution of a concrete Hamiltonian/Observables pair."""
    __metaclass__ = abc.ABCMeta

    def __init__(self, hamobj, ObsObj, obs_list, dirname, te_params, expm_params,
                 params_file=None):
        self.use_module = hamobj.use_module
        self.debug_verb = hamobj.debug_verb
        self.start_time = time.time()
        self.hamobj = hamobj
        self.device = hamobj.device
        self.dirname = dirname
        self.te_params = te_params
        self.expm_params = expm_params
        if expm_params.lanczos_order and expm_params.use_scaling:
            raise ValueError("Scaling and squaring cannot be used in conjunction with Lanczos.")
        if expm_params.use_scaling:
            raise NotImplementedError("The scaling-and-squaring expm_multiply branch is not part "
                                      "of this engine; use the Taylor series or Lanczos.")
        if te_params.use_two_streams:
            raise NotImplementedError("Using two streams does not appear to confer a benefit.")
        if te_params.tau != 0:
            raise NotImplementedError("Weighting functions (tau != 0) are not supported.")
        if te_params.garbage_tol >= 0:
            raise NotImplementedError("garbage_tol is not supported.")
        self.weighting_func = None
        if dirname is not None and not os.path.exists(dirname):
            raise FileNotFoundError(f"The results directory {dirname} does not exist.")

        real_valued = all(all(numpy.imag(v) == 0 for v in pd.values())
                          for pd in hamobj.use_terms.values())
        if not real_valued:
            raise NotImplementedError("Not all functions have been adapted to complex-valued"
                                      " off-diagonal Hamiltonian matrix elements.")

        self.obsobj = ObsObj(self, obs_list)
        self.ham_mat = None
        self.whoami = None
        self.numstates = None
        self.vector = None
        self.cso = None
        self.diag_vals = None
        self.wf_name_list = None
        self.log_ind = 0
        self.sparse_mats_dict = None
        self.last_expm_dbg = None
        self.last_dbg_dict = {}
        self._reducer = engine.Reducer(self.device)
        self._taylor_ws = _hip.Workspace(self.device)
        self._ws = hamobj._ws
        self._step_index = 0

        if dirname is None:
            self.params_file = None
        elif params_file is None:
            self.params_file = os.path.join(dirname, f"paces_params_run_{time.time()}.log")
        else:
            self.params_file = params_file
        self.initialize_params_file()

    # ---- params file (217-264) ----------------------------------------------------------------------
    def initialize_params_file(self):
        if self.params_file is None:
            return
        ver = f"### paces_amd version {__version__} ###"
        hashes = "#" * len(ver) + "\n"
        with open(self.params_file, "w", encoding="utf-8") as pf:
            pf.write(hashes + ver + "\n" + hashes + "\n")
            pf.write(devices.desc_str)
            pf.write("### Hilbert space constructed from"
                     f" {self.hamobj.__class__.__module__}.{self.hamobj.__class__.__name__}"
                     " with the following parameters:\n")
            pf.write(pprint.pformat(self.hamobj.input_args, compact=True) + "\n")
            pf.write("\n### Hamiltonian terms and their parameters:\n")
            pf.write(pprint.pformat(self.hamobj.use_terms, width=1) + "\n")
            pf.write("\n### Observables constructed from"
                     f" {self.obsobj.__class__.__module__}.{self.obsobj.__class__.__name__}"
                     " with the following observables:\n")
            pf.write(pprint.pformat([*self.obsobj.obs_dict]) + "\n")
            pf.write("\n### TimeEvolution constructed from"
                     f" {self.__class__.__module__}.{self.__class__.__name__}"
                     " with the following parameters:\n")
            pf.write(pprint.pformat(self.te_params) + "\n")
            pf.write(pprint.pformat(self.expm_params) + "\n")
            pf.write("\n### End of initialization parameter list\n"
                     "#################################################################\n")

    def write_current_params_to_file(self, frame, start_time=None):
        if self.params_file is None:
            return
        start_time = time.time() if start_time is None else start_time
        localtime = time.asctime(time.localtime(start_time))
        args, _, _, values = inspect.getargvalues(frame)
        arg_list = [f"{a}={values[a]}" for a in args if a != "self"]
        mod_obj = f"{self.__class__.__module__}.{self.__class__.__name__}"
        header = f"\nFunction call in {mod_obj} at {start_time} ({localtime} local):\n"
        header += f"  {frame.f_code.co_name}(\n    " + ",\n    ".join(arg_list) + ")\n"
        with open(self.params_file, "a", encoding="utf-8") as f:
            f.write(header)

    # ---- initial basis (271-359) ----------------------------------------------------------------------
    def load_basis_from_file(self, loadfile):
        """Load packed rows saved by this engine (uint64 words, one row per state)."""
        self.write_current_params_to_file(inspect.currentframe())
        arr = numpy.load(loadfile, allow_pickle=False)
        if arr.ndim != 2 or arr.shape[1] != self.hamobj.totwordwidth:
            raise ValueError("Loaded basis set doesn't match total number of bits given by hamobj.")
        if arr.dtype not in (numpy.uint64, numpy.int64):
            raise ValueError("Loaded basis set does not match the uint64 word format.")
        self.whoami = torch.from_numpy(arr.view(numpy.int64).copy()).to(self.device)
        self.numstates = self.whoami.shape[0]

    def grow_optimal_basis(self, n_max=numpy.inf, fillfac=1.0):
        """Enlarge the initial basis while it holds at most maxstates*fillfac states; keep the
        last one that did (time_evolution.py:290-319, including its n_max behaviour)."""
        self.write_current_params_to_file(inspect.currentframe())
        initsize = len(self.whoami)
        whoami = self.whoami
        n = 0
        while len(whoami) <= self.te_params.maxstates * fillfac:
            if n >= n_max:
                break
            n += 1
            previous_whoami = whoami
            whoami = self.hamobj.enlarge_basis_set(previous_whoami)
        self.whoami = previous_whoami  # noqa: F821 -- UnboundLocalError as in the reference
        self.numstates = self.whoami.shape[0]
        if self.debug_verb > INIT_VERBOSITY_LEVEL:
            print(f"Done! Performed {n} enlargements (grew from {initsize} to"
                  f" {len(self.whoami)} states).")
        return n

    def create_initial_vector(self, vector_coeffs, vector_coo, auto_normalize=True):
        """Place coefficients on the listed basis states (time_evolution.py:325-359)."""
        coo = numpy.asarray(vector_coo.cpu() if torch.is_tensor(vector_coo) else vector_coo)
        if coo.ndim != 2:
            raise ValueError("The coo list must be 2D"
                             " (even if it contains only the entry for one basis state).")
        if numpy.unique(coo, axis=0).shape != coo.shape:
            raise ValueError("There are duplicates in the specified vector coordinates.")
        coeffs = numpy.asarray(vector_coeffs, dtype=numpy.complex128)
        if coo.shape[0] != len(coeffs):
            raise ValueError("The number of coefficients does not match the number of coordinates.")
        if auto_normalize:
            coeffs = coeffs / numpy.linalg.norm(coeffs)
        rows = self.hamobj.compress_ind_arr(coo)
        idx = self.hamobj.searchsorted(self.whoami, rows, allow_escapes=False)
        self.vector = torch.zeros(len(self.whoami), dtype=torch.complex128, device=self.device)
        self.vector[idx.long()] = torch.from_numpy(coeffs).to(self.device)

    # ---- matrices (365-400) -------------------------------------------------------------------------
    def create_matrices(self, diag_params, mats_dict):
        self.sparse_mats_dict = mats_dict
        self.diag_vals = diag_params

    def _mats(self):
        return [self.sparse_mats_dict[t] for t in self.hamobj.offdiag_terms]

    def total_ham(self, vector, use_ham_mat=False):
        """H.x: diag*x plus every term's CSR product, in use_terms order (382-400)."""
        return engine.spmv(self.diag_vals, self._mats(), vector)

    def expectation_H(self, vector):
        return self._reducer.vdot(vector, self.total_ham(vector))

    # ---- timeline (426-458) ---------------------------------------------------------------------------
    def generate_timeline(self, t_array, coeff_save_obj):
        self.write_current_params_to_file(inspect.currentframe())
        self._compute_wf_name_list(t_array)
        self.cso = coeff_save_obj
        self.cso.max_i = len(t_array) - 1
        for i, t_curr in enumerate(t_array):
            self.log_ind = 0
            if i == 0:
                self._perform_first_step(t_array[0])
            else:
                self._evolve_single_step(i, t_curr, t_curr - t_array[i - 1])

    # ---- propagators (482-549) ------------------------------------------------------------------------
    def _simple_taylor(self, delta_t):
        ep = self.expm_params
        if ep.precision == "mixed":
            mm = engine.MixedMats(self.diag_vals, self._mats())
            tol = engine.TAYLOR_TOL_MIXED if ep.taylor_tol is None else ep.taylor_tol
            res, dbg = engine.simple_taylor_mixed(mm, self.vector, delta_t, ep.m_star, tol,
                                                  ws=self._taylor_ws, tuner=self._tuner())
        elif ep.precision == "fp64":
            tol = engine.TAYLOR_TOL if ep.taylor_tol is None else ep.taylor_tol
            res, dbg = engine.simple_taylor(self.diag_vals, self._mats(), self.vector, delta_t,
                                            ep.m_star, tol, ws=self._taylor_ws,
                                            tuner=self._tuner())
        else:
            raise ValueError(f"Unknown precision {ep.precision!r} (use 'fp64' or 'mixed').")
        if self.debug_verb > EXPM_EVO_LEVEL:
            print(f"         Taylor series: converged={dbg[0]} after j={dbg[1]}.")
        return res, dbg

    def _tuner(self):
        if not getattr(self.expm_params, "autotune", False):
            return None
        from . import autotune
        return autotune.default_tuner()

    def _lanczos(self, delta_t):
        """Krylov propagator without reorthogonalisation (time_evolution.py:515-549). The
        recurrence runs on the GPU; only the (m+1)^2 tridiagonal matrix is exponentiated on the
        host."""
        import scipy.linalg
        m = self.expm_params.lanczos_order
        red = self._reducer
        small = numpy.zeros((m + 1, m + 1), dtype=complex)
        orig = red.norm(self.vector)
        q = engine.axpby(1.0 / orig, self.vector, 0.0, torch.empty_like(self.vector))
        basis = [q]
        hv = self.total_ham(q)
        a = red.vdot(hv, q)
        engine.axpby(-a, q, 1.0, hv)
        small[0, 0] = a
        for i in range(1, m + 1):
            b = red.norm(hv)
            if b < 1e-15:
                break
            nq = engine.axpby(1.0 / b, hv, 0.0, torch.empty_like(hv))
            hv = self.total_ham(nq)
            engine.axpby(-b, q, 1.0, hv)
            a = red.vdot(hv, nq)
            engine.axpby(-a, nq, 1.0, hv)
            q = nq
            small[i, i] = a
            small[i - 1, i] = b
            small[i, i - 1] = b
            basis.append(q)
        left = scipy.linalg.expm(-1j * delta_t * small)[:, 0]
        res = torch.zeros_like(self.vector)
        for k, vk in enumerate(basis):
            engine.axpby(complex(left[k]) * orig, vk, 1.0, res)
        return res, [False, 0, 0.0, 0.0]

    # ---- steps (552-650) ------------------------------------------------------------------------------
    def _evolve_single_step(self, i, t_curr, delta_t, diag_coo_debug=None):
        if self.debug_verb > BASE_STEP_LEVEL:
            print(f"\nTimestep from {t_curr - delta_t} to {t_curr}.")
        dbg_dict, post = self._generate_new_hilbert_space(self.te_params.enlarge_steps)
        if self.expm_params.lanczos_order > 0:
            self.vector, expm_dbg = self._lanczos(delta_t)
        else:
            self.vector, expm_dbg = self._simple_taylor(delta_t)
        norm = self._reducer.norm(self.vector)
        if self.expm_params.renormalize:
            engine.axpby(1.0 / norm, self.vector, 0.0, self.vector)
            norm = self._reducer.norm(self.vector)
        elif norm > self.expm_params.explosion_cutoff:
            raise RuntimeError("Norm has exceeded preset explosion cutoff value"
                               f" ({self.expm_params.explosion_cutoff})!")
        self.last_expm_dbg = expm_dbg
        self._postprocess(t_curr, norm, i)
        self._save_diagnostic_data(t_curr, norm, post, expm_dbg, dbg_dict)
        self._step_index = i
        return None

    def _perform_first_step(self, t_curr):
        dbg_dict, post = self._generate_new_hilbert_space(self.te_params.enlarge_steps)
        norm = self._reducer.norm(self.vector)
        self.last_expm_dbg = [0, 0, 0, 0]
        self._postprocess(t_curr, norm, 0)
        self._save_diagnostic_data(t_curr, norm, post, [0] * 4, dbg_dict)
        return None

    def _postprocess(self, t_curr, norm, i):
        self.obsobj.primer_hook()
        self.obsobj.compute_all_and_write(self.vector, t_curr, norm, self.log_ind)
        if self.cso is not None and self.dirname is not None and self.cso.save_necessary(i):
            numpy.save(self.wf_name_list[i], self.vector.cpu().numpy())
            numpy.save(self.wf_name_list[i].replace("wf_file_", "whoami_"),
                       self.whoami.cpu().numpy().view(numpy.uint64))

    def _compute_wf_name_list(self, t_array):
        if self.dirname is None:
            self.wf_name_list = None
            return
        t_array = numpy.asarray(t_array)
        smallest_val = numpy.min(numpy.abs(t_array))
        smallest_diff = numpy.diff(numpy.abs(t_array)).min() if len(t_array) > 1 else 1.0
        if smallest_val == 0:
            decnum = numpy.abs(numpy.floor(numpy.log10(smallest_diff))) + 1
        else:
            decnum = numpy.abs(numpy.floor(numpy.log10(min(smallest_val, smallest_diff)))) + 1
        largest_val = numpy.max(numpy.abs(t_array))
        intnum = numpy.abs(numpy.ceil(numpy.log10(largest_val))) if largest_val > 0 else 1
        f_str = "wf_file_{0:0%i.%if}" % (intnum + decnum + 1, decnum)
        wf_dir = os.path.join(self.dirname, "wf_coeffs")
        if not os.path.exists(wf_dir):
            os.mkdir(wf_dir)
        self.wf_name_list = [os.path.join(wf_dir, f_str.format(t)) for t in t_array]

    def _save_diagnostic_data(self, t_curr, norm, norm_energy, expm_dbg, dbg_dict):
        self.last_dbg_dict = dbg_dict
        if not self.te_params.diagnostics or "diagnostics" not in self.obsobj.fname_dict:
            return
        row = [t_curr, norm, time.time() - self.start_time] + list(norm_energy)
        row += [float(x) for x in expm_dbg] + [self.numstates]
        row += [float(x) for x in flatten_dbg_dict(dbg_dict)]
        with open(self.obsobj.fname_dict["diagnostics"], "a", encoding="utf-8") as f:
            numpy.savetxt(f, row, newline=" ")
            f.write("\n")

    # ---- basis adaptation (735-910) ------------------------------------------------------------------
    def _generate_new_hilbert_space(self, enlarge_steps):
        select = self._determine_select_whoami(self.vector)
        self.last_select = select
        (diag_vals, new_inds), mats, dbg = self.hamobj.generate_mel(select,
                                                                    enlarge_steps=enlarge_steps)
        self.numstates = new_inds.shape[0]
        self.create_matrices(diag_vals, mats)
        self.vector = engine.reassign(self.whoami, self.vector, new_inds, self._ws)
        self.whoami = new_inds
        return dbg, self._post_assignment_diagnostics()

    # kept for callers that use the reference's name
    _generate_new_hilbert_space_cupy = _generate_new_hilbert_space

    def _post_assignment_diagnostics(self):
        if not self.te_params.diagnostics:
            return [-1, numpy.nan]
        norm = self._reducer.norm(self.vector)
        energy = self.expectation_H(self.vector)
        if energy.imag > 1e-12:
            raise RuntimeError(f"Energy has non-zero imaginary part: {energy.imag}.")
        return [norm, energy.real]

    def _determine_select_whoami(self, vector, use_weight_function=False, tau=0.0,
                                 garbage_tol=-1):
        """Rows of the maxstates largest |psi| (ties keep the highest indices), sorted."""
        ms = self.te_params.maxstates
        return engine.select_topk(vector, self.whoami, ms, self._ws)
