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
        """True when step i's coefficients are due: step 0 with save_first, step max_i with
        save_last, or any multiple of a nonzero save_every. max_i must have been set (the
        timeline sets it to the last step index)."""
        if not self.max_i:
            raise ValueError("CoeffSaveParams.max_i is 0: set it (generate_timeline does) "
                             "before asking which steps to save.")
        if i == 0 and self.save_first:
            return True
        if i == self.max_i and self.save_last:
            return True
        return bool(self.save_every) and i % self.save_every == 0


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
            raise ValueError("ExpmParams: lanczos_order > 0 and use_scaling are exclusive "
                             "propagators; choose one.")
        if te_params.use_two_streams:
            raise NotImplementedError("use_two_streams (the reference's concurrent basis on a "
                                      "second device) is not part of this one-GPU engine.")
        if te_params.tau != 0:
            raise NotImplementedError("Weighting functions (tau != 0) are not supported.")
        if te_params.garbage_tol >= 0:
            raise NotImplementedError("garbage_tol is not supported.")
        self.weighting_func = None
        if dirname is not None and not os.path.exists(dirname):
            raise FileNotFoundError(f"results directory {dirname!r} is missing; create it "
                                    "before building the TimeEvolution.")

        real_valued = all(all(numpy.imag(v) == 0 for v in pd.values())
                          for pd in hamobj.use_terms.values())
        if not real_valued:
            raise NotImplementedError("complex Hamiltonian parameters are not supported: "
                                      "the engine stores H as real CSR values.")

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
    def _params_sections(self):
        """(heading, text) blocks of the params log after its version banner and the device
        description; the headings and the pprint layouts are the log's format."""
        def origin(obj):
            return f"{obj.__class__.__module__}.{obj.__class__.__name__}"
        return [
            (f"### Hilbert space constructed from {origin(self.hamobj)} with the following "
             "parameters:", pprint.pformat(self.hamobj.input_args, compact=True)),
            ("\n### Hamiltonian terms and their parameters:",
             pprint.pformat(self.hamobj.use_terms, width=1)),
            (f"\n### Observables constructed from {origin(self.obsobj)} with the following "
             "observables:", pprint.pformat(list(self.obsobj.obs_dict))),
            (f"\n### TimeEvolution constructed from {origin(self)} with the following "
             "parameters:", pprint.pformat(self.te_params) + "\n"
             + pprint.pformat(self.expm_params)),
        ]

    def initialize_params_file(self):
        """Start the params log: a banner naming the version, framed by lines of '#' as long as
        the banner, a blank line, the device description, then one block per section and the
        closing marker."""
        if self.params_file is None:
            return
        banner = f"### paces_amd version {__version__} ###"
        frame = "#" * len(banner)
        body = []
        for heading, text in self._params_sections():
            body += [heading, text]
        body += ["\n### End of initialization parameter list", "#" * 65]
        with open(self.params_file, "w", encoding="utf-8") as pf:
            pf.write("\n".join([frame, banner, frame, "", ""]) + devices.desc_str
                     + "\n".join(body) + "\n")

    def write_current_params_to_file(self, frame, start_time=None):
        """Append a record of a set-up call to the params log: when it was made (epoch seconds
        and local time) and every argument except self as name=value, one per line."""
        if self.params_file is None:
            return
        when = time.time() if start_time is None else start_time
        info = inspect.getargvalues(frame)
        given = [f"{name}={info.locals[name]}" for name in info.args if name != "self"]
        owner = f"{self.__class__.__module__}.{self.__class__.__name__}"
        record = [f"\nFunction call in {owner} at {when} ({time.asctime(time.localtime(when))}"
                  " local):", f"  {frame.f_code.co_name}(", "    " + ",\n    ".join(given) + ")"]
        with open(self.params_file, "a", encoding="utf-8") as pf:
            pf.write("\n".join(record) + "\n")

    # ---- initial basis (271-359) ----------------------------------------------------------------------
    def load_basis_from_file(self, loadfile):
        """Load packed rows saved by this engine (uint64 words, one row per state)."""
        self.write_current_params_to_file(inspect.currentframe())
        arr = numpy.load(loadfile, allow_pickle=False)
        if arr.ndim != 2 or arr.shape[1] != self.hamobj.totwordwidth:
            raise ValueError(f"{loadfile}: rows of {arr.shape[-1]} words, the layout needs "
                             f"{self.hamobj.totwordwidth}.")
        if arr.dtype not in (numpy.uint64, numpy.int64):
            raise ValueError(f"{loadfile}: {arr.dtype} words, expected uint64.")
        self.whoami = torch.from_numpy(arr.view(numpy.int64).copy()).to(self.device)
        self.numstates = self.whoami.shape[0]

    def grow_optimal_basis(self, n_max=numpy.inf, fillfac=1.0):
        """Grow the initial basis by whole expansions through H (time_evolution.py:290-319).

        Expansions continue while the basis holds at most maxstates * fillfac rows and fewer
        than n_max have been made; the basis before the last expansion is kept, so the result is
        the largest grown basis within the limit, and with n_max binding it holds n_max - 1
        expansions. An initial basis already above the limit raises UnboundLocalError, as the
        reference's loop does. Returns the number of expansions made."""
        self.write_current_params_to_file(inspect.currentframe())
        limit = self.te_params.maxstates * fillfac
        start = int(self.whoami.shape[0])
        current, kept, made = self.whoami, None, 0
        while current.shape[0] <= limit and made < n_max:
            kept = current
            current = self.hamobj.enlarge_basis_set(kept)
            made += 1
        if kept is None:
            raise UnboundLocalError(f"the initial basis ({start} rows) already exceeds "
                                    f"maxstates * fillfac = {limit}")
        self.whoami = kept
        self.numstates = int(kept.shape[0])
        if self.debug_verb > INIT_VERBOSITY_LEVEL:
            print(f"grow_optimal_basis: {made} expansions made, basis of {start} rows grown to "
                  f"{self.numstates}.")
        return made

    def create_initial_vector(self, vector_coeffs, vector_coo, auto_normalize=True):
        """The state with the given coefficients on the listed basis states, zero elsewhere
        (time_evolution.py:325-359). vector_coo holds one field tuple per row; every tuple must
        be a row of the basis."""
        coo = numpy.asarray(vector_coo.cpu() if torch.is_tensor(vector_coo) else vector_coo)
        if coo.ndim != 2:
            raise ValueError(f"vector_coo must be two-dimensional, one field tuple per row "
                             f"(got {coo.ndim} dimensions)")
        if len(numpy.unique(coo, axis=0)) != len(coo):
            raise ValueError("vector_coo lists a basis state more than once")
        coeffs = numpy.asarray(vector_coeffs, dtype=numpy.complex128)
        if len(coeffs) != len(coo):
            raise ValueError(f"{len(coeffs)} coefficients for {len(coo)} basis states")
        if auto_normalize:
            coeffs = coeffs / numpy.linalg.norm(coeffs)
        idx = self.hamobj.searchsorted(self.whoami, self.hamobj.compress_ind_arr(coo),
                                       allow_escapes=False)
        self.vector = torch.zeros(len(self.whoami), dtype=torch.complex128, device=self.device)
        self.vector[idx.long()] = torch.from_numpy(coeffs).to(self.device)

    # ---- matrices (365-400) -------------------------------------------------------------------------
    def create_matrices(self, diag_params, mats_dict):
        self.sparse_mats_dict = mats_dict
        self.diag_vals = diag_params
        if self.expm_params.use_scaling:
            self.create_ham_mat()

    def create_ham_mat(self):
        """The total Hamiltonian as one operator (time_evolution.py:375-379): the diagonal and
        the term CSRs, applied together."""
        from ..aux.expm import HamiltonianOperator
        self.ham_mat = HamiltonianOperator(self.diag_vals, self._mats(), scale=1.0)

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

    def _scaling_and_squaring(self, delta_t):
        """exp(-i dt H) psi by the shifted scaling-and-squaring Taylor method (the reference's
        use_scaling branch, time_evolution.py:580-581, with the time direction of its other
        propagators: it passes 1j * H, which evolves backwards)."""
        from ..aux.expm import HamiltonianOperator, expm_multiply_simple
        op = HamiltonianOperator(self.diag_vals, self._mats(), scale=-1j)
        res, dbg = expm_multiply_simple(op, self.vector, t=delta_t, return_dbg=True,
                                        seed=self._step_index + 1)
        self.last_expm_choice = op.last_choice
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
        elif self.expm_params.use_scaling:
            self.vector, expm_dbg = self._scaling_and_squaring(delta_t)
        else:
            self.vector, expm_dbg = self._simple_taylor(delta_t)
        norm = self._norm_after_propagation()
        self.last_expm_dbg = expm_dbg
        self._postprocess(t_curr, norm, i)
        self._save_diagnostic_data(t_curr, norm, post, expm_dbg, dbg_dict)
        self._step_index = i
        return None

    def _norm_after_propagation(self):
        """The state's norm after a propagation: rescaled to 1 with renormalize, else checked
        against explosion_cutoff (time_evolution.py:585-589)."""
        norm = self._reducer.norm(self.vector)
        cutoff = self.expm_params.explosion_cutoff
        if self.expm_params.renormalize:
            engine.axpby(1.0 / norm, self.vector, 0.0, self.vector)
            return self._reducer.norm(self.vector)
        if norm > cutoff:
            raise RuntimeError(f"the propagated state's norm {norm} is above explosion_cutoff "
                               f"{cutoff}: the series has diverged")
        return norm

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

    @staticmethod
    def wf_time_format(t_array):
        """The fixed-point format of the times in wavefunction file names: enough decimals to
        tell the smallest nonzero |t| and the smallest spacing of |t| apart (one more than the
        order of magnitude of the smaller of the two), and enough width for the largest |t|,
        zero-padded. A single time, or times that are all zero, use one integer digit and the
        decimals of the time itself (the reference cannot name these)."""
        mags = numpy.abs(numpy.asarray(t_array, dtype=numpy.float64))
        small = [float(mags.min())] if mags.min() > 0 else []
        if mags.size > 1:
            small.append(float(numpy.diff(mags).min()))
        decimals = int(abs(math.floor(math.log10(min(small))))) + 1 if small else 1
        top = float(mags.max())
        int_digits = int(abs(math.ceil(math.log10(top)))) if top > 0 else 1
        return "{0:0%d.%df}" % (int_digits + decimals + 1, decimals)

    def _compute_wf_name_list(self, t_array):
        """Names of the wavefunction files of each step, dirname/wf_coeffs/wf_file_<t> with t in
        wf_time_format; the folder is created when missing."""
        if self.dirname is None:
            self.wf_name_list = None
            return
        folder = os.path.join(self.dirname, "wf_coeffs")
        os.makedirs(folder, exist_ok=True)
        fmt = self.wf_time_format(t_array)
        self.wf_name_list = [os.path.join(folder, "wf_file_" + fmt.format(t)) for t in t_array]

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
