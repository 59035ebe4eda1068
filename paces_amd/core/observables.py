"""ObservablesFramework: observable registry and file output (paces/core/observables.py:13-123).

Observables are ``calculate_<name>(vector)`` methods decorated with ``@obs_attrs(fname=, header=)``;
results are written as rows ``t norm values...`` to ``observables/<fname>.real`` and ``.imag``.
When the TimeEvolution has ``dirname=None`` nothing is written (the values are still kept in
``last_results``).
"""

import abc
import os

import numpy

from ..aux.helpers import obs_attrs
from ..config import INIT_VERBOSITY_LEVEL, OBSERVABLES_LEVEL


class ObservablesFramework:
    __metaclass__ = abc.ABCMeta

    def __init__(self, teobj, obs_list):
        if teobj.debug_verb > INIT_VERBOSITY_LEVEL:
            print("Initializing Observables object...")
        self.teobj = teobj
        self.hamobj = teobj.hamobj
        self.use_module = teobj.hamobj.use_module
        self.norm = None
        self.obs_dict = {}
        self.fname_dict = {}
        self.last_results = {}
        write = teobj.dirname is not None
        obs_dir = os.path.join(teobj.dirname, "observables") if write else None
        if write and not os.path.exists(obs_dir):
            os.mkdir(obs_dir)
        for key in obs_list:
            try:
                meth = getattr(self, "calculate_" + key)
            except AttributeError as exc:
                raise KeyError(f"Calculation method for {key} not found.") from exc
            self.obs_dict[key] = meth
            try:
                fname = meth.fname
            except AttributeError as exc:
                raise AttributeError(f"The observables function {meth} is missing a decorator"
                                     " specifying the fname for the observables file.") from exc
            if write:
                self.fname_dict[key] = os.path.join(obs_dir, fname)
                for part in ("real", "imag"):
                    with open(".".join([self.fname_dict[key], part]), "w", encoding="utf-8") as f:
                        f.write("#time norm " + meth.header + "\n")
        if write and teobj.te_params.diagnostics:
            extra = " ".join(self.hamobj.extra_dbg_list)
            if extra:
                extra = " " + extra
            self.fname_dict["diagnostics"] = os.path.join(obs_dir, "diagnostics.log")
            with open(self.fname_dict["diagnostics"], "w", encoding="utf-8") as f:
                f.write("#time norm seconds_since_start post_adapt_norm post_adapt_H"
                        " expm_converged final_m final_expm_term rel_error_expm numstates"
                        + extra + "\n")

    def primer_hook(self):
        """Per-pass preparation; subclasses override when needed (observables.py:84-92)."""

    @obs_attrs(fname="total_energy", header="energy")
    def calculate_total_energy(self, vector):
        """<psi|H|psi> (observables.py:95-99)."""
        return self.teobj.expectation_H(vector)

    def compute_all_and_write(self, vector, t, norm, log_index):
        """Evaluate every registered observable and append it to its files (101-123)."""
        for i, name in enumerate(self.obs_dict):
            result = self.obs_dict[name](vector)
            self.last_results[name] = result
            if self.teobj.debug_verb > OBSERVABLES_LEVEL:
                print(f"      {log_index}.{i + 1}a: Calculated {name}.")
            if name not in self.fname_dict:
                continue
            arr = numpy.asarray(result)
            for part in ("real", "imag"):
                with open(".".join([self.fname_dict[name], part]), "a", encoding="utf-8") as f:
                    numpy.savetxt(f, numpy.append([t, norm], getattr(arr, part)), newline=" ")
                    f.write("\n")
