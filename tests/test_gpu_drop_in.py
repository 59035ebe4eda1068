"""The reference's example script, sizes scaled down, through the ``paces`` import path: its
call sequence (main_example.py:22-139) runs unchanged and every file it writes matches the
oracle run of the same timeline.

Scaled: a 16-site chain with phonon dimension 8 (main_example: 25 sites, 128), maxstates 2e4
(8e6), 20 steps of 0.05 (100), coefficients saved every 10 steps. The observables are the
example's n_pho, n_exc and total_energy plus V_hop, V_vib_coupl and mu_to_0.
"""

import glob
import os

import numpy
import pytest

from oracle import paces_oracle as po

pytestmark = pytest.mark.gpu


def _rows(path):
    out = []
    with open(path, encoding="utf-8") as f:
        header = f.readline()
        for line in f:
            out.append([float(x) for x in line.split()])
    return header, numpy.array(out)


def test_main_example_sequence_files_match_oracle(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    # ---- the example's script, sizes scaled (main_example.py:8-139) -------------------------
    from paces.config import devices
    from paces.core.time_evolution import CoeffSaveParams, ExpmParams, TimeEvoParams
    from paces.models import holstein

    if not devices.configure_called:
        devices.configure()
    debug_verb = 0
    dirname = str(tmp_path)
    params_file = os.path.join(dirname, "paces_holstein_test_params.log")
    nchain = 16
    max_ho_dim = 8
    term_param_dict = {
        "diag": dict(eps_sys=-1.0, hbar_omega=1.0, delta_eps=0.0),
        "hopping": dict(J=1.0),
        "vib_coupling": dict(g=4.0),
    }
    initpos = nchain // 2
    fillfac = 0.1
    obs_list = ["n_pho", "n_exc", "total_energy", "hopping", "coupling", "mu_to_0"]
    te_params = TimeEvoParams(maxstates=int(2e4), enlarge_steps=1, diagnostics=True,
                              shuffle_seed=0)
    coeff_save_obj = CoeffSaveParams(save_every=10, save_first=True, save_last=True)
    t_array = numpy.arange(0.00, 1.05, 0.05)
    timeline_params = {"t_array": t_array, "coeff_save_obj": coeff_save_obj}
    expm_params = ExpmParams()
    with devices.vector_dev:
        hamobj = holstein.Hamiltonian(nchain=nchain, max_ho_dims=[max_ho_dim] * nchain,
                                      use_terms=term_param_dict, debug_verb=debug_verb)
        te = holstein.TimeEvolution(hamobj, holstein.Observables, obs_list, dirname, te_params,
                                    expm_params, params_file)
        te.create_nonuniform_basis([1] * nchain, minpos=initpos, maxpos=initpos + 1)
        te.grow_optimal_basis(fillfac=fillfac)
        te.create_initial_vector(vector_coeffs=[1], vector_coo=[[initpos] + nchain * [0]],
                                 auto_normalize=True)
        te.generate_timeline(**timeline_params)
    torch.cuda.synchronize()

    # ---- the oracle over the same timeline --------------------------------------------------
    ham = po.HolsteinOracle(nchain, [max_ho_dim] * nchain, use_terms=term_param_dict)
    ote = po.TimeEvolutionOracle(ham, int(2e4), 1)
    ote.create_single_state_basis(initpos)
    ote.grow_optimal_basis(fillfac=fillfac)
    raw = numpy.zeros((1, nchain + 1), dtype=numpy.int64)
    raw[0, 0] = initpos
    ote.create_initial_vector([1.0], raw)
    recs = []
    ote.generate_timeline(t_array, recs.append)
    assert len(recs) == len(t_array)

    def want(name, rec):
        csr = [rec.csrs[t] for t in ham.offdiag_terms]
        if name == "n_exc":
            return rec.obs["n_exc"]
        if name == "n_pho":
            return rec.obs["n_pho"]
        if name == "total_energy":
            return numpy.array([po.energy(rec.diag, csr, rec.vec)])
        if name == "V_hop":
            return po.hopping_bonds(ham, rec.whoami, rec.vec, rec.csrs["hopping"])
        if name == "V_vib_coupl":
            return po.per_position_products(ham, rec.whoami, rec.vec, rec.csrs["vib_coupling"])
        if name == "mu_to_0":
            return numpy.array([po.vacuum_amplitude_sum(ham, rec.whoami, rec.vec)])
        raise KeyError(name)

    obs_dir = os.path.join(dirname, "observables")
    for name in ("n_pho", "n_exc", "total_energy", "V_hop", "V_vib_coupl", "mu_to_0"):
        hr, re_ = _rows(os.path.join(obs_dir, name + ".real"))
        hi, im_ = _rows(os.path.join(obs_dir, name + ".imag"))
        assert hr.startswith("#time norm ") and hi == hr, name
        assert re_.shape[0] == len(t_array) and im_.shape == re_.shape, name
        for i, rec in enumerate(recs):
            assert re_[i, 0] == t_array[i] and im_[i, 0] == t_array[i], (name, i)
            assert abs(re_[i, 1] - rec.obs["norm"]) <= 1e-10, (name, i)
            w = numpy.asarray(want(name, rec), dtype=numpy.complex128)
            assert re_.shape[1] == 2 + w.size, name
            assert numpy.max(numpy.abs(re_[i, 2:] - w.real)) <= 1e-10, (name, i)
            assert numpy.max(numpy.abs(im_[i, 2:] - w.imag)) <= 1e-10, (name, i)

    hd, diag = _rows(os.path.join(obs_dir, "diagnostics.log"))
    assert hd.split() == ["#time", "norm", "seconds_since_start", "post_adapt_norm",
                          "post_adapt_H", "expm_converged", "final_m", "final_expm_term",
                          "rel_error_expm", "numstates", "ceiling_hits"]
    assert diag.shape == (len(t_array), 11)
    for i, rec in enumerate(recs):
        row = diag[i]
        assert row[0] == t_array[i] and abs(row[1] - rec.obs["norm"]) <= 1e-10, i
        assert abs(row[3] - rec.post_norm) <= 1e-10 and abs(row[4] - rec.post_energy) <= 1e-9, i
        conv, j, c12, ratio = rec.expm_dbg
        assert row[5] == float(conv) and row[6] == float(j), i
        assert abs(row[7] - c12) <= 1e-6 * abs(c12) + 1e-300, i
        assert abs(row[8] - ratio) <= 1e-6 * abs(ratio) + 1e-300, i
        assert row[9] == rec.whoami.shape[0], i
        assert row[10] == rec.dbg["vib_coupling"][0], i

    wf_dir = os.path.join(dirname, "wf_coeffs")
    saved = sorted(os.path.basename(p) for p in glob.glob(os.path.join(wf_dir, "*.npy")))
    assert saved == sorted(f"{k}_{t}.npy" for k in ("wf_file", "whoami")
                           for t in ("0.000", "0.500", "1.000")), saved
    for i, t in ((0, "0.000"), (10, "0.500"), (20, "1.000")):
        vec = numpy.load(os.path.join(wf_dir, f"wf_file_{t}.npy"))
        rows = numpy.load(os.path.join(wf_dir, f"whoami_{t}.npy"))
        numpy.testing.assert_array_equal(rows, recs[i].whoami)
        assert numpy.max(numpy.abs(vec - recs[i].vec)) <= 1e-10, i

    text = open(params_file, encoding="utf-8").read()
    assert "### End of initialization parameter list" in text
    for call in ("create_nonuniform_basis(", "grow_optimal_basis(", "generate_timeline("):
        assert call in text, call
