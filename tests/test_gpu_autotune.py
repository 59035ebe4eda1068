"""Autotuner (launch shapes) through the C ABI.

* Every launch shape of the Taylor-term kernels gives byte-identical results. Shapes are 128 to
  1024 threads per workgroup and 1, 2 or 4 rows per thread. The check covers plain FP64, plain
  mixed precision and batched (FP64 and mixed): partial sums, terms and series state.
* A run with autotuning equals the run without it, bit for bit, and the tuner records a choice
  among the candidates, with a timing for each.
* The scorer's inference tiles (64, 128 and 256 rows) give identical outputs.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 0.05


@pytest.fixture(scope="module")
def env():
    import torch
    from paces_amd.config import devices
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not devices.configure_called:
        devices.configure()
    from paces_amd import _hip
    from paces_amd.core import autotune, engine
    from paces_amd.models import holstein
    return dict(torch=torch, dev=devices.vector_dev, hip=_hip, engine=engine, holstein=holstein,
                autotune=autotune)


def _chain(env, autotune=False, precision="fp64", steps=8):
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    h = env["holstein"]
    ham = h.Hamiltonian(16, [8] * 16, use_terms={
        "diag": {"eps_sys": 0.0, "hbar_omega": 1.0, "delta_eps": 0.0},
        "hopping": {"J": -1.0}, "vib_coupling": {"g": 1.2}})
    te = h.TimeEvolution(ham, h.Observables, ["n_exc", "n_pho"], None,
                         TimeEvoParams(maxstates=20000, enlarge_steps=1, shuffle_seed=None,
                                       diagnostics=False),
                         ExpmParams(precision=precision, autotune=autotune))
    te.create_nonuniform_basis([1] * 16, minpos=8, maxpos=9)
    raw = np.zeros((1, 17), dtype=np.int64)
    raw[0, 0] = 8
    te.create_initial_vector([1.0], raw)
    te._perform_first_step(0.0)
    for i in range(1, steps + 1):
        te._evolve_single_step(i, DT * i, DT)
    return te


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plain_launch_shapes_are_bit_identical(env, precision):
    torch, hip, eng, at = env["torch"], env["hip"], env["engine"], env["autotune"]
    lib = hip.load()
    te = _chain(env)
    R = te.numstates
    mats = te._mats()
    mixed = precision == "mixed"
    if mixed:
        mm = eng.MixedMats(te.diag_vals, mats)
    s = eng.csr_set(mats)
    ws = hip.Workspace(env["dev"])
    wsp, wsn = ws.get(lib.paces_taylor_workspace_bytes(R))
    ref = None
    for shape in [None] + at.SHAPES:
        res = torch.empty_like(te.vector)
        state = torch.zeros(C.sizeof(hip.TaylorState), dtype=torch.uint8, device=env["dev"])
        vdt = torch.complex64 if mixed else torch.complex128
        bufs = [torch.empty(R, dtype=vdt, device=env["dev"]) for _ in range(3)]
        if mixed:
            hip.check(lib.paces_taylor_begin_mixed(
                R, hip.ptr(torch.view_as_real(te.vector)), hip.ptr(torch.view_as_real(res)),
                hip.ptr(torch.view_as_real(bufs[2])), hip.ptr(state), wsp, wsn, hip.stream()))
            x_in = bufs[2]
        else:
            hip.check(lib.paces_taylor_begin(R, hip.ptr(torch.view_as_real(te.vector)),
                                             hip.ptr(torch.view_as_real(res)), hip.ptr(state),
                                             wsp, wsn, hip.stream()))
            x_in = te.vector
        cfg = at.launch_cfg(shape)
        cref = C.byref(cfg) if cfg is not None else None
        for j in range(5):
            x_out = bufs[j % 2]
            c = eng.taylor_coeff(DT, j)
            if mixed:
                hip.check(lib.paces_taylor_term_mixed_tuned(
                    R, hip.ptr(mm.diag), C.byref(mm.cset), hip.ptr(torch.view_as_real(x_in)),
                    hip.ptr(torch.view_as_real(x_out)), hip.ptr(torch.view_as_real(res)), c, j,
                    -1.0, hip.ptr(state), wsp, wsn, cref, hip.stream()))
            else:
                hip.check(lib.paces_taylor_term_tuned(
                    R, hip.ptr(te.diag_vals), C.byref(s), hip.ptr(torch.view_as_real(x_in)),
                    hip.ptr(torch.view_as_real(x_out)), hip.ptr(torch.view_as_real(res)), c, j,
                    -1.0, hip.ptr(state), wsp, wsn, cref, hip.stream()))
            x_in = x_out
        got = (res.cpu().numpy().tobytes(), x_in.cpu().numpy().tobytes(),
               state.cpu().numpy().tobytes())
        if ref is None:
            ref = got
        else:
            assert got == ref, shape


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_launch_shapes_are_bit_identical(env, precision):
    torch, h, at = env["torch"], env["holstein"], env["autotune"]
    rng = np.random.default_rng(4)
    draws = [(int(rng.choice([6, 8, 11, 16])), float(rng.uniform(0.5, 1.5)),
              float(rng.uniform(0.5, 2.0)), int(10 ** rng.uniform(2.3, 3.3))) for _ in range(7)]
    out = []
    for tuned in (False, True):
        bham = h.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                    [d[2] for d in draws])
        bte = h.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                     precision=precision, autotune=tuned)
        bte.create_single_exciton_states([d[0] // 2 for d in draws])
        for i in range(10):
            bte.step(i, DT)
        out.append((bte.whoami.cpu().numpy().tobytes(), bte.vector.cpu().numpy().tobytes(),
                    bte.last_expm.tobytes()))
    assert out[0] == out[1]
    tuner = at.default_tuner()
    keys = [k for k in tuner.timings if k.startswith("btay")]
    assert keys and all(len(tuner.timings[k]) == len(at.SHAPES) for k in keys)
    assert all(tuple(tuner.choices[k]) in at.SHAPES for k in keys)


def test_autotuned_plain_run_equals_default(env):
    torch, at = env["torch"], env["autotune"]
    a = _chain(env, autotune=False, steps=10)
    b = _chain(env, autotune=True, steps=10)
    assert torch.equal(a.whoami, b.whoami)
    assert torch.equal(a.vector, b.vector)
    tuner = at.default_tuner()
    keys = [k for k in tuner.choices if k.startswith("tay64")]
    assert keys
    for k in keys:
        assert tuple(tuner.choices[k]) in at.SHAPES
        print(f"[autotune] {k}: {tuner.choices[k]} "
              f"(default {tuner.timings[k][str(at.DEFAULT_SHAPE)] * 1e6:.1f} us, best "
              f"{min(tuner.timings[k].values()) * 1e6:.1f} us)")


def test_scorer_tiles_identical(env):
    torch = env["torch"]
    from paces_amd.core import scorer
    s = scorer.CandidateScorer(env["dev"], seed=4)
    X = (torch.rand((5000, 64), device=env["dev"]) * 2 - 1).to(torch.bfloat16).view(torch.int16)
    ref = s.infer(X, tile_rows=64)
    for tile in (128, 256):
        assert torch.equal(ref, s.infer(X, tile_rows=tile)), tile
        assert torch.equal(s.infer(X, want_logit=True, tile_rows=64),
                           s.infer(X, want_logit=True, tile_rows=tile)), tile
