"""Benchmark of the adaptive-basis time step on one MI355X.

Default workload: 1D Holstein ring, N=64 sites, periodic, J=1, w0=1, g=1.5, phonon cap 15
(4 bits; rows of 6+256 bits in 5 uint64 words), exciton at site 32, maxstates=2e6, dt=0.02,
enlarge_steps=1, plain complex128/FP64 form. One step = select + expansion + CSR + reassignment +
Taylor propagation + norm + observables, with every buffer in HBM.

Prints one JSON line. ``--gpus`` above 1 exits: this engine runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "state-steps/s at ≤1e-4 density error: batched+learned vs plain, 1 GPU"

WORKLOADS = {
    "ring64": dict(N=64, dim=16, geo=1, periodic=True, J=1.0, g=1.5, pos0=32, ms=2_000_000,
                   dt=0.02, enlarge=1, fillfac=0.1,
                   desc="1D Holstein ring N=64 periodic, J=1, w0=1, g=1.5, cap 15 (5 words), "
                        "exciton at 32, maxstates=2e6, dt=0.02, enlarge_steps=1"),
    "chain16": dict(N=16, dim=8, geo=1, periodic=False, J=1.0, g=1.0, pos0=8, ms=20_000,
                    dt=0.05, enlarge=1, fillfac=0.1,
                    desc="1D Holstein chain N=16 open, J=1, w0=1, g=1, cap 7 (1 word), "
                         "exciton at 8, maxstates=2e4, dt=0.05, enlarge_steps=1"),
}


def use_terms(J, g, w0=1.0):
    return {"diag": {"eps_sys": 0.0, "hbar_omega": w0, "delta_eps": 0.0},
            "hopping": {"J": -J}, "vib_coupling": {"g": g * w0}}


def taylor_term_bytes(R, mats):
    """Algorithmic HBM bytes of one fused Taylor term: diag (8R), each term's indptr (4(R+1)),
    indices+values (12 per nonzero), x_in and x_out (16R each), res read+write (32R)."""
    nnz = sum(m.nnz for m in mats)
    return 8 * R + 4 * (R + 1) * len(mats) + 12 * nnz + 16 * R + 16 * R + 32 * R


def measure_taylor_kernel(te, reps=20):
    """Average duration of paces_taylor_term at the current basis, timed with HIP events on the
    stream the kernels run on."""
    import ctypes as C
    import torch
    from paces_amd import _hip
    from paces_amd.core import engine
    lib = _hip.load()
    mats = te._mats()
    s = engine.csr_set(mats)
    R = te.numstates
    v = te.vector
    res = torch.empty_like(v)
    bufs = (torch.empty_like(v), torch.empty_like(v))
    state = torch.zeros(C.sizeof(_hip.TaylorState), dtype=torch.uint8, device=v.device)
    st = torch.cuda.current_stream()
    ws = _hip.Workspace(v.device)
    wsp, wsn = ws.get(lib.paces_taylor_workspace_bytes(R))
    _hip.check(lib.paces_taylor_begin(R, _hip.ptr(torch.view_as_real(v)),
                                      _hip.ptr(torch.view_as_real(res)), _hip.ptr(state), wsp,
                                      wsn, C.c_void_p(st.cuda_stream)))
    x_in = v
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # tol < 0 never converges, so every launch does the full term
    for j in range(reps + 2):
        if j == 2:
            e0.record(st)
        x_out = bufs[j % 2]
        _hip.check(lib.paces_taylor_term(R, _hip.ptr(te.diag_vals), C.byref(s),
                                         _hip.ptr(torch.view_as_real(x_in)),
                                         _hip.ptr(torch.view_as_real(x_out)),
                                         _hip.ptr(torch.view_as_real(res)), 1e-3, j, -1.0,
                                         _hip.ptr(state), wsp, wsn, C.c_void_p(st.cuda_stream)))
        x_in = x_out
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return ms * 1e-3, taylor_term_bytes(R, mats)


def cpu_baseline(wl, budget_s=25.0):
    """The oracle (NumPy restatement of the reference step) on this host: the same workload from
    the same initial state; step 0 untimed, then whole steps until the budget is used."""
    from oracle import paces_oracle as po
    ham = po.HolsteinOracle(wl["N"], [wl["dim"]] * wl["N"], geo_dim=wl["geo"],
                            use_terms=use_terms(wl["J"], wl["g"]), periodic=wl["periodic"])
    te = po.TimeEvolutionOracle(ham, wl["ms"], wl["enlarge"])
    te.create_single_state_basis(wl["pos0"])
    te.grow_optimal_basis(fillfac=wl["fillfac"])
    raw = np.zeros((1, wl["N"] + 1), dtype=np.int64)
    raw[0, 0] = wl["pos0"]
    te.create_initial_vector([1.0], raw)
    te.step(0, 0.0, 0.0)
    units, elapsed, steps, last, last_prev = 0, 0.0, 0, 0.0, 1
    while steps < 50:
        prev = te.whoami.shape[0]
        # stop before a step predicted (linear in the basis size) to overrun the budget
        if steps and elapsed + last * prev / last_prev > budget_s:
            break
        t0 = time.perf_counter()
        te.step(steps + 1, wl["dt"] * (steps + 1), wl["dt"])
        last = time.perf_counter() - t0
        last_prev = prev
        elapsed += last
        units += min(wl["ms"], prev)
        steps += 1
    return {"value": units / elapsed, "unit": "state-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle steps 1..{steps} of the same workload from its initial state "
                      f"({units} retained state-steps in {elapsed:.1f} s, NumPy single thread)"}


def traj256_draws(n=256, seed=0):
    """Config-3 draws in order: N in {16,24,32,48}, J ~ U[0.5,1.5], g ~ U[0.5,2.0],
    maxstates = round(10^U[3, log10 2e5])."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.choice([16, 24, 32, 48]))
        J = float(rng.uniform(0.5, 1.5))
        g = float(rng.uniform(0.5, 2.0))
        ms = int(round(10 ** rng.uniform(3.0, np.log10(2e5))))
        out.append((N, J, g, ms))
    return out


def run_traj256(a):
    """256 independent 1D Holstein chains (OBC, cap 7, w0=1), exciton at N//2, dt=0.05,
    enlarge_steps=1. form "batched": one BatchedTimeEvolution (mixed precision); form "plain":
    every trajectory stepped by its own plain FP64 engine, one after the other."""
    import torch
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein
    draws = traj256_draws(a.ntraj)
    dt = 0.05
    if a.form == "batched":
        bham = holstein.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                           [d[2] for d in draws])
        bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                            precision=a.precision_batched)
        bte.create_single_exciton_states([d[0] // 2 for d in draws])
        bte.step(0, dt)
        for i in range(1, a.warmup + 1):
            bte.step(i, dt)
            print(f"warmup step {i}: basis {bte.numstates}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        units = 0
        t0 = time.perf_counter()
        for i in range(a.warmup + 1, a.warmup + 1 + a.steps):
            segs = bte.seg_off.cpu().numpy()
            prev = np.diff(segs)
            bte.step(i, dt)
            units += int(np.minimum(prev, bte.maxstates).sum())
            print(f"timed step {i}: basis {bte.numstates}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        rows = bte.numstates
        dtype = "c64 terms / f64 state" if a.precision_batched == "mixed" else "f64"
    else:
        tes = []
        for (N, J, g, ms) in draws:
            ham = holstein.Hamiltonian(N, [8] * N, use_terms=use_terms(J, g))
            te = holstein.TimeEvolution(ham, holstein.Observables, ["n_exc", "n_pho"], None,
                                        TimeEvoParams(maxstates=ms, enlarge_steps=1,
                                                      shuffle_seed=None, diagnostics=False),
                                        ExpmParams())
            te.create_nonuniform_basis([1] * N, minpos=N // 2, maxpos=N // 2 + 1)
            raw = np.zeros((1, N + 1), dtype=np.int64)
            raw[0, 0] = N // 2
            te.create_initial_vector([1.0], raw)
            te._perform_first_step(0.0)
            tes.append(te)
        for i in range(1, a.warmup + 1):
            for te in tes:
                te._evolve_single_step(i, dt * i, dt)
            print(f"warmup step {i}: basis {sum(te.numstates for te in tes)}", file=sys.stderr,
                  flush=True)
        torch.cuda.synchronize()
        units = 0
        t0 = time.perf_counter()
        for i in range(a.warmup + 1, a.warmup + 1 + a.steps):
            for te, d in zip(tes, draws):
                prev = te.numstates
                te._evolve_single_step(i, dt * i, dt)
                units += min(prev, d[3])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        rows = sum(te.numstates for te in tes)
        dtype = "f64"
    line = {"metric": METRIC, "value": units / elapsed, "unit": "state-steps/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic",
            "config": {"workload": f"traj256: {a.ntraj} 1D Holstein chains, N in {{16,24,32,48}}, "
                                   "J~U[0.5,1.5], g~U[0.5,2], cap 7, maxstates~10^U[3,5.3], "
                                   "dt=0.05, enlarge_steps=1 (seed 0)",
                       "form": a.form, "final_basis_rows": int(rows),
                       "sum_maxstates": int(sum(d[3] for d in draws))},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--workload", default="ring64", choices=sorted(WORKLOADS) + ["traj256"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--form", default="batched", choices=["batched", "plain"],
                    help="traj256 only")
    ap.add_argument("--ntraj", type=int, default=256)
    ap.add_argument("--precision-batched", default="mixed", choices=["fp64", "mixed"])
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(trajectories shard across GPUs, see DESIGN.md).", file=sys.stderr)
        return 2

    import torch
    from paces_amd.config import devices
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein

    import contextlib
    with contextlib.redirect_stdout(sys.stderr):  # stdout carries only the JSON line
        devices.configure()
    if a.workload == "traj256":
        return run_traj256(a)
    wl = WORKLOADS[a.workload]
    ham = holstein.Hamiltonian(wl["N"], [wl["dim"]] * wl["N"], geo_dim=wl["geo"],
                               periodic=wl["periodic"], use_terms=use_terms(wl["J"], wl["g"]))
    te = holstein.TimeEvolution(ham, holstein.Observables, ["n_exc", "n_pho"], None,
                                TimeEvoParams(maxstates=wl["ms"], enlarge_steps=wl["enlarge"],
                                              shuffle_seed=None),
                                ExpmParams(precision=a.precision))
    te.create_nonuniform_basis([1] * wl["N"], minpos=wl["pos0"], maxpos=wl["pos0"] + 1)
    te.grow_optimal_basis(fillfac=wl["fillfac"])
    raw = np.zeros((1, wl["N"] + 1), dtype=np.int64)
    raw[0, 0] = wl["pos0"]
    te.create_initial_vector([1.0], raw)

    dt = wl["dt"]
    te._perform_first_step(0.0)
    i = 0
    for _ in range(a.warmup):
        i += 1
        te._evolve_single_step(i, dt * i, dt)
        print(f"warmup step {i}: basis {te.numstates}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    units = 0
    r_units = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        prev = te.numstates
        i += 1
        te._evolve_single_step(i, dt * i, dt)
        units += min(wl["ms"], prev)
        r_units += te.numstates
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"timed {a.steps} steps in {elapsed:.3f} s, basis {te.numstates}", file=sys.stderr,
          flush=True)

    sec, nbytes = measure_taylor_kernel(te) if a.precision == "fp64" else (None, None)
    peak = 8000.0
    if sec is not None:
        achieved = nbytes / sec / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": peak, "unit": "GB/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "kernel": "k_tay_term (paces_taylor_term)", "avg_us": round(sec * 1e6, 2),
                "bytes_per_launch": int(nbytes), "rows": int(te.numstates)}
    else:
        roof = None
    cpu = None if a.no_cpu_baseline else cpu_baseline(wl, a.cpu_budget)

    line = {"metric": METRIC, "value": units / elapsed, "unit": "state-steps/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if a.precision == "fp64" else "c64 terms / f64 state",
            "data": "synthetic",
            "config": {"workload": f"{a.workload}: {wl['desc']}",
                       "form": "plain" if a.precision == "fp64" else "plain-mixed",
                       "unit_of_work": "retained basis state advanced one step "
                                       "(min(maxstates, R_prev) per step)",
                       "r_state_steps_per_s": r_units / elapsed,
                       "final_basis_rows": int(te.numstates)},
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
