"""Benchmark of the adaptive-basis time step on one MI355X.

Default workload (traj256): 256 independent 1D Holstein chains (open), N in {16,24,32,48},
J ~ U[0.5,1.5], w0=1, g ~ U[0.5,2], phonon cap 7, maxstates ~ 10^U[3,5.3], exciton at N//2,
dt=0.05, enlarge_steps=1, all advanced together by the batched form (complex64 Taylor terms,
complex128 state). A step is the whole step of the reference's timeline: select, expansions,
CSR, reassignment, Taylor series, then every trajectory's norm (checked against the explosion
cutoff) and site observables. Before the timed run the same run is made once, untimed, so that
the launch-shape tuner and the series graphs have seen every size the timed steps meet.

The line also carries: the plain FP64 form's throughput on the same trajectories over the same
steps (one engine per trajectory); the batched form's largest site-density error against the
FP64 oracle at every 10th step and the last, on the trajectories the oracle runs in seconds;
the Taylor term's HBM roofline; the oracle's throughput on this host's cores; and, under
"learned", the scorer-pruned form on its own workload (one 16-site chain at maxstates 2e6, the
scorer trained untimed on 1e8 labels logged from plain runs at three other couplings) against
the plain FP64 form over the same steps, with the scorer's MFMA roofline.

``--workload ring64|chain16`` runs one plain-form trajectory instead (64-site periodic ring,
cap 15, maxstates 2e6, dt 0.02; or 16-site open chain, cap 7, maxstates 2e4, dt 0.05);
``--workload learned`` the scorer-pruned form on the 16-site chain at maxstates 2e6.

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
    """Trajectory draws in order: N in {16,24,32,48}, J ~ U[0.5,1.5], g ~ U[0.5,2.0],
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


TRAJ_DT = 0.05
TRAJ_DESC = ("traj256: {n} independent 1D Holstein chains (open), N in {{16,24,32,48}}, "
             "J~U[0.5,1.5], w0=1, g~U[0.5,2], phonon cap 7, maxstates~10^U[3,5.3], exciton at "
             "N//2, dt=0.05, enlarge_steps=1 (draws seeded 0)")


def btay_term_bytes(R, mats, mixed, form="explicit"):
    """Algorithmic HBM bytes of one batched Taylor term over R rows: diagonal (4 or 8 per row),
    the matrix, x_in and x_out (8 or 16 per row each), and in the FP64 form the complex128
    partial sum, read on every term and written on odd terms (16 + 8 per row, the mean of an even
    and an odd term: the roofline loop times both alike); the mixed form's terms neither read nor
    write it. The matrix: slot form, 4 bytes per slot (nslots per row, every term); CSR forms,
    each term's indptr (4(R+1)) and indices + values (4 + 4 or 8 per nonzero; coded: 4 + 1 per
    vib nonzero and 4 per hopping nonzero)."""
    f = 4 if mixed else 8
    res = 0 if mixed else 24
    if form == "slots":
        return f * R + 4 * mats[0].slotmat.layout.nslots * R + 4 * f * R + res * R
    nnz = sum(m.nnz for m in mats)
    if form == "coded":
        vals = sum(m.nnz for m in mats if getattr(m, "code", None) is not None)
    else:
        vals = f * nnz
    return f * R + 4 * (R + 1) * len(mats) + 4 * nnz + vals + 4 * f * R + res * R


# HBM bytes per row of one k_btay_term<float, B, RR, NS> launch, per matrix form and launch
# shape (the tuner picks among shapes within a few percent of each other), from FETCH_SIZE and
# WRITE_SIZE counter passes over the roofline loop below with the shape fixed (--term-shape),
# each scaled by the counter's reading for a 1 GiB device copy
# (profiles/k_btay_term_traj256_hbm_traffic.txt). Other forms and shapes report null.
BTAY_TRAFFIC_BYTES_PER_ROW = {("mixed", (256, 4), "slots"): 42.27}


def measure_batched_term(bte, reps=20):
    """Average duration of paces_batch_taylor_term_tuned (k_btay_term + its per-trajectory
    reduction) at the current basis, in the launch shape the timed steps used, with the stopping
    test disabled (tol < 0) so every launch processes every row; HIP events on the stream the
    kernels run on."""
    import ctypes as C
    import torch
    from paces_amd import _hip
    from paces_amd.core import engine
    lib = _hip.load()
    h = bte.hamobj
    R = bte.numstates
    mixed = bte.precision == "mixed"
    mats = [bte.sparse_mats_dict[t] for t in h.offdiag_terms]
    vdt = torch.complex64 if mixed else torch.complex128
    res = torch.empty_like(bte.vector)
    bufs = [torch.empty(R, dtype=vdt, device=bte.device) for _ in range(3)]
    ts = torch.zeros(h.ntraj * C.sizeof(_hip.BatchTaylor), dtype=torch.uint8, device=bte.device)
    hdr = torch.zeros(16, dtype=torch.uint8, device=bte.device)
    ws = _hip.Workspace(bte.device)
    wsp, wsn = ws.get(lib.paces_batch_taylor_workspace_bytes(R, h.ntraj))
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)
    shape = bte.last_launch_shape
    cfg = _hip.LaunchCfg(*shape) if shape else None
    cref = C.byref(cfg) if cfg is not None else None
    vr = _hip.ptr(torch.view_as_real(bte.vector))
    _hip.check(lib.paces_batch_taylor_begin_tuned(R, vr, _hip.ptr(torch.view_as_real(res)),
                                                  _hip.ptr(torch.view_as_real(bufs[2])),
                                                  1 if mixed else 0, _hip.ptr(bte.seg_off),
                                                  h.ntraj, _hip.ptr(ts), wsp, wsn, cref, sp))
    # the same term operator (coded or explicit CSR values) the timed steps launched
    diag, keep, term, _ = bte._term_operator(mats, mixed)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    x_in = bufs[2]
    for j in range(reps + 2):
        if j == 2:
            e0.record(st)
        x_out = bufs[j % 2]
        term(R, _hip.ptr(torch.view_as_real(x_in)), _hip.ptr(torch.view_as_real(x_out)),
             _hip.ptr(torch.view_as_real(res)), engine.taylor_coeff(TRAJ_DT, j), j, -1.0,
             _hip.ptr(ts), _hip.ptr(hdr), wsp, wsn, cref, sp)
        x_in = x_out
    e1.record(st)
    e1.synchronize()
    del keep, diag
    return (e0.elapsed_time(e1) / reps * 1e-3,
            btay_term_bytes(R, mats, mixed, bte.csr_values), reps, shape)


def batched_leg(draws, precision, warmup, steps, label, graph=False, autotune=False,
                launch_shape=None, csr_values="slots", record_steps=(), graph_pool=None,
                verbose=True):
    """Batched run over t = TRAJ_DT * arange(warmup + steps + 1) (each step's dt is the
    difference of successive times, as in the reference's timeline): step 0 and ``warmup``
    steps untimed, then ``steps`` timed whole steps (select, expansions, CSR, reassignment,
    Taylor series, per-trajectory norm and site observables). Returns (state-steps, seconds,
    the evolution object, {step: n_exc} for the steps in record_steps)."""
    import torch
    from paces_amd.models import holstein
    recorded = {}
    record_steps = set(record_steps)

    def keep(i, n_exc, n_pho, nrm2):
        if i in record_steps:
            recorded[i] = n_exc.copy()

    bham = holstein.BatchedHamiltonian([d[0] for d in draws], 8, [-d[1] for d in draws],
                                       [d[2] for d in draws])
    bte = holstein.BatchedTimeEvolution(bham, [d[3] for d in draws], enlarge_steps=1,
                                        precision=precision, graph=graph, autotune=autotune,
                                        launch_shape=launch_shape, csr_values=csr_values,
                                        on_observables=keep, graph_pool=graph_pool)
    bte.create_single_exciton_states([d[0] // 2 for d in draws])
    t_arr = TRAJ_DT * np.arange(warmup + steps + 1)
    bte.step(0, 0.0)
    for i in range(1, warmup + 1):
        bte.step(i, t_arr[i] - t_arr[i - 1])
        if verbose:
            print(f"[{label}] warmup step {i}: basis {bte.numstates}", file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    prevs = []  # each step's incoming segment offsets, counted after the timed region
    w0 = time.monotonic_ns()
    t0 = time.perf_counter()
    for i in range(warmup + 1, warmup + 1 + steps):
        prevs.append(bte.seg_off)
        bte.step(i, t_arr[i] - t_arr[i - 1])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if verbose:
        # the timed window on CLOCK_MONOTONIC, to cut a kernel trace to the timed steps
        print(f"[{label}] timed window monotonic_ns {w0} {time.monotonic_ns()}",
              file=sys.stderr, flush=True)
    bte.observables()  # the last step's norms are checked and recorded
    units = sum(int(np.minimum(np.diff(p.cpu().numpy()), bte.maxstates).sum()) for p in prevs)
    if verbose:
        print(f"[{label}] timed {steps} steps in {elapsed:.3f} s, basis {bte.numstates}",
              file=sys.stderr, flush=True)
    return units, elapsed, bte, recorded


def plain_leg(draws, warmup, steps):
    """Every trajectory stepped by its own plain FP64 engine, one after the other, over the
    same steps as the batched leg (untimed step 0 and warmup steps, then ``steps`` timed whole
    steps with norm and observables)."""
    import torch
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein
    t_arr = TRAJ_DT * np.arange(warmup + steps + 1)
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
    for i in range(1, warmup + 1):
        for te in tes:
            te._evolve_single_step(i, t_arr[i], t_arr[i] - t_arr[i - 1])
        print(f"[plain] warmup step {i}: basis {sum(te.numstates for te in tes)}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    units = 0
    t0 = time.perf_counter()
    for i in range(warmup + 1, warmup + 1 + steps):
        for te, d in zip(tes, draws):
            prev = te.numstates
            te._evolve_single_step(i, t_arr[i], t_arr[i] - t_arr[i - 1])
            units += min(prev, d[3])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"[plain] timed {steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    return units, elapsed


def host_cpu():
    """(CPU model name, cores this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def oracle_runs(draws, nsteps, record, max_ms, workers):
    """The oracle on every draw with maxstates <= max_ms, one process per worker (spawned, so
    nothing of this process's GPU state is inherited). Returns ({draw index: {step: n_exc}},
    retained state-steps, wall seconds of the runs, worker count)."""
    import multiprocessing as mp
    from oracle import pool as opool
    idx = [k for k, d in enumerate(draws) if d[3] <= max_ms]
    # the largest first, so the pool ends evenly
    idx.sort(key=lambda k: -draws[k][3])
    jobs = [(*draws[k], TRAJ_DT, nsteps, tuple(record)) for k in idx]
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        pool.map(opool.ready, range(workers))  # the workers are up before the clock starts
        t0 = time.perf_counter()
        res = pool.map(opool.run_chain, jobs, chunksize=1)
        wall = time.perf_counter() - t0
    return ({k: r[0] for k, r in zip(idx, res)}, sum(r[1] for r in res), wall, len(jobs))


def run_traj256(a):
    """Headline line: the batched form (default complex64 terms) of the 256-trajectory workload,
    beside the plain FP64 form on the same trajectories over the same steps, the batched form's
    site-density error against the FP64 oracle, the Taylor term's roofline and the oracle's CPU
    throughput."""
    import torch
    from paces_amd.models import holstein
    draws = traj256_draws(a.ntraj)
    prec = a.precision_batched
    last = a.warmup + a.steps
    record = sorted(set(range(10, last + 1, 10)) | {last})
    pool = None
    if a.graph or a.autotune:
        # untimed preparation: the same run once through, so that the tuner has chosen the
        # Taylor term's launch shape for every size the run meets and the series graphs of
        # those sizes are captured (the timed run replays them)
        dev = torch.device("cuda", torch.cuda.current_device())
        pool = holstein.SeriesGraphPool(dev)
        t0 = time.perf_counter()
        *_, prep, _ = batched_leg(draws, prec, a.warmup, a.steps, "prepare", graph=a.graph,
                                  autotune=a.autotune, launch_shape=a.term_shape,
                                  csr_values=a.csr_values, graph_pool=pool, verbose=False)
        del prep
        torch.cuda.empty_cache()
        print(f"[prepare] tuned and captured in {time.perf_counter() - t0:.2f} s",
              file=sys.stderr, flush=True)
    units, elapsed, bte, recorded = batched_leg(
        draws, prec, a.warmup, a.steps, "batched-" + prec, graph=a.graph, autotune=a.autotune,
        launch_shape=a.term_shape, csr_values=a.csr_values, record_steps=record,
        graph_pool=pool)
    rows = bte.numstates
    value = units / elapsed
    bte_csr_values = bte.csr_values
    nslots = next(iter(bte.sparse_mats_dict.values())).slotmat.layout.nslots \
        if bte.csr_values == "slots" else 0
    sec, nbytes, reps, shape = measure_batched_term(bte)
    kshape = "%d, %d" % shape if shape else "256, 1"
    achieved = nbytes / sec / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4),
            "traffic": (int(BTAY_TRAFFIC_BYTES_PER_ROW[(prec, shape, bte.csr_values)] * rows)
                        if (prec, shape, bte.csr_values) in BTAY_TRAFFIC_BYTES_PER_ROW else None),
            "kernel": "k_btay_term<%s, %s, %d> (paces_batch_taylor_term_%s, %s, the timed "
                      "steps' launch shape, stopping test off)"
                      % ("float" if prec == "mixed" else "double", kshape,
                         nslots, {"slots": "slots", "coded": "coded"}.get(bte.csr_values, "tuned"),
                         {"slots": "slot matrix", "coded": "coded CSR values",
                          "explicit": "explicit CSR values"}[bte.csr_values]),
            "avg_us": round(sec * 1e6, 2), "bytes_per_launch": int(nbytes), "rows": int(rows),
            "launches_timed": reps}
    del bte, pool
    torch.cuda.empty_cache()
    extra = {}
    if not a.no_plain:
        pu, pe = plain_leg(draws, a.warmup, a.steps)
        extra["plain_gpu_state_steps_per_s"] = pu / pe
        extra["plain_ms_per_step"] = 1e3 * pe / a.steps
        extra["speedup_vs_plain"] = value / (pu / pe)
        torch.cuda.empty_cache()
    cpu = None
    if not (a.no_check and a.no_cpu_baseline):
        model, host_cores = host_cpu()
        workers = max(1, min(16, host_cores))
        ref, ounits, owall, njobs = oracle_runs(draws, last, record, a.oracle_max_ms, workers)
        if not a.no_check:
            err = 0.0
            for k, per_step in ref.items():
                N = draws[k][0]
                for i, n_exc in per_step.items():
                    err = max(err, float(np.max(np.abs(recorded[i][k, :N] - n_exc))))
            extra["density_err_vs_fp64_oracle"] = err
            extra["density_err_checked"] = (f"max over sites of |n_exc - n_exc(oracle)| at steps "
                                            f"{record}, trajectories with maxstates <= "
                                            f"{a.oracle_max_ms} ({njobs} of {len(draws)})")
        if not a.no_cpu_baseline:
            cpu = {"value": ounits / owall, "unit": "state-steps/s", "cores": workers,
                   "kind": "port", "cpu_model": model, "host_cores": host_cores,
                   "sample": f"oracle (NumPy) on the {njobs} draws with maxstates <= "
                             f"{a.oracle_max_ms}, steps 1..{last} from the initial state, one "
                             f"process per core: {ounits} retained state-steps in {owall:.1f} s "
                             f"wall"}
    if a.learned_leg:
        # the learned form on its own workload (one chain at maxstates 2e6), reported beside
        # the headline: its throughput against the plain FP64 form over the same steps
        torch.cuda.empty_cache()
        r = learned_leg(a.learned_ms, a.labels, a.log_steps, a.epochs, a.learned_warmup,
                        a.learned_steps, a.learned_precision, verbose=False, lr=a.lr)
        extra["learned"] = dict({"workload": LEARNED_DESC.format(ms=a.learned_ms),
                                 "form": "learned-" + a.learned_precision,
                                 "steps": a.learned_steps, "warmup": a.learned_warmup,
                                 "state_steps_per_s": r.pop("value")}, **r)
        torch.cuda.empty_cache()
    dtype = "c64 terms / f64 state" if prec == "mixed" else "f64"
    line = {"metric": METRIC, "value": value, "unit": "state-steps/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic",
            "config": dict({"workload": TRAJ_DESC.format(n=a.ntraj),
                            "form": "batched-" + prec + " (no scorer)",
                            "step": "select, expansions, CSR, reassignment, Taylor series, "
                                    "per-trajectory norm check and n_exc/n_pho",
                            "series_graph": bool(a.graph), "autotune": bool(a.autotune),
                            "term_shape": list(a.term_shape) if a.term_shape else "tuned",
                            "csr_values": bte_csr_values,
                            "unit_of_work": "retained basis state advanced one step "
                                            "(sum over trajectories of min(maxstates, R_prev))",
                            "final_basis_rows": int(rows),
                            "sum_maxstates": int(sum(d[3] for d in draws))}, **extra),
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def run_plain_form(a):
    """The plain form on its own (per-trajectory engines), for a separate line."""
    draws = traj256_draws(a.ntraj)
    pu, pe = plain_leg(draws, a.warmup, a.steps)
    line = {"metric": METRIC, "value": pu / pe, "unit": "state-steps/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * pe / a.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": TRAJ_DESC.format(n=a.ntraj), "form": "plain"},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


LEARNED_DESC = ("learned: 1D Holstein chain N=16 (open), J=1, w0=1, phonon cap 7, exciton at 8, "
                "maxstates={ms}, dt=0.05, enlarge_steps=1; g~U[0.5,2] seeded 0: three training "
                "couplings for the logged labels, the fourth for the timed runs")
# FLOPs of one scorer inference row: 2 * (64*256 + 256*256 + 256*256 + 256)
MLP_FLOPS_PER_ROW = 2 * (64 * 256 + 2 * 256 * 256 + 256)


def learned_draws(seed=0):
    return np.random.default_rng(seed).uniform(0.5, 2.0, 4)


def chain_te(cls, g, ms, precision="fp64", **kw):
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein
    ham = holstein.Hamiltonian(16, [8] * 16, use_terms=use_terms(1.0, g))
    te = cls(ham, holstein.Observables, ["n_exc", "n_pho"], None,
             TimeEvoParams(maxstates=ms, enlarge_steps=1, shuffle_seed=None, diagnostics=False),
             ExpmParams(precision=precision), **kw)
    te.create_nonuniform_basis([1] * 16, minpos=8, maxpos=9)
    raw = np.zeros((1, 17), dtype=np.int64)
    raw[0, 0] = 8
    te.create_initial_vector([1.0], raw)
    te._perform_first_step(0.0)
    return te


def measure_scorer(scorer, X, reps=10):
    """Average duration of paces_mlp_infer over X, HIP events on the current stream."""
    import torch
    st = torch.cuda.current_stream()
    scorer.infer(X)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        scorer.infer(X)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def learned_leg(ms, labels, log_steps, epochs, warmup, steps, precision="mixed",
                verbose=True, lr=2e-3):
    """The learned form (scorer pruning before the second expansion, complex64 Taylor terms by
    default) against the plain FP64 form on the same trajectory, after logging labels from plain
    runs at three couplings and training the scorer on them; both timed over the same steps.
    Returns a dict of the measurements (the plain FP64 form equals the FP64 oracle bit for bit,
    so its densities are the reference of the error)."""
    import torch
    from paces_amd.core.scorer import CandidateScorer
    from paces_amd.models import holstein
    gs = learned_draws()
    dev = torch.device("cuda", torch.cuda.current_device())
    log = holstein.LabelLog(dev, labels)
    t0 = time.perf_counter()
    t_arr = 0.05 * np.arange(max(log_steps, warmup + steps) + 1)
    for g in gs[:3]:
        te = chain_te(holstein.LearnedTimeEvolution, g, ms, label_log=log)
        for i in range(1, log_steps + 1):
            te._evolve_single_step(i, t_arr[i], t_arr[i] - t_arr[i - 1])
            if log.full:
                break
        if verbose:
            print(f"[learned] logged g={g:.3f}: {log.n} labels", file=sys.stderr, flush=True)
        del te
    torch.cuda.synchronize()
    log_s = time.perf_counter() - t0
    X, y = log.data()
    scorer = CandidateScorer(dev, seed=1)
    t0 = time.perf_counter()
    hist = scorer.train(X, y, epochs=epochs, batch=16384, lr=lr)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    if verbose:
        print(f"[learned] trained on {log.n} labels in {train_s:.1f} s, loss {hist}",
              file=sys.stderr, flush=True)
    nlabels = log.n
    del X, y, log
    torch.cuda.empty_cache()
    g = gs[3]
    plain = chain_te(holstein.TimeEvolution, g, ms)
    learned = chain_te(holstein.LearnedTimeEvolution, g, ms, precision=precision,
                       scorer=scorer)
    tp = tl = 0.0
    up = ul = 0
    err = 0.0
    rows_ratio = []
    windows = []
    for i in range(1, warmup + steps + 1):
        timed = i > warmup
        for name, te in (("plain", plain), ("learned", learned)):
            prev = te.numstates
            torch.cuda.synchronize()
            w0 = time.monotonic_ns()
            t0 = time.perf_counter()
            te._evolve_single_step(i, t_arr[i], t_arr[i] - t_arr[i - 1])
            torch.cuda.synchronize()
            dt_ = time.perf_counter() - t0
            if timed:
                windows.append((name, i, w0, time.monotonic_ns()))
            if timed:
                if name == "plain":
                    tp += dt_
                    up += min(prev, ms)
                else:
                    tl += dt_
                    ul += min(prev, ms)
        e = float(np.max(np.abs(learned.obsobj.last_results["n_exc"]
                                - plain.obsobj.last_results["n_exc"])))
        err = max(err, e)
        s = learned.stats[-1]
        if timed:
            rows_ratio.append(s["rows"] / plain.numstates)
        if verbose:
            print(f"[learned] step {i}: first expansion {s['rows_first']}, scored {s['scored']}, "
                  f"kept {s['kept']}, rows {s['rows']} (plain {plain.numstates}), fallback "
                  f"{s['fallback']}, density err {e:.2e}", file=sys.stderr, flush=True)
    if verbose:
        # per-step windows, for cutting a kernel trace of this run by step and form
        for name, i, w0, w1 in windows:
            print(f"[learned] step window monotonic_ns {name} {i} {w0} {w1}", file=sys.stderr,
                  flush=True)
    st = learned.stats[warmup:]
    # the scorer's roofline on as many rows as the last step scored
    nscore = max(int(learned.stats[-1]["scored"]), 1)
    Xb = torch.zeros((nscore, 64), dtype=torch.int16, device=dev)
    sec = measure_scorer(scorer, Xb)
    tile = scorer.tile_rows or 128  # 0: the library default, 128 rows
    achieved = MLP_FLOPS_PER_ROW * nscore / sec / 1e12
    return {"value": ul / tl, "ms_per_step": 1e3 * tl / steps,
            "plain_gpu_state_steps_per_s": up / tp, "plain_ms_per_step": 1e3 * tp / steps,
            "speedup_vs_plain": (ul / tl) / (up / tp), "density_err_vs_plain_fp64": err,
            "mean_kept_fraction_of_scored": float(np.mean([x["kept"] / max(x["scored"], 1)
                                                           for x in st])),
            "mean_rows_vs_plain": float(np.mean(rows_ratio)),
            "fallbacks": int(sum(x["fallback"] for x in st)), "labels": int(nlabels),
            "label_log_s": log_s, "train_s": train_s, "train_epochs": epochs, "train_lr": lr,
            "precision": precision,
            "scorer_roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": 2500.0,
                                "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4),
                                "traffic": None,
                                "kernel": "k_mlp_infer<%d> (paces_mlp_infer, bf16 MFMA, %d-row "
                                          "tiles)" % (tile // 64, tile),
                                "avg_us": round(sec * 1e6, 1), "rows": int(nscore)}}


def run_learned(a):
    """The learned form's own line (config-4 physics): value = its state-steps/s."""
    r = learned_leg(a.learned_ms, a.labels, a.log_steps, a.epochs, a.warmup, a.steps,
                    a.learned_precision, lr=a.lr)
    roof = r.pop("scorer_roofline")
    line = {"metric": METRIC, "value": r.pop("value"), "unit": "state-steps/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": r.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "c64 terms / f64 state + bf16 scorer" if a.learned_precision == "mixed"
            else "f64 + bf16 scorer",
            "data": "synthetic",
            "config": dict({"workload": LEARNED_DESC.format(ms=a.learned_ms),
                            "form": "learned-" + a.learned_precision}, **r),
            "roofline": roof, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE counter readings:
    each launch reads and writes ``nbytes``."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": reps}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps after step 0 (default 8 for traj256, 6 otherwise)")
    ap.add_argument("--workload", default="traj256", choices=sorted(WORKLOADS) + ["traj256", "learned"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--form", default="batched", choices=["batched", "plain"],
                    help="traj256 only: 'plain' times the per-trajectory engines alone")
    ap.add_argument("--ntraj", type=int, default=256)
    ap.add_argument("--precision-batched", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the site-density error against the oracle")
    ap.add_argument("--oracle-max-ms", type=int, default=5000,
                    help="traj256: trajectories with maxstates up to this run on the oracle "
                         "(density error and CPU baseline)")
    ap.add_argument("--graph", type=int, default=1,
                    help="traj256: run each step's Taylor series as one replayed HIP graph")
    ap.add_argument("--term-shape", default=None,
                    help="traj256: fix the Taylor term's launch shape, e.g. 256,4 (threads, rows "
                         "per thread), instead of autotuning it (for reproducible profiles)")
    ap.add_argument("--csr-values", default="slots", choices=["slots", "coded", "explicit"],
                    help="traj256: the batched series reads one slot matrix (a uint32 per row "
                         "and op), coded CSR values (a code byte per vib entry) or explicit "
                         "FP32/FP64 values (identical results)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="traj256: tune the Taylor-term launch shape")
    ap.add_argument("--learned-leg", type=int, default=1,
                    help="traj256: add the learned form's own workload to the line (label "
                         "logging and scorer training untimed, about half a minute)")
    ap.add_argument("--learned-warmup", type=int, default=12,
                    help="traj256: untimed steps of the learned leg")
    ap.add_argument("--learned-steps", type=int, default=10,
                    help="traj256: timed steps of the learned leg")
    ap.add_argument("--learned-ms", type=int, default=2_000_000)
    ap.add_argument("--learned-precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--labels", type=int, default=100_000_000,
                    help="labelled candidates logged for training (learned workload)")
    ap.add_argument("--log-steps", type=int, default=35,
                    help="steps per training coupling to log labels from (until --labels)")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3, help="scorer training: Adam step size")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration)")
    a = ap.parse_args(argv)
    if a.term_shape is not None:
        a.term_shape = tuple(int(v) for v in a.term_shape.split(","))
    if a.warmup is None:
        a.warmup = {"traj256": 8, "learned": 12}.get(a.workload, 6)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(independent trajectories shard one process per GPU).", file=sys.stderr)
        return 2

    import torch
    from paces_amd.config import devices
    from paces_amd.core.time_evolution import ExpmParams, TimeEvoParams
    from paces_amd.models import holstein

    import contextlib
    with contextlib.redirect_stdout(sys.stderr):  # stdout carries only the JSON line
        devices.configure()
    if a.calibrate_copy:
        return calibrate_copy()
    if a.workload == "traj256":
        return run_plain_form(a) if a.form == "plain" else run_traj256(a)
    if a.workload == "learned":
        return run_learned(a)
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
