"""Oracle trajectories in worker processes -- TEST AND BENCHMARK INFRASTRUCTURE ONLY.

The benchmark's CPU baseline and its site-density check run the oracle (paces_oracle) on many
independent Holstein chains at once, one process per core. The worker lives in this module so
that spawned processes import it by name.
"""

import time

import numpy as np


def chain_terms(J, g, w0=1.0):
    """The chain Hamiltonian -J sum(a+a + h.c.) + w0 sum b+b + g w0 sum n(b+ + b) in the
    reference's use_terms convention (hopping element +J_ref, vib element g_ref sqrt(n+1))."""
    return {"diag": {"eps_sys": 0.0, "hbar_omega": w0, "delta_eps": 0.0},
            "hopping": {"J": -J}, "vib_coupling": {"g": g * w0}}


def run_chain(args):
    """One open chain (N sites, phonon cap 7, exciton at N//2, enlarge_steps 1) on the oracle
    over t = dt * arange(nsteps + 1). Returns ({step: n_exc} for the steps in record, retained
    state-steps of steps 1..nsteps, seconds of those steps)."""
    N, J, g, ms, dt, nsteps, record = args
    from . import paces_oracle as po
    ham = po.HolsteinOracle(N, [8] * N, use_terms=chain_terms(J, g))
    te = po.TimeEvolutionOracle(ham, ms, 1)
    te.create_single_state_basis(N // 2)
    raw = np.zeros((1, N + 1), dtype=np.int64)
    raw[0, 0] = N // 2
    te.create_initial_vector([1.0], raw)
    t_arr = dt * np.arange(nsteps + 1)
    out, units, secs = {}, 0, 0.0
    for i in range(nsteps + 1):
        prev = te.whoami.shape[0]
        t0 = time.perf_counter()
        rec = te.step(i, float(t_arr[i]), float(t_arr[i] - t_arr[i - 1]) if i else 0.0)
        if i:
            secs += time.perf_counter() - t0
            units += min(prev, ms)
        if i in record:
            out[i] = rec.obs["n_exc"]
    return out, units, secs


def ready(_):
    """A no-op task: returns once a worker has started."""
    return 0
