"""Core classes: HamiltonianFramework, ObservablesFramework, TimeEvolutionFramework."""

from . import engine, hamiltonian, observables, time_evolution  # noqa: F401
