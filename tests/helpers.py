"""Test helpers: run a circuit on a runtime module and read results in a fixed ordering."""
import numpy as np

from qgate_amd.simulator.simulator import Simulator

# tolerances against the FP64 oracle
TOL_C128 = 1e-12
TOL_C64 = 5e-6


def run(module, ops, prep='one_static', dtype=np.float64, ordering=None):
    sim = Simulator(module, circuit_prep=prep, dtype=dtype)
    if ordering is not None:
        sim.qubits.set_ordering(ordering)
    sim.run(ops)
    return sim


def states(sim):
    return sim.qubits.states[:]


def tol(dtype):
    return TOL_C128 if dtype == np.float64 else TOL_C64


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.complex128) - np.asarray(b, np.complex128))))
