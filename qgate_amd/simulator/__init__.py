"""Simulator front-end and the HIP (MI355X) runtime.

`hip(**prefs)` returns a Simulator whose state vectors live in GPU memory and whose gate,
probability, measurement, readout and sampling work runs in hand-written HIP kernels
(the counterpart of qgate.simulator.py()/cpu()/cuda(), qgate/simulator/__init__.py:18-28).
"""
from . import simulator
from .qubits import null, abs2, prob, Qubits
from .value_store import ValueStore
from .observation import Observation, ObservationList, ObservationHistgram


def hip(**prefs):
    """Simulator on the HIP runtime; prefs: dtype (np.float32 / np.float64), circuit_prep."""
    from . import hipruntime
    return simulator.Simulator(hipruntime, **prefs)
