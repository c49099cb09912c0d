"""Large readouts through the pinned staging (several staging chunks of 64 MiB each).

A product state of RY rotations has the closed form a[x] = prod_k (cos(t_k/2) if bit k of x is
0 else sin(t_k/2)), so full and strided readouts of amplitudes and probabilities, in the
state's own lane order (the state copied chunk by chunk) and in a permuted order (the gather
kernel feeding the stager), are checked element by element at 25 qubits complex64 (256 MiB,
4 chunks) and 24 qubits complex128.
"""
import numpy as np
import pytest

from qgate_amd.script import new_qregs, Ry
from qgate_amd.simulator.simulator import Simulator

pytestmark = pytest.mark.gpu


def _product_state(thetas, dtype):
    a = np.ones(1, dtype)
    for t in thetas:   # qubit k is bit k: later qubits are higher bits
        a = np.concatenate([a * np.cos(t / 2.), a * np.sin(t / 2.)])
    return a


@pytest.mark.parametrize('n,dtype,tol', [(25, np.float32, 1e-6), (24, np.float64, 1e-14)])
def test_large_readouts(hip_module, n, dtype, tol):
    thetas = np.linspace(0.3, 2.9, n)
    q = new_qregs(n)
    sim = Simulator(hip_module, circuit_prep='one_static', dtype=dtype)
    try:
        sim.qubits.set_ordering(list(q))
        sim.run([Ry(float(t))(x) for t, x in zip(thetas, q)])
        ref = _product_state(thetas, np.float64)
        amps = sim.qubits.states[:]
        assert amps.shape == (1 << n,)
        assert np.max(np.abs(amps - ref)) < tol
        probs = sim.qubits.prob[:]
        assert np.max(np.abs(probs - ref * ref)) < tol
        sl = sim.qubits.states[7:(1 << n) - 5:3]
        assert np.max(np.abs(sl - ref[7:(1 << n) - 5:3])) < tol
        # permuted external order: x's bit k is qubit n-1-k
        sim.qubits.set_ordering(list(reversed(q)))
        rev = _product_state(thetas[::-1], np.float64)
        amps = sim.qubits.states[:]
        assert np.max(np.abs(amps - rev)) < tol
        probs = sim.qubits.prob[1:]
        assert np.max(np.abs(probs - rev[1:] ** 2)) < tol
    finally:
        sim.terminate()
