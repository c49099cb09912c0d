"""Launches split into pieces (the path every kernel takes at 33+ qubits, where one grid
would exceed 2^32 work items) give bit-identical results to single launches.

The piece size is lowered to 2^12 work items so that 14-16 qubit states run through the
piecewise path of every kernel: each gate variant, |0> initialisation, join, measurement
(decohere, decohere-and-separate), reset, probability arrays, the sampling pool's draws and
the adjoint Pauli-sum pass.
"""
import math

import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.script import new_qregs, new_references, measure, reset, prob, ctrl, H, T
from qgate_amd.simulator.simulator import Simulator

pytestmark = pytest.mark.gpu

SMALL_PIECE = 1 << 12
DEFAULT_PIECE = 1 << 30


def _circuit(n):
    q = new_qregs(n)
    _, ops = workloads.brickwork(n, 3, seed=21, qregs=q)
    _, more = workloads.random_circuit(n, 60, seed=22, qregs=q)
    ops += more
    ops += [T(q[0]), ctrl(q[2]).S(q[9]), ctrl(q[1], q[4]).T(q[0])]      # phase kernels
    refs = new_references(3)
    ops += [measure(refs[0], q[3]), measure(refs[1], q[0]), reset(q[0]), H(q[0]),
            prob(refs[2], q[n - 1])]
    return q, ops


def _run(hip_module, dtype, prep, piece):
    hip_module.set_launch_piece(piece)
    try:
        np.random.seed(77)
        n = 16
        q, ops = _circuit(n)
        sim = Simulator(hip_module, circuit_prep=prep, dtype=dtype)
        sim.qubits.set_ordering(list(q))
        sim.run(ops)
        states = np.array(sim.qubits.states[:])
        probs = np.array(sim.qubits.prob[:])
        u = np.random.RandomState(5).random_sample(20000)
        pool = sim.qubits.create_sampling_pool(list(q))
        samples = np.array(pool.sample(len(u), u).intarray)
        return states, probs, samples
    finally:
        hip_module.set_launch_piece(DEFAULT_PIECE)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('prep', ['dynamic', 'one_static'])
def test_pieces_bit_identical(hip_module, dtype, prep):
    one = _run(hip_module, dtype, prep, DEFAULT_PIECE)
    many = _run(hip_module, dtype, prep, SMALL_PIECE)
    for a, b in zip(one, many):
        assert a.shape == b.shape
        assert np.array_equal(a, b)


def test_piece_size_checked(hip_module):
    for bad in (1000, (1 << 12) + 1, 1 << 31):
        with pytest.raises(ValueError):
            hip_module.set_launch_piece(bad)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_adjoint_pieces_bit_identical(hip_module, dtype):
    from qgate_amd.adjoint import PauliSum, hea_circuit
    n, layers = 14, 2
    params = workloads.hea_params(n, layers, seed=9)
    a, b = workloads.hea_observable(n, seed=9)
    h = PauliSum.zz_x(a, b)
    out = []
    for piece in (DEFAULT_PIECE, SMALL_PIECE):
        hip_module.set_launch_piece(piece)
        try:
            c = hea_circuit(n, layers, dtype)
            value, grad = c.gradient(params, h)
            c.expectation(params, h)
            out.append((value, grad, c.state()))
            c.delete()
        finally:
            hip_module.set_launch_piece(DEFAULT_PIECE)
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
