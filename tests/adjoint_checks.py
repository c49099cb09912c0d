"""CPU checkers for the adjoint-differentiation tests.

Expectation values come from the CPU oracle's state vector (the reference CPU runtime
restated in C, driven through the qgate front-end) and a numpy Pauli-sum product; gradients
from the parameter-shift rule (uncontrolled rotations: d<H>/dtheta = (<H>(theta + pi/2) -
<H>(theta - pi/2)) / 2, exact because the generator G/2 has eigenvalues +-1/2) or from
central differences of oracle expectations (controlled rotations).
"""
import math

import numpy as np

from qgate_amd.model import ir
from qgate_amd.model import gate_type as gtype
from qgate_amd.script import new_qregs
from qgate_amd.simulator.simulator import Simulator

_ROT = {'RX': gtype.RX, 'RY': gtype.RY, 'RZ': gtype.RZ}


def to_qgate_ops(circuit, params, shifts=None):
    """qgate operators of a ParametricCircuit; shifts maps op position -> angle offset."""
    shifts = shifts or {}
    q = new_qregs(circuit.n_qubits)
    ops = []
    for pos, op in enumerate(circuit.ops):
        if op[0] == 'G':
            _, target, gate, ctrl, adjoint = op
            g = ir.Gate(gate)
            g.set_adjoint(adjoint)
        else:
            kind, target, k, ctrl = op
            g = ir.Gate(_ROT[kind](float(params[k]) + shifts.get(pos, 0.)))
        g.set_qreg(q[target])
        if ctrl:
            g.set_ctrllist([q[c] for c in ctrl])
        ops.append(g)
    return q, ops


def oracle_state(module, circuit, params, dtype=np.float64, shifts=None):
    """State vector of the circuit from the oracle runtime; qubit q = bit q."""
    q, ops = to_qgate_ops(circuit, params, shifts)
    sim = Simulator(module, circuit_prep='one_static', dtype=dtype)
    sim.qubits.set_ordering(list(q))
    sim.run(ops)
    return np.asarray(sim.qubits.states[:], np.complex128)


def pauli_apply(observable, psi):
    """H psi for a PauliSum, by index arithmetic (pinned against the dense matrix)."""
    x = np.arange(len(psi), dtype=np.uint64)
    out = np.zeros_like(psi, dtype=np.complex128)
    for c, xm, zm in zip(observable.coefs, observable.xmask, observable.zmask):
        src = x ^ np.uint64(xm)
        parity = np.zeros(len(psi), np.uint64)
        bits = src & np.uint64(zm)
        while np.any(bits):
            parity ^= bits & np.uint64(1)
            bits >>= np.uint64(1)
        sign = 1. - 2. * parity.astype(np.float64)
        n_y = bin(int(xm) & int(zm)).count('1')
        out += c * (1j ** n_y) * sign * psi[src.astype(np.int64)]
    return out


def expectation(observable, psi):
    return float(np.real(np.vdot(psi, pauli_apply(observable, psi))))


def oracle_expectation(module, circuit, params, observable, shifts=None):
    return expectation(observable, oracle_state(module, circuit, params, shifts=shifts))


def parameter_shift_gradient(module, circuit, params, observable):
    """Gradient from two shifted oracle runs per rotation (uncontrolled rotations only)."""
    grad = np.zeros(circuit.n_params)
    for pos, op in enumerate(circuit.ops):
        if op[0] == 'G':
            continue
        if op[3]:
            raise ValueError('parameter shift needs uncontrolled rotations')
        plus = oracle_expectation(module, circuit, params, observable, {pos: math.pi / 2})
        minus = oracle_expectation(module, circuit, params, observable, {pos: -math.pi / 2})
        grad[op[2]] += (plus - minus) / 2.
    return grad


def central_difference_gradient(module, circuit, params, observable, h=1e-5):
    grad = np.zeros(circuit.n_params)
    for pos, op in enumerate(circuit.ops):
        if op[0] == 'G':
            continue
        plus = oracle_expectation(module, circuit, params, observable, {pos: h})
        minus = oracle_expectation(module, circuit, params, observable, {pos: -h})
        grad[op[2]] += (plus - minus) / (2. * h)
    return grad


class CircuitSpec:
    """Host-only description with ParametricCircuit's builder methods (for CPU tests and
    for building the same circuit on the device with `build`)."""

    def __init__(self, n_qubits):
        self.n_qubits = n_qubits
        self.ops = []
        self.n_params = 0

    def _rot(self, kind, qubit, param, controls):
        self.ops.append((kind, qubit, int(param), tuple(controls)))
        self.n_params = max(self.n_params, int(param) + 1)
        return self

    def rx(self, qubit, param, controls=()):
        return self._rot('RX', qubit, param, controls)

    def ry(self, qubit, param, controls=()):
        return self._rot('RY', qubit, param, controls)

    def rz(self, qubit, param, controls=()):
        return self._rot('RZ', qubit, param, controls)

    def gate(self, gate, qubit, controls=(), adjoint=False):
        self.ops.append(('G', qubit, gate, tuple(controls), bool(adjoint)))
        return self

    def cz(self, a, b):
        return self.gate(gtype.Z(), b, controls=[a])

    def build(self, dtype):
        from qgate_amd.adjoint import ParametricCircuit
        c = ParametricCircuit(self.n_qubits, dtype)
        for op in self.ops:
            if op[0] == 'G':
                c.gate(op[2], op[1], op[3], op[4])
            else:
                c._rot(op[0][1], op[1], op[2], op[3])
        return c


def hea_spec(n, layers):
    """RY and RZ on every qubit then a CZ ring, per layer (parameter order of hea_ansatz)."""
    c = CircuitSpec(n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c.ry(q, k)
            k += 1
        for q in range(n):
            c.rz(q, k)
            k += 1
        for q in range(n):
            c.cz(q, (q + 1) % n)
    return c


def mixed_spec(n, n_ops, seed, n_params=None):
    """Random circuit: fixed gates (H, S, T, X, SH, U), rotations with 0..2 controls, and
    parameters shared between rotations when n_params < number of rotations."""
    rs = np.random.RandomState(seed)
    c = CircuitSpec(n)
    fixed = [lambda: gtype.H(), lambda: gtype.S(), lambda: gtype.T(), lambda: gtype.X(),
             lambda: gtype.SH(), lambda: gtype.U(*rs.uniform(0, 6.3, 3))]
    n_rot = 0
    for _ in range(n_ops):
        n_ctrl = int(rs.choice([0, 0, 0, 1, 1, 2]))
        lanes = [int(v) for v in rs.choice(n, n_ctrl + 1, replace=False)]
        target, ctrl = lanes[0], lanes[1:]
        if rs.randint(3) == 0:
            c.gate(fixed[rs.randint(len(fixed))](), target, ctrl, adjoint=bool(rs.randint(2)))
        else:
            k = n_rot if n_params is None else int(rs.randint(n_params))
            c._rot(['RX', 'RY', 'RZ'][rs.randint(3)], target, k, ctrl)
            n_rot += 1
    if n_params is not None:
        c.n_params = n_params
    return c


def random_observable(n, n_terms, seed):
    from qgate_amd.adjoint import PauliSum
    rs = np.random.RandomState(seed)
    terms = []
    for _ in range(n_terms):
        k = int(rs.randint(1, 4))
        qs = rs.choice(n, k, replace=False)
        terms.append((float(rs.normal()), {int(q): 'XYZ'[rs.randint(3)] for q in qs}))
    return PauliSum(terms)
