"""Synthetic benchmark circuits, built with the script DSL from a seed.

Circuit structure and angles come from numpy.random.RandomState(seed); draw order is fixed
below so that every runtime (and the reference) sees identical circuits.

  brickwork(n, layers, seed)   per layer: U3(t, p, l) on every qubit (t, p, l = uniform(0, 2pi, 3)
                               per qubit, in qubit order), then CX(2i, 2i+1) on even layers or
                               CX(2i+1, 2i+2) on odd layers.
  random_circuit(n, g, seed)   g gates; per gate: kind = randint(5) over (U3, H, Rz, CX, CCX),
                               qubits = choice(n, arity, replace=False) (leading ones are
                               controls), then the angles: U3 uniform(0, 2pi, 3), Rz uniform(0, 2pi).
  qft(qregs)                   H + controlled-U1(pi / 2^k) ladder per qubit, then the bit-reversal
                               swaps (each swap lowers to three CX).
  hea_ansatz(n, layers, params) per layer: RY then RZ on every qubit, then a CZ ring
                               (i, i+1 mod n); params has 2 n layers entries.
"""
import math

import numpy as np

from .script import new_qregs, H, U3, Rz, Ry, ctrl, Swap
from .script.qelib1 import cu1, cz

GATE_KINDS = ('U3', 'H', 'Rz', 'CX', 'CCX')
_ARITY = {'U3': 1, 'H': 1, 'Rz': 1, 'CX': 2, 'CCX': 3}


def brickwork(n, layers, seed, qregs=None):
    rs = np.random.RandomState(seed)
    q = new_qregs(n) if qregs is None else qregs
    ops = []
    for layer in range(layers):
        for i in range(n):
            t, p, l = rs.uniform(0., 2. * math.pi, 3)
            ops.append(U3(float(t), float(p), float(l))(q[i]))
        start = 0 if layer % 2 == 0 else 1
        for i in range(start, n - 1, 2):
            ops.append(ctrl(q[i]).X(q[i + 1]))
    return q, ops


def random_gate_list(n, n_gates, seed):
    """[(kind, qubits, angles)] with qubits as indices (controls first, target last)."""
    rs = np.random.RandomState(seed)
    out = []
    for _ in range(n_gates):
        kind = GATE_KINDS[rs.randint(len(GATE_KINDS))]
        qubits = [int(v) for v in rs.choice(n, _ARITY[kind], replace=False)]
        if kind == 'U3':
            angles = [float(v) for v in rs.uniform(0., 2. * math.pi, 3)]
        elif kind == 'Rz':
            angles = [float(rs.uniform(0., 2. * math.pi))]
        else:
            angles = []
        out.append((kind, qubits, angles))
    return out


def to_ops(gates, q):
    ops = []
    for kind, qubits, angles in gates:
        if kind == 'U3':
            ops.append(U3(*angles)(q[qubits[0]]))
        elif kind == 'H':
            ops.append(H(q[qubits[0]]))
        elif kind == 'Rz':
            ops.append(Rz(angles[0])(q[qubits[0]]))
        elif kind == 'CX':
            ops.append(ctrl(q[qubits[0]]).X(q[qubits[1]]))
        elif kind == 'CCX':
            ops.append(ctrl(q[qubits[0]], q[qubits[1]]).X(q[qubits[2]]))
        else:
            raise ValueError(kind)
    return ops


def random_circuit(n, n_gates, seed, qregs=None):
    q = new_qregs(n) if qregs is None else qregs
    return q, to_ops(random_gate_list(n, n_gates, seed), q)


def qft(q):
    ops = []
    n = len(q)
    for i in range(n - 1, -1, -1):
        ops.append(H(q[i]))
        for j in range(i - 1, -1, -1):
            ops.append(cu1(math.pi / float(1 << (i - j)), q[j], q[i]))
    for i in range(n // 2):
        ops.append(Swap(q[i], q[n - 1 - i]))
    return ops


def qft_random(n, n_random, seed, qregs=None):
    """H on all qubits, QFT, then n_random random U3/CX gates (large-state workload)."""
    rs = np.random.RandomState(seed)
    q = new_qregs(n) if qregs is None else qregs
    ops = [H(x) for x in q] + qft(q)
    for _ in range(n_random):
        if rs.randint(2) == 0:
            t, p, l = rs.uniform(0., 2. * math.pi, 3)
            ops.append(U3(float(t), float(p), float(l))(q[int(rs.randint(n))]))
        else:
            c, t = rs.choice(n, 2, replace=False)
            ops.append(ctrl(q[int(c)]).X(q[int(t)]))
    return q, ops


def hea_params(n, layers, seed):
    return np.random.RandomState(seed).uniform(-math.pi, math.pi, 2 * n * layers)


def hea_observable(n, seed):
    """(a, b): H = sum_{i<n-1} a_i Z_i Z_{i+1} + sum_i b_i X_i, a, b ~ N(0, 1)."""
    rs = np.random.RandomState(seed + 1000)
    return rs.normal(size=n - 1), rs.normal(size=n)


def hea_ansatz(n, layers, params, qregs=None):
    q = new_qregs(n) if qregs is None else qregs
    ops = []
    k = 0
    for _ in range(layers):
        for i in range(n):
            ops.append(Ry(float(params[k]))(q[i]))
            k += 1
        for i in range(n):
            ops.append(Rz(float(params[k]))(q[i]))
            k += 1
        for i in range(n):
            ops.append(cz(q[i], q[(i + 1) % n]))
    return q, ops
