"""Built-in gate types and their FP64 2x2 matrices.

Gate kinds and argument order follow qgate/model/gate_type.py:35-130. `pymat()` gives the
FP64 matrix with qgate's phase conventions (qgate/simulator/pymatrix.py:18-149,
qgate/simulator/src/GateMatrix.cpp:13-158):
  U(t, p, l) = [[e^{-i(l+p)/2} cos t/2, -e^{i(l-p)/2} sin t/2],
                [e^{-i(l-p)/2} sin t/2,  e^{i(l+p)/2} cos t/2]]   (differs from OpenQASM
               by e^{i(l+p)/2}), RZ(t) = diag(e^{-it/2}, e^{it/2}), ExpiI(t) = e^{it} I,
  ExpiZ(t) = diag(e^{it}, e^{-it}), SH = sqrt(1/2) [[1, 1], [i, -i]].
The device runtime builds the same matrices natively (csrc/gate_matrix.cpp); this module
is the host-side description used by the front-end and by tests.
"""
import cmath
import math

import numpy as np

from .ir import GateType


def _check_1bit(gate_type, gate):
    if gate.ctrllist is not None and gate.qreg in gate.ctrllist:
        raise RuntimeError('control and operand overlapped.')


def _check_expi(gate_type, macro):
    targets = set()
    for gate in macro.gatelist:
        if not isinstance(gate.gate_type, (ID, X, Y, Z)):
            raise RuntimeError('exp gate only accepts ID, X, Y and Z gates')
        if gate.ctrllist is not None:
            raise RuntimeError('control qreg(s) should not be set for exp gate parameters.')
        targets.add(gate.qreg)
    if macro.ctrllist is not None and targets & set(macro.ctrllist):
        raise RuntimeError('control bit and target should not overlap.')


def adjoint(mat):
    """Conjugate transpose of a gate matrix."""
    return np.conjugate(np.asarray(mat).T)


_R = math.sqrt(0.5)


def _u(theta, phi, lam):
    c, s = math.cos(theta * 0.5), math.sin(theta * 0.5)
    return [[cmath.exp(-0.5j * (lam + phi)) * c, -cmath.exp(0.5j * (lam - phi)) * s],
            [cmath.exp(-0.5j * (lam - phi)) * s, cmath.exp(0.5j * (lam + phi)) * c]]


def _rx(theta):
    c, s = math.cos(theta / 2.), math.sin(theta / 2.)
    return [[c, -1j * s], [-1j * s, c]]


def _ry(theta):
    c, s = math.cos(theta / 2.), math.sin(theta / 2.)
    return [[c, -s], [s, c]]


def _rz(theta):
    return [[cmath.exp(-0.5j * theta), 0.], [0., cmath.exp(0.5j * theta)]]


_MATRICES = {
    'U': _u,
    'U2': lambda phi, lam: [[cmath.exp(-0.5j * (lam + phi)) * _R, -cmath.exp(0.5j * (lam - phi)) * _R],
                            [cmath.exp(-0.5j * (lam - phi)) * _R, cmath.exp(0.5j * (lam + phi)) * _R]],
    'U1': lambda lam: [[1., 0.], [0., cmath.exp(1j * lam)]],
    'ID': lambda: [[1., 0.], [0., 1.]],
    'X': lambda: [[0., 1.], [1., 0.]],
    'Y': lambda: [[0., -1j], [1j, 0.]],
    'Z': lambda: [[1., 0.], [0., -1.]],
    'H': lambda: [[_R, _R], [_R, -_R]],
    'S': lambda: [[1., 0.], [0., 1j]],
    'T': lambda: [[1., 0.], [0., cmath.exp(0.25j * math.pi)]],
    'RX': _rx,
    'RY': _ry,
    'RZ': _rz,
    'SH': lambda: [[_R, _R], [1j * _R, -1j * _R]],
    'ExpiI': lambda t: [[cmath.exp(1j * t), 0.], [0., cmath.exp(1j * t)]],
    'ExpiZ': lambda t: [[cmath.exp(1j * t), 0.], [0., cmath.exp(-1j * t)]],
}


class _Unary(GateType):
    """Single-qubit gate type; `name` keys the matrix factory on both host and device."""
    name = None
    n_args = 0

    def __init__(self, *args):
        if len(args) != self.n_args:
            raise TypeError('{} takes {} argument(s)'.format(type(self).__name__, self.n_args))
        GateType.__init__(self, *args)

    def constraints(self, gate):
        _check_1bit(self, gate)

    def pymat(self):
        return np.array(_MATRICES[self.name](*self.args), np.complex128)


def _unary(name, n_args, pyname=None):
    cls = type(pyname or name, (_Unary,), {'name': name, 'n_args': n_args})
    return cls


U = _unary('U', 3)
U2 = _unary('U2', 2)
U1 = _unary('U1', 1)
ID = _unary('ID', 0)
X = _unary('X', 0)
Y = _unary('Y', 0)
Z = _unary('Z', 0)
H = _unary('H', 0)
S = _unary('S', 0)
T = _unary('T', 0)
RX = _unary('RX', 1)
RY = _unary('RY', 1)
RZ = _unary('RZ', 1)
SH = _unary('SH', 0)
ExpiI = _unary('ExpiI', 1)
ExpiZ = _unary('ExpiZ', 1)

UNARY_TYPES = (U, U2, U1, ID, X, Y, Z, H, S, T, RX, RY, RZ, SH, ExpiI, ExpiZ)


class Expi(GateType):
    """exp(i theta P) for a Pauli product P, expanded before execution."""

    def __init__(self, theta):
        GateType.__init__(self, theta)

    def constraints(self, macro):
        _check_expi(self, macro)


class SWAP(GateType):
    def __init__(self):
        GateType.__init__(self)

    def constraints(self, gate):
        pass
