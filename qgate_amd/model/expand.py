"""Lowering of SWAP, Expi and Pauli observables into 1-qubit / controlled gates.

Behaviour follows qgate/model/expand.py:1-90 and
qgate/model/pauli_gates_diagonalizer.py:9-146: a Pauli product is reduced per qubit to
i^k * P with P in {I, X, Y, Z}, each non-identity factor is rotated to Z (X by H, Y by SH),
the Z factors are chained with CX, and the rotation ends on the last Z qubit.
The per-qubit reduction here is exact Pauli algebra (phase tracked as a power of i).
"""
import math

from . import ir
from . import gate_type as gtype

# single-qubit Pauli product table: (a, b) -> (power of i, a*b)
_PAULI_MUL = {}
for _a in 'IXYZ':
    _PAULI_MUL[('I', _a)] = (0, _a)
    _PAULI_MUL[(_a, 'I')] = (0, _a)
for _a in 'XYZ':
    _PAULI_MUL[(_a, _a)] = (0, 'I')
_PAULI_MUL[('X', 'Y')] = (1, 'Z')
_PAULI_MUL[('Y', 'Z')] = (1, 'X')
_PAULI_MUL[('Z', 'X')] = (1, 'Y')
_PAULI_MUL[('Y', 'X')] = (3, 'Z')
_PAULI_MUL[('Z', 'Y')] = (3, 'X')
_PAULI_MUL[('X', 'Z')] = (3, 'Y')

_PAULI_OF = {gtype.ID: 'I', gtype.X: 'X', gtype.Y: 'Y', gtype.Z: 'Z'}


def _gate(gt, qreg, ctrl=None):
    g = ir.Gate(gt)
    if ctrl is not None:
        g.set_ctrllist([ctrl])
    g.set_qreg(qreg)
    return g


def cx(control, target):
    return _gate(gtype.X(), target, control)


def adjoint(gates):
    """Reverse a gate list and flip every adjoint flag (qgate/model/expand.py:6-12)."""
    out = []
    for gate in reversed(gates):
        g = gate.copy()
        g.adjoint = not gate.adjoint
        out.append(g)
    return out


class PauliGatesDiagonalizer:
    """Rotate a Pauli product onto Z of one qubit (qgate/model/pauli_gates_diagonalizer.py)."""

    def __init__(self, gatelist):
        self.gatelist = gatelist

    @staticmethod
    def reduce_lane(gates):
        """Product of the gates on one qubit as (power of i, Pauli letter)."""
        phase, letter = 0, 'I'
        for gate in gates:
            p = _PAULI_OF.get(type(gate.gate_type))
            if p is None:
                raise RuntimeError('input must be X, Y, Z or ID, but {} passed.'.format(gate.gate_type))
            k, letter = _PAULI_MUL[(letter, p)]
            phase += k
        return phase % 4, letter

    def diagonalize(self):
        by_qreg = {}
        for gate in self.gatelist:
            by_qreg.setdefault(gate.qreg, []).append(gate)
        zqregs, idqregs = [], []
        self.plist = []
        phase = 0
        for qreg in sorted(by_qreg, key=lambda q: q.id):
            k, letter = self.reduce_lane(by_qreg[qreg])
            phase += k
            if letter == 'I':
                idqregs.append(qreg)
                continue
            zqregs.append(qreg)
            if letter == 'X':
                self.plist.append(_gate(gtype.H(), qreg))
            elif letter == 'Y':
                self.plist.append(_gate(gtype.SH(), qreg))
        self.phase_offset_in_pi_2 = phase % 4
        self.cxchain = []
        if not zqregs:
            self.op_qreg = idqregs[0]
            return False
        self.cxchain = [cx(zqregs[i], zqregs[i + 1]) for i in range(len(zqregs) - 1)]
        self.op_qreg = zqregs[-1]
        return True

    def get_phase_offset_in_pi_2(self):
        return self.phase_offset_in_pi_2

    def get_phase_coef(self):
        angle = self.phase_offset_in_pi_2 * math.pi / 2.
        return complex(math.cos(angle), math.sin(angle))

    def get_pcx(self):
        return [op.copy() for op in self.plist + self.cxchain]


def pauli_string(gatelist):
    """The Pauli product of a PauliProb / PauliMeasure as [(qreg, letter)] over its
    non-identity qubits in qreg-id order, and the qubit the expansion would measure (the last
    of them, pauli_gates_diagonalizer.py:96-135). Per-qubit products are reduced exactly and
    their phases dropped, as the expansion drops them for observers (expand.py:53-77)."""
    by_qreg = {}
    for gate in gatelist:
        by_qreg.setdefault(gate.qreg, []).append(gate)
    terms = []
    for qreg in sorted(by_qreg, key=lambda q: q.id):
        _, letter = PauliGatesDiagonalizer.reduce_lane(by_qreg[qreg])
        if letter != 'I':
            terms.append((qreg, letter))
    if not terms:
        raise RuntimeError('measurement is not z-based.')
    return terms, terms[-1][0]


def expand_swap(op):
    q0, q1 = op.qreglist[0], op.qreglist[1]
    return [cx(q0, q1), cx(q1, q0), cx(q0, q1)]


def expand_exp(exp):
    """exp(i theta P) -> basis change + CX chain + ExpiZ/ExpiI + inverse chain.

    The reference raises a NameError for Expi(...).Adj (qgate/model/expand.py:45-47); here
    the adjoint is lowered as exp(-i theta P), its mathematical meaning."""
    diag = PauliGatesDiagonalizer(exp.gatelist)
    z_based = diag.diagonalize()
    pcx = diag.get_pcx()
    if diag.phase_offset_in_pi_2 % 2 != 0:
        raise RuntimeError('cannot expand, {}.'.format(repr(exp)))
    coef = diag.get_phase_coef()
    theta = exp.gate_type.args[0]
    if exp.adjoint:
        theta = -theta
    angle = (coef * theta).real
    core = _gate(gtype.ExpiZ(angle) if z_based else gtype.ExpiI(angle), diag.op_qreg)
    expanded = pcx + [core] + adjoint(pcx)
    if exp.ctrllist is not None:
        ctrls = set(exp.ctrllist)
        for gate in expanded:
            merged = ctrls | set(gate.ctrllist or [])
            gate.set_ctrllist(sorted(merged, key=lambda q: q.id))
            assert gate.qreg not in merged, 'control bits must not overlap targets.'
    expanded.reverse()
    return expanded


def _expand_observer(op, make):
    diag = PauliGatesDiagonalizer(op.gatelist)
    if not diag.diagonalize():
        raise RuntimeError('measurement is not z-based.')
    pcx = diag.get_pcx()
    expanded = pcx + [make(op.outref, diag.op_qreg)] + adjoint(pcx)
    expanded.reverse()
    return expanded


def expand_pmeasure(pmeasure):
    return _expand_observer(pmeasure, ir.Measure)


def expand_pprob(pprob):
    return _expand_observer(pprob, ir.Prob)


def expand(op):
    if isinstance(op, ir.PauliMeasure):
        return expand_pmeasure(op)
    if isinstance(op, ir.PauliProb):
        return expand_pprob(op)
    if isinstance(op.gate_type, gtype.SWAP):
        return expand_swap(op)
    if isinstance(op.gate_type, gtype.Expi):
        return expand_exp(op)
    raise RuntimeError('Unknown composed gate, {}.'.format(repr(op.gate_type)))
