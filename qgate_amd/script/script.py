"""User-facing circuit DSL: register/reference factories, gate factories, ctrl(), measure.

Names, call shapes and validation follow qgate/script/script.py:1-310:
  H(q), X.Adj(q), Rx(theta)(q), U3(t, p, l).Adj(q), ctrl(c0, c1).X(t), Expi(theta)(paulis),
  Swap(q0, q1), measure(ref, q), prob(ref, [X(q0), Z(q1)]), if_(refs, cond, clause).
"""
import numbers

from .. import model
from ..model import gate_type as gtype


def _flatten(args):
    out = []
    if isinstance(args, (list, tuple, set)):
        for child in args:
            out += _flatten(child)
    else:
        out.append(args)
    return out


def _number(obj):
    if not isinstance(obj, numbers.Number):
        raise RuntimeError('{} is not a number'.format(str(obj)))


def new_gatelist():
    return model.GateList()


def new_qreg():
    return model.Qreg()


def new_qregs(count):
    return [model.Qreg() for _ in range(count)]


def release_qreg(qreg):
    return model.ReleaseQreg(qreg)


def new_reference():
    return model.Reference()


def new_references(count):
    return [model.Reference() for _ in range(count)]


def measure(outref, args):
    if isinstance(args, model.Qreg):
        return model.Measure(outref, args)
    return model.PauliMeasure(outref, _flatten(args))


def prob(outref, args):
    if isinstance(args, model.Qreg):
        return model.Prob(outref, args)
    return model.PauliProb(outref, _flatten(args))


def barrier(*qregs):
    return model.Barrier(_flatten(qregs))


def reset(qreg):
    return model.Reset(qreg)


def if_(refs, cond, clause):
    clause_list = model.GateList()
    clause_list.set(clause)
    return model.IfClause(_flatten(refs), cond, clause_list)


class GateFactory:
    """Builds one Gate; `.Adj` flags it adjoint, call binds the target."""

    def __init__(self, gate_type, ctrllist=None):
        self.gate = model.Gate(gate_type)
        if ctrllist is not None:
            self.gate.set_ctrllist(ctrllist)

    @property
    def Adj(self):
        self.gate.set_adjoint(True)
        return self

    def __call__(self, qreg):
        self.gate.set_qreg(qreg)
        self.gate.check_constraints()
        return self.gate


class ConstGateFactory:
    """Parameterless gate (H, X, ...): every call makes a fresh Gate."""

    def __init__(self, gate_type):
        self.gate_type = gate_type

    @property
    def Adj(self):
        return GateFactory(self.gate_type).Adj

    def __call__(self, qreg):
        return GateFactory(self.gate_type)(qreg)


class GatelistMacroFactory:
    def __init__(self, gate_type, ctrllist=None):
        self.gate = model.GatelistMacro(gate_type)
        if ctrllist is not None:
            self.gate.set_ctrllist(ctrllist)

    @property
    def Adj(self):
        self.gate.set_adjoint(True)
        return self

    def __call__(self, *gatelist):
        self.gate.set_gatelist(_flatten(gatelist))
        self.gate.check_constraints()
        return self.gate


I = ConstGateFactory(gtype.ID())
H = ConstGateFactory(gtype.H())
S = ConstGateFactory(gtype.S())
T = ConstGateFactory(gtype.T())
X = ConstGateFactory(gtype.X())
Y = ConstGateFactory(gtype.Y())
Z = ConstGateFactory(gtype.Z())
SH = ConstGateFactory(gtype.SH())


def Rx(theta):
    _number(theta)
    return GateFactory(gtype.RX(theta))


def Ry(theta):
    _number(theta)
    return GateFactory(gtype.RY(theta))


def Rz(theta):
    _number(theta)
    return GateFactory(gtype.RZ(theta))


def U1(_lambda):
    _number(_lambda)
    return GateFactory(gtype.U1(_lambda))


def U2(phi, _lambda):
    _number(phi)
    _number(_lambda)
    return GateFactory(gtype.U2(phi, _lambda))


def U3(theta, phi, _lambda):
    for v in (theta, phi, _lambda):
        _number(v)
    return GateFactory(gtype.U(theta, phi, _lambda))


def Expii(theta):
    _number(theta)
    return GateFactory(gtype.ExpiI(theta))


def Expiz(theta):
    _number(theta)
    return GateFactory(gtype.ExpiZ(theta))


def Expi(theta):
    return GatelistMacroFactory(gtype.Expi(theta))


class ControlledGateFactory:
    """ctrl(c0, c1, ...).<Gate>: the same gate set with a control list attached."""

    def __init__(self, control):
        self.control = _flatten(control)
        if len(self.control) == 0:
            raise RuntimeError('control qreg list must not be empty.')

    def _make(self, gate_type):
        return GateFactory(gate_type, list(self.control))

    I = property(lambda self: self._make(gtype.ID()))
    H = property(lambda self: self._make(gtype.H()))
    S = property(lambda self: self._make(gtype.S()))
    T = property(lambda self: self._make(gtype.T()))
    X = property(lambda self: self._make(gtype.X()))
    Y = property(lambda self: self._make(gtype.Y()))
    Z = property(lambda self: self._make(gtype.Z()))
    SH = property(lambda self: self._make(gtype.SH()))

    def Rx(self, theta):
        _number(theta)
        return self._make(gtype.RX(theta))

    def Ry(self, theta):
        _number(theta)
        return self._make(gtype.RY(theta))

    def Rz(self, theta):
        _number(theta)
        return self._make(gtype.RZ(theta))

    def U1(self, _lambda):
        _number(_lambda)
        return self._make(gtype.U1(_lambda))

    def U2(self, phi, _lambda):
        _number(phi)
        _number(_lambda)
        return self._make(gtype.U2(phi, _lambda))

    def U3(self, theta, phi, _lambda):
        for v in (theta, phi, _lambda):
            _number(v)
        return self._make(gtype.U(theta, phi, _lambda))

    def Expii(self, theta):
        return self._make(gtype.ExpiI(theta))

    def Expiz(self, theta):
        return self._make(gtype.ExpiZ(theta))

    def Expi(self, theta):
        return GatelistMacroFactory(gtype.Expi(theta), list(self.control))


def controlled(*control):
    return ControlledGateFactory(control)


ctrl = controlled


def Swap(qreg0, qreg1):
    g = model.MultiQubitGate(gtype.SWAP())
    g.set_qreglist([qreg0, qreg1])
    return g
