"""Circuit IR: quantum registers, value references, operators and directives.

The classes keep the names, attributes and copy semantics of the reference model layer
(qgate/model/model.py:37-258, qgate/model/directive.py:1-34, qgate/model/gatelist.py:1-105)
because user circuits and the simulator front-end are written against them.
"""
import itertools


class Qreg:
    """One qubit register; identity is an integer id (qgate/model/model.py:38-54)."""
    _next_id = itertools.count()

    def __init__(self):
        self.id = next(Qreg._next_id)

    def __hash__(self):
        return self.id

    def __eq__(self, other):
        return isinstance(other, Qreg) and other.id == self.id

    def __repr__(self):
        return 'q[{}]'.format(self.id)


class Reference:
    """Value slot written by measure/prob (qgate/model/model.py:58-71)."""
    _next_id = itertools.count()

    def __init__(self):
        self.id = next(Reference._next_id)

    def __hash__(self):
        return self.id

    def __eq__(self, other):
        return isinstance(other, Reference) and other.id == self.id

    def __repr__(self):
        return 'r[{}]'.format(self.id)


def _fmt_qregs(qregs):
    if isinstance(qregs, Qreg):
        return repr(qregs)
    return ','.join(repr(q) for q in qregs)


class Operator:
    """Base of every op; `idx` is the execution order set by the preprocessor."""

    def set_idx(self, idx):
        self.idx = idx

    def get_idx(self):
        return self.idx

    def __repr__(self):
        return type(self).__name__


class GateType:
    """Gate kind plus its float parameters (qgate/model/model.py:85-87)."""

    def __init__(self, *args):
        self.args = args

    def __repr__(self):
        if self.args:
            return '{}({})'.format(type(self).__name__, ', '.join(repr(a) for a in self.args))
        return type(self).__name__


class Gate(Operator):
    """Single-target gate with optional control list (qgate/model/model.py:89-119)."""

    def __init__(self, gate_type):
        self.gate_type = gate_type
        self.adjoint = False
        self.qreg = None
        self.ctrllist = None

    def set_adjoint(self, adjoint):
        self.adjoint = adjoint

    def set_ctrllist(self, ctrllist):
        ctrllist = [ctrllist] if isinstance(ctrllist, Qreg) else ctrllist
        if not all(isinstance(q, Qreg) for q in ctrllist):
            raise RuntimeError('control arguments must be Qreg.')
        self.ctrllist = ctrllist

    def set_qreg(self, qreg):
        if not isinstance(qreg, Qreg):
            raise RuntimeError('{} is not a qreg'.format(repr(qreg)))
        self.qreg = qreg

    def check_constraints(self):
        self.gate_type.constraints(self)

    def copy(self):
        g = Gate(self.gate_type)
        g.adjoint = self.adjoint
        if self.ctrllist is not None:
            g.set_ctrllist(list(self.ctrllist))
        g.set_qreg(self.qreg)
        return g

    def __repr__(self):
        s = repr(self.gate_type)
        if self.adjoint:
            s += '.Adj'
        if self.ctrllist:
            s = 'ctrl({}).{}'.format(_fmt_qregs(self.ctrllist), s)
        return '{}({})'.format(s, _fmt_qregs(self.qreg))


class GatelistMacro(Operator):
    """Gate over a list of Pauli gates, e.g. Expi (qgate/model/model.py:121-153)."""

    def __init__(self, gate_type):
        self.gate_type = gate_type
        self.adjoint = False
        self.gatelist = None
        self.ctrllist = None

    def set_adjoint(self, adjoint):
        self.adjoint = adjoint

    def set_ctrllist(self, ctrllist):
        assert self.ctrllist is None, 'ctrl args already set.'
        self.ctrllist = [ctrllist] if isinstance(ctrllist, Qreg) else ctrllist

    def set_gatelist(self, gatelist):
        assert self.gatelist is None, 'gatelist already set.'
        assert isinstance(gatelist[0], Gate), 'gatelist must hold gates.'
        self.gatelist = gatelist

    def check_constraints(self):
        self.gate_type.constraints(self)

    def copy(self):
        g = GatelistMacro(self.gate_type)
        g.adjoint = self.adjoint
        if self.ctrllist is not None:
            g.set_ctrllist(list(self.ctrllist))
        g.set_gatelist([gate.copy() for gate in self.gatelist])
        return g


class MultiQubitGate(Operator):
    """Multi-qubit gate; only SWAP uses it (qgate/model/model.py:157-177)."""

    def __init__(self, gate_type):
        self.gate_type = gate_type
        self.adjoint = False
        self.qreglist = None

    def set_adjoint(self, adjoint):
        self.adjoint = adjoint

    def set_qreglist(self, qreglist):
        assert self.qreglist is None, 'qreglist already set.'
        self.qreglist = qreglist

    def check_constraints(self):
        self.gate_type.constraints(self)

    def copy(self):
        g = MultiQubitGate(self.gate_type)
        g.adjoint = self.adjoint
        g.set_qreglist(list(self.qreglist))
        return g


class Measure(Operator):
    """Projective Z measurement into a Reference (qgate/model/model.py:179-187)."""

    def __init__(self, ref, qreg):
        if not isinstance(qreg, Qreg) or not isinstance(ref, Reference):
            raise RuntimeError('Wrong argument for Measure, {}, {}.'.format(repr(qreg), repr(ref)))
        self.qreg, self.outref = qreg, ref

    def copy(self):
        return Measure(self.outref, self.qreg)

    def __repr__(self):
        return 'Measure({}, {})'.format(repr(self.outref), repr(self.qreg))


class Prob(Operator):
    """P(qreg == 0) into a Reference (qgate/model/model.py:189-197)."""

    def __init__(self, ref, qreg):
        if not isinstance(qreg, Qreg) or not isinstance(ref, Reference):
            raise RuntimeError('Wrong argument for Prob, {}, {}.'.format(repr(qreg), repr(ref)))
        self.qreg, self.outref = qreg, ref

    def copy(self):
        return Prob(self.outref, self.qreg)

    def __repr__(self):
        return 'Prob({}, {})'.format(repr(self.outref), repr(self.qreg))


class PauliObserver(Operator):
    """Observable given by a Pauli product (qgate/model/model.py:199-210)."""

    def __init__(self, ref, gatelist):
        if not isinstance(ref, Reference):
            raise RuntimeError('Wrong argument for Pmeasure, {}.'.format(repr(ref)))
        from . import gate_type as gtype
        for gate in gatelist:
            if not isinstance(gate.gate_type, (gtype.ID, gtype.X, gtype.Y, gtype.Z)):
                raise RuntimeError('Pmeasure only accepts ID, X, Y and Z gates')
            if gate.ctrllist is not None:
                raise RuntimeError('control qreg(s) should not be set for pauli operators.')
        self.gatelist, self.outref = gatelist, ref


class PauliMeasure(PauliObserver):
    def copy(self):
        return PauliMeasure(self.outref, [g.copy() for g in self.gatelist])


class PauliProb(PauliObserver):
    def copy(self):
        return PauliProb(self.outref, [g.copy() for g in self.gatelist])


class Barrier(Operator):
    def __init__(self, qregset):
        assert all(isinstance(q, Qreg) for q in qregset), 'arguments must be Qreg.'
        self.qregset = set(qregset)

    def copy(self):
        return Barrier(self.qregset)

    def __repr__(self):
        return 'Barrier({})'.format(_fmt_qregs(sorted(self.qregset, key=lambda q: q.id)))


class Reset(Operator):
    def __init__(self, qreg):
        assert isinstance(qreg, Qreg), 'arguments must be Qreg.'
        self.qreg = qreg

    def copy(self):
        return Reset(self.qreg)

    def __repr__(self):
        return 'Reset({})'.format(repr(self.qreg))


class IfClause(Operator):
    """Conditional clause on measured values (qgate/model/model.py:249-257)."""

    def __init__(self, refs, cond, clause):
        self.refs = refs
        self.cond = cond
        self.clause = clause

    def copy(self):
        return IfClause(self.refs, self.cond, self.clause)


# directives inserted by the preprocessor (qgate/model/directive.py:1-34)

class ClauseBegin(Operator):
    pass


class ClauseEnd(Operator):
    pass


class NewQreg(Operator):
    def __init__(self, qreg):
        self.qreg = qreg

    def __repr__(self):
        return 'NewQreg({})'.format(repr(self.qreg))


class ReleaseQreg(Operator):
    def __init__(self, qreg):
        self.qreg = qreg

    def copy(self):
        return ReleaseQreg(self.qreg)

    def __repr__(self):
        return 'ReleaseQreg({})'.format(repr(self.qreg))


class Join(Operator):
    def __init__(self, qregs):
        assert 1 < len(qregs)
        self.qreglist = list(qregs)

    def __repr__(self):
        return 'Join({})'.format(_fmt_qregs(self.qreglist))


class Separate(Operator):
    def __init__(self, qreg):
        self.qreg = qreg

    def __repr__(self):
        return 'Separate({})'.format(repr(self.qreg))


class GateList:
    """Ordered, nestable op list; ops are copied on insertion (qgate/model/gatelist.py:4-66)."""

    def __init__(self):
        self.ops = []

    @staticmethod
    def _copy_ops(ops):
        out = []
        for op in ops:
            if isinstance(op, Operator):
                out.append(op.copy())
            elif isinstance(op, (list, tuple)):
                inner = GateList()
                inner.ops = GateList._copy_ops(op)
                out.append(inner)
            elif isinstance(op, GateList):
                inner = GateList()
                inner.ops = GateList._copy_ops(op.ops)
                out.append(inner)
            else:
                raise RuntimeError('Unknown argument, {}'.format(repr(op)))
        return out

    def __getitem__(self, key):
        return self.ops[key]

    def __iter__(self):
        return iter(self.ops)

    def __len__(self):
        return len(self.ops)

    def copy(self):
        g = GateList()
        g.ops = GateList._copy_ops(self.ops)
        return g

    def set(self, ops):
        if isinstance(ops, Operator):
            self.ops = [ops.copy()]
        elif isinstance(ops, GateList):
            self.ops = GateList._copy_ops(ops.ops)
        else:
            self.ops = GateList._copy_ops(ops)

    def _set_no_copy(self, ops):
        self.ops = ops

    def append(self, op):
        assert isinstance(op, Operator), 'GateList.append() accepts Operator.'
        self.ops.append(op.copy())

    def __add__(self, other):
        if isinstance(other, GateList):
            other = other.ops
        elif isinstance(other, Operator):
            other = [other]
        g = GateList()
        g.ops = GateList._copy_ops(list(self.ops) + list(other))
        return g

    def __iadd__(self, other):
        if isinstance(other, GateList):
            self.ops += other.ops
        elif isinstance(other, Operator):
            self.ops.append(other.copy())
        else:
            self.ops += GateList._copy_ops(other)
        return self


class GateListIterator:
    """Depth-first walk through nested clauses; `prepend` enters a clause next
    (qgate/model/gatelist.py:69-105)."""

    def __init__(self, ops):
        self._stack = [iter(ops)]

    def prepend(self, op):
        ops = op.ops if isinstance(op, GateList) else op
        self._stack.append(iter(ops))

    def next(self):
        while self._stack:
            op = next(self._stack[-1], None)
            if op is None:
                self._stack.pop()
                continue
            if isinstance(op, GateList):
                self._stack.append(iter(op.ops))
                continue
            return op
        return None


def dump(ops, file=None, indent=0):
    """Print one op per line with its execution index (qgate/model/gatelist.py:109-117)."""
    import sys
    file = sys.stdout if file is None else file
    for op in ops:
        if isinstance(op, (GateList, list, tuple)):
            dump(op, file, indent + 1)
            continue
        idx = getattr(op, 'idx', '-')
        print('{}: {}{}'.format(idx, '  ' * indent, repr(op)), file=file)
        if isinstance(op, IfClause):
            dump(op.clause, file, indent + 1)
