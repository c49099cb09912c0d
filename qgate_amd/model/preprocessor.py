"""Circuit preprocessing: qreg grouping, Join/Separate insertion and macro lowering.

Follows qgate/model/preprocessor.py:9-142 and qgate/model/qreg_aggregator.py:1-71:
  * dynamic:    qregs start separate; a controlled gate joins the groups it touches
                (Join directive); every Measure is followed by a Separate.
  * static:     groups are formed by scanning the whole circuit first; one NewQreg/Join
                per group is prepended.
  * one_static: all qregs of the circuit form one state vector, joined up front.
"""
from . import ir
from .expand import expand, pauli_string
from .prefs import circuit_prep, dynamic, static, one_static


class QregAggregator:
    """Tracks which qregs share a state vector (qgate/model/qreg_aggregator.py)."""

    def __init__(self):
        self.reset()

    def reset(self):
        self.qregsetlist = []
        self.qregset = set()

    def find_qregset(self, qreg):
        for group in self.qregsetlist:
            if qreg in group:
                return group
        raise AssertionError('qreg {} is not registered.'.format(qreg.id))

    def add_qreg(self, qreg):
        if qreg in self.qregset:
            return False
        self.qregset.add(qreg)
        self.qregsetlist.append(frozenset([qreg]))
        return True

    def aggregate(self, qreglist):
        """Merge the groups of `qreglist`; returns the merged set, or None if already one."""
        for qreg in qreglist:
            self.add_qreg(qreg)
        groups = {self.find_qregset(q) for q in qreglist}
        if len(groups) == 1:
            return None
        merged = set()
        for group in groups:
            self.qregsetlist.remove(group)
            merged |= group
        self.qregsetlist.append(frozenset(merged))
        return merged

    def separate_qreg(self, qreg):
        group = self.find_qregset(qreg)
        self.qregsetlist.remove(group)
        self.qregsetlist.append(frozenset(group - {qreg}))
        self.qregsetlist.append(frozenset([qreg]))


class Preprocessor:

    def __init__(self, fuse_pauli=False, **prefdict):
        """fuse_pauli: keep PauliProb (and, in the static modes, PauliMeasure) as single ops
        for a runtime that evaluates Pauli observables directly; the qubits of the string
        join one state vector, as the expansion's CX chain would join them."""
        self.circ_prep = prefdict.get(circuit_prep, dynamic)
        self.dynamic = self.circ_prep == dynamic
        self.fuse_pauli = fuse_pauli
        self.aggregator = QregAggregator()
        self.reset()

    def reset(self):
        self.aggregator.reset()
        self._refset = set()

    def get_qregset(self):
        return self.aggregator.qregset

    def get_qregsetlist(self):
        return self.aggregator.qregsetlist

    def get_refset(self):
        return self._refset

    def preprocess_operator(self, op):
        out = []
        if isinstance(op, ir.Gate):
            if op.ctrllist is None:
                if self.aggregator.add_qreg(op.qreg) and self.dynamic:
                    out.append(ir.NewQreg(op.qreg))
            else:
                joined = self.aggregator.aggregate([op.qreg] + list(op.ctrllist))
                if joined is not None and self.dynamic:
                    out.append(ir.Join(joined))
            out.append(op)
        elif isinstance(op, (ir.Measure, ir.Prob)):
            if self.aggregator.add_qreg(op.qreg) and self.dynamic:
                out.append(ir.NewQreg(op.qreg))
            self._refset.add(op.outref)
            out.append(op)
            if isinstance(op, ir.Measure) and self.dynamic:
                self.aggregator.separate_qreg(op.qreg)
                out.append(ir.Separate(op.qreg))
        elif isinstance(op, ir.ReleaseQreg):
            if self.dynamic:
                if op.qreg not in self.aggregator.qregset:
                    raise RuntimeError('qreg{} is not in circuit.'.format(op.qreg.id))
                if len(self.aggregator.find_qregset(op.qreg)) != 1:
                    raise RuntimeError('qreg{} is not separated.'.format(op.qreg.id))
                out.append(op)
        elif isinstance(op, ir.Reset):
            if op.qreg not in self.aggregator.qregset:
                raise RuntimeError('unused qreg found, {}'.format(repr(op.qreg)))
            out.append(op)
        elif isinstance(op, ir.Barrier):
            out.append(op)
        elif isinstance(op, (ir.PauliMeasure, ir.PauliProb)) and self.fuse_pauli and \
                (isinstance(op, ir.PauliProb) or not self.dynamic):
            # a Measure in dynamic mode is followed by a Separate: that case keeps the expansion
            terms, _ = pauli_string(op.gatelist)
            qregs = [q for q, _ in terms]
            if len(qregs) == 1:
                if self.aggregator.add_qreg(qregs[0]) and self.dynamic:
                    out.append(ir.NewQreg(qregs[0]))
            else:
                joined = self.aggregator.aggregate(qregs)
                if joined is not None and self.dynamic:
                    out.append(ir.Join(joined))
            self._refset.add(op.outref)
            out.append(op)
        elif isinstance(op, (ir.MultiQubitGate, ir.GatelistMacro, ir.PauliMeasure, ir.PauliProb)):
            out += self.preprocess_operator_list(expand(op))
        elif isinstance(op, ir.GateList):
            out += self.preprocess_clause(op)
        elif isinstance(op, ir.IfClause):
            out.append(ir.IfClause(op.refs, op.cond, self.preprocess_clause(op.clause)))
        elif isinstance(op, (ir.ClauseBegin, ir.ClauseEnd)):
            out.append(op)
        else:
            raise RuntimeError('Unexpected op, {}'.format(repr(op)))
        return out

    def preprocess_operator_list(self, oplist):
        out = []
        for op in oplist:
            out += self.preprocess_operator(op)
        return out

    def preprocess_clause(self, clause):
        return [ir.ClauseBegin()] + self.preprocess_operator_list(clause.ops) + [ir.ClauseEnd()]

    def preprocess(self, clause):
        ops = self.preprocess_clause(clause)
        if not self.dynamic:
            groups = self.aggregator.qregsetlist if self.circ_prep == static \
                else ([self.aggregator.qregset] if self.aggregator.qregset else [])
            prologue = []
            for group in groups:
                if len(group) == 1:
                    prologue.append(ir.NewQreg(next(iter(group))))
                else:
                    prologue.append(ir.Join(group))
            ops = prologue + ops

        it = ir.GateListIterator(ops)
        idx = 0
        while True:
            op = it.next()
            if op is None:
                break
            op.set_idx(idx)
            idx += 1
            if isinstance(op, ir.IfClause):
                it.prepend(op.clause)

        out = ir.GateList()
        out.ops = ops
        return out
