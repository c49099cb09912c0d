"""Dispatch of preprocessed ops to a runtime's processor, in program order.

Behaviour follows qgate/simulator/model_executor.py:80-175 and rop_executor.py:28-53:
  * one uniform `numpy.random.random_sample()` is drawn per Measure, in program order;
    the outcome is 0 when u < P(qubit = 0), else 1;
  * a Measure immediately followed by its Separate (dynamic mode) collapses and splits the
    qubit off its state vector in one step, unless the vector already has one qubit;
    otherwise the vector is collapsed in place (decohere);
  * Reset is legal only on a measured qubit and flips it back to |0>.
The reference queues ops and drains the queue in order; ops here run as they arrive, with a
one-op lookahead after a Measure to see its Separate, which gives the same order of
runtime calls and random draws.
"""
import numpy.random as np_random

from .. import model
from ..model.expand import pauli_string


class ModelExecutor:

    def __init__(self, qubits_handler, value_store):
        self._qhandler = qubits_handler
        self._value_store = value_store
        self._pending_measure = None

    def enqueue(self, op):
        pending = self._pending_measure
        if pending is not None:
            self._pending_measure = None
            if isinstance(op, model.Separate):
                assert op.qreg == pending.qreg
                self._measure(pending, separate=True)
                return
            self._measure(pending, separate=False)
        if isinstance(op, model.Measure):
            self._pending_measure = op
            return
        self._dispatch(op)

    def flush(self):
        if self._pending_measure is not None:
            pending, self._pending_measure = self._pending_measure, None
            self._measure(pending, separate=False)

    def _measure(self, op, separate):
        lanes = self._qhandler.lanes
        if op.qreg not in lanes:
            self._qhandler.add_qubit_states([op.qreg])
        lane = lanes[op.qreg]
        qstates, local = lane.qstates, lane.local
        separate = separate and qstates.get_n_lanes() != 1
        randnum = np_random.random_sample()
        prob = qstates.processor.calc_probability(qstates, local)
        result = 0 if randnum < prob else 1
        self._value_store.set(op.outref, result)
        qstates.set_lane_state(local, result)
        if separate:
            self._qhandler.decohere_and_separate(op.qreg, result, prob)
        else:
            qstates.processor.decohere(result, prob, qstates, local)

    def _pauli_observe(self, op):
        """PauliProb / PauliMeasure kept whole by the preprocessor: the same value, random
        draw and collapse as their expansion (basis change, CX chain, Prob / Measure of the
        last qubit of the string, inverse gates; qgate/model/expand.py:53-77), in one pass."""
        terms, op_qreg = pauli_string(op.gatelist)
        lanes = self._qhandler.lanes
        qstates = lanes[op_qreg].qstates
        x_lanes = z_lanes = 0
        for qreg, letter in terms:
            lane = lanes[qreg]
            assert lane.qstates is qstates, 'Pauli observable spans several state vectors.'
            if letter in 'XY':
                x_lanes |= 1 << lane.local
            if letter in 'YZ':
                z_lanes |= 1 << lane.local
        proc = qstates.processor
        if isinstance(op, model.PauliProb):
            self._value_store.set(op.outref, proc.calc_pauli_probability(qstates, x_lanes, z_lanes))
            return
        randnum = np_random.random_sample()
        prob = proc.calc_pauli_probability(qstates, x_lanes, z_lanes)
        result = 0 if randnum < prob else 1
        self._value_store.set(op.outref, result)
        qstates.set_lane_state(lanes[op_qreg].local, result)
        proc.pauli_decohere(result, prob, qstates, x_lanes, z_lanes)

    def _dispatch(self, op):
        lanes = self._qhandler.lanes
        if isinstance(op, model.Gate):
            target = lanes[op.qreg]
            qstates = target.qstates
            if op.ctrllist is None:
                qstates.processor.apply_gate(op.gate_type, op.adjoint, qstates, target.local)
            else:
                ctrl_locals = [lanes[c].local for c in op.ctrllist]
                qstates.processor.apply_controlled_gate(op.gate_type, op.adjoint, qstates,
                                                        ctrl_locals, target.local)
        elif isinstance(op, model.NewQreg):
            self._qhandler.add_qubit_states([op.qreg])
        elif isinstance(op, model.Join):
            self._qhandler.join(op.qreglist)
        elif isinstance(op, model.ReleaseQreg):
            self._qhandler.deallocate_qubit_states(op.qreg)
        elif isinstance(op, model.Prob):
            lane = lanes[op.qreg]
            self._value_store.set(op.outref,
                                  lane.qstates.processor.calc_probability(lane.qstates, lane.local))
        elif isinstance(op, (model.PauliProb, model.PauliMeasure)):
            self._pauli_observe(op)
        elif isinstance(op, model.Reset):
            lane = lanes[op.qreg]
            bit = lane.qstates.get_lane_state(lane.local)
            if bit == -1:
                raise RuntimeError('Qubit is not measured.')
            if bit == 1:
                lane.qstates.processor.apply_reset(lane.qstates, lane.local)
                lane.qstates.set_lane_state(lane.local, -1)
        elif isinstance(op, (model.Barrier, model.ClauseBegin, model.ClauseEnd)):
            pass
        elif isinstance(op, model.Separate):
            raise RuntimeError('Separate without a preceding Measure.')
        else:
            raise RuntimeError('Unknown operator, {}.'.format(repr(op)))
