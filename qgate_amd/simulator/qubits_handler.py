"""State-vector lifecycle: allocate, join, measure-and-separate, release.

Follows qgate/simulator/qubits_handler.py:1-135. Local bit layout after a join: the joined
vectors are ordered by ascending lane count and the last (largest) one takes the lowest
local bits; brand-new qregs take the top bits.
"""


class QubitsHandler:

    def __init__(self, qubit_states_factory, qubits):
        self._create_qubit_states = qubit_states_factory
        self._qubits = qubits

    def reset(self):
        self._qubits.reset()

    @property
    def lanes(self):
        return self._qubits.lanes

    def _allocate_qubit_states(self, n_lanes):
        qstates = self._create_qubit_states(self._qubits.dtype)
        qstates.processor.initialize_qubit_states(qstates, n_lanes)
        self._qubits.qstates_list.append(qstates)
        return qstates

    def add_qubit_states(self, qregset):
        assert len(qregset) != 0, 'empty qreg set.'
        qstates = self._allocate_qubit_states(len(qregset))
        qstates.processor.reset_qubit_states(qstates)
        for local, qreg in enumerate(sorted(qregset, key=lambda q: q.id)):
            self.lanes.add_lane(qreg, qstates, local)

    def join(self, qregset):
        assert 1 < len(qregset), '2 or more qregs required.'
        by_qstates = {}
        new_qregs = []
        # qregs in id order: state vectors of equal lane count keep the order of their
        # smallest qreg id, so the local layout does not depend on set iteration order
        # (which follows the qregs' hashes in the reference, qubits_handler.py:41-60)
        for qreg in sorted(qregset, key=lambda q: q.id):
            if qreg in self.lanes:
                lane = self.lanes[qreg]
                by_qstates.setdefault(id(lane.qstates), (lane.qstates, []))[1].append(lane)
            else:
                new_qregs.append(qreg)
        new_qregs.sort(key=lambda q: q.id)

        if not by_qstates:
            self.add_qubit_states(new_qregs)
            return

        sources = sorted((entry[0] for entry in by_qstates.values()),
                         key=lambda qs: qs.get_n_lanes())
        for qs in sources:
            self._qubits.qstates_list.remove(qs)
        joined = self._allocate_qubit_states(len(qregset))
        joined.processor.join(joined, sources, len(new_qregs))

        offset = 0
        for qs in reversed(sources):
            moved = by_qstates[id(qs)][1]
            for lane in moved:
                new_local = lane.local + offset
                joined.set_lane_state(new_local, qs.get_lane_state(lane.local))
                lane.update(joined, new_local)
            offset += len(moved)
        for i, qreg in enumerate(new_qregs):
            self.lanes.add_lane(qreg, joined, offset + i)
        for qs in sources:
            qs.delete()

    def decohere_and_separate(self, qreg, value, prob):
        sep_lane = self.lanes[qreg]
        qstates, local = sep_lane.qstates, sep_lane.local
        n_lanes = qstates.get_n_lanes()
        qstates0 = self._allocate_qubit_states(n_lanes - 1)
        qstates1 = self._allocate_qubit_states(1)
        self._qubits.qstates_list.remove(qstates)

        qstates.processor.decohere_and_separate(value, prob, qstates0, qstates1, qstates, local)

        others = [ln for ln in self.lanes.get_by_qubit_states(qstates) if ln is not sep_lane]
        for lane in others:
            new_local = lane.local - 1 if local < lane.local else lane.local
            qstates0.set_lane_state(new_local, qstates.get_lane_state(lane.local))
        qstates1.set_lane_state(0, qstates.get_lane_state(local))
        for lane in others:
            lane.update(qstates0, lane.local - 1 if local < lane.local else lane.local)
        sep_lane.update(qstates1, 0)
        qstates.delete()

    def deallocate_qubit_states(self, qreg):
        lane = self.lanes[qreg]
        if lane.qstates.get_n_lanes() != 1:
            raise RuntimeError('qreg/lane is not separated.')
        if lane.qstates.get_lane_state(0) == -1:
            raise RuntimeError('qreg/lane is not measured.')
        self.lanes.pop(qreg)
        self._qubits.qstates_list.remove(lane.qstates)
        lane.qstates.delete()
        lane.qstates = None
