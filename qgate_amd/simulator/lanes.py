"""Qreg -> (state vector, local bit) bookkeeping (qgate/simulator/lanes.py:1-51).

A *lane* is one qubit of a state vector: `qstates` is the vector and `local` the bit index
of that qubit in the vector's amplitude index. The *external* index of a qreg is its bit
position in the user-visible ordering (Qubits.ordering).
"""


class Lane:
    def __init__(self, qstates, local):
        self.qstates = qstates
        self.local = local
        self.external = -1

    def update(self, qstates, local):
        self.qstates = qstates
        self.local = local

    def set_external(self, external):
        self.external = external


class Lanes(dict):
    """qreg -> Lane."""

    def add_lane(self, qreg, qstates, local):
        self[qreg] = Lane(qstates, local)

    def get_by_qubit_states(self, qs):
        return [lane for lane in self.values() if lane.qstates is qs]


def create_lane_transformation(lanes, qreg_ordering):
    """Per state vector, its lanes sorted by local bit with their external positions.

    Returns [(qstates, [Lane(qreg, local) with .external]), ...] sorted by the lowest local
    bit of each vector; qregs absent from `qreg_ordering` get external = -1
    (qgate/simulator/lanes.py:30-51)."""
    position = {qreg: i for i, qreg in enumerate(qreg_ordering)}
    grouped = {}
    order = []
    for qreg, lane in lanes.items():
        key = id(lane.qstates)
        if key not in grouped:
            grouped[key] = (lane.qstates, [])
            order.append(key)
        entry = Lane(qreg, lane.local)
        entry.set_external(position.get(qreg, -1))
        grouped[key][1].append(entry)
    trans = []
    for key in order:
        qs, lanelist = grouped[key]
        lanelist.sort(key=lambda ln: ln.local)
        trans.append((qs, lanelist))
    trans.sort(key=lambda tr: tr[1][0].local)
    return trans
