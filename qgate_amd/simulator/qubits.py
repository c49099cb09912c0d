"""User view of all state vectors: amplitudes, probabilities, marginals, sampling pools.

Follows qgate/simulator/qubits.py:1-214. The external index space is the product of all
qregs in `ordering` (given ordering first, then the remaining qregs by id); qregs of the
given ordering that hold no state are *empty lanes* and always read as |0>.
"""
import numpy as np

from . import lanes as lanes_mod
from .observation import ObservationList


def null(v):
    return v


def abs2(v):
    return v.real ** 2 + v.imag ** 2


prob = abs2


class StateGetter:
    """`sim.qubits.states[...]` / `sim.qubits.prob[...]` indexing helper."""

    def __init__(self, qubits, mathop, n_qregs):
        import weakref
        self._qubits = weakref.ref(qubits)
        self._mathop = mathop
        self._n_qregs = n_qregs

    @property
    def n_qregs(self):
        return self._n_qregs

    @property
    def mathop(self):
        return self._mathop

    def __getitem__(self, key):
        qubits = self._qubits()
        if qubits is None:
            return None
        return qubits.get_states(self._mathop, key)


class EmptySamplingPool:
    """Pool over qregs that hold no state: every sample is 0 (empty_sampling_pool.py:4-10)."""

    def __init__(self, qreg_ordering):
        self.qreg_ordering = qreg_ordering

    def sample(self, n_samples, randnum=None):
        return ObservationList(self.qreg_ordering, np.zeros([n_samples], np.int64), 0)


def _normalize_slice(key, n_states):
    """(start, step, count) for a slice over n_states entries (qgate/simulator/qubits.py:148-186)."""
    start, stop, step = key.start, key.stop, key.step
    if step is None:
        step = 1
    if step == 0:
        raise ValueError('slice step cannot be zero')
    if step > 0:
        start = 0 if start is None else (start + n_states if start < 0 else start)
        stop = n_states if stop is None else (stop + n_states if stop < 0 else stop)
        start = max(0, min(start, n_states))
        stop = max(0, min(stop, n_states))
        stop = max(start, stop)
    else:
        start = n_states - 1 if start is None else (start + n_states if start < 0 else start)
        stop = -1 if stop is None else (stop + n_states if stop < 0 else stop)
        start = max(-1, min(start, n_states - 1))
        stop = max(-1, min(stop, n_states - 1))
        stop = min(start, stop)
    count = (abs(stop - start + step) - 1) // abs(step)
    return start, step, count


class Qubits:

    def __init__(self, states_getter, dtype):
        self.states_getter = states_getter
        self.dtype = dtype
        self._given_ordering = None
        self._ordering = []
        self.lanes = lanes_mod.Lanes()
        self.qstates_list = []

    def reset(self):
        self.lanes.clear()
        for qs in self.qstates_list:
            qs.delete()
        self.qstates_list = []

    def get_n_lanes(self):
        return len(self.lanes)

    def get_n_qregs(self):
        qregs = set(self.lanes.keys())
        if self._given_ordering is not None:
            qregs |= set(self._given_ordering)
        return len(qregs)

    @property
    def ordering(self):
        return self._ordering

    def set_ordering(self, qreglist):
        self._given_ordering = qreglist
        self.update_external_layout()

    @property
    def states(self):
        return StateGetter(self, null, self.get_n_qregs())

    @property
    def prob(self):
        return StateGetter(self, abs2, self.get_n_qregs())

    def update_external_layout(self):
        given = list(self._given_ordering) if self._given_ordering is not None else []
        rest = set(self.lanes.keys()) - set(given)
        self._ordering = given + sorted(rest, key=lambda q: q.id)

    def calc_probability(self, qreg):
        from ..model import Qreg
        if not isinstance(qreg, Qreg):
            raise RuntimeError('qreg must be an instance of class Qreg.')
        lane = self.lanes[qreg]
        return lane.qstates.calc_probability(lane.local)

    def create_sampling_pool(self, qreg_ordering, sampling_pool_factory=None):
        if len(set(qreg_ordering)) != len(qreg_ordering):
            raise RuntimeError('qreg_ordering has duplicate qregs, {}.'.format(repr(qreg_ordering)))
        pool_ordering = [q for q in qreg_ordering if q in self.lanes]
        if not pool_ordering:
            return EmptySamplingPool(qreg_ordering)
        empty_lanes = [i for i, q in enumerate(qreg_ordering) if q not in self.lanes]
        lane_trans = lanes_mod.create_lane_transformation(self.lanes, pool_ordering)
        n_pool_lanes = len(pool_ordering)
        n_hidden_lanes = len(self.lanes) - n_pool_lanes
        return self.states_getter.create_sampling_pool(
            qreg_ordering, n_pool_lanes, n_hidden_lanes, lane_trans, empty_lanes,
            sampling_pool_factory)

    def get_states(self, mathop=null, key=None):
        if mathop is null:
            dtype = np.complex64 if self.dtype == np.float32 else np.complex128
        elif mathop is abs2:
            dtype = self.dtype
        else:
            raise RuntimeError('unsupported mathop, {}'.format(repr(mathop)))

        empty_lanes = []
        if self._given_ordering is not None:
            empty_lanes = [i for i, q in enumerate(self._given_ordering) if q not in self.lanes]
        lane_trans = lanes_mod.create_lane_transformation(self.lanes, self._ordering)

        n_states = 1 << self.get_n_qregs()
        if key is None:
            key = slice(0, n_states)
        if isinstance(key, slice):
            start, step, count = _normalize_slice(key, n_states)
            if count == 0:
                return np.ones([0], dtype)
            values = np.empty([count], dtype)
            self.states_getter.get_states(values, 0, mathop, lane_trans, empty_lanes,
                                          count, start, step)
            return values

        idx = int(key)
        if idx < 0:
            if idx <= -n_states:
                raise ValueError('list index out of range')
            idx += n_states
        if n_states <= idx:
            raise RuntimeError('list index out of range')
        values = np.empty([1], dtype)
        self.states_getter.get_states(values, 0, mathop, lane_trans, empty_lanes, 1, idx, 1)
        return values[0]
