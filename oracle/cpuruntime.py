"""ORACLE — test infrastructure only: a qgate runtime module over the CPU oracle libraries.

`port()` loads liboracle.so (the C restatement of the reference CPU runtime) and
`reference()` loads _ref/libqgate_ref.so (the reference's own C++ CPU core); both implement
the C API of qgate_oracle.h. The returned module object plugs into
qgate_amd.simulator.simulator.Simulator like any runtime module, with the reference CPU
runtime's conventions: hidden sampling-pool lanes in the least significant bits
(qgate/simulator/cpuruntime.py:13-24, native_qubits_states_getter.py:73-82).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
"""
import ctypes
import os
import subprocess
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, 'liboracle.so')
REF_LIB = os.path.join(HERE, '_ref', 'libqgate_ref.so')
REFERENCE_SRC = '/root/reference/qgate/simulator/src'


def build(reference=True):
    """Compile the oracle (and the reference core where its sources are present)."""
    subprocess.run(['make', '-s', '-B', '-C', HERE, 'liboracle.so'], check=True)
    if reference and os.path.isdir(REFERENCE_SRC):
        subprocess.run(['make', '-s', '-B', '-C', HERE, 'ref', 'REFSRC=' + REFERENCE_SRC], check=True)


_V = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_I64 = ctypes.c_int64
_SIG = {
    'qgo_last_error': (ctypes.c_char_p, []),
    'qgo_kind': (ctypes.c_char_p, []),
    'qgo_num_workers': (_I, []),
    'qgo_states_new': (_V, [_I]),
    'qgo_states_delete': (None, [_V]),
    'qgo_states_n_lanes': (_I, [_V]),
    'qgo_states_data': (_V, [_V]),
    'qgo_initialize_states': (_I, [_V, _I]),
    'qgo_reset_states': (_I, [_V]),
    'qgo_calc_probability': (_I, [_V, _I, ctypes.POINTER(_D)]),
    'qgo_join': (_I, [_V, ctypes.POINTER(_V), _I, _I]),
    'qgo_decohere': (_I, [_I, _D, _V, _I]),
    'qgo_decohere_and_separate': (_I, [_I, _D, _V, _V, _V, _I]),
    'qgo_apply_reset': (_I, [_V, _I]),
    'qgo_apply_gate': (_I, [ctypes.POINTER(_D), _V, _I]),
    'qgo_apply_controlled_gate': (_I, [ctypes.POINTER(_D), _V, ctypes.POINTER(_I), _I, _I]),
    'qgo_gate_matrix': (_I, [ctypes.c_char_p, ctypes.POINTER(_D), _I, _I, ctypes.POINTER(_D)]),
    'qgo_get_states': (_I, [_I, _V, _I, ctypes.POINTER(_I), ctypes.POINTER(_V), _I, ctypes.c_uint64, _I64, _I64, _I64]),
    'qgo_prepare_prob_array': (_I, [_I, _V, ctypes.POINTER(_I), ctypes.POINTER(_V), _I, _I, _I]),
    'qgo_pool_new': (_V, [_I, ctypes.POINTER(_I), ctypes.POINTER(_V), _I, _I, _I, ctypes.POINTER(_I), _I]),
    'qgo_pool_sample': (_I, [_V, ctypes.c_void_p, _I, ctypes.c_void_p]),
    'qgo_pool_delete': (None, [_V]),
    'qgo_pool_cdf': (_V, [_V, ctypes.POINTER(_I64)]),
}

_libs = {}


def load(path):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError('oracle library not built: {}'.format(path))
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIG.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = L
    return L


def _ints(v):
    v = list(v)
    return (ctypes.c_int * max(1, len(v)))(*v)


def _doubles(v):
    v = [float(x) for x in v]
    return (ctypes.c_double * max(1, len(v)))(*v)


def _handles(objs):
    return (ctypes.c_void_p * max(1, len(objs)))(*[o.ptr for o in objs])


def _prec(dtype):
    return 0 if dtype == np.float64 else 1


class _Runtime:
    """Runtime-module object bound to one oracle library."""

    def __init__(self, path):
        self.L = load(path)
        self.kind = self.L.qgo_kind().decode()

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.qgo_last_error().decode(errors='replace'))

    def gate_matrix(self, gate_type, adjoint=False):
        out = (ctypes.c_double * 8)()
        args = gate_type.args
        self.check(self.L.qgo_gate_matrix(gate_type.name.encode(), _doubles(args), len(args),
                                          1 if adjoint else 0, out))
        v = np.array(out[:])
        return (v[0::2] + 1j * v[1::2]).reshape(2, 2)

    def create_qubit_states(self, dtype):
        return OracleQubitStates(self, dtype)

    def create_qubits_states_getter(self, dtype):
        return OracleQubitsStatesGetter(self, dtype)

    def module(self):
        m = types.SimpleNamespace()
        m.create_qubit_states = self.create_qubit_states
        m.create_qubits_states_getter = self.create_qubits_states_getter
        m.runtime = self
        return m


class OracleQubitProcessor:
    def __init__(self, rt, dtype):
        self.rt, self.dtype = rt, dtype

    def synchronize(self):
        pass

    def reset(self):
        pass

    def delete(self):
        pass

    def initialize_qubit_states(self, qs, n_lanes):
        self.rt.check(self.rt.L.qgo_initialize_states(qs.ptr, n_lanes))
        qs.reset_lane_states()

    def reset_qubit_states(self, qs):
        self.rt.check(self.rt.L.qgo_reset_states(qs.ptr))
        qs.reset_lane_states()

    def calc_probability(self, qs, lane):
        out = ctypes.c_double()
        self.rt.check(self.rt.L.qgo_calc_probability(qs.ptr, lane, ctypes.byref(out)))
        return out.value

    def join(self, qs, qs_list, n_new):
        self.rt.check(self.rt.L.qgo_join(qs.ptr, _handles(qs_list), len(qs_list), n_new))

    def decohere(self, value, prob, qs, lane):
        self.rt.check(self.rt.L.qgo_decohere(value, prob, qs.ptr, lane))

    def decohere_and_separate(self, value, prob, qs0, qs1, qs, lane):
        self.rt.check(self.rt.L.qgo_decohere_and_separate(value, prob, qs0.ptr, qs1.ptr, qs.ptr, lane))

    def apply_reset(self, qs, lane):
        self.rt.check(self.rt.L.qgo_apply_reset(qs.ptr, lane))

    def _mat(self, gate_type, adjoint):
        out = (ctypes.c_double * 8)()
        args = gate_type.args
        self.rt.check(self.rt.L.qgo_gate_matrix(gate_type.name.encode(), _doubles(args), len(args),
                                                1 if adjoint else 0, out))
        return out

    def apply_gate(self, gate_type, adjoint, qs, lane):
        self.rt.check(self.rt.L.qgo_apply_gate(self._mat(gate_type, adjoint), qs.ptr, lane))

    def apply_controlled_gate(self, gate_type, adjoint, qs, ctrl, target):
        self.rt.check(self.rt.L.qgo_apply_controlled_gate(self._mat(gate_type, adjoint), qs.ptr,
                                                          _ints(ctrl), len(ctrl), target))

    def apply_matrix(self, mat, qs, ctrl, target):
        m = np.asarray(mat, np.complex128).reshape(4)
        flat = np.empty(8)
        flat[0::2], flat[1::2] = m.real, m.imag
        self.rt.check(self.rt.L.qgo_apply_controlled_gate(_doubles(flat), qs.ptr, _ints(ctrl),
                                                          len(ctrl), target))


class OracleQubitStates:
    def __init__(self, rt, dtype):
        self.rt = rt
        self.dtype = dtype
        self.ptr = rt.L.qgo_states_new(_prec(dtype))
        self.processor = OracleQubitProcessor(rt, dtype)
        self.lane_states = []

    def __del__(self):
        self.delete()

    def delete(self):
        if getattr(self, 'ptr', None):
            self.rt.L.qgo_states_delete(self.ptr)
        self.ptr = None

    def get_n_lanes(self):
        return self.rt.L.qgo_states_n_lanes(self.ptr)

    def reset_lane_states(self):
        self.lane_states = [-1] * self.get_n_lanes()

    def get_lane_state(self, lane):
        return self.lane_states[lane]

    def set_lane_state(self, lane, value):
        self.lane_states[lane] = value

    def calc_probability(self, lane):
        return self.processor.calc_probability(self, lane)

    def local_states(self):
        """Copy of the raw amplitudes in local lane order."""
        n = 1 << self.get_n_lanes()
        ctype = np.complex128 if self.dtype == np.float64 else np.complex64
        buf = (ctypes.c_char * (n * np.dtype(ctype).itemsize)).from_address(self.rt.L.qgo_states_data(self.ptr))
        return np.frombuffer(buf, ctype).copy()


def _translate(lane_trans):
    qstates, tables = [], []
    for qs, lanelist in lane_trans:
        table = [None] * qs.get_n_lanes()
        for lane in lanelist:
            table[lane.local] = lane.external
        qstates.append(qs)
        tables += table
    return qstates, tables


class OracleSamplingPool:
    def __init__(self, rt, ptr, qreg_ordering, mask):
        self.rt, self.ptr = rt, ptr
        self.qreg_ordering = qreg_ordering
        self.mask = mask

    def __del__(self):
        if getattr(self, 'ptr', None):
            self.rt.L.qgo_pool_delete(self.ptr)
        self.ptr = None

    def sample(self, n_samples, randnum=None):
        from qgate_amd.simulator.observation import ObservationList
        if randnum is None:
            randnum = np.random.random_sample([n_samples])
        randnum = np.ascontiguousarray(randnum, np.float64)
        obs = np.empty([n_samples], np.int64)
        self.rt.check(self.rt.L.qgo_pool_sample(self.ptr, obs.ctypes.data, n_samples, randnum.ctypes.data))
        return ObservationList(self.qreg_ordering, obs, self.mask)

    def cdf(self):
        n = ctypes.c_int64()
        addr = self.rt.L.qgo_pool_cdf(self.ptr, ctypes.byref(n))
        dtype = np.float64 if self.dtype == np.float64 else np.float32
        buf = (ctypes.c_char * (n.value * np.dtype(dtype).itemsize)).from_address(addr)
        return np.frombuffer(buf, dtype).copy()


class OracleQubitsStatesGetter:
    def __init__(self, rt, dtype):
        self.rt, self.dtype = rt, dtype

    def delete(self):
        pass

    def get_states(self, values, offset, mathop, lane_trans, empty_lanes, n_states, start, step):
        from qgate_amd.simulator import qubits
        op = 0 if mathop is qubits.null else 1
        qstates, tables = _translate(lane_trans)
        mask = 0
        for p in empty_lanes:
            mask |= 1 << p
        if not qstates:
            values[offset:offset + n_states] = 0
            if start == 0:
                values[offset] = 1
            return
        out = values[offset:offset + n_states]
        self.rt.check(self.rt.L.qgo_get_states(_prec(self.dtype), out.ctypes.data, op, _ints(tables),
                                               _handles(qstates), len(qstates), mask, n_states,
                                               start, step))

    def create_sampling_pool(self, qreg_ordering, n_lanes, n_hidden_lanes, lane_trans, empty_lanes,
                             sampling_pool_factory=None):
        # hidden lanes in the least significant bits (native_qubits_states_getter.py:73-82)
        hidden = 0
        for qs, lanelist in lane_trans:
            for lane in lanelist:
                if lane.external == -1:
                    lane.external = hidden
                    hidden += 1
                else:
                    lane.external += n_hidden_lanes
        qstates, tables = _translate(lane_trans)
        if sampling_pool_factory is not None:
            prob = np.empty([1 << n_lanes], self.dtype)
            self.rt.check(self.rt.L.qgo_prepare_prob_array(_prec(self.dtype), prob.ctypes.data, _ints(tables),
                                                           _handles(qstates), len(qstates), n_lanes,
                                                           n_hidden_lanes))
            return sampling_pool_factory(prob, empty_lanes, qreg_ordering)
        ptr = self.rt.L.qgo_pool_new(_prec(self.dtype), _ints(tables), _handles(qstates), len(qstates),
                                     n_lanes, n_hidden_lanes, _ints(empty_lanes), len(empty_lanes))
        if not ptr:
            raise RuntimeError(self.rt.L.qgo_last_error().decode(errors='replace'))
        mask = 0
        for p in empty_lanes:
            mask |= 1 << p
        pool = OracleSamplingPool(self.rt, ptr, qreg_ordering, mask)
        pool.dtype = self.dtype
        return pool


def port():
    """Runtime module over the C restatement."""
    return _Runtime(PORT_LIB).module()


def reference():
    """Runtime module over the reference's own CPU core (needs oracle/_ref built)."""
    return _Runtime(REF_LIB).module()


def have_reference():
    return os.path.exists(REF_LIB)
