"""Collapses (decohere) interleaved with the calls that read or change a state vector.

The HIP runtime records a collapse and launches it from the next call that touches the state
vector; when that call is calc_probability, the collapse and the probability sum run in one
pass. Each sequence below makes the same processor calls on the HIP runtime and on the CPU
oracle (which collapses at once): every probability and the final amplitudes agree within
1e-12 (complex128) and 5e-6 (complex64). Lanes cover the amplitude bit inside a complex64 unit
(0), bits inside a 128-byte line (1..3), inside a wavefront's units (4..9) and above (10..12).
"""
import numpy as np
import pytest

from qgate_amd.model import gate_type as gtype
from qgate_amd.simulator import qubits
from tests.helpers import tol
from tests.test_dense_fusion import _prepared, _read

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
N = 13


def _value(p):
    """The outcome with the larger probability (keeps the collapse's norm moderate)."""
    return 0 if p >= 0.5 else 1


def _measure(proc, qs, lane, probs):
    p = proc.calc_probability(qs, lane)
    probs.append(p)
    v = _value(p)
    proc.decohere(v, p, qs, lane)
    return v


def _seq_measure_all(proc, qs, probs):
    for lane in (0, 1, 2, 3, 4, 6, 9, 10, 12, 5, 11, 7, 8):
        _measure(proc, qs, lane, probs)


def _seq_gate_between(proc, qs, probs):
    for a, b in ((3, 0), (0, 7), (12, 2), (5, 11)):
        _measure(proc, qs, a, probs)
        proc.apply_gate(gtype.H(), False, qs, b)
        proc.apply_controlled_gate(gtype.X(), False, qs, [b], a)
        probs.append(proc.calc_probability(qs, a))
        probs.append(proc.calc_probability(qs, b))


def _seq_collapse_twice(proc, qs, probs):
    # a collapsed lane collapsed again onto the same outcome (probability 1): the second
    # decohere launches the first
    for lane in (1, 8, 0):
        v = _measure(proc, qs, lane, probs)
        proc.decohere(v, 1. if v == 0 else 0., qs, lane)
    probs.append(proc.calc_probability(qs, 4))


def _seq_reset_between(proc, qs, probs):
    for lane in (2, 10):
        _measure(proc, qs, lane, probs)
        proc.apply_reset(qs, lane)
        probs.append(proc.calc_probability(qs, lane))
        proc.apply_gate(gtype.H(), False, qs, lane)
    _measure(proc, qs, 6, probs)


def _seq_same_lane(proc, qs, probs):
    # the probability of the lane just collapsed (0 or 1), from the fused pass
    for lane in (0, 3, 9, 12):
        _measure(proc, qs, lane, probs)
        probs.append(proc.calc_probability(qs, lane))


SEQUENCES = {'measure_all': _seq_measure_all, 'gate_between': _seq_gate_between,
             'collapse_twice': _seq_collapse_twice, 'reset_between': _seq_reset_between,
             'same_lane': _seq_same_lane}


def _run(module, dtype, seq, read):
    qs = _prepared(module, N, dtype, seed=17)
    probs = []
    SEQUENCES[seq](qs.processor, qs, probs)
    amps = read(qs)
    qs.delete()
    return np.array(probs), amps


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('seq', sorted(SEQUENCES))
def test_collapse_sequences_match_oracle(hip_module, port_module, dtype, seq):
    got_p, got = _run(hip_module, dtype, seq, lambda qs: _read(hip_module, qs, N, dtype))
    want_p, want = _run(port_module, dtype, seq, lambda qs: qs.local_states().astype(np.complex128))
    assert got_p.shape == want_p.shape
    assert np.max(np.abs(got_p - want_p)) < tol(dtype), (got_p, want_p)
    assert np.max(np.abs(got - want)) < tol(dtype)


@pytest.mark.parametrize('dtype', DTYPES)
def test_collapse_then_queued_gates(hip_module, port_module, dtype):
    """With tile fusion on, gates queued after a recorded collapse run after it."""
    rs = np.random.RandomState(3)
    lanes = rs.randint(0, N, 40)

    def seq(module, proc, qs, probs):
        _measure(proc, qs, 5, probs)
        for lane in lanes:
            proc.apply_gate(gtype.U(*rs.uniform(0, 6.3, 3)), False, qs, int(lane))
        _measure(proc, qs, 1, probs)
        for lane in lanes[::-1]:
            proc.apply_gate(gtype.H(), False, qs, int(lane))
        probs.append(proc.calc_probability(qs, 12))

    results = []
    for module, read in ((hip_module, lambda qs: _read(hip_module, qs, N, dtype)),
                         (port_module, lambda qs: qs.local_states().astype(np.complex128))):
        rs = np.random.RandomState(3)
        qs = _prepared(module, N, dtype, seed=23)
        probs = []
        if module is hip_module:
            qs.processor.set_fusion('tile', 0, -1)
            try:
                seq(module, qs.processor, qs, probs)
            finally:
                qs.processor.set_fusion('off')
        else:
            seq(module, qs.processor, qs, probs)
        results.append((np.array(probs), read(qs)))
        qs.delete()
    (gp, g), (wp, w) = results
    assert np.max(np.abs(gp - wp)) < tol(dtype), (gp, wp)
    assert np.max(np.abs(g - w)) < tol(dtype)


def test_recorded_collapse_dropped_with_its_state(hip_module):
    """A state vector deleted with a collapse still recorded is forgotten: synchronize and
    later state vectors are unaffected."""
    qs = _prepared(hip_module, 10, np.float32, seed=5)
    p = qs.processor.calc_probability(qs, 4)
    qs.processor.decohere(_value(p), p, qs, 4)
    qs.delete()
    fresh = _prepared(hip_module, 10, np.float32, seed=5)
    fresh.processor.synchronize()   # applies the queued work of every listed state vector
    assert abs(fresh.processor.calc_probability(fresh, 4) - p) < 5e-6
    fresh.delete()


@pytest.mark.parametrize('n,dtype', [(28, np.float32), (27, np.float64)])
def test_measure_all_full_size(hip_module, n, dtype):
    """Every qubit of a 2 GiB random state measured in a shuffled order (each probability after
    the first from the fused pass): afterwards every P(lane = 0) is 1 or 0 as the outcome says,
    and the basis state the outcomes spell holds the whole norm (size-independent checks)."""
    from qgate_amd.simulator.lanes import Lane
    rs = np.random.RandomState(n)
    qs = hip_module.create_qubit_states(dtype)
    proc = qs.processor
    proc.initialize_qubit_states(qs, n)
    proc.reset_qubit_states(qs)
    for lane in range(n):
        proc.apply_gate(gtype.U(*rs.uniform(0, 6.3, 3)), False, qs, lane)
    for lane in range(n - 1):
        proc.apply_controlled_gate(gtype.X(), False, qs, [lane], lane + 1)
    index = 0
    for lane in rs.permutation(n):
        p = proc.calc_probability(qs, int(lane))
        v = 0 if rs.random_sample() < p else 1
        proc.decohere(v, p, qs, int(lane))
        index |= v << int(lane)
    bar = 5e-6 if dtype == np.float32 else 1e-12
    for lane in range(n):
        want = 0. if (index >> lane) & 1 else 1.
        assert abs(proc.calc_probability(qs, lane) - want) < bar, lane
    lanes = []
    for i in range(n):
        ln = Lane(None, i)
        ln.set_external(i)
        lanes.append(ln)
    out = np.zeros(1, np.complex128 if dtype == np.float64 else np.complex64)
    hip_module.create_qubits_states_getter(dtype).get_states(out, 0, qubits.null, [(qs, lanes)], [], 1, index, 1)
    assert abs(abs(complex(out[0])) - 1.) < bar, out
    qs.delete()
