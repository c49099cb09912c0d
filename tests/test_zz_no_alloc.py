"""qghip_qubit_processor_calc_probability_noalloc against qghip_qubit_processor_calc_probability.

Both entry points are called through the C ABI on the same state vectors: two calls in a row,
the second on a smaller state, in both precisions. With no recorded collapse the two share the
kernels and the summation order, so the probabilities must be equal bit for bit; after a
recorded collapse the no-allocation variant sums in a separate pass, so the two agree to FP64
rounding.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1
OK, EVALUE = 0, -2


def _lib():
    from qgate_amd.simulator import hipruntime
    return hipruntime.lib()


def _check(rc):
    assert rc == OK, _lib().qghip_last_error().decode()


def _new(fn, *args):
    h = ctypes.c_void_p()
    _check(fn(*args, ctypes.byref(h)))
    return h.value


def _doubles(v):
    return (ctypes.c_double * max(1, len(v)))(*v)


def _prepare(L, proc, n_lanes, seed):
    """A random state of n_lanes lanes: U3 on every lane, then CX ladders and more U3."""
    u = ctypes.c_int()
    _check(L.qghip_register_matrix_factory(b'U', ctypes.byref(u)))
    x = ctypes.c_int()
    _check(L.qghip_register_matrix_factory(b'X', ctypes.byref(x)))
    qs = _new(L.qghip_qubit_states_new, proc[1])
    _check(L.qghip_qubit_processor_initialize_qubit_states(proc[0], qs, n_lanes))
    _check(L.qghip_qubit_processor_reset_qubit_states(proc[0], qs))
    rng = np.random.default_rng(seed)
    for layer in range(3):
        for t in range(n_lanes):
            args = _doubles(list(rng.uniform(0, 2 * np.pi, 3)))
            _check(L.qghip_qubit_processor_apply_gate(proc[0], u.value, args, 3, 0, qs, t))
        for t in range(layer % 2, n_lanes - 1, 2):
            ctrl = (ctypes.c_int * 1)(t)
            _check(L.qghip_qubit_processor_apply_controlled_gate(proc[0], x.value, _doubles([]), 0, 0, qs,
                                                                 ctrl, 1, t + 1))
    return qs


def _probs(L, fn, proc, qs, n_lanes):
    out = []
    for lane in range(n_lanes):
        v = ctypes.c_double()
        _check(fn(proc, qs, lane, ctypes.byref(v)))
        out.append(v.value)
    return np.array(out)


@pytest.mark.parametrize('prec', [FP64, FP32])
def test_noalloc_matches_calc_probability(prec):
    L = _lib()
    _check(L.qghip_devices_initialize(0))
    proc = _new(L.qghip_qubit_processor_new, prec)
    try:
        for n_lanes, seed in ((22, 11), (13, 12)):      # second call on a smaller state
            qs = _prepare(L, (proc, prec), n_lanes, seed)
            ref = _probs(L, L.qghip_qubit_processor_calc_probability, proc, qs, n_lanes)
            got = _probs(L, L.qghip_qubit_processor_calc_probability_noalloc, proc, qs, n_lanes)
            print('noalloc prec={} n={}: max |diff| = {:.3g}'.format(prec, n_lanes,
                                                                     float(np.max(np.abs(got - ref)))),
                  flush=True)
            assert np.array_equal(got, ref)
            assert np.all((ref > -1e-12) & (ref < 1 + 1e-12))
            _check(L.qghip_qubit_states_delete(qs))
    finally:
        _check(L.qghip_qubit_processor_delete(proc))


@pytest.mark.parametrize('prec', [FP64, FP32])
def test_noalloc_after_recorded_collapse(prec):
    """decohere records a collapse; the variant applies it, then sums (FP64 rounding apart)."""
    L = _lib()
    _check(L.qghip_devices_initialize(0))
    proc = _new(L.qghip_qubit_processor_new, prec)
    try:
        for n_lanes, seed in ((20, 21), (12, 22)):
            res = []
            for fn in (L.qghip_qubit_processor_calc_probability,
                       L.qghip_qubit_processor_calc_probability_noalloc):
                qs = _prepare(L, (proc, prec), n_lanes, seed)
                p0 = ctypes.c_double()
                _check(L.qghip_qubit_processor_calc_probability(proc, qs, 3, ctypes.byref(p0)))
                _check(L.qghip_qubit_processor_decohere(proc, 1, p0.value, qs, 3))
                row = []
                for lane in (0, 3, n_lanes - 1):
                    v = ctypes.c_double()
                    _check(fn(proc, qs, lane, ctypes.byref(v)))
                    row.append(v.value)
                res.append(row)
                _check(L.qghip_qubit_states_delete(qs))
            ref, got = np.array(res[0]), np.array(res[1])
            assert got[1] == 0.0 and ref[1] == 0.0        # lane 3 collapsed to |1>
            assert np.max(np.abs(got - ref)) < 1e-13
    finally:
        _check(L.qghip_qubit_processor_delete(proc))


def test_noalloc_refuses_queued_gates():
    L = _lib()
    _check(L.qghip_devices_initialize(0))
    proc = _new(L.qghip_qubit_processor_new, FP32)
    try:
        qs = _prepare(L, (proc, FP32), 12, 31)
        _check(L.qghip_qubit_processor_set_fusion(proc, 1, 0, -1))     # tile fusion: gates queue
        h = ctypes.c_int()
        _check(L.qghip_register_matrix_factory(b'H', ctypes.byref(h)))
        _check(L.qghip_qubit_processor_apply_gate(proc, h.value, _doubles([]), 0, 0, qs, 5))
        v = ctypes.c_double()
        assert L.qghip_qubit_processor_calc_probability_noalloc(proc, qs, 5, ctypes.byref(v)) == EVALUE
        assert b'synchronize' in L.qghip_last_error()
        _check(L.qghip_qubit_processor_synchronize(proc))
        ref = ctypes.c_double()
        _check(L.qghip_qubit_processor_calc_probability_noalloc(proc, qs, 5, ctypes.byref(v)))
        _check(L.qghip_qubit_processor_calc_probability(proc, qs, 5, ctypes.byref(ref)))
        assert v.value == ref.value
        _check(L.qghip_qubit_states_delete(qs))
    finally:
        _check(L.qghip_qubit_processor_delete(proc))
