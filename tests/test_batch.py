"""Batched adjoint evaluation (one captured graph for many circuits) against single circuits
and the CPU oracle.

Each item of a batch must give the same <H> and gradient as its own circuit run alone
(ParametricCircuit, itself checked against the oracle): within 1e-10 (FP64) / 2e-4 (FP32)
for gradients and 1e-11 / 2e-5 for values; differences come only from the device-side
sincos of the rotation angles and from tile grouping. Graph replay and direct launches agree
bit for bit; a replay after new parameters matches the single-circuit path again.
"""
import math

import numpy as np
import pytest

from qgate_amd.adjoint import PauliSum, hea_circuit
from tests import adjoint_checks as ac

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
GTOL = {np.float64: 1e-10, np.float32: 2e-4}
VTOL = {np.float64: 1e-11, np.float32: 2e-5}


def _hea_batch(dtype, sizes, layers, seed):
    from qgate_amd.batch import AdjointBatch
    rs = np.random.RandomState(seed)
    circuits = {int(n): hea_circuit(int(n), layers, dtype) for n in sorted(set(sizes))}
    batch = AdjointBatch(dtype)
    items = []
    for n in sizes:
        n = int(n)
        params = rs.uniform(-math.pi, math.pi, 2 * n * layers)
        h = PauliSum.zz_x(rs.normal(size=n - 1), rs.normal(size=n))
        batch.add(circuits[n], params, h)
        items.append((n, params, h))
    return batch, circuits, items


def _check(batch_out, circuits, items, dtype):
    values, grads = batch_out
    worst_v = worst_g = 0.
    for (n, params, h), v, g in zip(items, values, grads):
        rv, rg = circuits[n].gradient(params, h)
        worst_v = max(worst_v, abs(v - rv))
        worst_g = max(worst_g, float(np.max(np.abs(g - rg))))
    print('worst value err', worst_v, 'worst grad err', worst_g)
    assert worst_v < VTOL[dtype] and worst_g < GTOL[dtype], (worst_v, worst_g)


@pytest.mark.parametrize('dtype', DTYPES)
def test_batch_matches_single_circuits(hip_module, dtype):
    sizes = np.random.RandomState(3).randint(2, 15, 30)
    batch, circuits, items = _hea_batch(dtype, sizes, 3, seed=4)
    out = batch.run()
    _check(out, circuits, items, dtype)
    groups = batch.tuning
    assert sum(g['items'] for g in groups) == len(items)
    assert all(g['tile_bits'] >= 1 for g in groups)


@pytest.mark.parametrize('dtype', DTYPES)
def test_graph_replay_and_new_params(hip_module, dtype):
    sizes = [4, 9, 13, 9, 4, 16]
    batch, circuits, items = _hea_batch(dtype, sizes, 2, seed=8)
    v1, g1 = batch.run(use_graph=True)
    v2, g2 = batch.run(use_graph=False)
    assert np.array_equal(v1, v2)
    assert all(np.array_equal(a, b) for a, b in zip(g1, g2))
    rs = np.random.RandomState(9)
    new_items = []
    for i, (n, params, h) in enumerate(items):
        p = rs.uniform(-math.pi, math.pi, len(params))
        batch.set_params(i, p)
        new_items.append((n, p, h))
    _check(batch.run(), circuits, new_items, dtype)


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('tuning', [(5, 128, 1), (7, 256, 2), (9, 128, 8)])
def test_fixed_tunings_multi_pass(hip_module, dtype, tuning):
    """Small tiles force several passes per layer and tiles per circuit; a small persistent
    grid makes every workgroup stride over many tiles."""
    sizes = [6, 11, 14, 11]
    batch, circuits, items = _hea_batch(dtype, sizes, 2, seed=12)
    batch.set_tuning(*tuning)
    _check(batch.run(), circuits, items, dtype)
    for g in batch.tuning:
        assert g['tile_bits'] == min(tuning[0], 12)
        assert g['threads'] == tuning[1] and g['grid_per_cu'] == tuning[2]


@pytest.mark.parametrize('dtype', DTYPES)
def test_mixed_circuits_and_observables(hip_module, dtype):
    """Controlled rotations, fixed gates, shared parameters, X/Y/Z terms, several structures."""
    from qgate_amd.batch import AdjointBatch
    batch = AdjointBatch(dtype)
    cases = []
    for seed, n in ((1, 5), (2, 9), (3, 12), (4, 13)):
        spec = ac.mixed_spec(n, 50, seed=seed, n_params=12)
        circ = spec.build(dtype)
        h = ac.random_observable(n, 8, seed=seed + 20)
        for k in range(3):   # three parameter sets per structure
            params = np.random.RandomState(seed * 10 + k).uniform(-math.pi, math.pi, spec.n_params)
            batch.add(circ, params, h)
            cases.append((circ, params, h))
    values, grads = batch.run()
    for (circ, params, h), v, g in zip(cases, values, grads):
        rv, rg = circ.gradient(params, h)
        assert abs(v - rv) < VTOL[dtype] * 10
        assert np.max(np.abs(g - rg)) < GTOL[dtype]


# full-size tile bits of the forward and sweep passes (complex128, complex64)
FWD_BITS = {np.float64: 11, np.float32: 12}
SWEEP_BITS = {np.float64: 10, np.float32: 11}


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('registers', [1, 0])
def test_register_and_lds_passes(hip_module, dtype, registers):
    """Register-resident passes (forced on, for groups with full-size tiles) and LDS passes
    (forced) both give the single-circuit results. The circuits mix controlled rotations and
    fixed gates, so controls fall on register, lane/wave and outer bits of the tiles."""
    from qgate_amd.batch import AdjointBatch
    batch = AdjointBatch(dtype)
    cases = []
    for seed, n in ((5, 10), (6, 11), (7, 12), (8, 14), (9, 16)):
        spec = ac.mixed_spec(n, 120, seed=seed, n_params=20)
        circ = spec.build(dtype)
        h = ac.random_observable(n, 6, seed=seed + 30)
        for k in range(2):
            params = np.random.RandomState(seed * 7 + k).uniform(-math.pi, math.pi, spec.n_params)
            batch.add(circ, params, h)
            cases.append((circ, params, h))
    batch.set_tuning(registers=registers)
    values, grads = batch.run()
    for (circ, params, h), v, g in zip(cases, values, grads):
        rv, rg = circ.gradient(params, h)
        assert abs(v - rv) < VTOL[dtype] * 10
        assert np.max(np.abs(g - rg)) < GTOL[dtype]
    for g in batch.tuning:
        n = g['qubits']
        assert g['forward_in_registers'] == int(registers == 1 and n >= FWD_BITS[dtype]), g
        assert g['sweep_in_registers'] == int(registers == 1 and n >= SWEEP_BITS[dtype]), g


@pytest.mark.parametrize('dtype', DTYPES)
def test_pass_caps(hip_module, dtype):
    """Dense circuits on the low qubits would fit one tile: passes still end at the register
    program's gate limit (forward) and at the gradient-slot limit (sweep)."""
    from qgate_amd.batch import AdjointBatch
    from qgate_amd.adjoint import ParametricCircuit
    n = 12
    circ = ParametricCircuit(n, dtype)
    k = 0
    for layer in range(40):
        for q in range(5):
            circ.ry(q, k)
            circ.rz(q, k + 1)
            k += 2
        circ.cz(layer % 5, (layer + 1) % 5)
    batch = AdjointBatch(dtype)
    rs = np.random.RandomState(21)
    h = ac.random_observable(n, 5, seed=22)
    items = []
    for _ in range(3):
        params = rs.uniform(-math.pi, math.pi, k)
        batch.add(circ, params, h)
        items.append(params)
    batch.set_tuning(registers=1)
    values, grads = batch.run()
    for params, v, g in zip(items, values, grads):
        rv, rg = circ.gradient(params, h)
        assert abs(v - rv) < VTOL[dtype] * 10
        assert np.max(np.abs(g - rg)) < GTOL[dtype]
    (grp,) = batch.tuning
    assert grp['sweep_in_registers'] == 1
    assert grp['sweep_passes'] >= math.ceil(k / 96)
    assert grp['forward_passes'] >= math.ceil(len(range(40)) * 11 / (40 if dtype == np.float32 else 24))


def test_batch_against_oracle_parameter_shift(hip_module, port_module):
    from qgate_amd.batch import AdjointBatch
    batch = AdjointBatch(np.float64)
    specs = []
    for n in (3, 6):
        spec = ac.hea_spec(n, 2)
        circ = spec.build(np.float64)
        params = np.random.RandomState(n).uniform(-math.pi, math.pi, spec.n_params)
        h = PauliSum.zz_x(np.random.RandomState(n + 1).normal(size=n - 1), np.random.RandomState(n + 2).normal(size=n))
        batch.add(circ, params, h)
        specs.append((spec, params, h))
    values, grads = batch.run()
    for (spec, params, h), v, g in zip(specs, values, grads):
        assert abs(v - ac.oracle_expectation(port_module, spec, params, h)) < 1e-12
        assert np.max(np.abs(g - ac.parameter_shift_gradient(port_module, spec, params, h))) < 1e-10


def test_batch_argument_errors(hip_module):
    from qgate_amd.batch import AdjointBatch
    batch = AdjointBatch(np.float32)
    c64 = hea_circuit(4, 1, np.float64)
    with pytest.raises(ValueError):
        batch.add(c64, np.zeros(8), PauliSum.zz_x(np.ones(3), np.ones(4)))   # precision mismatch
    c = hea_circuit(4, 1, np.float32)
    with pytest.raises(ValueError):
        batch.add(c, np.zeros(3), PauliSum.zz_x(np.ones(3), np.ones(4)))      # too few parameters
    with pytest.raises(ValueError):
        batch.set_tuning(13, 0, 0)
    with pytest.raises(ValueError):
        batch.set_tuning(registers=2)
