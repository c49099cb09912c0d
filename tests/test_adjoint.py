"""Adjoint differentiation: <H> and d<H>/dparams on the GPU against the CPU oracle.

Bars: <H> and the final state within 1e-12 (FP64) / 5e-6 per amplitude (FP32) of the FP64
oracle; gradients within 1e-10 (FP64) / 1e-4 (FP32) of the oracle's parameter-shift
gradient (uncontrolled rotations) or within 1e-7 (FP64) of its central differences with
h = 1e-5 (controlled rotations, truncation error O(h^2)). The checkers themselves are
pinned by closed forms and by the dense Pauli matrix in the CPU tests.
"""
import math

import numpy as np
import pytest

from qgate_amd.adjoint import PauliSum
from qgate_amd.model import gate_type as gtype
from tests import adjoint_checks as ac

X = np.array([[0, 1], [1, 0]], np.complex128)
Y = np.array([[0, -1j], [1j, 0]], np.complex128)
Z = np.diag([1., -1.]).astype(np.complex128)
I2 = np.eye(2)


# ---- CPU: the checkers ------------------------------------------------------------------

def test_pauli_matrix_known_answers():
    assert np.allclose(PauliSum([(1., {0: 'Y'})]).matrix(1), Y)
    # qubit 1 is the high bit of the index: kron(op on qubit 1, op on qubit 0)
    assert np.allclose(PauliSum([(2., {1: 'X'})]).matrix(2), 2. * np.kron(X, I2))
    assert np.allclose(PauliSum([(1., {0: 'Z', 1: 'Y'})]).matrix(2), np.kron(Y, Z))
    h = PauliSum.zz_x([0.5], [1., -1.])
    ref = 0.5 * np.kron(Z, Z) + np.kron(I2, X) - np.kron(X, I2)
    assert np.allclose(h.matrix(2), ref)


def test_pauli_apply_matches_dense_matrix():
    rs = np.random.RandomState(3)
    n = 4
    obs = ac.random_observable(n, 12, seed=5)
    psi = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    assert np.allclose(ac.pauli_apply(obs, psi), obs.matrix(n) @ psi, atol=1e-13)


@pytest.mark.parametrize('kind,obs,value,deriv', [
    ('ry', {0: 'Z'}, math.cos, lambda t: -math.sin(t)),
    ('ry', {0: 'X'}, math.sin, math.cos),
    ('rx', {0: 'Z'}, math.cos, lambda t: -math.sin(t)),
    ('rx', {0: 'Y'}, lambda t: -math.sin(t), lambda t: -math.cos(t)),
])
def test_checkers_closed_forms(port_module, kind, obs, value, deriv):
    """One rotation on |0>: <Z> = cos t for RX/RY, <X> = sin t for RY, <Y> = -sin t for RX."""
    spec = ac.CircuitSpec(1)
    getattr(spec, kind)(0, 0)
    h = PauliSum([(1., obs)])
    for t in (0.3, 1.7, -2.2):
        assert abs(ac.oracle_expectation(port_module, spec, [t], h) - value(t)) < 1e-14
        g = ac.parameter_shift_gradient(port_module, spec, [t], h)
        assert abs(g[0] - deriv(t)) < 1e-14
        g = ac.central_difference_gradient(port_module, spec, [t], h)
        assert abs(g[0] - deriv(t)) < 1e-9


def test_checkers_agree_on_hea(port_module):
    spec = ac.hea_spec(4, 2)
    a, b = np.random.RandomState(1).normal(size=3), np.random.RandomState(2).normal(size=4)
    h = PauliSum.zz_x(a, b)
    params = np.random.RandomState(4).uniform(-math.pi, math.pi, spec.n_params)
    ps = ac.parameter_shift_gradient(port_module, spec, params, h)
    cd = ac.central_difference_gradient(port_module, spec, params, h)
    assert np.max(np.abs(ps - cd)) < 1e-8
    # a rotation left out of the shift (a dropped term) is caught by the comparison
    assert np.max(np.abs(ps)) > 1e-3


# ---- GPU ---------------------------------------------------------------------------------

DTYPES = [np.float64, np.float32]
GRAD_TOL = {np.float64: 1e-10, np.float32: 1e-4}
VALUE_TOL = {np.float64: 1e-12, np.float32: 5e-5}


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', DTYPES)
def test_hea_expectation_and_state(hip_module, port_module, dtype):
    n, layers = 10, 3
    spec = ac.hea_spec(n, layers)
    a, b = np.random.RandomState(8).normal(size=n - 1), np.random.RandomState(9).normal(size=n)
    h = PauliSum.zz_x(a, b)
    params = np.random.RandomState(10).uniform(-math.pi, math.pi, spec.n_params)
    circ = spec.build(dtype)
    value = circ.expectation(params, h)
    ref_state = ac.oracle_state(port_module, spec, params)
    ref = ac.expectation(h, ref_state)
    print('value', value, 'ref', ref)
    assert abs(value - ref) < VALUE_TOL[dtype] * max(1., abs(ref))
    err = np.max(np.abs(circ.state() - ref_state))
    assert err < (1e-12 if dtype == np.float64 else 5e-6), err


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', DTYPES)
def test_hea_gradient_vs_parameter_shift(hip_module, port_module, dtype):
    n, layers = 6, 2
    spec = ac.hea_spec(n, layers)
    a, b = np.random.RandomState(18).normal(size=n - 1), np.random.RandomState(19).normal(size=n)
    h = PauliSum.zz_x(a, b)
    params = np.random.RandomState(20).uniform(-math.pi, math.pi, spec.n_params)
    circ = spec.build(dtype)
    value, grad = circ.gradient(params, h)
    ref = ac.parameter_shift_gradient(port_module, spec, params, h)
    err = np.max(np.abs(grad - ref))
    print('grad err', err)
    assert err < GRAD_TOL[dtype], err
    assert abs(value - ac.oracle_expectation(port_module, spec, params, h)) < VALUE_TOL[dtype] * 10


@pytest.mark.gpu
@pytest.mark.parametrize('seed,n', [(0, 7), (1, 7), (2, 7), (3, 13), (4, 14)])
def test_mixed_gradient_vs_central_differences(hip_module, port_module, seed, n):
    """Controlled rotations, fixed gates (adjoints too), shared parameters, X/Y/Z terms; at 13
    and 14 qubits the tiles (2^11 / 2^10 amplitudes) leave some targets and controls outside."""
    spec = ac.mixed_spec(n, 60, seed=seed, n_params=15)
    h = ac.random_observable(n, 10, seed=seed + 50)
    params = np.random.RandomState(seed + 70).uniform(-math.pi, math.pi, spec.n_params)
    circ = spec.build(np.float64)
    value, grad = circ.gradient(params, h)
    ref_value = ac.oracle_expectation(port_module, spec, params, h)
    assert abs(value - ref_value) < 1e-12
    ref = ac.central_difference_gradient(port_module, spec, params, h)
    err = np.max(np.abs(grad - ref))
    print('seed', seed, 'grad err', err)
    assert err < 1e-7, err


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('n', [5, 12, 15])
@pytest.mark.parametrize('registers', [True, False])
def test_tile_passes_match_per_gate_launches(hip_module, dtype, n, registers):
    """The tile-pass forward and backward sweep (full-size tiles in registers or in LDS)
    against one launch per gate and per step."""
    spec = ac.mixed_spec(n, 120, seed=n, n_params=30)
    h = ac.random_observable(n, 12, seed=n + 1)
    params = np.random.RandomState(n + 2).uniform(-math.pi, math.pi, spec.n_params)
    out = {}
    hip_module.set_register_tiles(registers)
    try:
        for mode in ('off', 'tile'):
            circ = spec.build(dtype).set_fusion(mode)
            value, grad = circ.gradient(params, h)
            circ.expectation(params, h)
            out[mode] = (value, grad, circ.state())
    finally:
        hip_module.set_register_tiles(True)
    tol = 1e-12 if dtype == np.float64 else 2e-4
    assert abs(out['off'][0] - out['tile'][0]) < tol
    assert np.max(np.abs(out['off'][1] - out['tile'][1])) < tol
    assert np.max(np.abs(out['off'][2] - out['tile'][2])) < (1e-12 if dtype == np.float64 else 5e-6)


@pytest.mark.gpu
def test_backward_after_expectation_and_repeat(hip_module):
    """backward reuses the forward state of the last expectation call; calling it again (or
    with other parameters) recomputes the forward pass and gives the same gradient."""
    spec = ac.mixed_spec(8, 40, seed=4)
    h = ac.random_observable(8, 6, seed=44)
    rs = np.random.RandomState(5)
    p1, p2 = rs.uniform(-3, 3, spec.n_params), rs.uniform(-3, 3, spec.n_params)
    circ = spec.build(np.float64)
    v1, g1 = circ.gradient(p1, h)
    assert circ.expectation(p1, h) == v1
    assert np.array_equal(circ.backward(p1), g1)
    assert np.array_equal(circ.backward(p1), g1)      # forward recomputed
    _, g2 = circ.gradient(p2, h)
    circ.expectation(p1, h)
    assert np.array_equal(circ.backward(p2), g2)      # params differ: recomputed for p2


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', DTYPES)
def test_hea24_gradient_by_shift_on_device(hip_module, dtype):
    """The 24-qubit, 12-layer ansatz (576 parameters): 16 seeded entries of the adjoint
    gradient against parameter-shift differences of device expectations (exact for
    uncontrolled rotations, so only rounding separates the two)."""
    n, layers = 24, 12
    spec = ac.hea_spec(n, layers)
    a, b = np.random.RandomState(1004).normal(size=n - 1), np.random.RandomState(1005).normal(size=n)
    h = PauliSum.zz_x(a, b)
    params = np.random.RandomState(4).uniform(-math.pi, math.pi, spec.n_params)
    circ = spec.build(dtype)
    _, grad = circ.gradient(params, h)
    tol = 1e-10 if dtype == np.float64 else 2e-3
    subset = np.random.RandomState(44).choice(spec.n_params, 16, replace=False)
    errs = []
    for k in subset:
        pp, pm = params.copy(), params.copy()
        pp[k] += math.pi / 2
        pm[k] -= math.pi / 2
        ref = (circ.expectation(pp, h) - circ.expectation(pm, h)) / 2.
        errs.append(abs(grad[k] - ref))
    print('max err', max(errs))
    assert max(errs) < tol, errs


@pytest.mark.gpu
def test_gradient_rows_of_many_passes(hip_module):
    """26 qubits (complex128: 2^16 tiles, 512 KiB of partial sums per rotation) with 1248
    rotations: the per-tile partials outgrow their 256 MiB buffer and are summed in several
    launches during the sweep. Seeded gradient entries against parameter-shift differences,
    and every entry against the batched path, which sums the rows of each pass by itself."""
    from qgate_amd.batch import AdjointBatch
    n, layers = 26, 24
    spec = ac.hea_spec(n, layers)
    h = PauliSum.zz_x(np.random.RandomState(2601).normal(size=n - 1), np.random.RandomState(2602).normal(size=n))
    params = np.random.RandomState(26).uniform(-math.pi, math.pi, spec.n_params)
    circ = spec.build(np.float64)
    value, grad = circ.gradient(params, h)
    for k in np.random.RandomState(27).choice(spec.n_params, 6, replace=False):
        pp, pm = params.copy(), params.copy()
        pp[k] += math.pi / 2
        pm[k] -= math.pi / 2
        ref = (circ.expectation(pp, h) - circ.expectation(pm, h)) / 2.
        assert abs(grad[k] - ref) < 1e-10, (k, grad[k], ref)
    batch = AdjointBatch(np.float64)
    batch.add(circ, params, h)
    batch.set_tuning(11, 256, 4, 1)
    values, grads = batch.run()
    batch.delete()
    assert abs(values[0] - value) < 1e-10
    assert np.max(np.abs(grads[0] - grad)) < 1e-10


@pytest.mark.gpu
def test_torch_autograd(hip_module):
    import torch
    from qgate_amd.adjoint import Expectation
    spec = ac.hea_spec(6, 2)
    h = PauliSum.zz_x(np.ones(5), np.full(6, 0.5))
    circ = spec.build(np.float64)
    params = np.random.RandomState(3).uniform(-1, 1, spec.n_params)
    t = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    loss = 3. * Expectation.apply(t, circ, h)
    loss.backward()
    value, grad = circ.gradient(params, h)
    assert abs(loss.item() - 3. * value) < 1e-13
    assert np.allclose(t.grad.numpy(), 3. * grad, atol=1e-13)


@pytest.mark.gpu
def test_torch_autograd_two_observables(hip_module, port_module):
    """A loss that sums two observables of one circuit: each backward uses its own
    forward's observable, checked against the oracle's parameter-shift gradients."""
    import torch
    from qgate_amd.adjoint import Expectation
    spec = ac.hea_spec(5, 2)
    h1 = PauliSum.zz_x(np.linspace(-1., 1., 4), np.zeros(5))
    h2 = PauliSum([(0.7, {0: 'X', 2: 'Y'}), (-1.3, {4: 'Z'}), (0.4, {1: 'Y'})])
    circ = spec.build(np.float64)
    params = np.random.RandomState(4).uniform(-1, 1, spec.n_params)
    t = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    loss = 2. * Expectation.apply(t, circ, h1) - 0.5 * Expectation.apply(t, circ, h2)
    loss.backward()
    g1 = ac.parameter_shift_gradient(port_module, spec, params, h1)
    g2 = ac.parameter_shift_gradient(port_module, spec, params, h2)
    err = float(np.max(np.abs(t.grad.numpy() - (2. * g1 - 0.5 * g2))))
    assert err < 1e-11, err
    # an expectation call between forward and backward does not change the gradient
    t.grad = None
    y = Expectation.apply(t, circ, h1)
    circ.expectation(params + 0.3, h2)
    y.backward()
    assert float(np.max(np.abs(t.grad.numpy() - g1))) < 1e-11


@pytest.mark.gpu
def test_argument_errors(hip_module):
    from qgate_amd.adjoint import ParametricCircuit
    circ = ParametricCircuit(4, np.float64)
    with pytest.raises(ValueError):
        circ.ry(4, 0)                          # target out of range
    with pytest.raises(ValueError):
        circ.gate(gtype.X(), 1, controls=[1])  # control == target
    circ.ry(0, 2)
    with pytest.raises(ValueError):
        circ.backward([0., 0., 0.])            # no observable yet
    with pytest.raises(ValueError):
        circ.expectation([0., 0.], PauliSum([(1., {0: 'Z'})]))   # too few parameters
    with pytest.raises(ValueError):
        circ.expectation([0., 0., 0.], PauliSum([(1., {5: 'Z'})]))  # term outside register
    assert abs(circ.expectation([0., 0., math.pi / 3], PauliSum([(1., {0: 'Z'})])) - 0.5) < 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('n', [1, 2, 3])
def test_small_registers(hip_module, port_module, dtype, n):
    """Registers smaller than a tile row: X, Y and Z terms on every qubit."""
    spec = ac.CircuitSpec(n)
    k = 0
    for layer in range(3):
        for q in range(n):
            spec.rx(q, k)
            spec.ry(q, k + 1)
            spec.rz(q, k + 2)
            k += 3
        for q in range(n - 1):
            spec.cz(q, q + 1)
    terms = []
    for q in range(n):
        terms += [(0.7, {q: 'X'}), (-0.4, {q: 'Y'}), (0.3, {q: 'Z'})]
    if n >= 2:
        terms.append((0.9, {0: 'Y', n - 1: 'X'}))
    h = PauliSum(terms)
    params = np.random.RandomState(n).uniform(-math.pi, math.pi, spec.n_params)
    value, grad = spec.build(dtype).gradient(params, h)
    ref = ac.parameter_shift_gradient(port_module, spec, params, h)
    assert abs(value - ac.oracle_expectation(port_module, spec, params, h)) < VALUE_TOL[dtype] * 10
    assert np.max(np.abs(grad - ref)) < GRAD_TOL[dtype]


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('registers', [True, False])
def test_rotations_at_special_angles(hip_module, port_module, dtype, registers):
    """RX, RY, RZ at angles where the matrix takes a special form (0: the identity, a real
    matrix for RX too; +-pi, pi / 2) and at random angles, on a 12-qubit register (register
    sweep tiles of 2^11 / 2^10 amplitudes, and LDS tiles): every rotation keeps its
    generator's gradient term whatever form its matrix has."""
    n = 12
    rs = np.random.RandomState(21)
    spec = ac.CircuitSpec(n)
    special = [0., math.pi, -math.pi, math.pi / 2]
    params = []
    for layer in range(2):
        for q in range(n):
            for rot in (spec.rx, spec.ry, spec.rz):
                rot(q, len(params))
                params.append(special[rs.randint(4)] if rs.randint(3) else rs.uniform(-math.pi, math.pi))
        for q in range(n):
            spec.cz(q, (q + 1) % n)
    params = np.array(params)
    params[:3] = 0.     # RX, RY, RZ of qubit 0 at exactly 0
    terms = [(0.8, {q: 'X'}) for q in range(0, n, 2)] + [(-0.5, {q: 'Y'}) for q in range(1, n, 3)]
    terms += [(0.6, {q: 'Z', q + 1: 'Z'}) for q in range(n - 1)]
    h = PauliSum(terms)
    hip_module.set_register_tiles(registers)
    try:
        value, grad = spec.build(dtype).set_fusion('tile').gradient(params, h)
    finally:
        hip_module.set_register_tiles(True)
    ref = ac.parameter_shift_gradient(port_module, spec, params, h)
    err = np.max(np.abs(grad - ref))
    print(dtype.__name__, registers, 'grad err', err, 'at', int(np.argmax(np.abs(grad - ref))))
    assert abs(value - ac.oracle_expectation(port_module, spec, params, h)) < VALUE_TOL[dtype] * 10
    assert err < GRAD_TOL[dtype], err
