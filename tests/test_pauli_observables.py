"""Pauli-string observables (prob(ref, paulis), measure(ref, paulis)) evaluated in one pass.

The reference lowers them into a basis change, a CX chain, Prob / Measure of the last qubit
and the inverse gates (qgate/model/expand.py:53-77, pauli_gates_diagonalizer.py:96-135); the
HIP runtime keeps them whole (calc_pauli_probability, pauli_decohere). Checks:
  * the per-qubit reduction of the string against the reference's own test case
    (tests/test_pauli_macro_expansion.py: X.Y on one qubit measures Z there);
  * the oracle's expansion against the closed form P(+1) = (|psi|^2 + Re<psi|P psi>) / 2,
    with (P psi)_y = i^#Y (-1)^|(y ^ x) & z| psi_(y ^ x), on random states (CPU);
  * the fused GPU path against the expansion on the oracle and against the closed form, for
    X/Y/Z strings with identities, lane 0, diagonal strings and strings across state vectors;
  * PauliMeasure: the same outcome (same random draw) and the same collapsed state as the
    expansion, and the unfused HIP path agreeing with the fused one.
"""
import math

import numpy as np
import pytest

from qgate_amd import model
from qgate_amd.model.expand import pauli_string
from qgate_amd.script import *
from qgate_amd.simulator.simulator import Simulator
from tests.helpers import max_abs, tol

LETTER = {'X': X, 'Y': Y, 'Z': Z, 'I': I}


def _random_ops(q, n_gates, seed):
    rs = np.random.RandomState(seed)
    ops = [H(x) for x in q]
    for _ in range(n_gates):
        k = rs.randint(3)
        if k == 0:
            ops.append(U3(*rs.uniform(0., 2. * math.pi, 3))(q[rs.randint(len(q))]))
        elif k == 1:
            c, t = rs.choice(len(q), 2, replace=False)
            ops.append(ctrl(q[c]).X(q[t]))
        else:
            ops.append(Rz(rs.uniform(0., 2. * math.pi))(q[rs.randint(len(q))]))
    return ops


def closed_form_prob(psi, x, z):
    """P(Pauli string = +1) from the amplitudes (external index: bit k = ordering[k])."""
    idx = np.arange(psi.size)
    partner = idx ^ x
    sign = np.where(np.array([bin(int(v)).count('1') & 1 for v in (partner & z)]) == 1, -1., 1.)
    c = 1j ** bin(x & z).count('1')
    p_psi = c * sign * psi[partner]
    return 0.5 * (np.vdot(psi, psi).real + np.vdot(psi, p_psi).real)


def _masks(string, q):
    x = z = 0
    for k, letter in string.items():
        b = q.index(k)
        if letter in 'XY':
            x |= 1 << b
        if letter in 'YZ':
            z |= 1 << b
    return x, z


def _string_ops(string):
    return [LETTER[letter](qr) for qr, letter in string.items()]


STRINGS = [   # qubit index -> letter, over 6 qubits
    {0: 'X'}, {0: 'Y'}, {0: 'Z'}, {5: 'X'}, {2: 'Y', 3: 'Y'},
    {0: 'Z', 3: 'Z', 5: 'Z'}, {0: 'X', 1: 'Y', 4: 'Z'}, {1: 'X', 2: 'I', 5: 'Y'},
    {0: 'Y', 1: 'Y', 2: 'Y', 3: 'Y', 4: 'Y', 5: 'Y'}, {0: 'X', 1: 'X', 2: 'X', 3: 'X', 4: 'X', 5: 'X'},
]


def test_reduction_matches_reference_case():
    q = new_qregs(3)
    terms, op_qreg = pauli_string([X(q[2]), Y(q[2]), Y(q[1]), Z(q[0])])
    assert [(t.id, letter) for t, letter in terms] == [(q[0].id, 'Z'), (q[1].id, 'Y'), (q[2].id, 'Z')]
    assert op_qreg is q[2]
    with pytest.raises(RuntimeError):
        pauli_string([X(q[0]), X(q[0]), I(q[1])])


def test_fused_preprocessing_keeps_the_op_and_joins():
    q = new_qregs(3)
    r = new_reference()
    pre = model.Preprocessor(fuse_pauli=True, circuit_prep='dynamic')
    gl = model.GateList()
    gl.set([H(q[0]), prob(r, [X(q[0]), Z(q[2])])])
    ops = pre.preprocess(gl).ops
    kinds = [type(op).__name__ for op in ops]
    assert 'PauliProb' in kinds and 'Join' in kinds
    assert not any(isinstance(op, model.Gate) and op.ctrllist for op in ops)   # no CX chain


@pytest.mark.parametrize('seed', [1, 2])
def test_oracle_expansion_matches_closed_form(port_module, seed):
    n = 6
    for string in STRINGS:
        q = new_qregs(n)
        r = new_reference()
        s = {q[k]: v for k, v in string.items()}
        sim = Simulator(port_module, circuit_prep='one_static', dtype=np.float64)
        sim.qubits.set_ordering(list(q))
        sim.run(_random_ops(q, 40, seed))
        psi = sim.qubits.states[:]
        sim.run([prob(r, _string_ops(s))])
        x, z = _masks(s, list(q))
        assert abs(sim.values.get(r) - closed_form_prob(psi, x, z)) < 1e-13, string


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('prep', ['one_static', 'static', 'dynamic'])
def test_fused_prob_matches_expansion(hip_module, port_module, dtype, prep):
    n = 6
    for si, string in enumerate(STRINGS):
        q = new_qregs(n)
        r = new_reference()
        s = {q[k]: v for k, v in string.items()}
        ops = _random_ops(q, 40, 10 + si) + [prob(r, _string_ops(s))]
        vals, psis = {}, {}
        for name, module in (('hip', hip_module), ('oracle', port_module)):
            sim = Simulator(module, circuit_prep=prep, dtype=dtype)
            sim.qubits.set_ordering(list(q))
            sim.run(ops)
            vals[name] = sim.values.get(r)
            psis[name] = sim.qubits.states[:]
        x, z = _masks(s, list(q))
        ref = closed_form_prob(psis['oracle'].astype(np.complex128), x, z)
        assert abs(vals['hip'] - vals['oracle']) < tol(dtype), (string, vals)
        assert abs(vals['hip'] - ref) < tol(dtype), (string, vals['hip'], ref)
        assert max_abs(psis['hip'], psis['oracle']) < tol(dtype)   # prob leaves the state alone


@pytest.mark.gpu
def test_string_across_state_vectors(hip_module, port_module):
    """Qubits in separate groups (static mode) are joined for the observable."""
    q = new_qregs(4)
    r = new_reference()
    ops = [H(q[0]), Ry(0.7)(q[1]), ctrl(q[2]).X(q[3]), Rx(1.1)(q[2]), prob(r, [X(q[0]), Y(q[1]), Z(q[3])])]
    vals = []
    for module in (hip_module, port_module):
        sim = Simulator(module, circuit_prep='static', dtype=np.float64)
        sim.qubits.set_ordering(list(q))
        sim.run(ops)
        vals.append(sim.values.get(r))
    assert abs(vals[0] - vals[1]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('prep', ['one_static', 'static'])
def test_fused_measure_matches_expansion(hip_module, port_module, dtype, prep):
    n = 6
    for si, string in enumerate(STRINGS):
        q = new_qregs(n)
        refs = new_references(2)
        s = {q[k]: v for k, v in string.items()}
        ops = _random_ops(q, 40, 30 + si) + [measure(refs[0], _string_ops(s)), H(q[1]),
                                             measure(refs[1], _string_ops(s))]
        out = {}
        for name, module, prefs in (('hip', hip_module, {}), ('hip unfused', hip_module, {'fuse_pauli': False}),
                                    ('oracle', port_module, {})):
            np.random.seed(1000 + si)
            sim = Simulator(module, circuit_prep=prep, dtype=dtype, **prefs)
            sim.qubits.set_ordering(list(q))
            sim.run(ops)
            out[name] = ([sim.values.get(x) for x in refs], sim.qubits.states[:])
        for name in ('hip', 'hip unfused'):
            assert out[name][0] == out['oracle'][0], (string, name, out[name][0], out['oracle'][0])
            assert max_abs(out[name][1], out['oracle'][1]) < tol(dtype) * 4, (string, name)
