"""Pins of the oracle's functions that the golden vectors leave open.

* Gate matrices (`qgo_gate_matrix`, and the HIP library's `qghip_gate_matrix`, a host
  function callable without a GPU): the known answers of the reference's own tests
  (tests/test_gate_matrix.py:27-211 and qgate/simulator/src/GateMatrixTest.cpp:33-324,
  restated), identities that the matrix definitions fix at random angles (U as an RZ RY RZ
  product, U2 = U(pi/2, .), S = T^2, SH = S H, ...), unitarity and the adjoint, and the
  port against the reference core bit for bit for every gate type, plain and adjoint.
* Threaded spans (`distribute`, Parallel.cpp:21-67): at 18 qubits every loop exceeds the
  2^16 single-worker threshold, so states, probabilities, probability arrays and pool
  draws depend on how spans split the index range; the port reproduces the reference core
  bit for bit there, with the default worker count and with an uneven one (3 workers, in a
  child process since the count is read once per process).
* Empty lanes in get_states: index bits of qregs without a state vector read 0. The
  reference CPU getter evaluates that test on the first 256-entry block's index only
  (CPUQubitsStatesGetter.cpp:68-78: `extSrcIdx` is not advanced), so empty-lane bits >= 8
  are ignored there; the reference's py and CUDA getters (pyruntime.py:228,
  DeviceGetStates.cu:188) and this build apply the test to every index.
"""
import cmath
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from qgate_amd.model import gate_type as gt
from qgate_amd.script import new_qregs, H, ctrl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
R = math.sqrt(0.5)
PI = math.pi
j = 1j


def _m(a00, a01, a10, a11, coef=1.):
    return coef * np.array([[a00, a01], [a10, a11]], np.complex128)


# (gate type name, args, adjoint, expected) from tests/test_gate_matrix.py:27-211 and
# GateMatrixTest.cpp:135-324 (the fixed gates and the RX/RY/RZ/ExpiI/ExpiZ tables)
KNOWN = [
    ('ID', (), False, _m(1, 0, 0, 1)),
    ('X', (), False, _m(0, 1, 1, 0)),
    ('Y', (), False, _m(0, -j, j, 0)),
    ('Z', (), False, _m(1, 0, 0, -1)),
    ('H', (), False, _m(1, 1, 1, -1, R)),
    ('S', (), False, _m(1, 0, 0, j)),
    ('S', (), True, _m(1, 0, 0, -j)),
    ('T', (), False, _m(1, 0, 0, cmath.exp(1j * PI / 4.))),
    ('T', (), True, _m(1, 0, 0, cmath.exp(-1j * PI / 4.))),
    ('RX', (0.,), False, _m(1, 0, 0, 1)),
    ('RX', (2. * PI,), False, _m(-1, 0, 0, -1)),
    ('RX', (PI,), False, _m(0, -j, -j, 0)),
    ('RX', (-PI,), False, _m(0, j, j, 0)),
    ('RX', (PI / 2.,), False, _m(1, -j, -j, 1, R)),
    ('RX', (-PI / 2.,), False, _m(1, j, j, 1, R)),
    ('RY', (0.,), False, _m(1, 0, 0, 1)),
    ('RY', (2. * PI,), False, _m(-1, 0, 0, -1)),
    ('RY', (PI,), False, _m(0, -1, 1, 0)),
    ('RY', (-PI,), False, _m(0, 1, -1, 0)),
    ('RY', (PI / 2.,), False, _m(1, -1, 1, 1, R)),
    ('RY', (-PI / 2.,), False, _m(1, 1, -1, 1, R)),
    ('RZ', (0.,), False, _m(1, 0, 0, 1)),
    ('RZ', (2. * PI,), False, _m(-1, 0, 0, -1)),
    ('RZ', (PI,), False, _m(-j, 0, 0, j)),
    ('RZ', (-PI,), False, _m(j, 0, 0, -j)),
    ('RZ', (PI / 2.,), False, _m(1 - j, 0, 0, 1 + j, R)),
    ('RZ', (-PI / 2.,), False, _m(1 + j, 0, 0, 1 - j, R)),
    ('ExpiI', (0.,), False, _m(1, 0, 0, 1)),
    ('ExpiI', (PI,), False, _m(-1, 0, 0, -1)),
    ('ExpiI', (PI / 2.,), False, _m(j, 0, 0, j)),
    ('ExpiI', (-PI / 2.,), False, _m(-j, 0, 0, -j)),
    ('ExpiI', (PI / 4.,), False, _m(1, 0, 0, 1, R * (1 + j))),
    ('ExpiI', (-PI / 4.,), False, _m(1, 0, 0, 1, R * (1 - j))),
    ('ExpiZ', (0.,), False, _m(1, 0, 0, 1)),
    ('ExpiZ', (PI,), False, _m(-1, 0, 0, -1)),
    ('ExpiZ', (PI / 2.,), False, _m(j, 0, 0, -j)),
    ('ExpiZ', (-PI / 2.,), False, _m(-j, 0, 0, j)),
    ('ExpiZ', (PI / 4.,), False, _m(1 + j, 0, 0, 1 - j, R)),
    ('ExpiZ', (-PI / 4.,), False, _m(1 - j, 0, 0, 1 + j, R)),
    # GateMatrixTest.cpp:33-120: U and U1 against the Pauli gates and H (up to the factor i of
    # qgate's U phase convention), U1(0) = I
    ('U', (0., 0., 0.), False, _m(1, 0, 0, 1)),
    ('U1', (0.,), False, _m(1, 0, 0, 1)),
    ('U', (PI, 0., PI), False, _m(0, 1, 1, 0, -j)),
    ('U', (PI, PI / 2., PI / 2.), False, _m(0, -j, j, 0, -j)),
    ('U', (0., 0., PI), False, _m(1, 0, 0, -1, -j)),
    ('U1', (PI,), False, _m(1, 0, 0, -1)),
    ('U', (PI / 2., 0., PI), False, _m(R, R, R, -R, -j)),
    ('U2', (0., PI), False, _m(R, R, R, -R, -j)),
]

# GateMatrixTest.cpp:40-93: U2(phi, lambda) = U(pi/2, phi, lambda) at these points
U2_POINTS = [(0., 0.), (PI, 0.), (-PI, 0.), (PI / 2., 0.), (-PI / 2., 0.), (PI / 4., 0.),
             (-PI / 4., 0.), (0., PI / 2.), (0., -PI / 2.), (0., PI / 4.), (0., -PI / 4.)]

N_ARGS = {'U': 3, 'U2': 2, 'U1': 1, 'RX': 1, 'RY': 1, 'RZ': 1, 'ExpiI': 1, 'ExpiZ': 1}
ALL_TYPES = ['U', 'U2', 'U1', 'ID', 'X', 'Y', 'Z', 'H', 'S', 'T', 'RX', 'RY', 'RZ', 'ExpiI', 'ExpiZ', 'SH']


def _gate(name, args):
    return getattr(gt, name)(*args)


def _runtimes(oracle):
    """name -> function(gate type name, args, adjoint) -> FP64 2x2 matrix."""
    from qgate_amd.simulator import hipruntime
    out = {}
    port = oracle.port().runtime
    out['port'] = lambda n, a, adj: port.gate_matrix(_gate(n, a), adj)
    if oracle.have_reference():
        ref = oracle.reference().runtime
        out['reference'] = lambda n, a, adj: ref.gate_matrix(_gate(n, a), adj)
    out['hip'] = lambda n, a, adj: hipruntime.gate_matrix(_gate(n, a), adj)
    return out


@pytest.fixture(scope='module')
def matrix_fns(oracle):
    from qgate_amd import build
    if not os.path.exists(build.lib_path()):
        build.build()
    return _runtimes(oracle)


def _close(a, b, tol=1e-14):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) < tol


@pytest.mark.parametrize('case', range(len(KNOWN)))
def test_gate_matrix_known_answers(matrix_fns, case):
    name, args, adj, expected = KNOWN[case]
    for rt, fn in matrix_fns.items():
        assert _close(fn(name, args, adj), expected), (rt, name, args, adj)


def test_u2_is_u_at_half_pi(matrix_fns):
    for rt, fn in matrix_fns.items():
        for phi, lam in U2_POINTS:
            assert _close(fn('U2', (phi, lam), False), fn('U', (PI / 2., phi, lam), False)), (rt, phi, lam)


def _angles(rs, n):
    return tuple(float(v) for v in rs.uniform(-2. * PI, 2. * PI, n))


def test_gate_matrix_identities(matrix_fns):
    """Relations the matrix definitions fix at any angle; a dropped sign, a swapped
    argument or a transposed entry in one definition breaks at least one of them."""
    rs = np.random.RandomState(11)
    for rt, f in matrix_fns.items():
        m = lambda n, *a: f(n, a, False)   # noqa: E731
        for _ in range(16):
            t, p, l = _angles(rs, 3)
            # U(t, p, l) = RZ(p) RY(t) RZ(l): qgate's U convention
            assert _close(m('U', t, p, l), m('RZ', p) @ m('RY', t) @ m('RZ', l)), rt
            assert _close(m('U2', p, l), m('U', PI / 2., p, l)), rt
            # U1(l) = e^{il/2} RZ(l); ExpiZ(t) = RZ(-2t); ExpiI(t) = e^{it} I
            assert _close(m('U1', l), cmath.exp(0.5j * l) * m('RZ', l)), rt
            assert _close(m('ExpiZ', t), m('RZ', -2. * t)), rt
            assert _close(m('ExpiI', t), cmath.exp(1j * t) * np.eye(2)), rt
            # RX = H RZ H, RY = S RX S^dagger
            assert _close(m('RX', t), m('H') @ m('RZ', t) @ m('H')), rt
            assert _close(m('RY', t), m('S') @ m('RX', t) @ f('S', (), True)), rt
            # rotations compose additively
            assert _close(m('RY', t) @ m('RY', p), m('RY', t + p)), rt
        # fixed gates: T^2 = S, S^2 = Z, XY = iZ, HZH = X, SH = S H, (SH)^dagger Y (SH) = Z
        assert _close(m('T') @ m('T'), m('S')), rt
        assert _close(m('S') @ m('S'), m('Z')), rt
        assert _close(m('X') @ m('Y'), 1j * m('Z')), rt
        assert _close(m('H') @ m('Z') @ m('H'), m('X')), rt
        assert _close(m('SH'), m('S') @ m('H')), rt
        assert _close(f('SH', (), True) @ m('Y') @ m('SH'), m('Z')), rt
        assert _close(m('ID'), np.eye(2)), rt


def test_gate_matrix_adjoint_and_unitarity(matrix_fns):
    rs = np.random.RandomState(12)
    for rt, f in matrix_fns.items():
        for name in ALL_TYPES:
            args = _angles(rs, N_ARGS.get(name, 0))
            a, b = f(name, args, False), f(name, args, True)
            assert np.array_equal(b, np.conjugate(a.T)), (rt, name)    # exact: conjugate transpose
            assert _close(a @ b, np.eye(2)), (rt, name)


def test_front_end_matrices_match_runtimes(matrix_fns):
    """The front-end's host description (gate_type.pymat) equals the runtimes' factories."""
    rs = np.random.RandomState(13)
    for name in ALL_TYPES:
        args = _angles(rs, N_ARGS.get(name, 0))
        assert _close(_gate(name, args).pymat(), matrix_fns['port'](name, args, False)), name


def test_port_gate_matrices_equal_reference_bitwise(matrix_fns):
    if 'reference' not in matrix_fns:
        pytest.skip('reference core not built (no reference sources on this machine)')
    rs = np.random.RandomState(14)
    for name in ALL_TYPES:
        for _ in range(8):
            args = _angles(rs, N_ARGS.get(name, 0))
            for adj in (False, True):
                ref = matrix_fns['reference'](name, args, adj)
                assert np.array_equal(matrix_fns['port'](name, args, adj), ref), (name, args, adj)
                assert np.array_equal(matrix_fns['hip'](name, args, adj), ref), (name, args, adj)


# ---- threaded spans ---------------------------------------------------------------------

def _outputs(module, dtype, n=18, n_gates=40, seed=7):
    from qgate_amd import workloads
    from qgate_amd.simulator.simulator import Simulator
    q, ops = workloads.random_circuit(n, n_gates, seed)
    sim = Simulator(module, circuit_prep='one_static', dtype=dtype)
    sim.qubits.set_ordering(list(q))
    sim.run(ops)
    out = {'states': np.array(sim.qubits.states[:]),
           'probs': np.array(sim.qubits.prob[:]),
           'p0': np.array([sim.qubits.calc_probability(x) for x in q])}
    u = np.random.RandomState(seed).random_sample(2048)
    out['samples'] = np.array(sim.qubits.create_sampling_pool(list(q)).sample(len(u), u).intarray)
    # a pool over some qubits (the others hidden): the threaded hidden-lane sums
    sub = [q[i] for i in (3, 17, 0, 9, 12)]
    out['samples_hidden'] = np.array(sim.qubits.create_sampling_pool(sub).sample(len(u), u).intarray)
    out['probs_strided'] = np.array(sim.qubits.prob[::7])
    sim.terminate()
    return out


def _compare_port_reference(oracle):
    outs = {}
    for dtype in (np.float64, np.float32):
        a = _outputs(oracle.port(), dtype)
        b = _outputs(oracle.reference(), dtype)
        for k in a:
            assert a[k].dtype == b[k].dtype, k
            assert np.array_equal(a[k], b[k]), (dtype, k)
        outs[dtype] = a
    return outs


def test_threaded_spans_port_equals_reference(oracle):
    if not oracle.have_reference():
        pytest.skip('reference core not built (no reference sources on this machine)')
    L = oracle.load(oracle.PORT_LIB)
    assert L.qgo_num_workers() == len(os.sched_getaffinity(0)) or os.environ.get('QGATE_NUM_WORKERS')
    _compare_port_reference(oracle)


def test_threaded_spans_uneven_worker_count(oracle):
    if not oracle.have_reference():
        pytest.skip('reference core not built (no reference sources on this machine)')
    env = dict(os.environ, QGATE_NUM_WORKERS='3')
    code = ('import sys; sys.path.insert(0, {!r});'
            'from oracle import cpuruntime as o; from tests import test_oracle_pins as t;'
            'assert o.load(o.PORT_LIB).qgo_num_workers() == 3;'
            't._compare_port_reference(o); print("ok")').format(ROOT)
    r = subprocess.run([sys.executable, '-c', code], env=env, cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith('ok'), r.stdout + r.stderr


# ---- empty lanes in get_states ------------------------------------------------------------

def _empty_lane_states(module, n_used, empty_positions, dtype=np.float64):
    """Amplitudes with qregs that have no state vector at `empty_positions` of the ordering."""
    from qgate_amd.simulator.simulator import Simulator
    n = n_used + len(empty_positions)
    used = new_qregs(n_used)
    unused = new_qregs(len(empty_positions))
    ordering, ui, ei = [], 0, 0
    for pos in range(n):
        if pos in empty_positions:
            ordering.append(unused[ei])
            ei += 1
        else:
            ordering.append(used[ui])
            ui += 1
    ops = [H(x) for x in used] + [ctrl(used[0]).X(used[-1])]
    sim = Simulator(module, circuit_prep='one_static', dtype=dtype)
    sim.qubits.set_ordering(ordering)
    sim.run(ops)
    st = np.array(sim.qubits.states[:])
    pr = np.array(sim.qubits.prob[:])
    sim.terminate()
    return st, pr


def _empty_mask(empty_positions):
    m = 0
    for p in empty_positions:
        m |= 1 << p
    return m


def test_empty_lanes_read_zero_port(oracle):
    """Closed form: with every used qubit under H (plus a CX), |amplitude| is 2^(-n_used/2)
    where no empty-lane bit is set and 0 elsewhere, at empty bits below and above bit 8."""
    for empty in ([2, 5], [3, 9], [9, 10]):
        n_used = 9
        st, pr = _empty_lane_states(oracle.port(), n_used, empty)
        idx = np.arange(len(st))
        on = (idx & _empty_mask(empty)) == 0
        assert np.all(st[~on] == 0) and np.all(pr[~on] == 0), empty
        assert np.allclose(np.abs(st[on]), 2. ** (-n_used / 2.), atol=1e-15), empty
        assert abs(pr.sum() - 1.) < 1e-12


def test_reference_getter_ignores_empty_lane_bits_above_7(oracle):
    """Where the reference CPU getter and the port differ, and where they agree."""
    if not oracle.have_reference():
        pytest.skip('reference core not built (no reference sources on this machine)')
    # empty lanes below bit 8: identical
    a = _empty_lane_states(oracle.port(), 9, [2, 5])
    b = _empty_lane_states(oracle.reference(), 9, [2, 5])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # an empty lane at bit 9 (>= 8): the reference reads nonzero values at indices with bit 9
    # set beyond the first 256-entry block; the port (and the HIP runtime) read 0 there
    a = _empty_lane_states(oracle.port(), 9, [3, 9])
    b = _empty_lane_states(oracle.reference(), 9, [3, 9])
    idx = np.arange(len(a[0]))
    bit9 = (idx >> 9) & 1 == 1
    no3 = (idx >> 3) & 1 == 0
    assert np.all(a[0][bit9] == 0)
    assert np.count_nonzero(b[0][bit9 & no3]) > 0
    # elsewhere they agree bit for bit
    assert np.array_equal(a[0][~bit9], b[0][~bit9])
    assert np.array_equal(a[1][~bit9], b[1][~bit9])


@pytest.mark.gpu
def test_empty_lanes_read_zero_hip(hip_module):
    """The HIP getter applies the empty-lane test to every index (states, probabilities,
    strided slices), as the reference's py and CUDA getters do."""
    for empty in ([2, 5], [3, 9], [9, 10]):
        for dtype in (np.float64, np.float32):
            n_used = 9
            st, pr = _empty_lane_states(hip_module, n_used, empty, dtype)
            idx = np.arange(len(st))
            on = (idx & _empty_mask(empty)) == 0
            assert np.all(st[~on] == 0) and np.all(pr[~on] == 0), (empty, dtype)
            tol = 1e-12 if dtype == np.float64 else 5e-6
            assert np.allclose(np.abs(st[on]), 2. ** (-n_used / 2.), atol=tol), (empty, dtype)
