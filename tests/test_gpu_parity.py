"""HIP runtime vs the FP64/FP32 CPU oracle on identical circuits and seeds.

Bar: amplitudes within 1e-12 (complex128) and 5e-6 max-abs (complex64) of the oracle run at
the same precision; FP32 is also compared with the FP64 oracle at 5e-6; probabilities within
the same bounds; measurement outcomes identical; sampled indices identical except where a
uniform lies within the prefix-sum error of a bin edge (counted and checked).
"""
import math

import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.script import *
from qgate_amd.simulator.simulator import Simulator
from tests.helpers import max_abs, tol, TOL_C64

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def _sim(module, ops, prep, dtype, ordering):
    sim = Simulator(module, circuit_prep=prep, dtype=dtype)
    sim.qubits.set_ordering(ordering)
    sim.run(ops)
    return sim


@pytest.mark.parametrize('dtype', DTYPES)
def test_every_gate_every_target(hip_module, port_module, dtype):
    """Each gate type, adjoint or not, with 0..2 controls, on every target of a 12-qubit state."""
    n = 12
    rs = np.random.RandomState(7)
    q, prep_ops = workloads.brickwork(n, 2, seed=11)
    kinds = [lambda: H, lambda: X, lambda: Y, lambda: Z, lambda: S, lambda: T, lambda: SH, lambda: I,
             lambda: Rx(rs.uniform(0, 6.3)), lambda: Ry(rs.uniform(0, 6.3)), lambda: Rz(rs.uniform(0, 6.3)),
             lambda: U1(rs.uniform(0, 6.3)), lambda: U2(*rs.uniform(0, 6.3, 2)),
             lambda: U3(*rs.uniform(0, 6.3, 3)), lambda: Expii(rs.uniform(0, 6.3)),
             lambda: Expiz(rs.uniform(0, 6.3))]
    ops = list(prep_ops)
    for t in range(n):
        for make in kinds:
            for n_ctrl in (0, 1, 2):
                g = make()
                g = g.Adj if rs.randint(2) else g
                if n_ctrl == 0:
                    ops.append(g(q[t]))
                    continue
                others = [i for i in range(n) if i != t]
                cs = [q[i] for i in rs.choice(others, n_ctrl, replace=False)]
                cf = ctrl(cs)
                name = {H: 'H', X: 'X', Y: 'Y', Z: 'Z', S: 'S', T: 'T', SH: 'SH', I: 'I'}.get(g)
                if name is not None:
                    ops.append(getattr(cf, name)(q[t]))
                else:
                    gate = g.gate
                    f = cf._make(gate.gate_type)
                    f = f.Adj if gate.adjoint else f
                    ops.append(f(q[t]))
    a = _sim(hip_module, ops, 'one_static', dtype, q)
    b = _sim(port_module, ops, 'one_static', dtype, q)
    assert max_abs(a.qubits.states[:], b.qubits.states[:]) < tol(dtype) * 10
    c = _sim(port_module, ops, 'one_static', np.float64, q)
    assert max_abs(a.qubits.states[:], c.qubits.states[:]) < TOL_C64 * 10


@pytest.mark.parametrize('dtype', DTYPES)
def test_controls_on_low_bits(hip_module, port_module, dtype):
    """Every (control, target) pair of an 11-qubit state for a pair gate (U3, CX) and a phase
    gate (cu1, CZ), plus two-control gates whose controls sit on the lowest bits: controls on
    the lowest unit-index bits are tested per work item over whole lines, the others folded
    into the work-item index, across the unit, wavefront and strided kernels."""
    n = 11
    rs = np.random.RandomState(12)
    q, prep_ops = workloads.brickwork(n, 2, seed=13)
    ops = list(prep_ops)
    for c in range(n):
        for t in range(n):
            if c == t:
                continue
            ops.append(ctrl(q[c]).U3(*rs.uniform(0, 6.3, 3))(q[t]))
            ops.append(ctrl(q[c]).X(q[t]))
            ops.append(ctrl(q[c]).U1(rs.uniform(0, 6.3))(q[t]))
            ops.append(ctrl(q[c]).Z(q[t]))
    for c0 in range(4):
        for c1 in range(c0 + 1, 5):
            for t in range(n):
                if t in (c0, c1):
                    continue
                ops.append(ctrl(q[c0], q[c1]).U3(*rs.uniform(0, 6.3, 3))(q[t]))
                ops.append(ctrl(q[c0], q[c1]).S(q[t]))
    # partners 64 / 128 units away (LDS chunk exchange): one control inside the chunk (tested
    # per unit or per lane), one above it (folded into the chunk index)
    for t in (6, 7, 8):
        for c0 in (0, 1, 3, 5, 6):
            for c1 in (9, 10):
                if c0 == t:
                    continue
                ops.append(ctrl(q[c0], q[c1]).U3(*rs.uniform(0, 6.3, 3))(q[t]))
                ops.append(ctrl(q[c0], q[c1]).X(q[t]))
    a = _sim(hip_module, ops, 'one_static', dtype, q)
    b = _sim(port_module, ops, 'one_static', dtype, q)
    assert max_abs(a.qubits.states[:], b.qubits.states[:]) < tol(dtype) * 10


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('prep', ['dynamic', 'one_static'])
def test_brickwork_16q(hip_module, port_module, dtype, prep):
    """16-qubit random brickwork of 20 layers (U3 on all qubits, CX on alternating pairs):
    state, per-qubit probabilities and 100000 pool samples."""
    q, ops = workloads.brickwork(16, 20, seed=1)
    a = _sim(hip_module, ops, prep, dtype, q)
    b = _sim(port_module, ops, prep, np.float64, q)
    assert max_abs(a.qubits.states[:], b.qubits.states[:]) < tol(dtype)
    pa = np.array([a.qubits.calc_probability(x) for x in q])
    pb = np.array([b.qubits.calc_probability(x) for x in q])
    assert np.max(np.abs(pa - pb)) < tol(dtype)
    u = np.random.RandomState(1001).random_sample(100000)
    oa = a.qubits.create_sampling_pool(q).sample(len(u), u).intarray
    pool_b = b.qubits.create_sampling_pool(q)
    ob = pool_b.sample(len(u), u).intarray
    mismatch = np.nonzero(oa != ob)[0]
    if len(mismatch):
        # only draws that sit on a bin edge of the FP64 CDF may differ
        cdf = pool_b.cdf()
        edge = np.abs(cdf[np.minimum(ob[mismatch], oa[mismatch])] - u[mismatch])
        assert np.all(edge < (1e-12 if dtype == np.float64 else 1e-5)), edge
        # complex64 amplitudes carry ~1e-7 relative error, so the CDF moves by ~1e-6 at most:
        # a few draws in 10^5 land on the other side of an edge
        print(prep, dtype.__name__, 'draws at a bin edge:', len(mismatch))
        assert len(mismatch) <= (0 if dtype == np.float64 else 100)


@pytest.mark.parametrize('dtype', DTYPES)
def test_random_circuit_14q(hip_module, port_module, dtype):
    """Random {U3, H, Rz, CX, CCX} circuit, 400 gates on 14 qubits, then measure all."""
    q, ops = workloads.random_circuit(14, 400, seed=2)
    a = _sim(hip_module, ops, 'one_static', dtype, q)
    b = _sim(port_module, ops, 'one_static', dtype, q)
    c = _sim(port_module, ops, 'one_static', np.float64, q)
    assert max_abs(a.qubits.states[:], b.qubits.states[:]) < tol(dtype)
    assert max_abs(a.qubits.states[:], c.qubits.states[:]) < tol(dtype)
    refs = new_references(14)
    meas = [measure(r, x) for r, x in zip(refs, q)]
    np.random.seed(1002)
    ma = _sim(hip_module, ops + meas, 'dynamic', dtype, q)
    np.random.seed(1002)
    mb = _sim(port_module, ops + meas, 'dynamic', np.float64, q)
    assert ma.values.get(refs) == mb.values.get(refs)
    assert max_abs(ma.qubits.states[:], mb.qubits.states[:]) < tol(dtype)


@pytest.mark.parametrize('dtype', DTYPES)
def test_qft_random_12q(hip_module, port_module, dtype):
    q, ops = workloads.qft_random(12, 60, seed=3)
    a = _sim(hip_module, ops, 'one_static', dtype, q)
    b = _sim(port_module, ops, 'one_static', np.float64, q)
    assert max_abs(a.qubits.states[:], b.qubits.states[:]) < tol(dtype)


@pytest.mark.parametrize('dtype', DTYPES)
def test_hidden_lanes_and_orderings(hip_module, port_module, dtype):
    """Marginal probability arrays with hidden qubits and permuted orderings."""
    q, ops = workloads.brickwork(12, 6, seed=5)
    a = _sim(hip_module, ops, 'dynamic', dtype, q)
    b = _sim(port_module, ops, 'dynamic', np.float64, q)

    class Capture:
        def __init__(self, prob, empty_lanes, qreg_ordering):
            self.prob = prob

    rs = np.random.RandomState(9)
    for n_keep in (12, 9, 5, 1):
        order = [q[i] for i in rs.permutation(12)[:n_keep]]
        pa = a.qubits.create_sampling_pool(order, Capture).prob
        pb = b.qubits.create_sampling_pool(order, Capture).prob
        assert np.max(np.abs(pa.astype(np.float64) - pb)) < tol(dtype)
        perm = [q[i] for i in rs.permutation(12)]
        a.qubits.set_ordering(perm)
        b.qubits.set_ordering(perm)
        assert max_abs(a.qubits.states[:], b.qubits.states[:]) < tol(dtype)
        assert np.max(np.abs(a.qubits.prob[:] - b.qubits.prob[:])) < tol(dtype)
        assert max_abs(a.qubits.states[3:1000:7], b.qubits.states[3:1000:7]) < tol(dtype)
        assert max_abs(a.qubits.states[4000:100:-13], b.qubits.states[4000:100:-13]) < tol(dtype)


def test_sampling_pool_parity_fp64(hip_module, port_module):
    """FP64 pools: identical indices except on bin edges, with and without empty lanes."""
    q, ops = workloads.random_circuit(16, 200, seed=21)
    a = _sim(hip_module, ops, 'one_static', np.float64, q)
    b = _sim(port_module, ops, 'one_static', np.float64, q)
    extra = new_qregs(3)
    order = [q[0], extra[0]] + q[1:8] + [extra[1]] + q[8:] + [extra[2]]
    u = np.random.RandomState(1003).random_sample(1 << 20)
    for ordering in (q, order):
        oa = a.qubits.create_sampling_pool(ordering).sample(len(u), u).intarray
        ob = b.qubits.create_sampling_pool(ordering).sample(len(u), u).intarray
        assert np.count_nonzero(oa != ob) == 0


def test_random_unitary_round_trip_24q(hip_module):
    """Circuit followed by its inverse returns |0...0> (24 qubits, complex64 and complex128)."""
    for dtype in DTYPES:
        q, ops = workloads.random_circuit(24, 120, seed=31)
        inverse = []
        for op in reversed(ops):
            g = op.copy()
            g.set_adjoint(not op.adjoint)
            inverse.append(g)
        sim = _sim(hip_module, ops + inverse, 'one_static', dtype, q)
        st = sim.qubits.states[:]
        assert abs(st[0] - 1.) < (1e-11 if dtype == np.float64 else 5e-5)
        assert np.max(np.abs(st[1:])) < (1e-11 if dtype == np.float64 else 5e-5)


@pytest.mark.parametrize('dtype', DTYPES)
def test_sampling_pool_sparse_states(hip_module, dtype):
    """Pools of 2^24 bins (4096 tiles of 4096 bins in 4 super-tiles of 2^22 bins) over states
    whose nonzero bins sit at tile and super-tile edges. Every draw, including u = 0 and
    u = 1 - 2^-53, returns a nonzero bin: the first one above u on the CDF, or the last nonzero
    bin when rounding leaves the CDF's end just below u (the zero bins after it are never
    returned). With two bins of probability 1/2, draws below 1/2 return the first one."""
    n = 24
    q = new_qregs(n)
    u = np.concatenate([[0., 5e-324, 1e-300, 0.25, 0.5 - 1e-5, 0.5 + 1e-5, 0.75, 1. - 2. ** -52, 1. - 2. ** -53],
                        np.random.RandomState(5).random_sample(4096)])

    def draws(ops):
        sim = _sim(hip_module, ops, 'one_static', dtype, q)
        nz = np.flatnonzero(np.asarray(sim.qubits.prob[:]) > 0.25)
        obs = np.asarray(sim.qubits.create_sampling_pool(q).sample(len(u), u).intarray)
        sim.terminate()
        return nz, obs

    for k in (0, 4095, 4096, (1 << 22) - 1, (1 << 22) + 7, (1 << 24) - 1):
        nz, obs = draws([X(q[b]) for b in range(n) if (k >> b) & 1])
        assert len(nz) == 1, (k, nz)
        assert np.all(obs == nz[0]), (k, nz, np.unique(obs))
    k0 = 4095
    nz, obs = draws([X(q[b]) for b in range(12)] + [H(q[22])])
    assert len(nz) == 2 and k0 in nz, nz
    away = np.abs(u - 0.5) > 1e-6
    want = np.where(u < 0.5, nz[0], nz[1])
    assert np.all(obs[away] == want[away]), np.unique(obs[away] - want[away])
    assert np.all(np.isin(obs, nz))
