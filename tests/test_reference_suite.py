"""Closed-form behaviour checks taken from the reference test-suite, run on every backend.

Each test restates the assertions of the reference test named in its docstring
(qgate tests/…); the same bodies run on the CPU oracle (CPU suite) and on the HIP runtime
(marked gpu), as the reference instantiates its suite per runtime (tests/test_base.py:28-54).
"""
import cmath
import math

import numpy as np
import pytest

import qgate_amd
from qgate_amd.script import *
from qgate_amd.script.qelib1 import cx, ccx, cz, cu1, sdg, tdg
from qgate_amd.simulator import prob as PROB


def _run(simulator, circuit, isolate=True, ordering=None, **prefs):
    prep = qgate_amd.prefs.dynamic if isolate else qgate_amd.prefs.one_static
    sim = simulator(circuit_prep=prep, **prefs)
    if ordering is not None:
        sim.qubits.set_ordering(ordering)
    sim.run(circuit)
    return sim


def _probs(simulator, circuit, ordering, isolate=True):
    # the state getters hold the simulator's qubits weakly: keep the simulator alive
    sim = _run(simulator, circuit, isolate=isolate, ordering=ordering)
    return sim.qubits.prob[:]


# ---- tests/test_unary_gate.py ------------------------------------------------------------

def test_unary_gates(simulator):
    q = new_qreg()
    assert _run(simulator, I(q)).qubits.get_states(PROB)[0] == 1.
    assert _run(simulator, X(q)).qubits.get_states(PROB)[1] == 1.
    assert _run(simulator, [X(q), X(q)]).qubits.get_states(PROB)[0] == 1.
    p = _run(simulator, H(q)).qubits.get_states(PROB)
    assert p[0] == pytest.approx(0.5) and p[1] == pytest.approx(0.5)
    sim = _run(simulator, [H(q), H(q)])
    sim.qubits.set_ordering([q])
    p = sim.qubits.prob[:]
    assert p[0] == pytest.approx(1.) and p[1] == pytest.approx(0., abs=1e-7)


@pytest.mark.parametrize('n', range(1, 11))
def test_unary_multi_qubits(simulator, n):
    q = new_qregs(n)
    p = _run(simulator, [X(x) for x in q]).qubits.get_states(PROB)
    assert p[(1 << n) - 1] == pytest.approx(1.)
    p = _run(simulator, [H(x) for x in q]).qubits.get_states(PROB)
    assert np.allclose(p, 1. / (1 << n))


# ---- tests/test_control_gate.py ----------------------------------------------------------

def test_cx_2qubits(simulator):
    q = new_qregs(2)
    cases = [([], (0, 1), 0), ([X(q[0])], (0, 1), 3), ([X(q[1])], (0, 1), 2), ([X(q[0]), X(q[1])], (0, 1), 1),
             ([], (1, 0), 0), ([X(q[0])], (1, 0), 1), ([X(q[1])], (1, 0), 3), ([X(q[0]), X(q[1])], (1, 0), 2)]
    for pre, (c, t), expected in cases:
        p = _probs(simulator, pre + [ctrl(q[c]).X(q[t])], q)
        assert p[expected] == pytest.approx(1.)


def test_cx_multibits(simulator):
    for n in range(1, 9):
        for c in range(n):
            for t in range(n):
                if c == t:
                    continue
                q = new_qregs(n)
                p = _probs(simulator, [[I(x) for x in q], ctrl(q[c]).X(q[t])], q)
                assert p[0] == pytest.approx(1.)
                p = _probs(simulator, [[I(x) for x in q], X(q[c]), ctrl(q[c]).X(q[t])], q)
                assert p[(1 << c) | (1 << t)] == pytest.approx(1.)


def test_n_controlled_x(simulator):
    for n in range(2, 10):
        q = new_qregs(n)
        p = _probs(simulator, [[I(x) for x in q], [X(x) for x in q[:-1]], ctrl(q[:-1]).X(q[-1])], q)
        assert len(p) == 1 << n and p[-1] == pytest.approx(1.)
    for n in range(3, 10):
        q = new_qregs(n)
        p = _probs(simulator, [I(q[0]), [X(x) for x in q[1:-1]], ctrl(q[:-1]).X(q[-1])], q)
        assert p[-1] == pytest.approx(0., abs=1e-12)


# ---- tests/test_calc_prob.py -------------------------------------------------------------

def test_calc_prob(simulator):
    q = new_qreg()
    assert _run(simulator, I(q), isolate=False).qubits.calc_probability(q) == 1.
    assert _run(simulator, X(q), isolate=False).qubits.calc_probability(q) == pytest.approx(0.)
    for idx in range(10):
        for gate, value in ((I, 1.), (X, 0.), (H, 0.5)):
            q = new_qregs(10)
            sim = _run(simulator, [[I(x) for x in q], gate(q[idx])], isolate=False)
            for obs in range(10):
                expected = value if obs == idx else 1.
                assert sim.qubits.calc_probability(q[obs]) == pytest.approx(expected)


def test_calc_prob_cx_isolated(simulator):
    q = new_qregs(10)
    c = [X(q[i]) for i in range(0, 10, 2)] + [ctrl(q[i]).X(q[i + 1]) for i in range(0, 10, 2)]
    sim = _run(simulator, c, isolate=True)
    for x in q:
        assert sim.qubits.calc_probability(x) == pytest.approx(0.)


# ---- tests/test_measure.py / test_reset.py / test_if.py ---------------------------------

def test_measure(simulator):
    q, r = new_qreg(), new_reference()
    assert _run(simulator, [I(q), measure(r, q)]).values.get(r) == 0
    q, r = new_qreg(), new_reference()
    assert _run(simulator, [X(q), measure(r, q)]).values.get(r) == 1
    for qi in range(10):
        for ri in range(10):
            q, r = new_qregs(10), new_references(10)
            sim = _run(simulator, [ctrl(q[:-1]).I(q[-1]), measure(r[ri], q[qi])])
            assert sim.values.get(r[ri]) == 0
        q, r = new_qregs(10), new_references(10)
        sim = _run(simulator, [ctrl(q[:-1]).I(q[-1]), X(q[qi]), measure(r[ri], q[qi])])
        assert sim.values.get(r[ri]) == 1
        assert sim.values.get_packed_value(r) == 1 << ri


def test_reset(simulator):
    for pre, first in ((I, 0), (X, 1)):
        q, r = new_qreg(), new_references(2)
        sim = _run(simulator, [pre(q), measure(r[0], q), reset(q), measure(r[1], q)])
        assert sim.values.get(r[0]) == first and sim.values.get(r[1]) == 0
    for idx in range(10):
        for pre, first in ((I, 0), (X, 1)):
            q, r = new_qregs(10), new_references(2)
            sim = _run(simulator, [pre(q[idx]), measure(r[0], q[idx]), reset(q[idx]), measure(r[1], q[idx])])
            assert sim.values.get(r[0]) == first and sim.values.get(r[1]) == 0


def test_reset_not_allowed(simulator):
    q = new_qregs(10)
    c = [ctrl(q[:-1]).X(q[-1])]
    for idx in range(10):
        c.append(reset(q[idx]))
        with pytest.raises(RuntimeError):
            _run(simulator, c)


def test_if(simulator):
    q, r = new_qregs(2), new_reference()
    sim = _run(simulator, [X(q[0]), measure(r, q[0]), if_(r, 1, X(q[1]))])
    assert sim.qubits.prob[3] == pytest.approx(1.)
    for pre, expected in (([], 0), ([X], 1)):
        q, r = new_qregs(2), new_references(2)
        c = [p(q[0]) for p in pre] + [measure(r[0], q[0]), if_(r, 1, X(q[1])), measure(r[1], q[1])]
        sim = _run(simulator, c)
        assert sim.values.get(r[0]) == expected and sim.values.get(r[1]) == expected
        q, r = new_qregs(2), new_references(2)
        c = [p(q[0]) for p in pre] + [measure(r[0], q[0]),
                                      if_(r, lambda v0, v1: v0 == 1 and v1 is None, X(q[1])),
                                      measure(r[1], q[1])]
        sim = _run(simulator, c)
        assert sim.values.get(r[0]) == expected and sim.values.get(r[1]) == expected


# ---- tests/test_join.py ------------------------------------------------------------------

def _join_size(simulator):
    return 15 if simulator.runtime == 'port' else 20


def test_join_1_by_1(simulator):
    for n in range(2, _join_size(simulator)):
        q = new_qregs(n)
        c = [H(x) for x in q] + [ctrl(q[i]).X(q[i + 1]) for i in range(n - 1)]
        p = _probs(simulator, c, q)
        assert np.allclose(p, 1. / (1 << n))


def test_join_n_n(simulator):
    for n in range(2, _join_size(simulator) // 2):
        q0, q1 = new_qregs(n), new_qregs(n)
        c = [ctrl(q0[i]).I(q0[i + 1]) for i in range(n - 1)] + [H(x) for x in q0]
        c += [ctrl(q1[i]).I(q1[i + 1]) for i in range(n - 1)] + [H(x) for x in q1]
        c += [ctrl(q0[:-1]).I(q1[0])]
        p = _probs(simulator, c, q0 + q1)
        assert np.allclose(p, 1. / (1 << (2 * n)))


def test_join_n_1(simulator):
    q = new_qregs(_join_size(simulator))
    for target in q:
        c = [H(x) for x in q] + [ctrl([x for x in q if x is not target]).I(target)]
        p = _probs(simulator, c, q)
        assert np.allclose(p, 1. / (1 << len(q)))


# ---- tests/test_get_states.py ------------------------------------------------------------

@pytest.mark.parametrize('n', [4, 20])
def test_get_states(simulator, n):
    q = new_qregs(n)
    st = _run(simulator, [I(x) for x in q]).qubits.get_states()
    assert st[0] == 1 + 0j and np.all(st[1:] == 0.)
    st = _run(simulator, [I(x) for x in q]).qubits.get_states(key=slice(1, None))
    assert np.all(st == 0.)
    sim = _run(simulator, [X(x) for x in q])
    st = sim.qubits.get_states()
    assert st[-1] == 1 and np.all(st[:-1] == 0.)
    assert np.all(sim.qubits.get_states(key=slice(0, -1)) == 0.)
    assert sim.qubits.get_states(key=-1) == 1
    p = _run(simulator, [H(x) for x in q]).qubits.get_states(mathop=PROB)
    assert np.allclose(p, 1. / len(p))

    sim = _run(simulator, [H(x) for x in q], isolate=False)
    assert len(sim.qubits.get_states(key=slice(10, 0))) == 0
    assert len(sim.qubits.get_states(key=slice(10, 0, -1))) == 10
    full = sim.qubits.get_states()
    for key in (slice(0, 10, 3), slice(0, 10, 7), slice(10, None, -3), slice(10, None, -7),
                slice(None, None, -1), slice(10, None, -1)):
        part = sim.qubits.get_states(key=key)
        assert np.allclose(part, full[key])
    assert np.all(sim.qubits.states[:] == full)
    assert np.all(sim.qubits.states[10:] == sim.qubits.get_states(key=slice(10, None)))
    assert np.all(sim.qubits.prob[:10] == sim.qubits.get_states(PROB, key=slice(None, 10)))


# ---- tests/test_qreg_ordering.py ---------------------------------------------------------

def test_qreg_ordering(simulator):
    q = new_qregs(3)
    sim = simulator()
    sim.qubits.set_ordering(q)
    sim.run([])
    p = sim.qubits.prob[:]
    assert len(p) == 8 and p[0] == 1. and np.all(p[1:] == 0)
    q0, q1 = new_qregs(2)
    sim = simulator()
    sim.qubits.set_ordering([q1])
    sim.run([I(q0)])
    p = sim.qubits.prob[:]
    assert len(p) == 4 and p[0] == 1. and np.all(p[1:] == 0)


# ---- tests/test_swap_gate.py -------------------------------------------------------------

def test_swap(simulator):
    q0, q1 = q = new_qregs(2)
    for pre, expected in (([], 0), ([X(q0)], 2), ([X(q1)], 1), ([X(q0), X(q1)], 3)):
        p = _probs(simulator, pre + [Swap(q0, q1)], q)
        assert p[expected] == pytest.approx(1.)
    for n in range(3, 9):
        q = new_qregs(n)
        c = [X(q[0])] + [Swap(q[i], q[i + 1]) for i in range(n - 1)]
        p = _probs(simulator, c, q)
        for i in range(n - 1):
            assert p[1 << i] == pytest.approx(0., abs=1e-12)
        assert p[1 << (n - 1)] == pytest.approx(1.)


# ---- tests/test_exp_gate.py --------------------------------------------------------------

def test_exp_gates(simulator):
    q = new_qregs(4)
    c = [H(x) for x in q]
    v = 1. / (math.sqrt(2) ** 4)
    st = _run(simulator, c, isolate=False).qubits.get_states()
    assert np.allclose(st, v)
    st = _run(simulator, c + [Expii(math.pi / 8)(q[0])], isolate=False).qubits.get_states()
    assert np.allclose(st, cmath.exp(1.j * math.pi / 8) * v)
    st = _run(simulator, c + [Expiz(math.pi / 8)(q[0])], isolate=False).qubits.get_states()
    w = cmath.exp(1.j * math.pi / 8) * v
    assert np.allclose(st[0::2], w) and np.allclose(st[1::2], np.conjugate(w))


# ---- tests/test_simple_calls.py ----------------------------------------------------------

def test_simple_calls(simulator):
    q0, q1 = new_qregs(2)
    plain = [I(q0), H(q0), S(q0), T(q0), X(q0), Y(q0), Z(q0), Rx(0.)(q0), Ry(0.)(q0), Rz(0.)(q0),
             U1(0.)(q0), U2(0., 0.)(q0), U3(0., 0., 0.)(q0), Expii(0.)(q0), Expiz(0.)(q0),
             Expi(0.)(Z(q0)), SH(q0), Swap(q0, q1), barrier(q0, q1)]
    adj = [I.Adj(q0), H.Adj(q0), S.Adj(q0), T.Adj(q0), X.Adj(q0), Y.Adj(q0), Z.Adj(q0),
           Rx(0.).Adj(q0), Ry(0.).Adj(q0), Rz(0.).Adj(q0), U1(0.).Adj(q0), U2(0., 0.).Adj(q0),
           U3(0., 0., 0.).Adj(q0), Expii(0.).Adj(q0), Expiz(0.).Adj(q0), Expi(0.)(Z(q0)),
           SH.Adj(q0), Swap(q0, q1), barrier(q0, q1)]
    c = ctrl(q0)
    ctrl_ops = [c.I(q1), c.H(q1), c.S(q1), c.T(q1), c.X(q1), c.Y(q1), c.Z(q1), c.Rx(0.)(q1),
                c.Ry(0.)(q1), c.Rz(0.)(q1), c.U1(0.)(q1), c.U2(0., 0.)(q1), c.U3(0., 0., 0.)(q1),
                c.Expii(0.)(q1), c.Expiz(0.)(q1), c.Expi(0.)(Z(q1)), c.SH(q1)]
    ctrl_adj = [c.I.Adj(q1), c.H.Adj(q1), c.S.Adj(q1), c.T.Adj(q1), c.X.Adj(q1), c.Y.Adj(q1),
                c.Z.Adj(q1), c.Rx(0.).Adj(q1), c.Ry(0.).Adj(q1), c.Rz(0.).Adj(q1), c.U1(0.).Adj(q1),
                c.U2(0., 0.).Adj(q1), c.U3(0., 0., 0.).Adj(q1), c.Expii(0.).Adj(q1),
                c.Expiz(0.).Adj(q1), c.Expi(0.)(Z(q1)), c.SH.Adj(q1)]
    for circuit in (plain, adj, ctrl_ops, ctrl_adj, [cx(q0, q1), ccx(q0, q1, new_qreg()), cz(q0, q1),
                                                     cu1(0.3, q0, q1), sdg(q0), tdg(q1)]):
        qgate_amd.dump(circuit, file=open('/dev/null', 'w'))
        p = _run(simulator, circuit).qubits.get_states(PROB)
        assert np.sum(p) == pytest.approx(1.)


# ---- tests/test_big_circuits.py ----------------------------------------------------------

def test_big_circuits(simulator):
    n = 16 if simulator.runtime == 'port' else 20
    q = new_qregs(n)
    sim = _run(simulator, [H(x) for x in q], isolate=False)
    for x in q:
        assert sim.qubits.calc_probability(x) == pytest.approx(0.5)
    sim = _run(simulator, [X(x) for x in q], isolate=False)
    for x in q:
        assert sim.qubits.calc_probability(x) == pytest.approx(0.)
    c = [X(q[0])] + [ctrl(q[i]).X(q[i + 1]) for i in range(n - 1)]
    sim = _run(simulator, c, isolate=False)
    for x in q:
        assert sim.qubits.calc_probability(x) == pytest.approx(0.)


# ---- tests/test_sampling_pool.py ---------------------------------------------------------

class DummySamplingPool:
    def __init__(self, prob, empty_lanes, qreg_ordering):
        self.prob = prob
        self.empty_lanes = empty_lanes
        self.qreg_ordering = qreg_ordering


def _pool_size(simulator):
    return 16 if simulator.runtime == 'port' else 24


def test_sampling_pool_probs(simulator):
    n = _pool_size(simulator)
    q = new_qregs(n)
    sim = _run(simulator, [H(x) for x in q])
    sp = sim.qubits.create_sampling_pool(q, DummySamplingPool)
    assert len(sp.prob) == 1 << n and np.allclose(sp.prob, 1. / (1 << n))
    with pytest.raises(RuntimeError):
        sim.qubits.create_sampling_pool(q + q, DummySamplingPool)
    # hidden qregs
    for order in (q[::2], q[1::2]):
        sp = sim.qubits.create_sampling_pool(order, DummySamplingPool)
        assert len(sp.prob) == 1 << len(order) and np.allclose(sp.prob, 1. / (1 << len(order)))
    # empty lanes
    qo = new_qregs(n)
    sim = _run(simulator, [H(x) for x in qo[::2]])
    sp = sim.qubits.create_sampling_pool(qo, DummySamplingPool)
    assert sp.qreg_ordering == qo
    for x in qo[1::2]:
        assert qo.index(x) in sp.empty_lanes


def test_sampling_pool_x_patterns(simulator):
    n = 10
    q = new_qregs(n)
    for xl in range(n):
        c = [H(x) for x in q[:xl]] + [X(q[xl])] + [H(x) for x in q[xl + 1:]]
        sp = _run(simulator, c).qubits.create_sampling_pool(q, DummySamplingPool)
        idx = [i for i in range(1024) if (i & (1 << xl)) == 0]
        assert np.allclose(sp.prob[idx], 0.)
    xq, hq = q[0], q[1:]
    sim = _run(simulator, [X(xq)] + [H(x) for x in hq])
    for xl in range(9):
        sp = sim.qubits.create_sampling_pool(hq[:xl] + [xq] + hq[xl:], DummySamplingPool)
        idx = [i for i in range(1024) if (i & (1 << xl)) == 0]
        assert np.allclose(sp.prob[idx], 0.)
    hr, xr = q[::2], q[1::2]
    sim = _run(simulator, [H(x) for x in hr] + [X(x) for x in xr])
    sp = sim.qubits.create_sampling_pool(hr, DummySamplingPool)
    assert np.allclose(sp.prob, 1. / (1 << len(hr)))
    sp = sim.qubits.create_sampling_pool(xr, DummySamplingPool)
    assert sp.prob[-1] == pytest.approx(1.)


@pytest.mark.parametrize('one_static', [False, True])
def test_sampling_pool_known_answer(simulator, one_static):
    """u at the centre of every bin of a uniform distribution returns every bin in order
    (tests/test_sampling_pool.py:166-179, 205-220)."""
    n = _pool_size(simulator)
    if one_static:
        target, q = new_qreg(), new_qregs(n - 1)
        sim = _run(simulator, [X(target)] + [H(x) for x in q], isolate=False)
    else:
        q = new_qregs(n)
        sim = _run(simulator, [H(x) for x in q])
    sp = sim.qubits.create_sampling_pool(q)
    n_probs = 1 << len(q)
    p = 1. / n_probs
    u = np.linspace(0. + p / 2., 1., n_probs, endpoint=False, dtype=np.float64)
    obs = sp.sample(n_probs, u)
    ref = np.linspace(0. + p / 2., n_probs, n_probs, endpoint=False, dtype=np.int64)
    assert np.array_equal(obs.intarray, ref)
    obs = sp.sample(1024)
    assert len(obs) == 1024


def test_sampling_pool_empty_lane(simulator):
    qo = new_qregs(_pool_size(simulator))
    qs = qo[::2]
    sim = _run(simulator, [X(x) for x in qs])
    obs = sim.qubits.create_sampling_pool(qo).sample(1024)
    arr = obs.intarray
    for x in qo:
        bit = 1 << qo.index(x)
        if x in qs:
            assert all(v == 1 for v in obs(x)) and np.all((arr & bit) != 0)
        else:
            assert all(v is None for v in obs(x)) and np.all((arr & bit) == 0)
