"""Marginal probability arrays (hidden qubits) and bounded sampling pools on the HIP runtime.

Probability arrays with hidden qubits are compared with the oracle's (the reference CPU getter
restated) at 13 and 20 qubits, for hidden sets that take every path of the marginal kernels:
hidden qubits above the kept ones, below them (summed across the lanes of a wavefront), inside
the 16-byte unit (complex64 bit 0), all but one qubit hidden (sliced sums), random subsets and
permuted orders, one state vector (static preparation) and several (dynamic preparation). The
sum is over the same terms in another fixed order, so the bar is 1e-12 (FP64) / 5e-6 (FP32).

Bounded pools keep prefix sums only and read |a|^2 from the state when they draw; their draws
must equal the full pool's bit for bit, and the reference's known answers and edge cases must
hold for them too. Drawing after the state changed must fail.
"""
import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.script import new_qregs, H, X, ctrl
from qgate_amd.simulator.simulator import Simulator

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def tol(dtype):
    return 1e-12 if dtype == np.float64 else 5e-6


class Capture:
    def __init__(self, prob, empty_lanes, qreg_ordering):
        self.prob = prob


def _sim(module, ops, prep, dtype, q):
    sim = Simulator(module, circuit_prep=prep, dtype=dtype)
    sim.qubits.set_ordering(list(q))
    sim.run(ops)
    return sim


def _orders(q, rs):
    n = len(q)
    yield 'top hidden', q[:n - 3]
    yield 'one below top kept', q[:n - 1]
    yield 'low hidden', q[6:]
    yield 'bit 0 hidden', q[1:]
    yield 'lane bits hidden', [q[0]] + q[7:]
    yield 'one kept', [q[3]]
    yield 'one high kept', [q[n - 2]]
    yield 'two kept', [q[n - 1], q[0]]
    for k in (n, n - 5, n // 2, 4):
        yield 'random {}'.format(k), [q[i] for i in rs.permutation(n)[:k]]


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('n,prep', [(13, 'one_static'), (20, 'one_static'), (20, 'dynamic')])
def test_prob_arrays_with_hidden_qubits(hip_module, port_module, dtype, n, prep):
    if prep == 'dynamic':
        # two brickwork layers without the first two CX: state vectors {0}, {1, 2}, {3 .. n-1}
        q, ops = workloads.brickwork(n, 2, seed=7)
        ops = ops[:n] + ops[n + 2:]
    else:
        q, ops = workloads.random_circuit(n, 4 * n, seed=7)
    a = _sim(hip_module, ops, prep, dtype, q)
    b = _sim(port_module, ops, prep, np.float64, q)
    rs = np.random.RandomState(n)
    for name, order in _orders(q, rs):
        pa = np.asarray(a.qubits.create_sampling_pool(order, Capture).prob, np.float64)
        pb = np.asarray(b.qubits.create_sampling_pool(order, Capture).prob, np.float64)
        err = float(np.max(np.abs(pa - pb)))
        assert err < tol(dtype), (name, err)
        # the device pool over the same order: FP64 draws identical except at bin edges
        if dtype == np.float64:
            u = np.random.RandomState(7).random_sample(1 << 14)
            oa = np.asarray(a.qubits.create_sampling_pool(order).sample(len(u), u).intarray)
            ob = np.asarray(b.qubits.create_sampling_pool(order).sample(len(u), u).intarray)
            assert np.count_nonzero(oa != ob) == 0, name
    a.terminate()
    b.terminate()


def test_prob_array_repeatable(hip_module):
    """Fixed summation order: two calls give the same bits."""
    q, ops = workloads.random_circuit(20, 60, seed=3)
    a = _sim(hip_module, ops, 'one_static', np.float64, q)
    for order in (q[:17], [q[2]], q[5:]):
        p1 = a.qubits.create_sampling_pool(order, Capture).prob
        p2 = a.qubits.create_sampling_pool(order, Capture).prob
        assert np.array_equal(p1, p2)
    a.terminate()


# ---- bounded sampling pools -----------------------------------------------------------------

@pytest.fixture
def bounded(hip_module):
    hip_module.set_sampling_pool_workspace('bounded')
    yield hip_module
    hip_module.set_sampling_pool_workspace('auto')


@pytest.mark.parametrize('dtype', DTYPES)
def test_bounded_pool_draws_equal_full_pool(hip_module, dtype):
    q, ops = workloads.random_circuit(20, 200, seed=23)
    sim = _sim(hip_module, ops, 'one_static', dtype, q)
    u = np.random.RandomState(24).random_sample(1 << 20)
    u[:4] = [0., 1. - 2. ** -53, 0.5, 1e-300]
    hip_module.set_sampling_pool_workspace('full')
    full = np.asarray(sim.qubits.create_sampling_pool(q).sample(len(u), u).intarray)
    hip_module.set_sampling_pool_workspace('bounded')
    try:
        pool = sim.qubits.create_sampling_pool(q)
        bnd = np.asarray(pool.sample(len(u), u).intarray)
        assert np.array_equal(full, bnd)
        # a second batch from the same pool, then the state changes: drawing must fail
        assert np.array_equal(np.asarray(pool.sample(1024, u[:1024]).intarray), full[:1024])
        sim.run([H(q[3])])
        with pytest.raises((ValueError, RuntimeError)):
            pool.sample(16, u[:16])
        # a new pool sees the new state
        again = np.asarray(sim.qubits.create_sampling_pool(q).sample(1024, u[:1024]).intarray)
        assert again.min() >= 0 and again.max() < (1 << 20)
    finally:
        hip_module.set_sampling_pool_workspace('auto')
        sim.terminate()


def test_bounded_pool_after_state_freed(bounded):
    q, ops = workloads.random_circuit(12, 40, seed=25)
    sim = _sim(bounded, ops, 'one_static', np.float64, q)
    pool = sim.qubits.create_sampling_pool(q)
    pool.sample(8, np.linspace(0.05, 0.95, 8))
    sim.terminate()
    with pytest.raises((ValueError, RuntimeError)):
        pool.sample(8, np.linspace(0.05, 0.95, 8))
    pool = None


@pytest.mark.parametrize('dtype', DTYPES)
def test_bounded_pool_known_answer(bounded, dtype):
    """u at the centre of every bin of a uniform distribution returns every bin in order
    (tests/test_sampling_pool.py:166-179), on one state vector in its own order (bounded)."""
    n = 22
    q = new_qregs(n)
    sim = _sim(bounded, [H(x) for x in q], 'one_static', dtype, q)
    n_probs = 1 << n
    p = 1. / n_probs
    u = np.linspace(p / 2., 1., n_probs, endpoint=False, dtype=np.float64)
    obs = sim.qubits.create_sampling_pool(q).sample(n_probs, u)
    ref = np.linspace(p / 2., n_probs, n_probs, endpoint=False, dtype=np.int64)
    assert np.array_equal(obs.intarray, ref)
    sim.terminate()


def test_bounded_pool_empty_lane(bounded):
    """tests/test_sampling_pool.py:181-196 on one state vector with X on every other qubit,
    the others absent from the simulation: their bits read 0, the X bits 1."""
    qo = new_qregs(16)
    qs = qo[::2]
    sim = Simulator(bounded, circuit_prep='one_static', dtype=np.float64)
    sim.run([X(x) for x in qs])
    obs = sim.qubits.create_sampling_pool(qo).sample(1024)
    arr = np.asarray(obs.intarray)
    for x in qo:
        bit = 1 << qo.index(x)
        if x in qs:
            assert all(v == 1 for v in obs(x)) and np.all((arr & bit) != 0)
        else:
            assert all(v is None for v in obs(x)) and np.all((arr & bit) == 0)
    sim.terminate()


def test_bounded_pool_parity_fp64(bounded, port_module):
    """FP64 bounded pools against the oracle: identical indices, with and without empty lanes."""
    q, ops = workloads.random_circuit(16, 200, seed=21)
    a = _sim(bounded, ops, 'one_static', np.float64, q)
    b = _sim(port_module, ops, 'one_static', np.float64, q)
    extra = new_qregs(3)
    order = [q[0], extra[0]] + q[1:8] + [extra[1]] + q[8:] + [extra[2]]
    u = np.random.RandomState(1003).random_sample(1 << 20)
    for ordering in (q, order):
        oa = a.qubits.create_sampling_pool(ordering).sample(len(u), u).intarray
        ob = b.qubits.create_sampling_pool(ordering).sample(len(u), u).intarray
        assert np.count_nonzero(np.asarray(oa) != np.asarray(ob)) == 0
    a.terminate()
    b.terminate()


@pytest.mark.parametrize('n', [3, 11, 12, 13])
def test_bounded_pool_small_and_sparse(bounded, n):
    """Pools smaller than one 4096-bin tile (padding bins read no amplitude) and sparse states
    (most tiles zero): every draw lands on a nonzero bin; bin centres give every bin."""
    q = new_qregs(n)
    k = min(n - 1, 3)
    ops = [X(q[n - 1])] + [H(x) for x in q[:k]]
    sim = _sim(bounded, ops, 'one_static', np.float64, q)
    nz = (1 << (n - 1)) + np.arange(1 << k)
    p = 1. / (1 << k)
    u = np.linspace(p / 2., 1., 1 << k, endpoint=False)
    obs = np.asarray(sim.qubits.create_sampling_pool(q).sample(len(u), u).intarray)
    assert np.array_equal(obs, nz)
    u = np.random.RandomState(n).random_sample(4096)
    u[:3] = [0., 1. - 2. ** -53, 1.]
    obs = np.asarray(sim.qubits.create_sampling_pool(q).sample(len(u), u).intarray)
    assert np.all(np.isin(obs, nz))
    sim.terminate()
