"""Large single-GPU states: 34 qubits complex64 and 33 qubits complex128 (128 GiB each).

No CPU oracle reaches these sizes, so the checks are properties that hold at any size:
  * H on every qubit followed by the QFT (controlled-phase ladder and swaps) maps |0...0> back
    to |0...0> (the QFT of the uniform state is |0>): every P(qubit = 0) must be 1;
  * 200 random U3 / CX gates followed by their adjoints in reverse order return to |0...0>;
  * a sampling pool over 20 qubits with 14 hidden ones sees a normalised distribution (the
    pool rejects a total off by more than 0.05) and its draws fall on nonzero bins.
These runs take every kernel through the piecewise launches that grids beyond 2^32 work items
need, and the complex64 case runs the random part through fused tile passes.
"""
import gc
import math

import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.script import new_qregs, H, U3, ctrl
from qgate_amd.simulator.simulator import Simulator

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _random_and_inverse(q, n_gates, seed):
    rs = np.random.RandomState(seed)
    fwd, inv = [], []
    n = len(q)
    for _ in range(n_gates):
        if rs.randint(2) == 0:
            t, p, l = (float(v) for v in rs.uniform(0., 2. * math.pi, 3))
            k = int(rs.randint(n))
            fwd.append(U3(t, p, l)(q[k]))
            inv.append(U3(t, p, l).Adj(q[k]))
        else:
            c, t = (int(v) for v in rs.choice(n, 2, replace=False))
            fwd.append(ctrl(q[c]).X(q[t]))
            inv.append(ctrl(q[c]).X(q[t]))
    return fwd, inv[::-1]


@pytest.mark.parametrize('n,dtype,fusion', [(34, np.float32, 'tile'), (33, np.float64, 'off')])
def test_large_state_round_trips(hip_module, n, dtype, fusion):
    info = hip_module.device_info()
    need = (8 if dtype == np.float32 else 16) << n
    if info['total_mem'] < need + (24 << 30):
        pytest.skip('device memory {} < {} + headroom'.format(info['total_mem'], need))
    tol = 2e-5 if dtype == np.float32 else 1e-10
    q = new_qregs(n)
    sim = Simulator(hip_module, circuit_prep='one_static', dtype=dtype)
    sim.qubits.set_ordering(list(q))
    try:
        print('[large] H + QFT', n, dtype.__name__, flush=True)
        sim.run([H(x) for x in q] + workloads.qft(q))
        p0 = [sim.qubits.calc_probability(x) for x in q]
        print('[large] after QFT: min P(0) =', min(p0), flush=True)
        assert min(p0) > 1. - tol, min(p0)
        fwd, inv = _random_and_inverse(q, 200, seed=3)
        hip_module.set_fusion(fusion)
        try:
            sim.run(fwd)
            pool = sim.qubits.create_sampling_pool(list(q[:20]))
            u = np.random.RandomState(1003).random_sample(4096)
            obs = np.asarray(pool.sample(len(u), u).intarray)
            assert obs.min() >= 0 and obs.max() < (1 << 20)
            pool = None
            print('[large] random part + pool done', flush=True)
            sim.run(inv)
        finally:
            hip_module.set_fusion('off')
        p0 = [sim.qubits.calc_probability(x) for x in q]
        print('[large] after the inverse: min P(0) =', min(p0), flush=True)
        assert min(p0) > 1. - tol, min(p0)
    finally:
        sim.terminate()
        sim = None
        gc.collect()


class _Capture:
    def __init__(self, prob, empty_lanes, qreg_ordering):
        self.prob = prob


def test_large_state_pools_and_marginals_34q(hip_module):
    """34 qubits, complex64 (128 GiB): bounded sampling pools over all 34 qubits (prefix sums
    only, 8 GiB) on a sparse state (1024 nonzero bins among 2^34: every bin centre returns its
    bin, random draws land on nonzero bins) and on the uniform state (2^20 bin centres spread
    over all 2^34 bins return their bins: the known answer of tests/test_sampling_pool.py:
    166-179 on a subset); then, after random gates, the 30-of-34 marginal (4 hidden qubits)
    against sums of the 16 amplitudes read from the state at 1024 seeded indices, and the
    1-of-34 marginal against P(qubit = 0)."""
    from qgate_amd.script import X
    n, dtype = 34, np.float32
    info = hip_module.device_info()
    if info['total_mem'] < (8 << n) + (24 << 30):
        pytest.skip('device memory {} < {} + headroom'.format(info['total_mem'], 8 << n))
    q = new_qregs(n)
    sim = Simulator(hip_module, circuit_prep='one_static', dtype=dtype)
    sim.qubits.set_ordering(list(q))
    hip_module.set_sampling_pool_workspace('bounded')
    try:
        print('[large34] sparse state', flush=True)
        sim.run([X(q[n - 1])] + [H(x) for x in q[:10]])
        nz = (1 << (n - 1)) + np.arange(1 << 10, dtype=np.int64)
        pool = sim.qubits.create_sampling_pool(list(q))
        u = (np.arange(1 << 10) + 0.5) / (1 << 10)
        assert np.array_equal(np.asarray(pool.sample(len(u), u).intarray), nz)
        u = np.random.RandomState(34).random_sample(1 << 16)
        u[:3] = [0., 1. - 2. ** -53, 1.]
        assert np.all(np.isin(np.asarray(pool.sample(len(u), u).intarray), nz))
        pool = None
        print('[large34] uniform state', flush=True)
        sim.run([X(q[n - 1])] + [H(x) for x in q[10:]])
        pool = sim.qubits.create_sampling_pool(list(q))
        j = np.arange(1 << 20, dtype=np.int64)
        k = (j << 14) + (j * 7919) % (1 << 14)
        u = (k.astype(np.float64) + 0.5) / float(1 << n)
        obs = np.asarray(pool.sample(len(u), u).intarray)
        assert np.array_equal(obs, k), int(np.count_nonzero(obs != k))
        pool = None
        hip_module.set_sampling_pool_workspace('auto')
        print('[large34] marginals', flush=True)
        fwd, _ = _random_and_inverse(q, 40, seed=35)
        sim.run(fwd)
        prob = np.asarray(sim.qubits.create_sampling_pool(list(q[:30]), _Capture).prob)
        assert prob.shape == (1 << 30,)
        assert abs(float(np.sum(prob, dtype=np.float64)) - 1.) < 1e-4
        xs = np.random.RandomState(36).randint(0, 1 << 30, 1024)
        worst = 0.
        for x in xs:
            amps = np.asarray(sim.qubits.states[int(x)::1 << 30], np.complex128)
            assert len(amps) == 16
            ref = float(np.sum(np.abs(amps) ** 2))
            worst = max(worst, abs(float(prob[x]) - ref) / max(ref, 1e-30))
        print('[large34] 30-of-34 marginal: max relative error at 1024 indices {:.3g}'.format(worst), flush=True)
        assert worst < 1e-5
        one = np.asarray(sim.qubits.create_sampling_pool([q[5]], _Capture).prob, np.float64)
        p0 = sim.qubits.calc_probability(q[5])
        assert abs(one[0] - p0) < 1e-6 and abs(one[0] + one[1] - 1.) < 1e-5, (one, p0)
    finally:
        hip_module.set_sampling_pool_workspace('auto')
        sim.terminate()
        sim = None
        gc.collect()
