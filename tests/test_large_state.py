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
