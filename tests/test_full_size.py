"""The benchmarked workloads at their full sizes, checked by properties that hold at any size.

No CPU oracle finishes a 30-qubit, 2000-gate circuit in test time, so the full-size checks are:

* 30-qubit random circuit of 2000 gates from {U3, H, Rz, CX, CCX}, seed 2 (the gate
  benchmark's workload), complex64 and complex128, with fusion off, in tile passes and in dense
  blocks of k = 2..6 qubits: the circuit followed by its inverse returns to |0...0> (every
  P(qubit = 0) is 1); after the forward half, a strided slice of 2^20 amplitudes agrees with
  the unfused run (5e-6 complex64, 1e-11 complex128) and so does P(qubit = 0) of every qubit
  (2e-5 complex64: a sum over 2^29 complex64 probabilities carries the FP32 norm drift of
  2000 gates, about sqrt(2000) x 2^-24 x a few; 1e-11 complex128); the complex64 state
  against the complex128 state of the same circuit is the error figure printed per mode.
* The end of that workload: a sampling pool over all 30 qubits and 2^20 seeded draws, each
  draw checked against the cumulative sum of the same FP64 |a|^2 on the host, accumulated in
  extended precision (the draw is the first bin whose normalised prefix exceeds the uniform;
  the device's FP64 sums differ from it by rounding only, so a draw may land in the
  neighbouring bin only within 1e-12 of an edge, and such draws are counted);
  then a measurement of
  every qubit with seeded uniforms, each outcome checked against the conditional probability
  from the same host array, and the collapsed state checked to be the measured basis state.
* The 24-qubit, 12-layer RY/RZ/CZ-ring ansatz (576 parameters): the complex64 adjoint
  gradient against the complex128 one on all 576 entries (1e-4), central differences with
  h = 1e-6 of complex128 forwards (1e-7) and h = 1e-3 (1e-4 against the complex64
  gradient), and a parameter-shift gradient from the oracle (the reference CPU runtime
  restated) on seeded entries.
* 4096 batched ansatz circuits of 8..20 qubits (10 layers, seed 5: the batch benchmark's
  workload, tuned by the autotuner) against each circuit run alone, every item.
"""
import gc
import math

import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.script import new_qregs, new_reference, measure, U3, H, Rz, ctrl
from qgate_amd.simulator.simulator import Simulator

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N = 30
N_GATES = 2000
SEED = 2
FUSIONS = [('off', 0), ('tile', 0), ('dense', 2), ('dense', 3), ('dense', 4), ('dense', 5), ('dense', 6)]
SLICE_STEP = (1 << 10) + 1          # 2^20 amplitudes spread over the whole state


def _need_memory(hip_module, nbytes):
    info = hip_module.device_info()
    if info['total_mem'] < nbytes + (16 << 30):
        pytest.skip('device memory {} < {} + headroom'.format(info['total_mem'], nbytes))


def _inverse_ops(gates, q):
    ops = []
    for kind, qubits, angles in reversed(gates):
        if kind == 'U3':
            ops.append(U3(*angles).Adj(q[qubits[0]]))
        elif kind == 'H':
            ops.append(H(q[qubits[0]]))
        elif kind == 'Rz':
            ops.append(Rz(angles[0]).Adj(q[qubits[0]]))
        elif kind == 'CX':
            ops.append(ctrl(q[qubits[0]]).X(q[qubits[1]]))
        else:
            ops.append(ctrl(q[qubits[0]], q[qubits[1]]).X(q[qubits[2]]))
    return ops


def _forward(hip_module, dtype, fusion):
    """Simulator with the 30-qubit workload applied under `fusion`."""
    gates = workloads.random_gate_list(N, N_GATES, SEED)
    q = new_qregs(N)
    hip_module.set_fusion(*fusion)
    sim = Simulator(hip_module, circuit_prep='one_static', dtype=dtype)
    sim.qubits.set_ordering(list(q))
    sim.run(workloads.to_ops(gates, q))
    return sim, q, gates


def _summary(sim, q):
    p0 = np.array([sim.qubits.calc_probability(x) for x in q])
    amps = np.asarray(sim.qubits.states[::SLICE_STEP], np.complex128)
    return p0, amps


def _close_sim(sim, hip_module):
    hip_module.set_fusion('off')
    sim.terminate()
    gc.collect()


@pytest.mark.parametrize('dtype', [np.float32, np.float64], ids=['c64', 'c128'])
@pytest.mark.parametrize('fusion', FUSIONS, ids=['{}{}'.format(m, k or '') for m, k in FUSIONS])
def test_random30_round_trip(hip_module, dtype, fusion):
    amp = 8 if dtype == np.float32 else 16
    _need_memory(hip_module, amp << N)
    tol_amp = 5e-6 if dtype == np.float32 else 1e-11
    tol_p0 = 2e-5 if dtype == np.float32 else 1e-11
    tol_p = 1e-4 if dtype == np.float32 else 1e-10
    print('[full] random30 {} fusion {}'.format(np.dtype(dtype).name, fusion), flush=True)
    ref = None
    if fusion[0] != 'off':
        sim, q, _ = _forward(hip_module, dtype, ('off', 0))
        ref = _summary(sim, q)
        _close_sim(sim, hip_module)
    sim, q, gates = _forward(hip_module, dtype, fusion)
    try:
        p0, amps = _summary(sim, q)
        assert abs(float(np.sum(np.abs(amps) ** 2)) * SLICE_STEP - 1.) < 0.05   # a spread slice
        if ref is not None:
            dp = float(np.max(np.abs(p0 - ref[0])))
            da = float(np.max(np.abs(amps - ref[1])))
            print('[full] vs unfused: max |dP0| = {:.3g}, max |d amp| = {:.3g}'.format(dp, da), flush=True)
            assert da < tol_amp and dp < tol_p0, (da, dp)
        sim.run(_inverse_ops(gates, q))
        back = np.array([sim.qubits.calc_probability(x) for x in q])
        dev = float(np.max(1. - back))
        print('[full] after the inverse: max (1 - P0) = {:.3g}'.format(dev), flush=True)
        assert dev < tol_p, dev
    finally:
        _close_sim(sim, hip_module)


def test_random30_complex64_error_vs_complex128(hip_module):
    """The complex64 state of the workload against the complex128 state, per fusion mode, on
    a strided slice of 2^20 amplitudes (the error figure of the complex64 path at 30 qubits)."""
    _need_memory(hip_module, 24 << N)
    sim, q, _ = _forward(hip_module, np.float64, ('off', 0))
    ref_p0, ref_amps = _summary(sim, q)
    _close_sim(sim, hip_module)
    for fusion in (('off', 0), ('tile', 0), ('dense', 6)):
        sim, q, _ = _forward(hip_module, np.float32, fusion)
        p0, amps = _summary(sim, q)
        _close_sim(sim, hip_module)
        da = float(np.max(np.abs(amps - ref_amps)))
        dp = float(np.max(np.abs(p0 - ref_p0)))
        print('[full] complex64 vs complex128, fusion {}: max |d amp| = {:.3g}, max |dP0| = {:.3g}'
              .format(fusion, da, dp), flush=True)
        assert da < 5e-6 and dp < 2e-5, (fusion, da, dp)


def test_random30_pool_and_measure_all(hip_module):
    """2^20 pool draws and a measurement of every qubit at the end of the workload (complex64),
    each checked against the host's FP64 prefix sums of the probability array."""
    _need_memory(hip_module, 8 << N)
    sim, q, _ = _forward(hip_module, np.float32, ('off', 0))
    try:
        amps = np.asarray(sim.qubits.states[:])                 # complex64
        probs = np.square(amps.real, dtype=np.float64) + np.square(amps.imag, dtype=np.float64)
        del amps
        cdf = np.cumsum(probs, dtype=np.longdouble)
        del probs
        total = float(cdf[-1])
        assert abs(total - 1.) < 1e-4
        cdf /= cdf[-1]
        pool = sim.qubits.create_sampling_pool(list(q))
        u = np.random.RandomState(1020).random_sample(1 << 20)
        obs = np.asarray(pool.sample(len(u), u).intarray, np.int64)
        pool = None
        assert obs.min() >= 0 and obs.max() < (1 << N)
        lo = np.where(obs > 0, cdf[np.maximum(obs - 1, 0)], 0.).astype(np.float64)
        hi = cdf[obs].astype(np.float64)
        assert np.all(hi > lo)                       # every draw on a bin with p > 0
        bad = ~((lo <= u) & (u < hi))
        near = np.minimum(np.abs(u - lo), np.abs(u - hi)) < 1e-12
        print('[full] pool: {} draws, {} outside their bin, all of them within 1e-12 of an edge: {}'
              .format(len(u), int(np.count_nonzero(bad)), bool(np.all(near[bad]))), flush=True)
        assert not np.any(bad & ~near)
        # monotone: sorted uniforms give non-decreasing bins (upper_bound on one CDF)
        order = np.argsort(u, kind='stable')
        assert np.all(np.diff(obs[order]) >= 0)

        del cdf
        amps = np.asarray(sim.qubits.states[:])
        probs = np.square(amps.real, dtype=np.float64) + np.square(amps.imag, dtype=np.float64)
        del amps
        # measurement of every qubit in order, with seeded uniforms
        refs = [new_reference() for _ in q]
        np.random.seed(1021)
        draws = np.random.RandomState(1021).random_sample(N)
        sim.run([measure(r, x) for r, x in zip(refs, q)])
        outcomes = [sim.values.get(r) for r in refs]
        arr = probs
        ties = 0
        for i in range(N):
            cond0 = float(np.sum(arr[0::2])) / float(np.sum(arr))
            expect = 0 if draws[i] < cond0 else 1
            if abs(draws[i] - cond0) < 1e-5:
                ties += 1
            else:
                assert outcomes[i] == expect, (i, draws[i], cond0, outcomes[i])
            arr = arr[outcomes[i]::2]
        print('[full] measure-all: outcomes {}, near ties {}'.format(''.join(map(str, outcomes)), ties),
              flush=True)
        after = np.array([sim.qubits.calc_probability(x) for x in q])
        assert np.max(np.abs(after - (1. - np.array(outcomes, np.float64)))) < 1e-5
        index = sum(int(b) << i for i, b in enumerate(outcomes))
        assert abs(abs(complex(sim.qubits.states[index])) - 1.) < 1e-5
    finally:
        _close_sim(sim, hip_module)


# ---- 24-qubit ansatz: adjoint gradient ---------------------------------------------------

def _zz_x_expectation(psi, a, b):
    """<psi| sum_i a_i Z_i Z_i+1 + sum_i b_i X_i |psi> by reshaping (qubit i = bit i)."""
    n = int(round(math.log2(len(psi))))
    p = np.abs(psi) ** 2
    value = 0.
    idx = np.arange(len(psi), dtype=np.int64)
    for i in range(n - 1):
        parity = ((idx >> i) ^ (idx >> (i + 1))) & 1
        value += a[i] * float(np.sum(p * (1. - 2. * parity)))
    for i in range(n):
        v = psi.reshape(-1, 2, 1 << i)
        value += b[i] * 2. * float(np.real(np.vdot(v[:, 0, :], v[:, 1, :])))
    return value


def _hea24():
    from qgate_amd.adjoint import PauliSum
    from tests import adjoint_checks as ac
    n, layers = 24, 12
    spec = ac.hea_spec(n, layers)
    a = np.random.RandomState(1004).normal(size=n - 1)
    b = np.random.RandomState(1005).normal(size=n)
    params = np.random.RandomState(4).uniform(-math.pi, math.pi, spec.n_params)
    return spec, PauliSum.zz_x(a, b), params, a, b


def test_hea24_gradients_full(hip_module):
    spec, h, params, _, _ = _hea24()
    c64, c128 = spec.build(np.float32), spec.build(np.float64)
    v32, g32 = c64.gradient(params, h)
    v64, g64 = c128.gradient(params, h)
    dg = float(np.max(np.abs(g32 - g64)))
    print('[full] hea24: <H> c64 - c128 = {:.3g}; max |grad c64 - grad c128| over {} entries = {:.3g}'
          .format(v32 - v64, len(g64), dg), flush=True)
    assert abs(v32 - v64) < 1e-4 and dg < 1e-4
    subset = np.random.RandomState(45).choice(spec.n_params, 8, replace=False)
    err64 = err32 = err32own = 0.
    for k in subset:
        for hstep, circ, which in ((1e-6, c128, 'fp64'), (1e-3, c128, 'fp32'), (1e-3, c64, 'fp32own')):
            pp, pm = params.copy(), params.copy()
            pp[k] += hstep
            pm[k] -= hstep
            fd = (circ.expectation(pp, h) - circ.expectation(pm, h)) / (2. * hstep)
            if which == 'fp64':
                err64 = max(err64, abs(g64[k] - fd))
            elif which == 'fp32':
                err32 = max(err32, abs(g32[k] - fd))
            else:
                err32own = max(err32own, abs(g32[k] - fd))
    print('[full] hea24 central differences: complex128 h=1e-6 {:.3g}; complex64 gradient vs '
          'complex128 differences h=1e-3 {:.3g}; vs complex64 differences h=1e-3 {:.3g}'
          .format(err64, err32, err32own), flush=True)
    assert err64 < 1e-7 and err32 < 1e-4
    c64.delete()
    c128.delete()


def test_hea24_gradient_vs_oracle_parameter_shift(hip_module, port_module):
    """Parameter shift from oracle forwards (the reference CPU runtime restated, threaded) on
    4 seeded entries of the 576."""
    from tests import adjoint_checks as ac
    spec, h, params, a, b = _hea24()
    _, g64 = spec.build(np.float64).gradient(params, h)
    _, g32 = spec.build(np.float32).gradient(params, h)
    rot_pos = {op[2]: pos for pos, op in enumerate(spec.ops) if op[0] != 'G'}
    ks = np.random.RandomState(46).choice(spec.n_params, 4, replace=False)
    e64 = e32 = 0.
    for k in ks:
        pos = rot_pos[int(k)]
        vals = []
        for s in (math.pi / 2, -math.pi / 2):
            psi = ac.oracle_state(port_module, spec, params, shifts={pos: s})
            vals.append(_zz_x_expectation(psi, a, b))
        ref = (vals[0] - vals[1]) / 2.
        e64 = max(e64, abs(g64[k] - ref))
        e32 = max(e32, abs(g32[k] - ref))
    print('[full] hea24 vs oracle parameter shift: complex128 {:.3g}, complex64 {:.3g}'.format(e64, e32),
          flush=True)
    assert e64 < 1e-9 and e32 < 1e-4


# ---- 4096 batched circuits -----------------------------------------------------------------

def test_batch4096_every_item_against_single_circuits(hip_module):
    from qgate_amd.adjoint import PauliSum, hea_circuit
    from qgate_amd.batch import AdjointBatch
    layers = 10
    rs = np.random.RandomState(5)
    sizes = rs.randint(8, 21, 4096)
    circuits = {int(n): hea_circuit(int(n), layers, np.float32) for n in sorted(set(sizes))}
    circuits64 = {int(n): hea_circuit(int(n), layers, np.float64) for n in sorted(set(sizes))}
    batch = AdjointBatch(np.float32)
    items = []
    for n in sizes:
        n = int(n)
        params = rs.uniform(-math.pi, math.pi, 2 * n * layers)
        h = PauliSum.zz_x(rs.normal(size=n - 1), rs.normal(size=n))
        batch.add(circuits[n], params, h)
        items.append((n, params, h))
    values, grads = batch.run()
    tuning = batch.tuning
    batch.delete()
    assert {g['qubits'] for g in tuning} == set(range(8, 21))
    worst_v = worst_g = worst64_v = worst64_g = 0.
    for (n, params, h), v, g in zip(items, values, grads):
        rv, rg = circuits[n].gradient(params, h)
        worst_v = max(worst_v, abs(v - rv))
        worst_g = max(worst_g, float(np.max(np.abs(g - rg))))
        rv64, rg64 = circuits64[n].gradient(params, h)
        worst64_v = max(worst64_v, abs(v - rv64))
        worst64_g = max(worst64_g, float(np.max(np.abs(g - rg64))))
    print('[full] batch 4096: vs single complex64 circuits value {:.3g} grad {:.3g}; '
          'vs complex128 value {:.3g} grad {:.3g}'.format(worst_v, worst_g, worst64_v, worst64_g), flush=True)
    for c in list(circuits.values()) + list(circuits64.values()):
        c.delete()
    assert worst_v < 2e-5 and worst_g < 2e-4
    assert worst64_v < 1e-4 and worst64_g < 1e-4


# ---- complex64 error against the FP64 oracle, per workload ---------------------------------

def test_complex64_error_vs_oracle_per_config(hip_module, port_module):
    """The complex64 path against the FP64 oracle (the reference CPU runtime restated) on the
    same circuits: the 16-qubit, 20-layer brickwork of the CPU workload at full size, and the
    2000-gate random circuit of the gate benchmark at 22 qubits (the largest the oracle runs in
    seconds) with fusion off, in tile passes and in 6-qubit dense blocks. Max |amplitude error|
    (north_star bar 5e-6) and max |P(qubit = 0) error| are printed per case."""
    cases = [('brickwork16x20', lambda: workloads.brickwork(16, 20, seed=1), [('off', 0)])]
    q22 = new_qregs(22)
    cases.append(('random22x2000', lambda: (q22, workloads.to_ops(workloads.random_gate_list(22, 2000, 2), q22)),
                  [('off', 0), ('tile', 0), ('dense', 6)]))
    for name, make, fusions in cases:
        q, ops = make()
        ref = Simulator(port_module, circuit_prep='one_static', dtype=np.float64)
        ref.qubits.set_ordering(list(q))
        ref.run(ops)
        want = np.asarray(ref.qubits.states[:], np.complex128)
        want_p0 = np.array([ref.qubits.calc_probability(x) for x in q])
        ref.terminate()
        for fusion in fusions:
            hip_module.set_fusion(*fusion)
            sim = Simulator(hip_module, circuit_prep='one_static', dtype=np.float32)
            sim.qubits.set_ordering(list(q))
            try:
                sim.run(ops)
                got = np.asarray(sim.qubits.states[:], np.complex128)
                p0 = np.array([sim.qubits.calc_probability(x) for x in q])
            finally:
                _close_sim(sim, hip_module)
            da = float(np.max(np.abs(got - want)))
            dp = float(np.max(np.abs(p0 - want_p0)))
            print('[error] {} complex64 fusion {}: max |amp - FP64 oracle| = {:.3g} (max |amp| {:.3g}), '
                  'max |P0 - oracle| = {:.3g}'.format(name, fusion, da, float(np.max(np.abs(want))), dp), flush=True)
            assert da < 5e-6 and dp < 5e-5, (name, fusion, da, dp)
