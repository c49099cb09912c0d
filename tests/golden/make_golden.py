"""Writes the golden vectors of tests/golden/ from the reference's own CPU runtime.

The vectors come from the reference's C++ CPU core (qgate/simulator/src/CPU*.cpp and the
files they include), compiled from the sources where they lie by oracle/Makefile into
oracle/_ref/libqgate_ref.so and driven through this repository's qgate front-end (the
reference's Python package is not importable here: it needs `ply`). Each case stores the
final state vector, the probability array, the sampling-pool draws for seeded uniforms and
the measurement outcomes for seeded np.random draws, all in the reference's FP64 or FP32
arithmetic. manifest.json records each file's SHA-256 and how it was made.

    python tests/golden/make_golden.py     (needs the reference sources under /root/reference)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from qgate_amd import workloads                                    # noqa: E402
from qgate_amd.script import new_qregs, new_references, measure    # noqa: E402
from qgate_amd.simulator.simulator import Simulator                # noqa: E402


def _brickwork(n, layers, seed):
    def make():
        return workloads.brickwork(n, layers, seed)
    return make


def _random(n, gates, seed):
    def make():
        return workloads.random_circuit(n, gates, seed)
    return make


def _qft_random(n, gates, seed):
    def make():
        return workloads.qft_random(n, gates, seed)
    return make


# name: (circuit, prep, dtype, pool ordering (None: all qubits), draws, seed of the uniforms,
#        measure at the end: True for every qubit, or a list of qubit indices, in that order)
CASES = {
    'brickwork10_c128_dynamic': (_brickwork(10, 20, 1), 'dynamic', np.float64, None, 4096, 1001, False),
    'brickwork10_c128_static': (_brickwork(10, 20, 1), 'one_static', np.float64, None, 4096, 1001, False),
    'brickwork10_c64_static': (_brickwork(10, 20, 1), 'one_static', np.float32, None, 4096, 1001, False),
    'random12_c128': (_random(12, 300, 2), 'one_static', np.float64, None, 4096, 1002, True),
    'random12_c64': (_random(12, 300, 2), 'one_static', np.float32, None, 4096, 1002, True),
    'qft_random10_c128': (_qft_random(10, 50, 3), 'one_static', np.float64, None, 4096, 1003, False),
    'random11_c128_hidden': (_random(11, 200, 4), 'dynamic', np.float64, [9, 2, 5, 0, 7], 4096, 1004, False),
    # dynamic preparation with measurements of some qubits: each Measure is followed by a
    # Separate, so these pin the collapse into a smaller state vector at 0 < p < 1 (the
    # unmeasured qubits keep the rescaled amplitudes)
    'brickwork10_c128_dynamic_measure': (_brickwork(10, 20, 5), 'dynamic', np.float64, None, 4096, 1005,
                                         [7, 2, 4, 9, 0]),
    'brickwork10_c64_dynamic_measure': (_brickwork(10, 20, 5), 'dynamic', np.float32, None, 4096, 1005,
                                        [7, 2, 4, 9, 0]),
    'random11_c128_dynamic_measure': (_random(11, 150, 6), 'dynamic', np.float64, None, 4096, 1006,
                                      [1, 10, 5, 6]),
}


def run_case(module, name):
    """Outputs of one case on a runtime module (also used by the tests)."""
    make, prep, dtype, ordering, draws, useed, measure_all = CASES[name]
    q, ops = make()
    sim = Simulator(module, circuit_prep=prep, dtype=dtype)
    sim.qubits.set_ordering(list(q))
    sim.run(ops)
    out = {'states': np.array(sim.qubits.states[:]), 'probs': np.array(sim.qubits.prob[:])}
    pool_q = list(q) if ordering is None else [q[i] for i in ordering]
    u = np.random.RandomState(useed).random_sample(draws)
    out['uniforms'] = u
    out['samples'] = np.array(sim.qubits.create_sampling_pool(pool_q).sample(len(u), u).intarray, np.int64)
    if measure_all:   # a second run of the same circuit followed by measurements
        q, ops = make()
        which = list(range(len(q))) if measure_all is True else list(measure_all)
        refs = new_references(len(which))
        msim = Simulator(module, circuit_prep=prep, dtype=dtype)
        msim.qubits.set_ordering(list(q))
        np.random.seed(useed + 1)
        msim.run(ops + [measure(refs[k], q[i]) for k, i in enumerate(which)])
        out['outcomes'] = np.array(msim.values.get(refs), np.int64)
        out['states_after_measure'] = np.array(msim.qubits.states[:])
        msim.terminate()
    sim.terminate()
    return out


def main():
    from oracle import cpuruntime
    cpuruntime.build(reference=True)
    if not cpuruntime.have_reference():
        raise SystemExit('the reference core is not built (no reference sources?)')
    module = cpuruntime.reference()
    manifest = {'made_by': 'tests/golden/make_golden.py with oracle/_ref/libqgate_ref.so '
                           '(reference qgate/simulator/src CPU core, g++ -O2 -ffp-contract=off)',
                'files': {}}
    for name in CASES:
        out = run_case(module, name)
        path = os.path.join(HERE, name + '.npz')
        np.savez_compressed(path, **out)
        with open(path, 'rb') as f:
            digest = hashlib.sha256(f.read()).hexdigest()
        manifest['files'][name + '.npz'] = {'sha256': digest, 'keys': sorted(out)}
        print(name, {k: v.shape for k, v in out.items()})
    with open(os.path.join(HERE, 'manifest.json'), 'w') as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == '__main__':
    main()
