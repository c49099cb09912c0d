"""Golden vectors written by the reference's own C++ CPU core (tests/golden/make_golden.py).

CPU: the oracle (the C restatement of the reference CPU runtime) reproduces every golden
vector bit for bit, in FP64 and FP32: states, probability arrays, sampling-pool draws and
measurement outcomes; where the reference core is built here (oracle/_ref), it reproduces
them too, which shows the fixtures are current. The fixture files match the manifest's
SHA-256.

GPU: the HIP runtime matches the same vectors within 1e-12 (complex128) / 5e-6 (complex64)
per amplitude and probability, with identical measurement outcomes, and identical sampled
indices except where a uniform lies within the prefix-sum rounding of a bin edge.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from tests.golden import make_golden
from tests.helpers import TOL_C128, TOL_C64

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CASES = sorted(make_golden.CASES)


def _golden(name):
    with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_manifest_hashes():
    with open(os.path.join(HERE, 'manifest.json')) as f:
        manifest = json.load(f)
    assert sorted(manifest['files']) == sorted(c + '.npz' for c in CASES)
    for fname, info in manifest['files'].items():
        with open(os.path.join(HERE, fname), 'rb') as f:
            assert hashlib.sha256(f.read()).hexdigest() == info['sha256'], fname


@pytest.mark.parametrize('name', CASES)
def test_oracle_reproduces_golden(port_module, name):
    gold = _golden(name)
    out = make_golden.run_case(port_module, name)
    assert sorted(out) == sorted(gold)
    for k in gold:
        assert out[k].dtype == gold[k].dtype, k
        assert np.array_equal(out[k], gold[k]), k


@pytest.mark.parametrize('name', CASES)
def test_reference_core_reproduces_golden(oracle, name):
    if not oracle.have_reference():
        pytest.skip('reference core not built (no reference sources on this machine)')
    gold = _golden(name)
    out = make_golden.run_case(oracle.reference(), name)
    for k in gold:
        assert np.array_equal(out[k], gold[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize('name', CASES)
def test_hip_matches_golden(hip_module, name):
    gold = _golden(name)
    out = make_golden.run_case(hip_module, name)
    dtype = make_golden.CASES[name][2]
    tol = TOL_C128 if dtype == np.float64 else TOL_C64
    for k in ('states', 'probs', 'states_after_measure'):
        if k in gold:
            err = float(np.max(np.abs(out[k].astype(np.complex128) - gold[k].astype(np.complex128))))
            print(name, k, 'max abs err', err)
            assert err < tol, (k, err)
    if 'outcomes' in gold:
        assert np.array_equal(out['outcomes'], gold['outcomes'])
    bad = np.nonzero(out['samples'] != gold['samples'])[0]
    print(name, 'sample mismatches', len(bad))
    if len(bad):
        # every differing draw must sit within the rounding of the CDF at the edge between the
        # two bins it fell into
        make, prep, dtype, ordering, draws, useed, _ = make_golden.CASES[name]
        ordering_probs = _pool_probs(gold, ordering)
        cdf = np.cumsum(ordering_probs)
        cdf /= cdf[-1]
        u = gold['uniforms']
        edge_tol = 1e-5 if dtype == np.float32 else 1e-12
        for i in bad:
            lo, hi = sorted((int(out['samples'][i]), int(gold['samples'][i])))
            assert hi == lo + 1 or np.all(ordering_probs[lo + 1:hi] == 0.), (i, lo, hi)
            assert abs(u[i] - cdf[lo]) < edge_tol, (i, u[i], cdf[lo])
        assert len(bad) <= 8


def _pool_probs(gold, ordering):
    """Probabilities of the pool's bins (qubits in `ordering`, the rest summed over)."""
    p = gold['probs'].astype(np.float64)
    if ordering is None:
        return p
    n = int(np.log2(len(p)))
    t = p.reshape([2] * n)                     # axis k <-> bit n-1-k
    keep_axes = [n - 1 - b for b in ordering]  # output bit j = qubit ordering[j]
    drop = tuple(ax for ax in range(n) if ax not in keep_axes)
    t = t.sum(axis=drop)
    remaining = [ax for ax in range(n) if ax not in drop]
    # transpose so that the last axis is output bit 0
    order = [remaining.index(keep_axes[j]) for j in reversed(range(len(ordering)))]
    return np.transpose(t, order).reshape(-1)
