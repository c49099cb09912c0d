"""Dense k-qubit unitaries: qghip_qubit_processor_apply_unitary and the 'dense' fusion mode.

apply_unitary is checked against the same unitary applied in numpy (FP64) to the state read
back before it: within 1e-12 (complex128) and 5e-6 (complex64, the matrix-core kernel for 5 and
6 lanes accumulates in FP32). Lane sets are low, high, spread and unsorted (lanes[j] is bit j
of the matrix index), on states large enough for the matrix-core kernel and just too small for
it. The fusion mode, which builds each block's unitary from its gates in FP64 on the host, is
checked against the CPU oracle on the same circuits with the bars of the unfused path.
"""
import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.model import gate_type as gtype
from qgate_amd.simulator import qubits
from tests.helpers import max_abs, tol
from tests.test_fusion import _Fusion, _mixed_circuit, _sim

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def _random_unitary(k, rs):
    z = rs.normal(size=(1 << k, 1 << k)) + 1j * rs.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _prepared(module, n, dtype, seed):
    """A state of n lanes with every amplitude nonzero (U3 on every lane, CX chains)."""
    rs = np.random.RandomState(seed)
    qs = module.create_qubit_states(dtype)
    proc = qs.processor
    proc.initialize_qubit_states(qs, n)
    proc.reset_qubit_states(qs)
    for _ in range(2):
        for lane in range(n):
            proc.apply_gate(gtype.U(*rs.uniform(0, 6.3, 3)), False, qs, lane)
        for lane in range(n - 1):
            proc.apply_controlled_gate(gtype.X(), False, qs, [lane], lane + 1)
    return qs


def _read(module, qs, n, dtype):
    """The amplitudes, index bit i = lane i."""
    from qgate_amd.simulator.lanes import Lane
    lanes = []
    for i in range(n):
        ln = Lane(None, i)
        ln.set_external(i)
        lanes.append(ln)
    out = np.empty(1 << n, np.complex128 if dtype == np.float64 else np.complex64)
    module.create_qubits_states_getter(dtype).get_states(out, 0, qubits.null, [(qs, lanes)], [], 1 << n, 0, 1)
    return out.astype(np.complex128)


def _apply_numpy(psi, n, lanes, u):
    k = len(lanes)
    rest = [b for b in range(n) if b not in lanes]
    grp = np.arange(1 << (n - k))
    base = np.zeros_like(grp)
    for i, b in enumerate(rest):
        base |= ((grp >> i) & 1) << b
    member = np.array([sum(((c >> j) & 1) << lanes[j] for j in range(k)) for c in range(1 << k)])
    idx = base[:, None] | member[None, :]
    out = psi.copy()
    out[idx] = psi[idx] @ u.T
    return out


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('n,lanes', [
    (12, [0]), (12, [11, 3]), (12, [0, 1, 2]), (12, [9, 2, 5, 7]),
    (12, [0, 1, 2, 3, 4]), (12, [11, 10, 9, 8, 7]), (12, [3, 8, 0, 11, 5]),
    (12, [0, 1, 2, 3, 4, 5]), (12, [6, 7, 8, 9, 10, 11]), (12, [10, 1, 7, 4, 0, 9]),
    (10, [9, 0, 4, 6, 2, 7]),   # 16 groups: one matrix-core task
    (9, [8, 0, 4, 6, 2, 7]),    # 8 groups: the vector kernel
    (9, [1, 3, 5, 7, 8]),
    # several LDS-staged tiles (blocks of 2..4 lanes, complex128 also 5): rows below the lowest
    # block lane, block lanes above and below bit 6
    (16, [15, 2]), (16, [13, 6, 0, 9]), (15, [14, 12, 1, 3, 7]), (16, [4, 5, 6]),
    # matrix cores over many tasks (complex128 on the FP64 matrix cores too)
    (16, [15, 2, 9, 4, 11, 0]), (17, [16, 3, 8, 12, 1]),
    # complex128 4- and 5-lane blocks on matrix-core tiles of 2^11 / 2^10 amplitudes: one
    # tile, the block at the bottom of every tile, the block split between the lowest and the
    # highest lanes
    (10, [9, 0, 4, 6, 2]), (14, [0, 1, 2, 3, 4]), (14, [13, 12, 0, 1, 2]),
    (11, [10, 1, 5, 3]), (14, [0, 1, 2, 3]), (13, [12, 0, 11, 1]),
])
def test_apply_unitary_matches_numpy(hip_module, dtype, n, lanes):
    rs = np.random.RandomState(len(lanes) * 31 + n)
    qs = _prepared(hip_module, n, dtype, seed=n + len(lanes))
    before = _read(hip_module, qs, n, dtype)
    u = _random_unitary(len(lanes), rs)
    qs.processor.apply_unitary(qs, lanes, u)
    after = _read(hip_module, qs, n, dtype)
    want = _apply_numpy(before, n, lanes, u)
    err = np.max(np.abs(after - want))
    print(dtype.__name__, n, lanes, err)
    assert err < (1e-12 if dtype == np.float64 else 5e-6), err
    qs.delete()


def test_apply_unitary_checks_arguments(hip_module):
    qs = _prepared(hip_module, 8, np.float32, seed=1)
    with pytest.raises(ValueError):
        qs.processor.apply_unitary(qs, [0, 0], np.eye(4))
    with pytest.raises(ValueError):
        qs.processor.apply_unitary(qs, [0, 1], np.eye(2))
    with pytest.raises(ValueError):
        qs.processor.apply_unitary(qs, [0, 1, 2, 3, 4, 5, 6], np.eye(128))
    qs.delete()


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('width', [2, 3, 4, 5, 6])
def test_dense_fusion_random_circuit(hip_module, port_module, dtype, width):
    n = 14
    q, ops = workloads.random_circuit(n, 400, seed=37)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    with _Fusion(hip_module, 'dense', width):
        got = _sim(hip_module, ops, q, 'one_static', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    print(width, dtype.__name__, err)
    assert err < tol(dtype), err


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('width', [2, 4, 6])
def test_dense_fusion_every_gate_kind(hip_module, port_module, dtype, width):
    """Every gate kind with 0..2 controls (a doubly controlled gate is wider than a 2-lane
    block and runs alone); probabilities as well as amplitudes."""
    n = 13
    q, ops = _mixed_circuit(n, 400, seed=width + 11)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    with _Fusion(hip_module, 'dense', width):
        got = _sim(hip_module, ops, q, 'one_static', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
        perr = max_abs(got.qubits.prob[:], ref.qubits.prob[:])
    assert err < tol(dtype) and perr < tol(dtype), (err, perr)


@pytest.mark.parametrize('dtype', DTYPES)
def test_dense_fusion_brickwork_qft_dynamic(hip_module, port_module, dtype):
    n = 12
    q, ops = workloads.brickwork(n, 6, seed=4)
    ops += workloads.qft(q)
    ref = _sim(port_module, ops, q, 'dynamic', np.float64)
    with _Fusion(hip_module, 'dense', 5):
        got = _sim(hip_module, ops, q, 'dynamic', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    assert err < tol(dtype), err


def test_dense_mode_checked(hip_module):
    qs = hip_module.create_qubit_states(np.float32)
    proc = qs.processor
    for bad in ((1, -1), (7, -1)):
        with pytest.raises(ValueError):
            proc.set_fusion('dense', *bad)
    proc.set_fusion('dense', 0, -1)
    proc.set_fusion('off')
    qs.delete()
