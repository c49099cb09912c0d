"""Gate fusion in tile passes against the CPU oracle.

With fusion on, the HIP processor queues gates and applies them in tile passes (k_tile.hip).
Bars as for the unfused path: amplitudes within 1e-12 (complex128) / 5e-6 (complex64) of
the oracle, identical measurement outcomes. Tile shapes are varied (tiny tiles, tiles with one
contiguous bit, the defaults) so that gates land as in-tile pairs, in-tile and out-of-tile
diagonals, with controls inside and outside the tile; circuits interleave measurement,
probability, reset and readout, which must see every queued gate.
"""
import math

import numpy as np
import pytest

from qgate_amd import workloads
from qgate_amd.model import gate_type as gtype
from qgate_amd.script import new_qregs, new_references, measure, prob, reset, H
from qgate_amd.simulator.simulator import Simulator
from tests.helpers import max_abs, tol

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
SHAPES = [(0, -1), (4, 1), (6, 2), (8, 3), (10, 6), (5, 5)]   # (5, 5): no bit beyond the low ones


def _sim(module, ops, q, prep, dtype, seed=5):
    np.random.seed(seed)
    sim = Simulator(module, circuit_prep=prep, dtype=dtype)
    sim.qubits.set_ordering(list(q))
    sim.run(ops)
    return sim


class _Fusion:
    def __init__(self, module, *args):
        self.module, self.args = module, args

    def __enter__(self):
        self.module.set_fusion(*self.args)

    def __exit__(self, *exc):
        self.module.set_fusion('off')


def _mixed_circuit(n, n_gates, seed):
    """Every gate kind (adjoint or not) with 0..2 controls on random lanes."""
    from qgate_amd.model import ir
    rs = np.random.RandomState(seed)
    q = new_qregs(n)
    kinds = [(gtype.H, 0), (gtype.X, 0), (gtype.Y, 0), (gtype.Z, 0), (gtype.S, 0), (gtype.T, 0),
             (gtype.SH, 0), (gtype.ID, 0), (gtype.RX, 1), (gtype.RY, 1), (gtype.RZ, 1), (gtype.U1, 1),
             (gtype.U2, 2), (gtype.U, 3), (gtype.ExpiI, 1), (gtype.ExpiZ, 1)]
    ops = [H(x) for x in q]
    for _ in range(n_gates):
        cls, n_args = kinds[rs.randint(len(kinds))]
        g = ir.Gate(cls(*[float(v) for v in rs.uniform(0, 6.3, n_args)]))
        g.set_adjoint(bool(rs.randint(2)))
        n_ctrl = int(rs.choice([0, 0, 1, 1, 2]))
        lanes = [int(v) for v in rs.choice(n, n_ctrl + 1, replace=False)]
        g.set_qreg(q[lanes[0]])
        if n_ctrl:
            g.set_ctrllist([q[c] for c in lanes[1:]])
        ops.append(g)
    return q, ops


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('shape', SHAPES)
def test_every_gate_kind_fused(hip_module, port_module, dtype, shape):
    n = 13
    q, ops = _mixed_circuit(n, 400, seed=sum(shape) + 3)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    with _Fusion(hip_module, 'tile', *shape):
        got = _sim(hip_module, ops, q, 'one_static', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    print(shape, dtype.__name__, err)
    assert err < tol(dtype), err


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('shape', SHAPES)
def test_random_circuit_fused(hip_module, port_module, dtype, shape):
    n = 14
    q, ops = workloads.random_circuit(n, 400, seed=31)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    with _Fusion(hip_module, 'tile', *shape):
        got = _sim(hip_module, ops, q, 'one_static', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    print(shape, dtype.__name__, err)
    assert err < tol(dtype), err


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('prep', ['dynamic', 'one_static'])
def test_brickwork_and_qft_fused(hip_module, port_module, dtype, prep):
    n = 12
    q, ops = workloads.brickwork(n, 6, seed=3)
    ops += workloads.qft(q)
    ref = _sim(port_module, ops, q, prep, np.float64)
    with _Fusion(hip_module, 'tile'):
        got = _sim(hip_module, ops, q, prep, dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
        perr = max_abs(got.qubits.prob[:], ref.qubits.prob[:])
    assert err < tol(dtype) and perr < tol(dtype), (err, perr)


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('n', [1, 2, 3, 5, 8])
def test_small_states_one_pass(hip_module, port_module, dtype, n):
    q, ops = workloads.random_circuit(n, 60, seed=40 + n) if n >= 3 else workloads.brickwork(n, 5, seed=n)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    with _Fusion(hip_module, 'tile'):
        got = _sim(hip_module, ops, q, 'one_static', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    assert err < tol(dtype), err


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('prep', ['dynamic', 'one_static'])
def test_measure_reset_readout_see_queued_gates(hip_module, port_module, dtype, prep):
    """Measurements, probabilities and resets between gate runs flush the queue first."""
    n = 10
    q = new_qregs(n)
    refs = new_references(8)
    _, a = workloads.brickwork(n, 3, seed=9, qregs=q)
    _, b = workloads.random_circuit(n, 80, seed=10, qregs=q)
    _, c = workloads.random_circuit(n, 80, seed=11, qregs=q)
    ops = a + [measure(refs[0], q[2]), prob(refs[1], q[7])] + b + [measure(refs[2], q[0]), reset(q[0]),
                                                                      H(q[0])] + c + [measure(refs[3], q[9])]
    ref = _sim(port_module, ops, q, prep, np.float64, seed=77)
    with _Fusion(hip_module, 'tile', 8, 3):
        got = _sim(hip_module, ops, q, prep, dtype, seed=77)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
        values = got.values.get(refs[:4])
    expected = ref.values.get(refs[:4])
    assert [values[i] for i in (0, 2, 3)] == [expected[i] for i in (0, 2, 3)]   # measurements
    assert abs(values[1] - expected[1]) < tol(dtype)                           # prob
    assert err < tol(dtype), err


@pytest.mark.parametrize('dtype', DTYPES)
def test_pieces_with_fusion(hip_module, port_module, dtype):
    n = 16
    q, ops = workloads.random_circuit(n, 200, seed=12)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    hip_module.set_launch_piece(1 << 12)
    try:
        with _Fusion(hip_module, 'tile'):
            got = _sim(hip_module, ops, q, 'one_static', dtype)
            err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    finally:
        hip_module.set_launch_piece(1 << 30)
    assert err < tol(dtype), err


def test_fusion_mode_checked(hip_module):
    with pytest.raises(ValueError):
        hip_module.set_fusion('blocks')
    with pytest.raises(ValueError):
        hip_module.create_qubit_states(np.float64).processor.set_fusion('tile', 20, -1)


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('registers,bits', [(True, 0), (True, 13), (False, 0)])
def test_register_and_lds_tiles(hip_module, port_module, dtype, registers, bits):
    """Full-size tiles in registers (phases of up to 4 / 3 register-bit targets), the wide
    register tiles (13 / 12 bits, 512 threads; the default from 28 qubits, asked for here by
    width), and tiles in LDS."""
    n = 15
    q, ops = _mixed_circuit(n, 500, seed=77)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    hip_module.set_register_tiles(registers)
    try:
        with _Fusion(hip_module, 'tile', bits):
            got = _sim(hip_module, ops, q, 'one_static', dtype)
            err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    finally:
        hip_module.set_register_tiles(True)
    print('registers', registers, bits, dtype.__name__, err)
    assert err < tol(dtype), err


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('shape', [(0, -1), (6, 2)])
def test_single_lane_runs_merged(hip_module, port_module, dtype, shape):
    """Tile passes fold runs of uncontrolled gates on one lane into one 2x2 (FP64 product) and
    keep a real and a diagonal gate apart. Runs here are long and concentrated on three lanes
    (one below the contiguous tile bits, one in the middle, the top one): products that come
    back to the identity (H H, X X), diagonal (RZ S T U1), real (RY H) and general ones, cut
    by controlled gates whose controls or targets are those lanes."""
    from qgate_amd.model import ir
    n = 13
    rs = np.random.RandomState(11 + shape[0])
    q = new_qregs(n)
    hot = [0, 6, n - 1]
    one = [(gtype.H, 0), (gtype.X, 0), (gtype.S, 0), (gtype.T, 0), (gtype.Z, 0), (gtype.RZ, 1),
           (gtype.U1, 1), (gtype.RY, 1), (gtype.RX, 1), (gtype.U, 3), (gtype.SH, 0)]
    ops = [H(x) for x in q]
    for _ in range(600):
        if rs.randint(8) == 0:
            h = hot[rs.randint(3)]
            other = int(rs.choice([x for x in range(n) if x != h]))
            lanes = [h, other] if rs.randint(2) else [other, h]   # [target, control]
            g = ir.Gate(gtype.X())
            g.set_qreg(q[lanes[0]])
            g.set_ctrllist([q[lanes[1]]])
        else:
            cls, n_args = one[rs.randint(len(one))]
            g = ir.Gate(cls(*[float(v) for v in rs.uniform(0, 6.3, n_args)]))
            g.set_adjoint(bool(rs.randint(2)))
            g.set_qreg(q[hot[rs.randint(3)]])
        ops.append(g)
    ref = _sim(port_module, ops, q, 'one_static', np.float64)
    with _Fusion(hip_module, 'tile', *shape):
        got = _sim(hip_module, ops, q, 'one_static', dtype)
        err = max_abs(got.qubits.states[:], ref.qubits.states[:])
    print(shape, dtype.__name__, err)
    assert err < tol(dtype), err
