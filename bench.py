"""Benchmark: gate HBM throughput of the HIP state-vector runtime on one MI355X.

Workload (one step): the 2000-gate random circuit on 30 qubits, complex64 — gate kinds drawn
uniformly from {U3, H, Rz, CX, CCX} with seed 2 (qgate_amd.workloads.random_circuit) —
applied to a state vector already resident in HBM, through the runtime-module calls the
simulator front-end makes for it (processor.apply_gate / apply_controlled_gate). After the
timed steps the same state is used for the end-of-circuit work of the workload (sampling
pool + 2^20 draws, per-qubit measurement), timed separately and reported as extra fields.

value = algorithmic HBM bytes of the step / step time (GB/s). Algorithmic bytes per gate:
uncontrolled 2x2 gate 2 S 2^n, c-controlled 2 S 2^(n-c), diagonal gates with m00 = 1
(only the |1> half changes) S 2^n / 2^c; S = 8 (complex64).

roofline: the streaming 2x2 pair kernel on uncontrolled U3 gates at 30 qubits, timed with
HIP events on the library stream over every target bit; achieved = 2 S 2^30 bytes / mean
launch time, against 8 TB/s; traffic = HBM bytes per launch measured with the FETCH_SIZE and
WRITE_SIZE counters (profiles/traffic_gate_30q_c64.md); copy_GBs = the library's own
non-temporal streaming copy kernel over 4 GiB (read + write bytes, mean of 5 launches timed
with HIP events), and frac_of_copy = achieved / copy_GBs.

complex64_vs_complex128: the workload applied twice to |0...0> in complex64 and in complex128;
max |difference| over a strided slice of 2^20 amplitudes and over P(qubit = 0) of every qubit
(the error figure of the complex64 path at 30 qubits).

marginals_30q: probability arrays with hidden qubits on the 30-qubit state (the sum over
hidden qubits that pools over some qubits build): 1 of 30 kept, and 26 of 30 kept with the 4
lowest hidden; read bytes 8 * 2^30 per build.

large_state_34q: a 34-qubit complex64 state (128 GiB): per-target gate GB/s split at target 7,
P(0) throughput, a 30-of-34 and a 1-of-34 sampling-pool build (marginal over the hidden
qubits, then prefix sums), and a bounded pool over all 34 qubits (prefix sums only) with
2^20 draws.

batched_adjoint: 4096 ansatz circuits of 8..20 qubits (10 layers, complex64), each with its
own parameters and observable, <H> and gradients in one graph replay, against a loop of
single-circuit evaluations.

readout_30q: the whole state (8 GiB of amplitudes) and its probabilities (4 GiB) copied to
host memory through pinned double-buffered staging (best of two).

pauli_observable_30q: P(+1) of a 30-factor X/Y/Z string on the same state in one read pass,
against the reference's lowering of that PauliProb into gates (basis change, CX chain, P(0)).

openqasm_qft_30q: the 30-qubit QFT as OpenQASM text: host translate + load time, then the
run on the GPU (complex64, fusion on, first run of a fresh state).

adjoint_24q: forward (<H>) and backward (all 576 gradient entries) of the 24-qubit, 12-layer
RY/RZ/CZ-ring ansatz, each call synchronised; backward_over_forward is the cost of the gradient
once the forward pass has run, gradient_over_forward that of a gradient call from scratch.

complex128_step: the same unfused step on a complex128 state (16 GiB), one timed run.

fused_dense: the same step with runs of gates fused into dense unitaries of at most k = 2..6
qubits (built in FP64, applied as one kernel each; complex64 blocks of 5-6 qubits on FP32
matrix cores, smaller blocks and complex128 blocks of up to 5 qubits staged through LDS), and
single blocks of k = 2..6 qubits on the 30-qubit state in complex64 and complex128 with GB/s,
TFLOP/s and the fraction of max(bytes / 8 TB/s, flops / dense peak of the precision).

cpu_baseline: the reference's own C++ CPU runtime (oracle/_ref, or the C restatement when
that is not built) applying the first gates of the same circuit to a 30-qubit complex64
state in host memory, for about 15 s of work, all host cores; then one worker
(QGATE_NUM_WORKERS=1, in a child process: the count is read once per process) for about 10 s,
the CPU model and the host memory copy rate (numpy copy of 2 GiB, read + write, best of 3).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batch] [--no-large]
    python bench.py --batch-only      # the batched adjoint alone, one JSON object
    python bench.py --dense-only      # dense 5- and 6-qubit blocks alone, one JSON object
    python bench.py --marginal-only   # marginals and pools on the 30- and 34-qubit states
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = 'gate HBM GB/s (% peak) at 30q c64; adjoint grad time / forward time, 24q'


def gate_bytes(kind, n, amp):
    """Algorithmic HBM bytes of one gate of the workload."""
    if kind in ('U3', 'H', 'Rz'):
        return 2 * amp * (1 << n)
    if kind == 'CX':
        return 2 * amp * (1 << (n - 1))
    if kind == 'CCX':
        return 2 * amp * (1 << (n - 2))
    raise ValueError(kind)


class Runner:
    """Applies a workload gate list through a runtime module's processor, lane i = qubit i."""

    def __init__(self, module, n, dtype):
        self.module = module
        self.qs = module.create_qubit_states(dtype)
        self.proc = self.qs.processor
        self.proc.initialize_qubit_states(self.qs, n)
        self.proc.reset_qubit_states(self.qs)
        from qgate_amd.model import gate_type as gt
        self.H = gt.H()
        self.X = gt.X()
        self.gt = gt

    def apply(self, gates):
        proc, qs, gt = self.proc, self.qs, self.gt
        for kind, qubits, angles in gates:
            if kind == 'U3':
                proc.apply_gate(gt.U(*angles), False, qs, qubits[0])
            elif kind == 'H':
                proc.apply_gate(self.H, False, qs, qubits[0])
            elif kind == 'Rz':
                proc.apply_gate(gt.RZ(angles[0]), False, qs, qubits[0])
            else:
                proc.apply_controlled_gate(self.X, False, qs, qubits[:-1], qubits[-1])

    def sync(self):
        self.proc.synchronize()

    def delete(self):
        self.qs.delete()


def cpu_baseline(gates, n, amp, budget_s=15.0):
    """Gate GB/s of the reference CPU core on the first gates of the workload (see module doc)."""
    from oracle import cpuruntime
    if cpuruntime.have_reference():
        module, kind = cpuruntime.reference(), 'reference'
    else:
        if not os.path.exists(cpuruntime.PORT_LIB):
            cpuruntime.build(reference=False)
        module, kind = cpuruntime.port(), 'port'
    cores = module.runtime.L.qgo_num_workers()
    r = Runner(module, n, np.float32)
    done, nbytes = 0, 0
    t0 = time.perf_counter()
    while done < len(gates) and time.perf_counter() - t0 < budget_s:
        r.apply(gates[done:done + 1])
        nbytes += gate_bytes(gates[done][0], n, amp)
        done += 1
    dt = time.perf_counter() - t0
    r.delete()
    return {'value': round(nbytes / dt / 1e9, 3), 'unit': 'GB/s', 'cores': cores, 'kind': kind,
            'sample': 'first {} of the {} gates ({} qubits, complex64, host memory), {:.1f} s, '
                      '{:.3f} s/gate'.format(done, len(gates), n, dt, dt / max(done, 1))}


def host_facts():
    """CPU model and the host memory copy rate (numpy copy of 2 GiB, read + write, best of 3)."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    a = np.ones(1 << 28)
    b = np.empty_like(a)
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        np.copyto(b, a)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {'cpu_model': model, 'host_copy_GBs': round(2 * a.nbytes / best / 1e9, 1),
            'host_threads_visible': len(os.sched_getaffinity(0))}


def cpu_baseline_one_worker(budget_s=10.0):
    """The CPU baseline with QGATE_NUM_WORKERS=1, run in a child process."""
    import subprocess
    env = dict(os.environ, QGATE_NUM_WORKERS='1')
    cmd = [sys.executable, os.path.abspath(__file__), '--cpu-only', '--cpu-budget', str(budget_s)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {'error': r.stderr[-400:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def measured_traffic():
    """HBM bytes per launch of the gate kernels at 30 qubits complex64, from the FETCH_SIZE /
    WRITE_SIZE counter runs summarised in profiles/traffic_gate_30q_c64.md (scaled by a copy
    kernel of known size); None if that summary is absent."""
    path = os.path.join(ROOT, 'profiles', 'traffic_gate_30q_c64.md')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith('traffic_bytes_per_launch:'):
                return int(line.split(':')[1])
    return None


def roofline(runner, n, amp, reps=4):
    """Mean launch time of the uncontrolled 2x2 pair kernel over every target bit."""
    from qgate_amd.simulator import hipruntime
    mat = hipruntime.gate_matrix(runner.gt.U(0.3, 1.1, 2.3))
    per_target = []
    for t in range(n):
        runner.proc.profile_gate(mat, runner.qs, [], t, 1)    # untimed warm-up launch
        per_target.append(runner.proc.profile_gate(mat, runner.qs, [], t, reps))
    avg_ms = float(np.mean(per_target))
    launch_bytes = 2 * amp * (1 << n)
    achieved = launch_bytes / (avg_ms * 1e-3) / 1e9
    split = {}
    for name, rng in (('t=0 unit pair', range(0, 1)), ('1<=t<=6 wavefront exchange', range(1, 7)),
                      ('t>=7 strided', range(7, n))):
        ms = float(np.mean([per_target[t] for t in rng]))
        split[name] = round(launch_bytes / (ms * 1e-3) / 1e9, 1)
    # amplitude-bit classes: partner inside the wavefront's 64 amplitudes, inside a 2^13 block,
    # or further away
    classes = {}
    for name, rng in (('t<6', range(0, 6)), ('6<=t<13', range(6, 13)), ('t>=13', range(13, n))):
        ms = float(np.mean([per_target[t] for t in rng]))
        classes[name] = round(launch_bytes / (ms * 1e-3) / 1e9, 1)
    return {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': measured_traffic() if (n, amp) == (30, 8) else None,
            'kernel': 'uncontrolled 2x2 gate (k_unit_pair / k_wave_pair / k_strided_pair), complex64, 30 qubits',
            'launch_bytes': launch_bytes, 'mean_launch_ms': round(avg_ms, 4),
            'by_kernel_GBs': split,
            'by_target_class_GBs': classes,
            'per_target_GBs': [round(launch_bytes / (ms * 1e-3) / 1e9) for ms in per_target]}


MFMA_FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix-core peak


def _strided_slice(module, r, n, dtype, count=1 << 20):
    """count amplitudes of runner r's state, spread over all 2^n (stride 2^n // count + 1)."""
    from qgate_amd.simulator import qubits as qmod
    from qgate_amd.simulator.lanes import Lane
    getter = module.create_qubits_states_getter(dtype)
    lanes = []
    for i in range(n):
        ln = Lane(None, i)
        ln.set_external(i)
        lanes.append(ln)
    step = (1 << n) // count + 1
    cnt = ((1 << n) - 1) // step + 1
    out = np.empty(cnt, np.complex64 if dtype == np.float32 else np.complex128)
    getter.get_states(out, 0, qmod.null, [(r.qs, lanes)], [], cnt, 0, step)
    return out.astype(np.complex128)


def c128_step_timing(module, n, gates):
    """The same unfused step on a complex128 state (16 GiB at 30 qubits): one warm-up, one
    timed run; algorithmic bytes with S = 16; then the step fused into tile passes (one
    warm-up, one timed run). The complex64 state of the two unfused runs from |0...0> is
    compared against the complex128 one (complex64_vs_complex128)."""
    r = Runner(module, n, np.float64)
    r.apply(gates)
    r.sync()
    t0 = time.perf_counter()
    r.apply(gates)
    r.sync()
    dt = time.perf_counter() - t0
    ref = _strided_slice(module, r, n, np.float64)
    ref_p0 = [r.proc.calc_probability(r.qs, q) for q in range(n)]
    # the same step fused into tile passes: one warm-up, one timed run
    module.set_fusion('tile')
    r.proc.set_fusion('tile')
    r.apply(gates)
    r.sync()
    t0 = time.perf_counter()
    r.apply(gates)
    r.sync()
    fused_dt = time.perf_counter() - t0
    module.set_fusion('off')
    r.proc.set_fusion('off')
    r.delete()
    r32 = Runner(module, n, np.float32)
    r32.apply(gates)
    r32.apply(gates)
    r32.sync()
    got = _strided_slice(module, r32, n, np.float32)
    p0 = [r32.proc.calc_probability(r32.qs, q) for q in range(n)]
    r32.delete()
    step_bytes = sum(gate_bytes(k, n, 16) for k, _, _ in gates)
    err = {'max_abs_amplitude_diff': float(np.max(np.abs(got - ref))),
           'amplitudes_compared': len(ref),
           'max_abs_p0_diff': float(np.max(np.abs(np.array(p0) - np.array(ref_p0)))),
           'max_abs_amplitude': float(np.max(np.abs(ref)))}
    return {'ms_per_step': round(dt * 1e3, 1), 'GBs': round(step_bytes / dt / 1e9, 1),
            'frac_of_hbm_peak': round(step_bytes / dt / 1e9 / HBM_PEAK_GBS, 4),
            'gates_per_s': round(len(gates) / dt, 1),
            'fused_tile_ms_per_step': round(fused_dt * 1e3, 1)}, err


def dense_timing(r, gates, step_s, step_bytes, n, widths=(2, 3, 4, 5, 6)):
    """The step with dense-block fusion at the given widths (one warm-up, one timed run each),
    then single dense blocks (dense_block_timing)."""
    from qgate_amd.simulator import hipruntime
    steps = {}
    for k in widths:
        hipruntime.set_fusion('dense', k)
        r.proc.set_fusion('dense', k)
        r.apply(gates)
        r.sync()
        t0 = time.perf_counter()
        r.apply(gates)
        r.sync()
        dt = time.perf_counter() - t0
        steps[k] = {'ms_per_step': round(dt * 1e3, 1), 'gate_bytes_per_s_GBs': round(step_bytes / dt / 1e9, 1),
                    'speedup_vs_unfused': round(step_s / dt, 2)}
    hipruntime.set_fusion('off')
    r.proc.set_fusion('off')
    return {'steps': steps, 'blocks': dense_block_timing(n, r)}


DENSE_LANES = {2: [3, 17], 3: [2, 14, 25], 4: [0, 9, 18, 27], 5: [2, 9, 15, 21, 27], 6: [1, 6, 12, 17, 23, 28]}
PEAK_TFLOPS = {np.float32: 157.3, np.float64: 78.6}   # MI355X dense FP32 / FP64


def dense_block_timing(n, r32=None, widths=(2, 3, 4, 5, 6)):
    """Single dense blocks of k = 2..6 qubits on the n-qubit state in complex64 and complex128
    (best of 3 synchronised calls): 2 S 2^n bytes and 8 2^k 2^n flops per block, against
    bound = max(bytes / 8 TB/s, flops / peak) with the MI355X dense peak of the precision
    (FP32 157.3, FP64 78.6 TFLOP/s)."""
    from qgate_amd.simulator import hipruntime
    out = {}
    for dtype, amp in ((np.float32, 8), (np.float64, 16)):
        r = r32 if (dtype == np.float32 and r32 is not None) else Runner(hipruntime, n, dtype)
        rs = np.random.RandomState(17)
        res = {}
        for k in widths:
            lanes = DENSE_LANES[k]
            z = rs.normal(size=(1 << k, 1 << k)) + 1j * rs.normal(size=(1 << k, 1 << k))
            u, _ = np.linalg.qr(z)
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                r.proc.apply_unitary(r.qs, lanes, u)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            nbytes, flops = 2 * amp * (1 << n), 8 * (1 << k) * (1 << n)
            bound = max(nbytes / (HBM_PEAK_GBS * 1e9), flops / (PEAK_TFLOPS[dtype] * 1e12))
            res[k] = {'lanes': lanes, 'ms': round(best * 1e3, 3), 'GBs': round(nbytes / best / 1e9, 1),
                      'TFLOPs': round(flops / best / 1e12, 1), 'bound_ms': round(bound * 1e3, 3),
                      'bound': 'hbm' if nbytes / (HBM_PEAK_GBS * 1e9) >= flops / (PEAK_TFLOPS[dtype] * 1e12)
                      else 'flops', 'frac_of_bound': round(bound / best, 3)}
        if r is not r32:
            r.delete()
        out['complex64' if dtype == np.float32 else 'complex128'] = res
    return out




def copy_rate_GBs(gib=4, reps=5):
    """The library's non-temporal streaming copy kernel over gib GiB (mean of reps launches,
    HIP events on the library stream; bytes = read + write): the streaming ceiling of HBM."""
    from qgate_amd.simulator import hipruntime
    nbytes = gib << 30
    ms = hipruntime.profile_copy(nbytes, reps)
    return round(2 * nbytes / (ms * 1e-3) / 1e9, 1)


def device_copy_ms(module, n):
    """A plain copy kernel of known traffic for counter calibration: joining one 2^n-amplitude
    complex64 vector into another (k_join with one source) reads and writes 8 * 2^n bytes."""
    src = Runner(module, n, np.float32)
    dst = module.create_qubit_states(np.float32)
    dst.processor.initialize_qubit_states(dst, n)
    src.sync()
    t0 = time.perf_counter()
    dst.processor.join(dst, [src.qs], 0)
    dst.processor.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    dst.delete()
    src.delete()
    return {'bytes_each_way': 8 << n, 'ms': round(ms, 3)}


def adjoint_timing(dtype, reps=9):
    """Forward (<H>) and backward (full gradient) times of the 24-qubit, 12-layer ansatz
    (576 parameters; H = sum a_i Z_i Z_i+1 + sum b_i X_i, a, b ~ N(0, 1), seed 4)."""
    from qgate_amd import workloads
    from qgate_amd.adjoint import PauliSum, hea_circuit
    n, layers = 24, 12
    params = workloads.hea_params(n, layers, seed=4)
    a, b = workloads.hea_observable(n, seed=4)
    h = PauliSum.zz_x(a, b)
    c = hea_circuit(n, layers, dtype)
    c.expectation(params, h)
    c.backward(params)
    fwd, bwd = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        c.expectation(params, h)
        t1 = time.perf_counter()
        c.backward(params)
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    c.delete()
    f, g = float(np.median(fwd)), float(np.median(bwd))
    return {'forward_ms': round(f * 1e3, 2), 'backward_ms': round(g * 1e3, 2),
            'backward_over_forward': round(g / f, 3), 'gradient_over_forward': round((f + g) / f, 3),
            'gates': 3 * n * layers, 'params': 2 * n * layers}


def batch_timing(n_circuits=4096, layers=10, seed=5, loop_sample=512):
    """Batched adjoint: n_circuits ansatz circuits (RY, RZ on every qubit then a CZ ring per
    layer) with qubit counts uniform in 8..20, each with its own parameters and ZZ + X
    observable, evaluated by one graph replay; against a loop of single-circuit evaluations
    (timed over the first loop_sample circuits and scaled to all)."""
    from qgate_amd.adjoint import PauliSum, hea_circuit
    from qgate_amd.batch import AdjointBatch
    rs = np.random.RandomState(seed)
    sizes = rs.randint(8, 21, n_circuits)
    circuits = {int(n): hea_circuit(int(n), layers, np.float32) for n in sorted(set(sizes))}
    batch = AdjointBatch(np.float32)
    items = []
    for n in sizes:
        n = int(n)
        params = rs.uniform(-math.pi, math.pi, 2 * n * layers)
        h = PauliSum.zz_x(rs.normal(size=n - 1), rs.normal(size=n))
        batch.add(circuits[n], params, h)
        items.append((n, params, h))
    t0 = time.perf_counter()
    batch.run()                       # tunes, builds and captures the graph
    first_s = time.perf_counter() - t0
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        values, grads = batch.run()
    graph_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for n, params, h in items[:loop_sample]:
        circuits[n].gradient(params, h)
    loop_s = (time.perf_counter() - t0) * len(items) / loop_sample
    tuning = batch.tuning
    batch.delete()
    for c in circuits.values():
        c.delete()
    return {'circuits': n_circuits, 'qubits': '8..20', 'layers': layers, 'dtype': 'complex64',
            'graph_ms': round(graph_s * 1e3, 2), 'first_run_with_tuning_ms': round(first_s * 1e3, 1),
            'loop_ms_scaled': round(loop_s * 1e3, 1), 'loop_sample': loop_sample,
            'speedup_vs_loop': round(loop_s / graph_s, 2), 'circuits_per_s': round(n_circuits / graph_s, 1),
            'groups': [{k: g[k] for k in ('qubits', 'items', 'tile_bits', 'threads', 'grid_per_cu', 'registers',
                                          'forward_in_registers', 'sweep_in_registers')} for g in tuning]}


def pauli_timing(r, n, reps=5):
    """One Pauli-string observable on the n-qubit complex64 state of runner r: X on lanes
    0, 3, 6, ..., Y on 1, 4, 7, ..., Z on 2, 5, 8, ... P(+1) in one read pass
    (calc_pauli_probability; algorithmic bytes 8 * 2^n) against the reference's lowering of
    the same PauliProb: basis change, CX chain, P(last lane = 0), inverse gates."""
    from qgate_amd.model import gate_type as gt
    x = z = 0
    for lane in range(n):
        if lane % 3 in (0, 1):
            x |= 1 << lane
        if lane % 3 in (1, 2):
            z |= 1 << lane
    fused = r.proc.calc_pauli_probability(r.qs, x, z)
    t0 = time.perf_counter()
    for _ in range(reps):
        fused = r.proc.calc_pauli_probability(r.qs, x, z)
    t_fused = (time.perf_counter() - t0) / reps

    def expanded():
        basis = [(lane, lane % 3) for lane in range(n)]
        for lane, k in basis:   # X -> H, Y -> SH^dagger (the lowering's first gates)
            if k == 0:
                r.proc.apply_gate(gt.H(), False, r.qs, lane)
            elif k == 1:
                r.proc.apply_gate(gt.SH(), True, r.qs, lane)
        for lane in range(n - 1):
            r.proc.apply_controlled_gate(gt.X(), False, r.qs, [lane], lane + 1)
        p = r.proc.calc_probability(r.qs, n - 1)
        for lane in reversed(range(n - 1)):
            r.proc.apply_controlled_gate(gt.X(), False, r.qs, [lane], lane + 1)
        for lane, k in reversed(basis):
            if k == 0:
                r.proc.apply_gate(gt.H(), False, r.qs, lane)
            elif k == 1:
                r.proc.apply_gate(gt.SH(), False, r.qs, lane)
        r.sync()
        return p

    t1 = time.perf_counter()
    p_exp = expanded()
    t_exp = time.perf_counter() - t1
    return {'qubits': n, 'dtype': 'complex64', 'string': 'X,Y,Z on every third qubit (30 factors)',
            'fused_ms': round(t_fused * 1e3, 3), 'fused_GBs': round(8 * (1 << n) / t_fused / 1e9, 1),
            'expanded_ms': round(t_exp * 1e3, 2), 'speedup': round(t_exp / t_fused, 1),
            'prob': fused, 'abs_diff_vs_expanded': abs(fused - p_exp)}


def qasm_qft_timing(n=30):
    """The QFT on n qubits written as OpenQASM 2.0 (h, cu1 and swap lines after an h layer on
    every qubit): translate + load time on the host, then the circuit on the HIP runtime in
    complex64 with fusion (tile passes), state resident in HBM."""
    from qgate_amd import openqasm
    from qgate_amd.simulator import hipruntime
    from qgate_amd.simulator.simulator import Simulator
    lines = ['OPENQASM 2.0;', 'include "qelib1.inc";', 'qreg q[{}];'.format(n), 'h q;']
    for j in reversed(range(n)):
        lines.append('h q[{}];'.format(j))
        for k in reversed(range(j)):
            lines.append('cu1(pi/{}) q[{}], q[{}];'.format(2 ** (j - k), k, j))
    for j in range(n // 2):
        lines.append('swap q[{}], q[{}];'.format(j, n - 1 - j))
    text = '\n'.join(lines) + '\n'
    t0 = time.perf_counter()
    mod = openqasm.load_circuit(text)
    t1 = time.perf_counter()
    hipruntime.set_fusion('tile')
    sim = Simulator(hipruntime, circuit_prep='one_static', dtype=np.float32)
    sim.qubits.set_ordering(list(mod.q))
    try:
        sim.run(mod.circuit)   # first run: allocation and |0> included
        t2 = time.perf_counter()
        p0 = sim.qubits.calc_probability(mod.q[0])
        t3 = time.perf_counter()
    finally:
        hipruntime.set_fusion('off')
        sim.terminate()
    return {'qubits': n, 'qasm_lines': len(lines), 'load_ms': round((t1 - t0) * 1e3, 2),
            'run_ms': round((t2 - t1) * 1e3, 2), 'p0_qubit0': p0, 'p0_ms': round((t3 - t2) * 1e3, 2)}


def pool_build_ms(getter, qs, n, kept, reps=3):
    """Build time (best of reps, ms) of a device sampling pool over the lanes `kept` of an
    n-lane state vector, in that order, the other lanes hidden."""
    from qgate_amd.simulator.lanes import Lane
    pos = {lane: i for i, lane in enumerate(kept)}
    best = None
    for _ in range(reps):
        lanes = []
        for i in range(n):
            ln = Lane(None, i)
            ln.set_external(pos.get(i, -1))
            lanes.append(ln)
        t0 = time.perf_counter()
        pool = getter.create_sampling_pool(list(range(len(kept))), len(kept), n - len(kept), [(qs, lanes)], [])
        dt = time.perf_counter() - t0
        pool.delete()
        best = dt if best is None else min(best, dt)
    return best * 1e3


def marginal_timing(r, n, amp):
    """Pools over some qubits of runner r's n-qubit state (the hidden qubits summed by the
    marginal kernels, then prefix sums): GB/s = amp * 2^n read bytes / build time."""
    from qgate_amd.simulator import hipruntime
    getter = hipruntime.create_qubits_states_getter(np.float32)
    out = {}
    for name, kept in (('1_of_{}'.format(n), [5]),
                       ('{}_of_{}_low_hidden'.format(n - 4, n), list(range(4, n))),
                       ('{}_of_{}_high_hidden'.format(n - 4, n), list(range(n - 4)))):
        ms = pool_build_ms(getter, r.qs, n, kept)
        out[name] = {'ms': round(ms, 2), 'read_GBs': round((amp << n) / (ms * 1e-3) / 1e9, 1)}
    return out


def large_state_timing(n=34):
    """Config of the large single-GPU state: 34 qubits complex64 (128 GiB). H on every qubit,
    then per-target timing of an uncontrolled gate (HIP events, 2 launches per target), P(0)
    of every qubit, and a sampling pool over 30 qubits (4 hidden) as prefix-sum throughput."""
    from qgate_amd.simulator import hipruntime
    from qgate_amd.simulator.lanes import Lane
    info = hipruntime.device_info()
    amp = 8
    if info['total_mem'] < (amp << n) + (40 << 30):
        return {'skipped': 'device memory {} GB'.format(info['total_mem'] // (1 << 30))}
    r = Runner(hipruntime, n, np.float32)
    t0 = time.perf_counter()
    for q in range(n):
        r.proc.apply_gate(r.H, False, r.qs, q)
    r.sync()
    h_s = time.perf_counter() - t0
    mat = hipruntime.gate_matrix(r.gt.U(0.3, 1.1, 2.3))
    per_target = [r.proc.profile_gate(mat, r.qs, [], t, 2) for t in range(n)]
    launch_bytes = 2 * amp * (1 << n)
    gbs = [launch_bytes / (ms * 1e-3) / 1e9 for ms in per_target]
    t0 = time.perf_counter()
    probs = [r.proc.calc_probability(r.qs, q) for q in range(n)]
    prob_s = (time.perf_counter() - t0) / n
    getter = hipruntime.create_qubits_states_getter(np.float32)
    lanes = []
    for i in range(n):
        ln = Lane(None, i)
        ln.set_external(i if i < 30 else -1)
        lanes.append(ln)
    t0 = time.perf_counter()
    pool = getter.create_sampling_pool(list(range(30)), 30, n - 30, [(r.qs, lanes)], [])
    pool_s = time.perf_counter() - t0
    pool.delete()
    pool_s = min(pool_s, pool_build_ms(getter, r.qs, n, list(range(30)), reps=2) * 1e-3)
    one_ms = pool_build_ms(getter, r.qs, n, [5], reps=2)
    # a bounded pool over all 34 qubits: prefix sums only (8 GiB), |a|^2 read at draw time
    hipruntime.set_sampling_pool_workspace('bounded')
    try:
        all_lanes = []
        for i in range(n):
            ln = Lane(None, i)
            ln.set_external(i)
            all_lanes.append(ln)
        t0 = time.perf_counter()
        pool = getter.create_sampling_pool(list(range(n)), n, 0, [(r.qs, all_lanes)], [])
        bounded_s = time.perf_counter() - t0
        u = np.random.RandomState(1034).random_sample(1 << 20)
        t0 = time.perf_counter()
        pool.sample(len(u), u)
        draws_s = time.perf_counter() - t0
        pool.delete()
    finally:
        hipruntime.set_sampling_pool_workspace('auto')
    r.delete()
    return {'qubits': n, 'dtype': 'complex64', 'state_GiB': (amp << n) >> 30,
            'h_all_ms': round(h_s * 1e3, 1),
            'gate_GBs_targets_lt7': round(float(np.mean(gbs[:7])), 1),
            'gate_GBs_targets_ge7': round(float(np.mean(gbs[7:])), 1),
            'per_target_GBs': [round(v) for v in gbs],
            'prob_ms_per_qubit': round(prob_s * 1e3, 2),
            'prob_GBs': round((amp << (n - 1)) / prob_s / 1e9, 1),
            'prob0_state': 'H on every qubit, then the 68 timing launches of U(0.3, 1.1, 2.3)',
            'min_prob0': round(min(probs), 6), 'max_prob0': round(max(probs), 6),
            'pool_build_30_of_34_ms': round(pool_s * 1e3, 1),
            'pool_build_GBs': round((amp << n) / pool_s / 1e9, 1),
            'pool_build_1_of_34_ms': round(one_ms, 1),
            'pool_build_1_of_34_GBs': round((amp << n) / (one_ms * 1e-3) / 1e9, 1),
            'bounded_pool_34_of_34': {'build_ms': round(bounded_s * 1e3, 1),
                                      'build_GBs': round((amp << n) / bounded_s / 1e9, 1),
                                      'workspace_GiB': round(((1 << n) // 16 * 8) / 2 ** 30, 2),
                                      'draws_2p20_ms': round(draws_s * 1e3, 2)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--qubits', type=int, default=30)
    ap.add_argument('--gates', type=int, default=2000)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-batch', action='store_true')
    ap.add_argument('--no-large', action='store_true')
    ap.add_argument('--batch-only', action='store_true',
                    help='only the batched adjoint (tuning run, three graph replays, a short single-circuit '
                         'loop); prints its JSON object (for profiler runs)')
    ap.add_argument('--dense-only', action='store_true',
                    help='only the dense 5- and 6-qubit blocks on the 30-qubit state; prints their JSON object '
                         '(for profiler runs)')
    ap.add_argument('--marginal-only', action='store_true',
                    help='only the marginal and pool timings on the 30-qubit and 34-qubit states')
    ap.add_argument('--tile-only', action='store_true',
                    help='only the workload step with tile-pass fusion (one warm-up, one timed step); prints '
                         'its JSON object (for counter runs)')
    ap.add_argument('--adjoint-only', action='store_true',
                    help='only the 24-qubit adjoint forward and gradient (complex64 and complex128); prints '
                         'their JSON object (for counter runs)')
    ap.add_argument('--cpu-only', action='store_true',
                    help='only the CPU baseline (no GPU use); prints its JSON object')
    ap.add_argument('--cpu-budget', type=float, default=15.0)
    ap.add_argument('--roofline-only', action='store_true',
                    help='only the per-target gate timing loop and a 2 GiB copy kernel (for counter runs)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: the state-vector backend does not distribute qubits '
              'across GPUs (got --gpus {}).'.format(args.gpus), file=sys.stderr)
        sys.exit(2)

    if args.cpu_only:
        from qgate_amd import workloads
        gates = workloads.random_gate_list(args.qubits, args.gates, seed=2)
        print(json.dumps(cpu_baseline(gates, args.qubits, 8, budget_s=args.cpu_budget)), flush=True)
        return
    if args.batch_only:
        print(json.dumps(batch_timing(loop_sample=32)), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({name: adjoint_timing(dt, reps=3) for name, dt in
                          (('complex64', np.float32), ('complex128', np.float64))}), flush=True)
        return

    from qgate_amd import workloads
    from qgate_amd.simulator import hipruntime

    n, amp = args.qubits, 8
    if args.dense_only:
        print(json.dumps(dense_block_timing(n)), flush=True)
        return
    gates = workloads.random_gate_list(n, args.gates, seed=2)
    step_bytes = sum(gate_bytes(k, n, amp) for k, _, _ in gates)
    info = hipruntime.device_info()

    r = Runner(hipruntime, n, np.float32)
    if args.tile_only:
        hipruntime.set_fusion('tile')
        r.proc.set_fusion('tile')
        r.apply(gates)
        r.sync()
        t0 = time.perf_counter()
        r.apply(gates)
        r.sync()
        dt = time.perf_counter() - t0
        hipruntime.set_fusion('off')
        r.delete()
        print(json.dumps({'fused_tile_ms': round(dt * 1e3, 2), 'gates': len(gates)}), flush=True)
        return
    if args.marginal_only:
        r.apply(gates[:200])
        r.sync()
        out = {'marginals_30q': marginal_timing(r, n, amp)}
        r.delete()
        out['large_state_34q'] = large_state_timing()
        print(json.dumps(out), flush=True)
        return
    if args.roofline_only:
        roof = roofline(r, n, amp)
        r.delete()
        copy = device_copy_ms(hipruntime, 28)
        print(json.dumps({'roofline': roof, 'copy': copy, 'stream_copy_GBs': copy_rate_GBs()}))
        return
    print('[bench] device {} ({} CUs), {} qubits complex64, {} gates, {:.1f} GB per step'
          .format(info['device_id'], info['n_cus'], n, len(gates), step_bytes / 1e9), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        r.apply(gates)
    r.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.apply(gates)
        r.sync()
        print('[bench] step {} done, {:.3f} s'.format(i, time.perf_counter() - t0), file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    value = step_bytes / step_s / 1e9

    # the same step with gate fusion in tile passes (several gates per trip through HBM)
    hipruntime.set_fusion('tile')
    r.proc.set_fusion('tile')
    r.apply(gates)
    r.sync()
    tf0 = time.perf_counter()
    for _ in range(args.steps):
        r.apply(gates)
        r.sync()
    fused_s = (time.perf_counter() - tf0) / args.steps
    hipruntime.set_fusion('off')
    r.proc.set_fusion('off')
    print('[bench] fused step {:.3f} s'.format(fused_s), file=sys.stderr, flush=True)
    dense = dense_timing(r, gates, step_s, step_bytes, n)
    print('[bench] dense fusion: {}'.format(dense), file=sys.stderr, flush=True)

    # a Pauli observable on the circuit's output state (before the measurements collapse it)
    pauli = pauli_timing(r, n)
    print('[bench] pauli observable: {}'.format(pauli), file=sys.stderr, flush=True)

    # end-of-circuit work of the workload on the same state: readout, pool + 2^20 draws, measures
    from qgate_amd.simulator import qubits as qmod
    from qgate_amd.simulator.lanes import Lane
    getter = hipruntime.create_qubits_states_getter(np.float32)
    lanes = []
    for i in range(n):
        ln = Lane(None, i)
        ln.set_external(i)
        lanes.append(ln)
    readout = {}
    for name, op, dt in (('amplitudes', qmod.null, np.complex64), ('probabilities', qmod.abs2, np.float32)):
        out = np.empty(1 << n, dt)
        ts = []
        for _ in range(2):
            tr0 = time.perf_counter()
            getter.get_states(out, 0, op, [(r.qs, lanes)], [], 1 << n, 0, 1)
            ts.append(time.perf_counter() - tr0)
        readout[name] = {'GiB': out.nbytes / 2 ** 30, 'ms': round(min(ts) * 1e3, 1),
                         'GBs': round(out.nbytes / min(ts) / 1e9, 1)}
        del out
    print('[bench] readout: {}'.format(readout), file=sys.stderr, flush=True)
    marginals = marginal_timing(r, n, amp)
    print('[bench] marginals: {}'.format(marginals), file=sys.stderr, flush=True)
    u = np.random.RandomState(1002).random_sample(1 << 20)
    t2 = time.perf_counter()
    pool = getter.create_sampling_pool(list(range(n)), n, 0, [(r.qs, lanes)], [])
    t3 = time.perf_counter()
    pool.sample(len(u), u)
    t4 = time.perf_counter()
    pool.delete()
    um = np.random.RandomState(1003).random_sample(n)
    for lane in range(n):
        p = r.proc.calc_probability(r.qs, lane)
        r.proc.decohere(0 if um[lane] < p else 1, p, r.qs, lane)
    r.sync()
    t5 = time.perf_counter()

    roof = roofline(r, n, amp)
    r.delete()
    roof['copy_GBs'] = copy_rate_GBs()
    c128, c64_err = c128_step_timing(hipruntime, n, gates)
    print('[bench] complex128 step: {}'.format(c128), file=sys.stderr, flush=True)
    roof['frac_of_copy'] = round(roof['achieved'] / roof['copy_GBs'], 4)
    print('[bench] roofline: {:.1f} GB/s, device copy {:.1f} GB/s'.format(roof['achieved'], roof['copy_GBs']),
          file=sys.stderr, flush=True)

    adjoint = {}
    for name, dt in (('complex64', np.float32), ('complex128', np.float64)):
        adjoint[name] = adjoint_timing(dt)
        print('[bench] adjoint 24q {}: {}'.format(name, adjoint[name]), file=sys.stderr, flush=True)
    batched = batch_timing() if not args.no_batch else None
    print('[bench] batch: {}'.format(batched), file=sys.stderr, flush=True)
    large = large_state_timing() if not args.no_large else None
    print('[bench] large state: {}'.format(large), file=sys.stderr, flush=True)
    qasm = qasm_qft_timing()
    print('[bench] openqasm qft: {}'.format(qasm), file=sys.stderr, flush=True)

    result = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'GB/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(step_s * 1e3, 2),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'complex64 (f32)', 'data': 'synthetic',
        'config': {'workload': 'random circuit, {} qubits, {} gates from {{U3,H,Rz,CX,CCX}}, '
                               'seed 2, complex64, fusion off'.format(n, len(gates)),
                   'qubits': n, 'gates': len(gates), 'fusion': 'off',
                   'parallelism': 'single GPU'},
        'frac_of_hbm_peak': round(value / HBM_PEAK_GBS, 4),
        'gates_per_s': round(len(gates) / step_s, 1),
        'pool_build_ms': round((t3 - t2) * 1e3, 2),
        'pool_2p20_draws_ms': round((t4 - t3) * 1e3, 2),
        'measure_all_ms': round((t5 - t4) * 1e3, 2),
        'fused_tile': {'ms_per_step': round(fused_s * 1e3, 2), 'gates_per_s': round(len(gates) / fused_s, 1),
                       'gate_bytes_per_s_GBs': round(step_bytes / fused_s / 1e9, 1),
                       'speedup_vs_unfused': round(step_s / fused_s, 2)},
        'fused_dense': dense,
        'complex128_step': c128,
        'complex64_vs_complex128': c64_err,
        'marginals_30q': marginals,
        'pauli_observable_30q': pauli,
        'readout_30q': readout,
        'adjoint_24q': adjoint,
        'batched_adjoint': batched,
        'large_state_34q': large,
        'openqasm_qft_30q': qasm,
        'roofline': roof,
    }
    if not args.no_cpu:
        print('[bench] CPU baseline ...', file=sys.stderr, flush=True)
        cb = cpu_baseline(gates, n, amp)
        cb.update(host_facts())
        cb['one_worker'] = cpu_baseline_one_worker()
        result['cpu_baseline'] = cb
    print(json.dumps(result))


if __name__ == '__main__':
    main()
