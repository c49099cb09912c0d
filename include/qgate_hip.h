/*
 * C ABI of the qgate HIP runtime (libqgate_hip.so): the MI355X state-vector backend behind
 * qgate's runtime-module protocol.
 *
 * The entry points are the ones the reference binds from Python through its CPython glue
 * module and backend factories (qgate/simulator/src/glue.cpp:605-631, cpuext.cpp:73-79,
 * cudaext.cpp:127-136) over the native interfaces of qgate/simulator/src/Interfaces.h:7-81,
 * restated with plain C types: opaque handles instead of numpy.uint64-boxed pointers, and
 * caller-owned buffers with explicit sizes instead of ndarrays.
 *
 * Conventions
 *   - Every function returns QGHIP_OK (0) on success, QGHIP_ERROR (-1) on a runtime error or
 *     QGHIP_EVALUE (-2) on an invalid argument; qghip_last_error() returns the message of the
 *     last failure on the calling thread. No C++ exception crosses the boundary (the
 *     reference glue lets them terminate the process).
 *   - prec: QGHIP_FP64 (complex128 amplitudes) or QGHIP_FP32 (complex64 amplitudes).
 *   - Lanes are local qubit indices of one state vector: bit `lane` of the amplitude index.
 *   - All device work of the library is ordered on one HIP stream; calls that return host
 *     data (probabilities, states, samples) synchronise it.
 *   - Not thread-safe per handle; calls on different handles may come from one thread at a
 *     time per process (the reference holds the GIL for every call).
 */
#ifndef QGATE_HIP_H
#define QGATE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QGHIP_OK = 0, QGHIP_ERROR = -1, QGHIP_EVALUE = -2 };
enum { QGHIP_FP64 = 0, QGHIP_FP32 = 1 };
enum { QGHIP_MATHOP_NULL = 0, QGHIP_MATHOP_PROB = 1 };
enum { QGHIP_FUSION_OFF = 0, QGHIP_FUSION_TILE = 1, QGHIP_FUSION_DENSE = 2 };
enum { QGHIP_POOL_AUTO = 0, QGHIP_POOL_BOUNDED = 1, QGHIP_POOL_FULL = 2 };

typedef struct qghip_qubit_states qghip_qubit_states;
typedef struct qghip_qubit_processor qghip_qubit_processor;
typedef struct qghip_qubits_states_getter qghip_qubits_states_getter;
typedef struct qghip_sampling_pool qghip_sampling_pool;

/* Message of the last failed call on this thread ("" if none). */
const char *qghip_last_error(void);

/* ---- module / device ------------------------------------------------------------------ */

/* Select the GPU and create the library stream; replaces devices_initialize
 * (cudaext.cpp:32-58, called from cudaruntime.py:74-76 module_init). One device only. */
int qghip_devices_initialize(int device_id);
/* Apply queued work (fused gates, recorded collapses), then free cached device memory and the
 * stream; replaces devices_clear (cudaext.cpp:60-68, cudaruntime.py:78-84 module_finalize). */
int qghip_devices_clear(void);
/* Device id, compute-unit count and HBM size of the initialised device. */
int qghip_device_info(int *device_id, int *n_cus, uint64_t *total_mem);

/* ---- gate matrices -------------------------------------------------------------------- */

/* Factory id for a gate type name (U, U2, U1, ID, X, Y, Z, H, S, T, RX, RY, RZ, ExpiI, ExpiZ,
 * SH); replaces register_matrix_factory (glue.cpp:92-134). */
int qghip_register_matrix_factory(const char *gate_type, int *factory_id);
/* FP64 2x2 matrix of a gate, conjugate-transposed when `adjoint` != 0, as row-major
 * (m00, m01, m10, m11) (re, im) pairs; the matrix built inside qubit_processor_apply_gate
 * (glue.cpp:390-398, GateMatrix.cpp:13-158). */
int qghip_gate_matrix(int factory_id, const double *args, int n_args, int adjoint, double *mat8);

/* ---- qubit states --------------------------------------------------------------------- */

/* qubit_states_new (cpuext.cpp:11-27, cudaext.cpp:70-88). */
int qghip_qubit_states_new(int prec, qghip_qubit_states **out);
/* qubit_states_delete (glue.cpp:193-205); frees the device memory. */
int qghip_qubit_states_delete(qghip_qubit_states *qs);
/* qubit_states_deallocate (glue.cpp:180-190). */
int qghip_qubit_states_deallocate(qghip_qubit_states *qs);
/* qubit_states_get_n_lanes (glue.cpp:237-246). */
int qghip_qubit_states_get_n_lanes(const qghip_qubit_states *qs, int *n_lanes);

/* ---- qubit processor (Interfaces.h:22-50) ------------------------------------------- */

/* qubit_processor_new (cpuext.cpp:31-49, cudaext.cpp:90-107). */
int qghip_qubit_processor_new(int prec, qghip_qubit_processor **out);
/* qubit_processor_delete (glue.cpp:208-219). */
int qghip_qubit_processor_delete(qghip_qubit_processor *p);
/* Addition, no reference counterpart: gate fusion for the gates this processor applies.
 * QGHIP_FUSION_TILE queues apply_gate / apply_controlled_gate / apply_matrix calls per state
 * vector and applies them in tile passes (many gates per trip through HBM) when anything else
 * touches that state vector, on synchronize, or every 4096 gates. tile_bits (<= 0: default
 * 12 for FP32, 11 for FP64) bounds the tile to 2^tile_bits amplitudes; low_bits (< 0: default
 * 5 / 4) is the number of lowest bits every tile keeps contiguous. QGHIP_FUSION_DENSE queues
 * the same way and applies runs of consecutive gates on at most tile_bits (2..6; <= 0: 6)
 * lanes as one dense unitary each, built in FP64 on the host (complex64 blocks of 5 or 6
 * lanes on FP32 matrix cores, the others on the vector units; a run of one gate on the
 * streaming gate kernels). QGHIP_FUSION_OFF applies each gate when called, as the reference
 * does. */
int qghip_qubit_processor_set_fusion(qghip_qubit_processor *p, int mode, int tile_bits,
                                     int low_bits);
/* qubit_processor_synchronize (glue.cpp:251-260): completes all queued device work. */
int qghip_qubit_processor_synchronize(qghip_qubit_processor *p);
/* qubit_processor_reset (glue.cpp:263-272). */
int qghip_qubit_processor_reset(qghip_qubit_processor *p);
/* qubit_processor_initialize_qubit_states (glue.cpp:275-286): allocate 2^n_lanes amplitudes. */
int qghip_qubit_processor_initialize_qubit_states(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                                  int n_lanes);
/* qubit_processor_reset_qubit_states (glue.cpp:289-299): set |0...0>. */
int qghip_qubit_processor_reset_qubit_states(qghip_qubit_processor *p, qghip_qubit_states *qs);
/* qubit_processor_calc_probability (glue.cpp:302-311): P(bit local_lane == 0), FP64. */
int qghip_qubit_processor_calc_probability(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                           int local_lane, double *prob);
/* Addition, no reference counterpart: qubit_processor_calc_probability for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same kernels and summation order, so the same bits as
 * qghip_qubit_processor_calc_probability on a state with no recorded collapse; the partial
 * sums go to a fixed 32 MiB device array the library holds per GPU for the life of the
 * process, so states of at most 35 lanes are accepted (QGHIP_EVALUE beyond). A recorded
 * collapse is applied first in its own pass (the result then differs from the fused path of
 * calc_probability by FP64 rounding only); queued fused gates are refused (QGHIP_EVALUE:
 * synchronize first). Every call on a GPU shares that GPU's array; the library orders its
 * calls on one stream under one lock, and work that overlapped a call on the array would
 * race. Calls no allocator. */
int qghip_qubit_processor_calc_probability_noalloc(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                                   int local_lane, double *prob);
/* qubit_processor_join (glue.cpp:314-334): qs <- kron(list[0], ..., list[n-1]) with the last
 * entry in the lowest bits; the n_new_lanes top lanes start in |0>. */
int qghip_qubit_processor_join(qghip_qubit_processor *p, qghip_qubit_states *qs,
                               qghip_qubit_states *const *qstates_list, int n_qstates,
                               int n_new_lanes);
/* qubit_processor_decohere (glue.cpp:337-348): collapse local_lane to `value` in place. The
 * collapse is recorded and launched by the next call that reads or writes qs (a following
 * calc_probability runs it in the same pass as its sum); results are those of collapsing at
 * once, and a device failure is reported by that next call. */
int qghip_qubit_processor_decohere(qghip_qubit_processor *p, int value, double prob,
                                   qghip_qubit_states *qs, int local_lane);
/* qubit_processor_decohere_and_separate (glue.cpp:351-365): collapse local_lane of qs to
 * `value`; qs0 receives the other lanes (n-1), qs1 the measured lane. */
int qghip_qubit_processor_decohere_and_separate(qghip_qubit_processor *p, int value, double prob,
                                                qghip_qubit_states *qs0, qghip_qubit_states *qs1,
                                                qghip_qubit_states *qs, int local_lane);
/* Pauli-string observable in one pass: the probability that measuring P = (x) sigma_k gives
 * +1, where lane k carries X if bit k of x_lanes is set, Z if bit k of z_lanes is set, Y if
 * both. Replaces what qgate/model/expand.py:66-77 (expand_pprob) lowers a PauliProb into:
 * the basis change and CX chain of pauli_gates_diagonalizer.py:96-135, calcProbability
 * (CPUQubitProcessor.cpp:75-125) of the last lane, and the inverse gates. FP64 result. */
int qghip_qubit_processor_calc_pauli_probability(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                                 uint64_t x_lanes, uint64_t z_lanes, double *prob);
/* Collapse onto the eigenspace P = +1 (value 0) or -1 (value 1): psi <- (1 +- P) psi / (2 sqrt(p)),
 * p = prob or 1 - prob. Replaces the decohere (CPUQubitProcessor.cpp:217-246) between the
 * basis changes that expand.py:53-64 (expand_pmeasure) lowers a PauliMeasure into. */
int qghip_qubit_processor_pauli_decohere(qghip_qubit_processor *p, int value, double prob,
                                         qghip_qubit_states *qs, uint64_t x_lanes, uint64_t z_lanes);
/* qubit_processor_apply_reset (glue.cpp:368-379): move the |1> half of local_lane to |0>. */
int qghip_qubit_processor_apply_reset(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                      int local_lane);
/* qubit_processor_apply_gate (glue.cpp:383-405): matrix from factory + args (+ adjoint). */
int qghip_qubit_processor_apply_gate(qghip_qubit_processor *p, int factory_id, const double *args,
                                     int n_args, int adjoint, qghip_qubit_states *qs,
                                     int local_lane);
/* qubit_processor_apply_controlled_gate (glue.cpp:409-431). */
int qghip_qubit_processor_apply_controlled_gate(qghip_qubit_processor *p, int factory_id,
                                                const double *args, int n_args, int adjoint,
                                                qghip_qubit_states *qs,
                                                const int *local_control_lanes, int n_controls,
                                                int local_target_lane);
/* Same as apply_controlled_gate with an explicit FP64 matrix (n_controls may be 0);
 * QubitProcessor::applyControlledGate (Interfaces.h:48-49). */
int qghip_qubit_processor_apply_matrix(qghip_qubit_processor *p, const double *mat8,
                                       qghip_qubit_states *qs, const int *local_control_lanes,
                                       int n_controls, int local_target_lane);
/* Addition, no reference counterpart: apply a dense 2^k x 2^k unitary (row-major (re, im)
 * pairs, FP64, 1 <= k <= 6) to lanes[0..k-1]; lanes[j] is bit j of the matrix index. */
int qghip_qubit_processor_apply_unitary(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                        const int *lanes, int k, const double *mat);

/* ---- states getter (Interfaces.h:53-72) ---------------------------------------------- */

/* qubits_states_getter_new (cpuext.cpp:52-70, cudaext.cpp:109-125). */
int qghip_qubits_states_getter_new(int prec, qghip_qubits_states_getter **out);
/* qubits_states_getter_delete (glue.cpp:222-233). */
int qghip_qubits_states_getter_delete(qghip_qubits_states_getter *g);
/* qubits_states_getter_get_states (glue.cpp:448-506): for j < n_states, ext = start + j*step,
 * array[array_offset + j] = 0 if ext & empty_lane_mask, else the product over the state
 * vectors of f(amplitude at the local index of ext); f = identity (QGHIP_MATHOP_NULL, complex
 * output) or |.|^2 (QGHIP_MATHOP_PROB, real output) in the state precision.
 * lane_tables concatenates, for each state vector in order, its local->external lane table.
 * With n_qstates == 0 the output is 1 at index 0 and 0 elsewhere. Host output buffer. */
int qghip_qubits_states_getter_get_states(qghip_qubits_states_getter *g, void *array,
                                          int64_t array_size, int64_t array_offset, int mathop,
                                          const int *lane_tables,
                                          qghip_qubit_states *const *qstates_list, int n_qstates,
                                          uint64_t empty_lane_mask, int n_qregs,
                                          int64_t n_states, int64_t start, int64_t step);
/* qubits_states_getter_prepare_prob_array (glue.cpp:509-536): prob[x] = sum over hidden
 * lanes of the probability of external index x; external lanes >= n_lanes are hidden
 * (placed in the most significant bits). Output in the state precision, host buffer. */
int qghip_qubits_states_getter_prepare_prob_array(qghip_qubits_states_getter *g, void *prob,
                                                  int64_t array_size, const int *lane_tables,
                                                  qghip_qubit_states *const *qstates_list,
                                                  int n_qstates, int n_lanes, int n_hidden_lanes);
/* qubits_states_getter_create_sampling_pool (glue.cpp:539-559): device-resident FP64 CDF of
 * the same marginal distribution; empty_lanes are external bit positions that read as 0.
 * A pool over every lane of one state vector in its own order may be bounded (see
 * qghip_set_sampling_pool_workspace): it then keeps only prefix sums per 16 bins (2^n / 16
 * doubles) and reads |a|^2 of the found 16 bins from the state vector at draw time, so
 * sampling after that state vector is written or freed fails with QGHIP_EVALUE. */
int qghip_qubits_states_getter_create_sampling_pool(qghip_qubits_states_getter *g,
                                                    const int *lane_tables,
                                                    qghip_qubit_states *const *qstates_list,
                                                    int n_qstates, int n_lanes,
                                                    int n_hidden_lanes, const int *empty_lanes,
                                                    int n_empty_lanes,
                                                    qghip_sampling_pool **out);

/* ---- sampling pool (Interfaces.h:75-81) ---------------------------------------------- */

/* sampling_pool_sample (glue.cpp:562-589): observations[i] = first bin whose normalised
 * cumulative probability exceeds randnum[i] (upper_bound), with zero bits inserted at the
 * empty lanes. A randnum at or above the normalised total maps to the last nonzero bin. */
int qghip_sampling_pool_sample(qghip_sampling_pool *pool, int64_t *observations,
                               int64_t array_size, int64_t n_samples, const double *randnum);
/* sampling_pool_delete (glue.cpp:592-602). */
int qghip_sampling_pool_delete(qghip_sampling_pool *pool);
/* Addition, no reference counterpart: workspace of the sampling pools created from now on.
 * QGHIP_POOL_FULL keeps the FP64 probability of every bin (8 bytes per bin plus the prefix
 * sums), as the reference's pools do (CPUSamplingPool.cpp:8-63); QGHIP_POOL_BOUNDED keeps
 * only the prefix sums (1/16 of that) for pools over a whole state vector in its own lane
 * order, and reads the amplitudes at draw time; QGHIP_POOL_AUTO (default) is FULL unless the
 * probability array would not fit in free device memory. Draws are identical in all modes. */
int qghip_set_sampling_pool_workspace(int mode);

/* ---- adjoint differentiation (no reference counterpart) ------------------------------ */

/* A parameterised circuit on n_qubits qubits (qubit q = bit q of the amplitude index) for
 * expectation values <psi|H|psi> of Pauli sums and their gradients by adjoint (reverse-mode)
 * differentiation: two device-resident state vectors, no checkpoints. Callers: autograd
 * wrappers and variational workloads. */
typedef struct qghip_circuit qghip_circuit;
int qghip_circuit_new(int prec, int n_qubits, qghip_circuit **out);
int qghip_circuit_delete(qghip_circuit *c);
/* Append a fixed (controlled) 2x2 gate, FP64 row-major (re, im) pairs. */
int qghip_circuit_add_gate(qghip_circuit *c, const double *mat8, const int *ctrl, int n_ctrl,
                           int target);
/* Append a (controlled) rotation exp(-i theta G / 2), theta = params[param_index],
 * G = X (pauli 1), Y (2) or Z (3): qgate's RX / RY / RZ. */
int qghip_circuit_add_rotation(qghip_circuit *c, int pauli, int param_index, const int *ctrl,
                               int n_ctrl, int target);
/* QGHIP_FUSION_TILE (default): forward and backward run in tile passes (many gates per trip
 * through HBM; the backward sweep carries phi and lambda together); QGHIP_FUSION_OFF: one
 * launch per gate and per backward step. */
int qghip_circuit_set_fusion(qghip_circuit *c, int mode);
int qghip_circuit_n_params(const qghip_circuit *c, int *n_params);
/* <H> for H = sum_k coefs[k] P_k, P_k the Pauli string with X on xmask[k] bits, Z on
 * zmask[k] bits (both bits: Y). Leaves the final state and H|psi> on the device. */
int qghip_circuit_expectation(qghip_circuit *c, const double *params, int n_params, int n_terms,
                              const double *coefs, const uint64_t *xmask, const uint64_t *zmask,
                              double *value);
/* d<H>/dparams for the observable of the last expectation call (re-running the forward
 * pass if `params` differ from that call's); grad has n_params entries. */
int qghip_circuit_backward(qghip_circuit *c, const double *params, int n_params, double *grad);
/* expectation + backward. */
int qghip_circuit_gradient(qghip_circuit *c, const double *params, int n_params, int n_terms,
                           const double *coefs, const uint64_t *xmask, const uint64_t *zmask,
                           double *value, double *grad);
/* Copy of the final state of the last expectation call (2^n complex amplitudes). */
int qghip_circuit_state(qghip_circuit *c, void *out, int64_t array_size);

/* ---- batches of parameterised circuits (no reference counterpart) --------------------- */

/* Many circuits of uneven qubit counts, each with its own parameters and Pauli-sum
 * coefficients, evaluated together: <H> and the full adjoint gradient of every item in one
 * captured HIP graph of tile-pass kernels. Items added with the same circuit handle, qubit
 * count and observable terms (masks; the coefficients may differ) share a group. The first
 * run tunes per group the tile shape (register-resident or LDS passes, tile size) and the
 * workgroup and grid size of the LDS passes by timing candidates, then captures the graph;
 * later runs replay it with the current parameters.
 * Callers: variational workloads that evaluate many small circuits at once. */
typedef struct qghip_batch qghip_batch;
int qghip_batch_new(int prec, qghip_batch **out);
int qghip_batch_delete(qghip_batch *b);
/* Append an item: the gate structure of circuit `c` (copied; c must stay alive while items
 * built from it are added), its parameters and observable. *item receives its index. */
int qghip_batch_add(qghip_batch *b, const qghip_circuit *c, const double *params, int n_params, int n_terms,
                    const double *coefs, const uint64_t *xmask, const uint64_t *zmask, int *item);
/* Replace the parameters of one item (used by the next run). */
int qghip_batch_set_params(qghip_batch *b, int item, const double *params, int n_params);
/* Fix tuning knobs instead of searching them (0, or -1 for registers: search): tile bits
 * (1..12), threads per workgroup of the LDS passes (128 or 256), their persistent workgroups
 * per compute unit, and registers (1: forward and sweep passes keep their tiles in registers
 * where the group's qubit count allows full-size tiles, 0: LDS passes only). Re-tunes on the
 * next run. */
int qghip_batch_set_tuning(qghip_batch *b, int tile_bits, int threads, int grid_per_cu, int registers);
/* Evaluate every item: values[i] = <H>_i; grads (if not null) receives each item's gradient
 * (its circuit's n_params entries) in item order. use_graph = 0 launches the same kernels
 * without the graph. */
int qghip_batch_run(qghip_batch *b, int use_graph, double *values, double *grads, int64_t grads_size);
/* JSON list of the groups with their chosen tuning. */
int qghip_batch_describe(const qghip_batch *b, char *buf, int size);

/* ---- launch shape (no reference counterpart) ----------------------------------------- */

/* Kernels over more than `items` work items (one per amplitude, unit or pair) are launched
 * in pieces of `items` (a power of two in [2^12, 2^30]; default 2^30), keeping every grid
 * below the 2^32 work-item limit at 33+ qubits. Results do not depend on it; tests lower it
 * to run the piecewise path on small states. */
int qghip_set_launch_piece(int64_t items);
/* Full-size tile passes keep the tile in registers, going through LDS once per phase of
 * gates (on, the default), or apply every gate to the tile in LDS (off). Results agree to
 * rounding; tests compare the two. */
int qghip_set_register_tiles(int on);

/* ---- measurement helper (no reference counterpart) ------------------------------------ */

/* Time `reps` applications of one (controlled) gate with HIP events on the library stream;
 * avg_ms receives the mean kernel duration. For roofline reporting. */
int qghip_profile_gate(qghip_qubit_processor *p, qghip_qubit_states *qs, const double *mat8,
                       const int *local_control_lanes, int n_controls, int local_target_lane,
                       int reps, double *avg_ms);

/* Time `reps` runs of the library's non-temporal streaming copy (one 16-byte unit per
 * thread, one wavefront per block, as the fastest gate kernels) over two fresh device buffers of `bytes` bytes (a multiple of 16),
 * HIP events on the library stream; avg_ms receives the mean kernel duration. The copy moves
 * 2 * bytes per run: the HBM streaming ceiling the gate kernels are compared with. */
int qghip_profile_copy(int64_t bytes, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* QGATE_HIP_H */
