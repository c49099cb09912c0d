// Launchers of the qgate HIP kernels. All launches go to the runtime stream.
#pragma once

#include <vector>

#include "gate_matrix.hpp"
#include "qg_common.hpp"

namespace qg {

// ---- gates -----------------------------------------------------------------------------
// Apply a 2x2 matrix to `target`, on the subspace where every bit of `ctrl_mask` is 1.
// Dispatches to the streaming pair kernel, the bit-0 pair kernel, or (diagonal matrix with
// m00 == 1) the phase kernel that only touches amplitudes with the target bit set.
template <class R>
void launch_gate(Cx<R> *a, int n_lanes, const GateMatrix &m, const std::vector<int> &ctrl,
                 int target);

// Apply a dense 2^k x 2^k matrix (row-major complex, FP64) to the lanes `lanes` (lanes[j] is
// bit j of the matrix index). k <= 6.
template <class R>
void launch_unitary(Cx<R> *a, int n_lanes, const std::vector<int> &lanes,
                    const std::vector<std::complex<double>> &mat);

// A gate waiting in a fusion queue: FP64 matrix, control lanes, target lane.
struct QueuedGate {
    GateMatrix m;
    std::vector<int> ctrl;
    int target;
};

// Apply `gates` in order in tile passes (k_tile.hip): each pass loads tiles of up to 2^max_bits
// amplitudes (<= 0: the precision's default) whose lowest `low_bits` bits are contiguous
// (< 0: default) and applies every gate it can hold before writing them back.
template <class R>
void apply_gate_sequence(Cx<R> *a, int n_lanes, const std::vector<QueuedGate> &gates, int max_bits,
                         int low_bits);

// Apply `gates` in order as dense unitaries of runs of consecutive gates on at most `width`
// (2..6) lanes, each built in FP64 on the host (k_gates.hip); a run of one gate, or a gate on
// more lanes than `width`, goes to the streaming gate kernels.
template <class R>
void apply_dense_sequence(Cx<R> *a, int n_lanes, const std::vector<QueuedGate> &gates, int width);

// One step of the adjoint backward sweep: U^dagger (FP64) with its controls and target, and
// the generator of a rotation (pauli 1 X, 2 Y, 3 Z; 0 for a fixed gate).
struct SweepGate {
    GateMatrix udag;
    int pauli;
    std::vector<int> ctrl;
    int target;
};

// Backward sweep in tile passes over phi and lambda: for gate i in order, d_terms[i] =
// Im<lambda| C G |phi> (rotations only; FP64, summed over tiles in a fixed order), then
// phi, lambda <- U^dagger. Synchronises.
template <class R>
void adjoint_sweep(Cx<R> *phi, Cx<R> *lam, int n_lanes, const std::vector<SweepGate> &gates, double *d_terms,
                   int max_bits, int low_bits);

// Pauli-sum observable: terms with the same X mask form a group; cre/cim carry the term's
// coefficient times (-i)^(number of Y factors); xlocal is filled in by the launcher.
struct PauliTerm {
    double cre, cim;
    u64 zmask;
};
struct PauliGroup {
    u64 xmask;
    int first, count;
    unsigned xlocal;
};

// lambda = H psi in tile passes (groups whose X bits fit one tile share a pass) and
// <psi|H|psi> = Re<psi|lambda> (FP64, fixed order) from the last pass. Synchronises.
template <class R>
double pauli_expectation(const Cx<R> *psi, Cx<R> *lam, int n_lanes, const std::vector<PauliGroup> &groups,
                         const std::vector<PauliTerm> &terms);

// ---- state -----------------------------------------------------------------------------
template <class R> void launch_set_zero_state(Cx<R> *a, int n_lanes);
// dst = src (bytes a multiple of 16): the non-temporal streaming copy used as the HBM ceiling.
void launch_stream_copy(void *dst, const void *src, u64 bytes);
// P(bit `lane` == 0), FP64 accumulation; synchronises.
template <class R> double calc_prob(const Cx<R> *a, int n_lanes, int lane);
// Doubles of partial-sum buffer calc_prob_in needs for this lane (the partials, run sums and
// total), and calc_prob with that buffer given instead of taken from the allocator (no
// allocation; the buffer must hold at least that many doubles).
template <class R> u64 calc_prob_buffer_len(int n_lanes, int lane);
template <class R> double calc_prob_in(const Cx<R> *a, int n_lanes, int lane, double *buf, u64 buf_len);
template <class R> void launch_decohere(Cx<R> *a, int n_lanes, int lane, int value, R norm);
// launch_decohere, then calc_prob(prob_lane) of the collapsed state, in one pass; synchronises.
template <class R> double decohere_prob(Cx<R> *a, int n_lanes, int lane, int value, R norm, int prob_lane);
// Pauli string P over lanes x | z (X: x, Z: z, Y: both): P(measuring P gives +1) in FP64,
// i.e. sum over pairs of |((1 + P) psi)_y|^2 / 2; synchronises.
template <class R> double calc_pauli_prob(const Cx<R> *a, int n_lanes, u64 x, u64 z);
// Collapse onto the eigenspace P = (-1)^value: psi <- (1 + (-1)^value P) psi / (2 sqrt(p_value)),
// p_0 = prob, p_1 = 1 - prob.
template <class R> void launch_pauli_decohere(Cx<R> *a, int n_lanes, u64 x, u64 z, int value, double prob);
template <class R> void launch_decohere_and_separate(Cx<R> *dst, const Cx<R> *src, int n_lanes,
                                                     int lane, int value, R norm);
template <class R> void launch_reset(Cx<R> *a, int n_lanes, int lane);
template <class R> void launch_set_basis1(Cx<R> *a, int value);
// Kronecker product: srcs[last] occupies the lowest bits; the rest of dst is zeroed.
template <class R> void launch_join(Cx<R> *dst, int dst_lanes, const std::vector<const Cx<R> *> &srcs,
                                   const std::vector<int> &src_lanes);

// ---- readout ---------------------------------------------------------------------------
struct StateView {                // one state vector as seen from the external index space
    const void *ptr;
    int n_lanes;
    std::vector<int> local_to_ext; // external bit of each local lane
};

// out[j] = 0 if ext & empty_mask else prod_q f(a_q[local_q(ext)]), ext = start + j * step.
// mathop 0: complex amplitude product, 1: |amplitude|^2 product. Host output.
template <class R>
void get_states(void *host_out, int mathop, const std::vector<StateView> &qs, u64 empty_mask,
                i64 n_states, i64 start, i64 step);

// p[x] = sum over hidden h of prod_q |a_q(x + h << n_lanes)|^2 (hidden lanes in the MSBs).
// Writes `out_dtype_double ? double : R` into device memory `d_out`.
template <class R>
void launch_prob_array(void *d_out, bool out_double, const std::vector<StateView> &qs,
                       int n_lanes, int n_hidden);

// ---- sampling pool ---------------------------------------------------------------------
struct PoolData {
    i64 n = 0;                 // number of bins (2^n_lanes)
    double *p = nullptr;       // bin probabilities (unnormalised), padded to whole chunks;
                               // null for a bounded pool, which reads them from `amps`
    const void *amps = nullptr;   // bounded pool: the state vector, |a|^2 taken in FP64
    int amp_kind = 0;             // 0: p array, 1: complex64 amplitudes, 2: complex128
    double *e_inc = nullptr;   // per 16-bin chunk: inclusive prefix within its tile
    double *m_inc = nullptr;   // per 4096-bin tile: inclusive prefix within its super-tile
    double *h_inc = nullptr;   // per 2^22-bin super-tile: inclusive prefix over the pool
    i64 *d_last_nonzero = nullptr;   // last nonzero bin + 1 (0: none)
    i64 *t_last = nullptr;           // per tile / super-tile: last nonzero bin, -1 if none
    i64 *s_last = nullptr;
    i64 n_chunks = 0, n_tiles = 0, n_supers = 0;
    double total = 0., norm = 0.;
    i64 last_nonzero = -1;
    size_t bytes = 0;
    void *block = nullptr;     // single device allocation holding the arrays
};
// bounded: keep no probability array (2^n_lanes / 16 doubles of prefix sums instead of
// 2^n_lanes + that); draws then read |a|^2 of the found chunk from the state vector, which
// must not change while the pool is used.
template <class R>
void pool_build_from_state(PoolData &pd, const Cx<R> *a, int n_lanes, bool bounded);
void pool_build_from_probs(PoolData &pd, int n_lanes);   // pd.p already filled
// Device bytes of a pool over n_lanes lanes, with or without its probability array.
size_t pool_bytes(int n_lanes, bool with_p);
void pool_alloc(PoolData &pd, int n_lanes, bool with_p = true);
void pool_free(PoolData &pd);
void pool_sample(const PoolData &pd, i64 *host_obs, i64 n_samples, const double *host_u,
                 const std::vector<int> &empty_lanes);

}  // namespace qg
