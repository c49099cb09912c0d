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

// ---- state -----------------------------------------------------------------------------
template <class R> void launch_set_zero_state(Cx<R> *a, int n_lanes);
// P(bit `lane` == 0), FP64 accumulation; synchronises.
template <class R> double calc_prob(const Cx<R> *a, int n_lanes, int lane);
template <class R> void launch_decohere(Cx<R> *a, int n_lanes, int lane, int value, R norm);
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
    double *p = nullptr;       // bin probabilities (unnormalised), padded to whole chunks
    double *e_inc = nullptr;   // per 16-bin chunk: inclusive prefix within its tile
    double *m_inc = nullptr;   // per 4096-bin tile: inclusive prefix within its super-tile
    double *h_inc = nullptr;   // per 2^22-bin super-tile: inclusive prefix over the pool
    i64 *d_last_nonzero = nullptr;
    i64 n_chunks = 0, n_tiles = 0, n_supers = 0;
    double total = 0., norm = 0.;
    i64 last_nonzero = -1;
    size_t bytes = 0;
    void *block = nullptr;     // single device allocation holding the arrays
};
template <class R>
void pool_build_from_state(PoolData &pd, const Cx<R> *a, int n_lanes);
void pool_build_from_probs(PoolData &pd, int n_lanes);   // pd.p already filled
void pool_alloc(PoolData &pd, int n_lanes);
void pool_free(PoolData &pd);
void pool_sample(const PoolData &pd, i64 *host_obs, i64 n_samples, const double *host_u,
                 const std::vector<int> &empty_lanes);

}  // namespace qg
