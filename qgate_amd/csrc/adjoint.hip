// Adjoint differentiation of parameterised circuits: <H> = <psi|H|psi> for a Pauli sum H and
// its full gradient, with |psi> = U_N ... U_1 |0...0>. (No reference counterpart: the
// reference runtime has no gradients or expectation values.)
//
// Method (adjoint / "reverse-mode" state-vector differentiation):
//   forward:  psi = U_N ... U_1 |0>,  lambda = H psi,  <H> = Re <psi|lambda>
//   backward: for k = N .. 1, with phi = psi_k (state after U_k) and lambda_k:
//               if U_k = exp(-i theta G / 2) (optionally controlled, G in {X, Y, Z}):
//                 d<H>/dtheta += 2 Re <lambda_k| dU_k/dtheta |psi_{k-1}>
//                              = Im <lambda_k| C G |phi>     (C: projector on the controls)
//               phi <- U_k^dagger phi, lambda <- U_k^dagger lambda
// Two state vectors (phi, lambda) stay resident; each backward step is one pass that reads
// the pair in both vectors, adds the gradient term in FP64 and writes both un-computed
// pairs. Gradient terms of one step are reduced per block into a partial buffer and summed
// into the parameter's slot on the device in a fixed order.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "../../include/qgate_hip.h"
#include "circuit.hpp"
#include "gate_matrix.hpp"
#include "kernels.hpp"
#include "qg_runtime.hpp"

using namespace qg;

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(u64 n, int b) { return (unsigned)((n + b - 1) / b); }

typedef CircuitOp Op;

// Row-major (re, im) entries of the Pauli matrix X (1), Y (2) or Z (3); zeros otherwise.
void set_generator(double *gen, int pauli) {
    for (int i = 0; i < 8; ++i) gen[i] = 0.;
    if (pauli == 1) {          // [[0, 1], [1, 0]]
        gen[2] = 1.; gen[4] = 1.;
    } else if (pauli == 2) {   // [[0, -i], [i, 0]]
        gen[3] = -1.; gen[5] = 1.;
    } else if (pauli == 3) {   // [[1, 0], [0, -1]]
        gen[0] = 1.; gen[6] = -1.;
    }
}

// Pauli-sum observable grouped by X mask: lambda[x] = sum_g psi[x ^ xm_g] * sum_{k in g}
// c_k (-1)^popcount(x & zm_k), where c_k already holds the (-i)^(number of Y) factor.
struct TermDev {
    double cre, cim;
    u64 zmask;
};
struct GroupDev {
    u64 xmask;
    int first, count;
};

template <class R>
__global__ __launch_bounds__(kBlock) void k_apply_pauli_sum(Cx<R> *__restrict__ lam, const Cx<R> *__restrict__ psi,
                                                            const GroupDev *__restrict__ groups, int n_groups,
                                                            const TermDev *__restrict__ terms, u64 n,
                                                            u64 first) {
    const u64 x = first + (u64)blockIdx.x * kBlock + threadIdx.x;
    if (x >= n) return;
    double ar = 0., ai = 0.;
    for (int g = 0; g < n_groups; ++g) {
        const GroupDev gr = groups[g];
        double cr = 0., ci = 0.;
        for (int k = gr.first; k < gr.first + gr.count; ++k) {
            const TermDev t = terms[k];
            const double sgn = (__popcll(x & t.zmask) & 1) ? -1. : 1.;
            cr += sgn * t.cre;
            ci += sgn * t.cim;
        }
        const Cx<R> v = psi[x ^ gr.xmask];
        ar += cr * (double)v.re - ci * (double)v.im;
        ai += cr * (double)v.im + ci * (double)v.re;
    }
    lam[x] = cx<R>((R)ar, (R)ai);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kBlock / 64];
    v = wave_sum(v);
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = v;
    __syncthreads();
    double s = 0.;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) s += part[i];
    __syncthreads();
    return s;
}

constexpr int kItemsPerThread = 8;

// partial[b] = sum over this block's amplitudes of Re(conj(a) * b)
template <class R>
__global__ __launch_bounds__(kBlock) void k_dot_re(const Cx<R> *__restrict__ a, const Cx<R> *__restrict__ b,
                                                   u64 n, double *partial) {
    double s = 0.;
    const u64 base = (u64)blockIdx.x * (kBlock * kItemsPerThread) + threadIdx.x;
    for (int r = 0; r < kItemsPerThread; ++r) {
        const u64 i = base + (u64)r * kBlock;
        if (i < n) {
            const Cx<R> u = a[i], v = b[i];
            s += (double)u.re * (double)v.re + (double)u.im * (double)v.im;
        }
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0] (+)= sum of partial[0..n) in a fixed order.
__global__ __launch_bounds__(1024) void k_collect(const double *partial, u64 n, double *out, int accumulate) {
    double s = 0.;
    for (u64 i = threadIdx.x; i < n; i += 1024) s += partial[i];
    s = wave_sum(s);
    __shared__ double part[16];
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int i = 0; i < 16; ++i) t += part[i];
        out[0] = accumulate ? out[0] + t : t;
    }
}

template <class R>
struct BackArgs {
    Mat2<R> udag;          // U_k^dagger, applied to both vectors
    BitInsert ins;         // sorted control + target positions
    u64 ctrl, tbit, n_pairs;
    double gen[8];         // generator G of a rotation, row-major (re, im) pairs
    int has_grad;          // 1 for rotations, 0 for fixed gates
};

// One backward step: gradient term Im <lambda| C G |phi> from the current pairs, then both
// vectors are un-computed by U^dagger.
template <class R>
__global__ __launch_bounds__(kBlock) void k_backward_step(Cx<R> *__restrict__ phi, Cx<R> *__restrict__ lam,
                                                          const BackArgs<R> g, double *partial) {
    double s = 0.;
    const u64 base = (u64)blockIdx.x * (kBlock * kItemsPerThread) + threadIdx.x;
    for (int r = 0; r < kItemsPerThread; ++r) {
        const u64 k = base + (u64)r * kBlock;
        if (k >= g.n_pairs) break;
        const u64 i0 = insert_zero_bits(k, g.ins) | g.ctrl, i1 = i0 | g.tbit;
        const Cx<R> p0 = phi[i0], p1 = phi[i1], l0 = lam[i0], l1 = lam[i1];
        if (g.has_grad) {
            // q = G phi on the pair, in FP64 (G has entries 0, +-1, +-i), as a matrix product
            // rather than a branch per Pauli kind
            const double p0r = p0.re, p0i = p0.im, p1r = p1.re, p1i = p1.im;
            const double q0r = g.gen[0] * p0r - g.gen[1] * p0i + g.gen[2] * p1r - g.gen[3] * p1i;
            const double q0i = g.gen[0] * p0i + g.gen[1] * p0r + g.gen[2] * p1i + g.gen[3] * p1r;
            const double q1r = g.gen[4] * p0r - g.gen[5] * p0i + g.gen[6] * p1r - g.gen[7] * p1i;
            const double q1i = g.gen[4] * p0i + g.gen[5] * p0r + g.gen[6] * p1i + g.gen[7] * p1r;
            // Im(conj(l0) q0 + conj(l1) q1)
            s += (double)l0.re * q0i - (double)l0.im * q0r + (double)l1.re * q1i - (double)l1.im * q1r;
        }
        phi[i0] = cadd(cmul(g.udag.m00, p0), cmul(g.udag.m01, p1));
        phi[i1] = cadd(cmul(g.udag.m10, p0), cmul(g.udag.m11, p1));
        lam[i0] = cadd(cmul(g.udag.m00, l0), cmul(g.udag.m01, l1));
        lam[i1] = cadd(cmul(g.udag.m10, l0), cmul(g.udag.m11, l1));
    }
    if (g.has_grad) {
        s = block_sum(s);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

}  // namespace

namespace {

template <class F>
int guarded_adj(F &&f) {
    try {
        std::lock_guard<std::recursive_mutex> lk(runtime().mtx);
        f();
        return QGHIP_OK;
    } catch (const std::invalid_argument &e) {
        last_error() = e.what();
        return QGHIP_EVALUE;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return QGHIP_ERROR;
    }
}

void require_arg(bool ok, const char *msg) {
    if (!ok) throw std::invalid_argument(msg);
}

void ensure_states(qghip_circuit *c) {
    const size_t bytes = amp_bytes(c->prec) << c->n_qubits;
    if (c->psi && c->state_bytes == bytes) return;
    if (c->psi) dev_free(c->psi, c->state_bytes);
    if (c->lam) dev_free(c->lam, c->state_bytes);
    c->psi = c->lam = nullptr;
    c->psi = dev_alloc(bytes);
    c->lam = dev_alloc(bytes);
    c->state_bytes = bytes;
}

GateMatrix op_matrix(const Op &op, const double *params) {
    if (op.kind == 0) return op.mat;
    return rotation(op.pauli, params[op.param]);
}

struct Observable {
    std::vector<GroupDev> groups;
    std::vector<TermDev> terms;
};

Observable make_observable(int n_terms, const double *coefs, const uint64_t *xmask, const uint64_t *zmask) {
    Observable o;
    std::vector<int> order(n_terms);
    for (int i = 0; i < n_terms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return xmask[a] < xmask[b]; });
    for (int idx : order) {
        // (-i)^(number of Y), Y = bits set in both masks
        const int ny = __builtin_popcountll(xmask[idx] & zmask[idx]) & 3;
        static const double pr[4] = {1., 0., -1., 0.}, pi[4] = {0., -1., 0., 1.};
        TermDev t;
        t.cre = coefs[idx] * pr[ny];
        t.cim = coefs[idx] * pi[ny];
        t.zmask = zmask[idx];
        if (o.groups.empty() || o.groups.back().xmask != xmask[idx]) {
            GroupDev g;
            g.xmask = xmask[idx];
            g.first = (int)o.terms.size();
            g.count = 0;
            o.groups.push_back(g);
        }
        o.terms.push_back(t);
        o.groups.back().count += 1;
    }
    return o;
}

template <class R>
double forward_impl(qghip_circuit *c, const double *params, const Observable &obs) {
    ensure_states(c);
    Cx<R> *psi = static_cast<Cx<R> *>(c->psi), *lam = static_cast<Cx<R> *>(c->lam);
    hipStream_t s = stream();
    launch_set_zero_state<R>(psi, c->n_qubits);
    if (c->fusion == QGHIP_FUSION_TILE) {
        std::vector<QueuedGate> gates;
        gates.reserve(c->ops.size());
        for (const Op &op : c->ops) gates.push_back(QueuedGate{op_matrix(op, params), op.ctrl, op.target});
        apply_gate_sequence<R>(psi, c->n_qubits, gates, 0, -1);
    } else {
        for (const Op &op : c->ops) launch_gate<R>(psi, c->n_qubits, op_matrix(op, params), op.ctrl, op.target);
    }
    if (c->fusion == QGHIP_FUSION_TILE) {
        std::vector<PauliGroup> groups;
        std::vector<PauliTerm> terms;
        for (const GroupDev &g : obs.groups) groups.push_back(PauliGroup{g.xmask, g.first, g.count, 0u});
        for (const TermDev &t : obs.terms) terms.push_back(PauliTerm{t.cre, t.cim, t.zmask});
        return pauli_expectation<R>(psi, lam, c->n_qubits, groups, terms);
    }
    const u64 n = u64(1) << c->n_qubits;
    const size_t obs_bytes = sizeof(GroupDev) * obs.groups.size() + sizeof(TermDev) * obs.terms.size();
    char *d_obs = static_cast<char *>(dev_alloc(obs_bytes + 16));
    GroupDev *d_groups = reinterpret_cast<GroupDev *>(d_obs);
    TermDev *d_terms = reinterpret_cast<TermDev *>(d_obs + sizeof(GroupDev) * obs.groups.size());
    const u64 per_block = (u64)kBlock * kItemsPerThread;
    const u64 n_blocks = (n + per_block - 1) / per_block;
    const size_t pbytes = sizeof(double) * (n_blocks + 1);
    double *partial = static_cast<double *>(dev_alloc(pbytes));
    Runtime &rt = runtime();
    hipError_t e = hipMemcpyAsync(d_groups, obs.groups.data(), sizeof(GroupDev) * obs.groups.size(),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_terms, obs.terms.data(), sizeof(TermDev) * obs.terms.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        for_each_piece(n, [&](u64 first, u64 count) {
            hipLaunchKernelGGL((k_apply_pauli_sum<R>), blocks_for(count, kBlock), kBlock, 0, s, lam, psi,
                               d_groups, (int)obs.groups.size(), d_terms, n, first);
        });
        hipLaunchKernelGGL((k_dot_re<R>), (unsigned)n_blocks, kBlock, 0, s, psi, lam, n, partial);
        hipLaunchKernelGGL(k_collect, 1, 1024, 0, s, partial, n_blocks, partial + n_blocks, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(rt.h_pinned, partial + n_blocks, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // obs tables are host temporaries
    dev_free(partial, pbytes);
    dev_free(d_obs, obs_bytes + 16);
    QG_HIP(e);
    return rt.h_pinned[0];
}

template <class R>
void backward_tiles(qghip_circuit *c, const double *params, double *grad) {
    Cx<R> *phi = static_cast<Cx<R> *>(c->psi), *lam = static_cast<Cx<R> *>(c->lam);
    const int n_ops = (int)c->ops.size();
    std::vector<SweepGate> gates;
    gates.reserve(n_ops);
    for (int k = n_ops - 1; k >= 0; --k) {   // the sweep runs from the last gate back
        const Op &op = c->ops[k];
        gates.push_back(SweepGate{dagger(op_matrix(op, params)), op.kind == 1 ? op.pauli : 0, op.ctrl, op.target});
    }
    const size_t bytes = sizeof(double) * (size_t)std::max(1, n_ops);
    double *d_terms = static_cast<double *>(dev_alloc(bytes));
    std::vector<double> terms(std::max(1, n_ops));
    hipError_t e = hipSuccess;
    try {
        adjoint_sweep<R>(phi, lam, c->n_qubits, gates, d_terms, 0, -1);
        e = hipMemcpyAsync(terms.data(), d_terms, bytes, hipMemcpyDeviceToHost, stream());
        if (e == hipSuccess) e = hipStreamSynchronize(stream());
    } catch (...) {
        dev_free(d_terms, bytes);
        throw;
    }
    dev_free(d_terms, bytes);
    QG_HIP(e);
    for (int i = 0; i < c->n_params; ++i) grad[i] = 0.;
    for (int i = 0; i < n_ops; ++i) {   // sweep order: last gate first
        const Op &op = c->ops[n_ops - 1 - i];
        if (op.kind == 1) grad[op.param] += terms[i];
    }
}

template <class R>
void backward_impl(qghip_circuit *c, const double *params, double *grad) {
    if (c->fusion == QGHIP_FUSION_TILE) {
        backward_tiles<R>(c, params, grad);
        return;
    }
    Cx<R> *phi = static_cast<Cx<R> *>(c->psi), *lam = static_cast<Cx<R> *>(c->lam);
    hipStream_t s = stream();
    const int n = c->n_qubits;
    double *d_grad = static_cast<double *>(dev_alloc(sizeof(double) * (c->n_params + 1)));
    const u64 max_pairs = u64(1) << (n - 1);
    const u64 per_block = (u64)kBlock * kItemsPerThread;
    const u64 max_blocks = (max_pairs + per_block - 1) / per_block;
    double *partial = static_cast<double *>(dev_alloc(sizeof(double) * (max_blocks + 1)));
    hipError_t e = hipMemsetAsync(d_grad, 0, sizeof(double) * (c->n_params + 1), s);
    for (int k = (int)c->ops.size() - 1; k >= 0 && e == hipSuccess; --k) {
        const Op &op = c->ops[k];
        BackArgs<R> g;
        g.udag = round_matrix<R>(dagger(op_matrix(op, params)));
        std::vector<int> pos(op.ctrl);
        pos.push_back(op.target);
        std::sort(pos.begin(), pos.end());
        g.ins.n = (int)pos.size();
        for (int i = 0; i < g.ins.n; ++i) g.ins.pos[i] = pos[i];
        g.ctrl = 0;
        for (int q : op.ctrl) g.ctrl |= u64(1) << q;
        g.tbit = u64(1) << op.target;
        g.n_pairs = u64(1) << (n - g.ins.n);
        g.has_grad = op.kind == 1 ? 1 : 0;
        set_generator(g.gen, op.kind == 1 ? op.pauli : 0);
        const u64 n_blocks = (g.n_pairs + per_block - 1) / per_block;
        hipLaunchKernelGGL((k_backward_step<R>), (unsigned)n_blocks, kBlock, 0, s, phi, lam, g, partial);
        if (op.kind == 1)   // grad[param] += this step's term, in op order
            hipLaunchKernelGGL(k_collect, 1, 1024, 0, s, partial, n_blocks, d_grad + op.param, 1);
        e = hipGetLastError();
    }
    std::vector<double> h(c->n_params + 1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), d_grad, sizeof(double) * c->n_params, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(partial, sizeof(double) * (max_blocks + 1));
    dev_free(d_grad, sizeof(double) * (c->n_params + 1));
    QG_HIP(e);
    for (int i = 0; i < c->n_params; ++i) grad[i] = h[i];
}

}  // namespace

extern "C" {

int qghip_circuit_new(int prec, int n_qubits, qghip_circuit **out) {
    return guarded_adj([&] {
        require_arg(out != nullptr, "null argument");
        require_arg(prec == QGHIP_FP64 || prec == QGHIP_FP32, "bad precision");
        require_arg(1 <= n_qubits && n_qubits <= 34, "n_qubits out of range (1..34)");
        qghip_circuit *c = new qghip_circuit;
        c->prec = prec == QGHIP_FP64 ? precFP64 : precFP32;
        c->n_qubits = n_qubits;
        *out = c;
    });
}

int qghip_circuit_delete(qghip_circuit *c) {
    return guarded_adj([&] {
        if (c == nullptr) return;
        if (c->psi) dev_free(c->psi, c->state_bytes);
        if (c->lam) dev_free(c->lam, c->state_bytes);
        delete c;
    });
}

static void check_lanes(qghip_circuit *c, const int *ctrl, int n_ctrl, int target) {
    require_arg(0 <= target && target < c->n_qubits, "target out of range");
    require_arg(n_ctrl >= 0 && (n_ctrl == 0 || ctrl != nullptr), "bad control list");
    std::vector<int> all(ctrl, ctrl + n_ctrl);
    all.push_back(target);
    std::sort(all.begin(), all.end());
    for (size_t i = 0; i < all.size(); ++i) {
        require_arg(0 <= all[i] && all[i] < c->n_qubits, "control out of range");
        require_arg(i == 0 || all[i] != all[i - 1], "control and target overlap");
    }
}

int qghip_circuit_add_gate(qghip_circuit *c, const double *mat8, const int *ctrl, int n_ctrl, int target) {
    return guarded_adj([&] {
        require_arg(c != nullptr && mat8 != nullptr, "null argument");
        check_lanes(c, ctrl, n_ctrl, target);
        Op op;
        op.kind = 0;
        op.pauli = 0;
        op.param = -1;
        op.mat = gate_matrix_from_array(mat8);
        op.ctrl.assign(ctrl, ctrl + n_ctrl);
        op.target = target;
        c->ops.push_back(op);
    });
}

int qghip_circuit_add_rotation(qghip_circuit *c, int pauli, int param_index, const int *ctrl, int n_ctrl,
                               int target) {
    return guarded_adj([&] {
        require_arg(c != nullptr, "null argument");
        require_arg(1 <= pauli && pauli <= 3, "pauli must be 1 (X), 2 (Y) or 3 (Z)");
        require_arg(param_index >= 0, "negative parameter index");
        check_lanes(c, ctrl, n_ctrl, target);
        Op op;
        op.kind = 1;
        op.pauli = pauli;
        op.param = param_index;
        op.ctrl.assign(ctrl, ctrl + n_ctrl);
        op.target = target;
        c->ops.push_back(op);
        c->n_params = std::max(c->n_params, param_index + 1);
    });
}

int qghip_circuit_set_fusion(qghip_circuit *c, int mode) {
    return guarded_adj([&] {
        require_arg(c != nullptr, "null argument");
        require_arg(mode == QGHIP_FUSION_OFF || mode == QGHIP_FUSION_TILE, "unknown fusion mode");
        c->fusion = mode;
        c->forward_valid = false;
    });
}

int qghip_circuit_n_params(const qghip_circuit *c, int *n_params) {
    return guarded_adj([&] {
        require_arg(c != nullptr && n_params != nullptr, "null argument");
        *n_params = c->n_params;
    });
}

int qghip_circuit_expectation(qghip_circuit *c, const double *params, int n_params, int n_terms,
                              const double *coefs, const uint64_t *xmask, const uint64_t *zmask,
                              double *value) {
    return guarded_adj([&] {
        require_arg(c != nullptr && value != nullptr, "null argument");
        require_arg(n_params >= c->n_params && (c->n_params == 0 || params != nullptr), "too few parameters");
        require_arg(n_terms >= 1 && coefs && xmask && zmask, "empty observable");
        const u64 lim = c->n_qubits >= 64 ? ~u64(0) : (u64(1) << c->n_qubits) - 1;
        for (int i = 0; i < n_terms; ++i)
            require_arg((xmask[i] & ~lim) == 0 && (zmask[i] & ~lim) == 0, "Pauli term outside the register");
        c->forward_valid = false;
        Observable obs = make_observable(n_terms, coefs, xmask, zmask);
        if (c->prec == precFP32)
            *value = forward_impl<float>(c, params, obs);
        else
            *value = forward_impl<double>(c, params, obs);
        c->forward_params.assign(params, params + c->n_params);
        c->obs_coefs.assign(coefs, coefs + n_terms);
        c->obs_xmask.assign(xmask, xmask + n_terms);
        c->obs_zmask.assign(zmask, zmask + n_terms);
        c->forward_valid = true;
    });
}

int qghip_circuit_backward(qghip_circuit *c, const double *params, int n_params, double *grad) {
    return guarded_adj([&] {
        require_arg(c != nullptr && grad != nullptr, "null argument");
        require_arg(n_params >= c->n_params && (c->n_params == 0 || params != nullptr), "too few parameters");
        require_arg(!c->obs_coefs.empty(), "backward needs a preceding expectation call");
        const bool same = c->forward_valid &&
                          std::equal(c->forward_params.begin(), c->forward_params.end(), params);
        if (!same) {   // recompute the forward state for these parameters
            Observable obs = make_observable((int)c->obs_coefs.size(), c->obs_coefs.data(),
                                             c->obs_xmask.data(), c->obs_zmask.data());
            if (c->prec == precFP32)
                forward_impl<float>(c, params, obs);
            else
                forward_impl<double>(c, params, obs);
        }
        c->forward_valid = false;   // the sweep un-computes both vectors
        if (c->prec == precFP32)
            backward_impl<float>(c, params, grad);
        else
            backward_impl<double>(c, params, grad);
    });
}

int qghip_circuit_gradient(qghip_circuit *c, const double *params, int n_params, int n_terms,
                           const double *coefs, const uint64_t *xmask, const uint64_t *zmask,
                           double *value, double *grad) {
    int rc = qghip_circuit_expectation(c, params, n_params, n_terms, coefs, xmask, zmask, value);
    if (rc != QGHIP_OK) return rc;
    return qghip_circuit_backward(c, params, n_params, grad);
}

int qghip_circuit_state(qghip_circuit *c, void *out, int64_t array_size) {
    return guarded_adj([&] {
        require_arg(c != nullptr && out != nullptr, "null argument");
        require_arg(c->forward_valid, "no forward state (call expectation first)");
        const int64_t n = int64_t(1) << c->n_qubits;
        require_arg(array_size >= n, "array size too small.");
        QG_HIP(hipMemcpyAsync(out, c->psi, c->state_bytes, hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

}  // extern "C"
