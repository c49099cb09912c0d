// Marginal probability arrays: p[x] = sum over the hidden lanes h of prod_q |a_q(x, h)|^2, the
// array behind Qubits.prob with hidden qregs and behind sampling pools over some of the
// qubits (reference: CPUQubitsStatesGetter.cpp:159-254 prepareProbArray,
// DeviceProbArrayCalculator.cu; hidden lanes sit in the most significant external bits).
//
// The sum over hidden lanes factorises over state vectors (each lane belongs to one vector),
// so each vector with hidden lanes is first reduced on its own to m_q[k] over its kept lanes
// k (in local order), by k_marginal; then p[x] = prod_q m_q[k_q(x)] (or |a_q(local_q(x))|^2
// for a vector with no hidden lane) by k_combine, the index maps as byte tables. A single
// vector whose kept lanes are external bits 0, 1, ... in local order writes p directly.
//
// k_marginal streams the state in its own order: the 64 lanes of a wavefront read 64
// consecutive 16-byte units (the lowest L = 6 + log2(amplitudes per unit) index bits), and
// each lane walks the upper index bits, hidden ones fastest, accumulating |a|^2 in FP64 with
// four loads in flight. Hidden bits among the lowest L are summed across lanes by a fixed
// butterfly; hidden ranges too long for one wavefront are split into slices whose partial
// sums a second kernel adds in a fixed order. Every sum has a fixed order, so repeated calls
// return the same bits.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"
#include "qg_runtime.hpp"

namespace qg {

namespace {

constexpr int kMargThreads = 256;
constexpr int kBatch = 4;

// next subset of `mask` after `cur` in increasing order (wraps to 0 after mask itself)
__host__ __device__ __forceinline__ u64 next_sub(u64 cur, u64 mask) { return ((cur | ~mask) + 1) & mask; }

// the low bits of v deposited at the set bits of mask, lowest first
__host__ __device__ inline u64 deposit(u64 v, u64 mask) {
    u64 r = 0;
    for (u64 m = mask; m != 0; m &= m - 1) {
        if (v & 1) r |= m & (~m + 1);
        v >>= 1;
    }
    return r;
}

// the bits of v at the set bits of mask, packed from bit 0
__host__ __device__ inline u64 extract(u64 v, u64 mask) {
    u64 r = 0;
    int k = 0;
    for (u64 m = mask; m != 0; m &= m - 1, ++k)
        if (v & (m & (~m + 1))) r |= u64(1) << k;
    return r;
}

__device__ __forceinline__ double wave_total(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct MargArgs {
    u64 ku_mask, hu_mask;   // kept / hidden amplitude-index bits at or above L
    u64 kl_mask;            // kept amplitude-index bits below L
    u64 n_waves;
    u64 n_out;              // 2^(kept lanes): outputs, and the stride of partial slices
    int hl_lane;            // hidden bits of the lane index (amplitude bits L-6 .. L-1)
    int h_in_unit;          // complex64: amplitude bit 0 is hidden
    int nh_u, nk_low, c;    // popcount(hu_mask), popcount(kl_mask), log2 items per wavefront
    int sliced;             // 1: the hidden range of a kept group spans several wavefronts
};

// One wavefront per 2^c items; item i = (kept-upper group, hidden-upper index), hidden fastest.
// The sums of one unit's two kept amplitudes (k even): one 16-byte non-temporal store for FP64
// output, so a wavefront's store covers 1 KiB contiguous (two 8-byte stores a lane each covered
// half of every 16 bytes).
template <class O>
__device__ __forceinline__ void store_pair(O *out, u64 k, double s0, double s1) {
    if constexpr (std::is_same<O, double>::value) {
        __builtin_nontemporal_store((f64x2){s0, s1}, reinterpret_cast<f64x2 *>(out + k));
    } else {
        out[k] = (O)s0;
        out[k + 1] = (O)s1;
    }
}

// Unsliced: out[k] for each whole group. Sliced: partial[slice * n_out + k].
template <class R, class O>
__global__ __launch_bounds__(kMargThreads) void k_marginal(const Cx<R> *__restrict__ a, const MargArgs g,
                                                           O *__restrict__ out, double *__restrict__ partial) {
    constexpr int W = Unit<R>::W, WB = W == 2 ? 1 : 0;
    const u64 w = (u64)blockIdx.x * (kMargThreads / 64) + threadIdx.x / 64;
    if (w >= g.n_waves) return;
    const int lane = threadIdx.x & 63;
    const u64 hu_count = u64(1) << g.nh_u;
    const u64 per = g.sliced ? (u64(1) << g.c) : hu_count;   // items per group
    const u64 u0 = w << g.c;
    u64 grp = u0 >> g.nh_u;
    u64 ku_bits = deposit(grp, g.ku_mask);
    u64 hu_bits = deposit(u0 & (hu_count - 1), g.hu_mask);
    const u64 slice = (u0 & (hu_count - 1)) >> g.c;
    const u64 k_low = extract((u64)lane << WB, g.kl_mask);
    const bool writer = (lane & g.hl_lane) == 0;
    double acc0 = 0., acc1 = 0.;
    const u64 n_items = u64(1) << g.c;
    for (u64 t = 0; t < n_items; t += kBatch) {
        u64 unit[kBatch];
        bool end[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            unit[b] = ((ku_bits | hu_bits) >> WB) + (u64)lane;
            end[b] = ((t + b + 1) & (per - 1)) == 0;
            hu_bits = next_sub(hu_bits, g.hu_mask);
            if (end[b]) ku_bits = next_sub(ku_bits, g.ku_mask);
        }
        Amps<R> v[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) v[b] = load_unit(a, unit[b]);
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            acc0 += abs2d(v[b].v[0]);
            if (W == 2) acc1 += abs2d(v[b].v[W - 1]);
            if (end[b]) {
                double s0 = acc0, s1 = acc1;
                if (W == 2 && g.h_in_unit) s0 = s0 + s1;
#pragma unroll
                for (int bit = 0; bit < 6; ++bit) {
                    if ((g.hl_lane >> bit) & 1) {
                        s0 += __shfl_xor(s0, 1 << bit, 64);
                        if (W == 2 && !g.h_in_unit) s1 += __shfl_xor(s1, 1 << bit, 64);
                    }
                }
                if (writer) {
                    const u64 k = (grp << g.nk_low) | k_low;
                    if (g.sliced) {
                        partial[slice * g.n_out + k] = s0;
                        if (W == 2 && !g.h_in_unit) partial[slice * g.n_out + k + 1] = s1;
                    } else if (W == 2 && !g.h_in_unit) {
                        store_pair(out, k, s0, s1);
                    } else {
                        out[k] = (O)s0;
                    }
                }
                ++grp;
                acc0 = acc1 = 0.;
            }
        }
    }
}

// The unsliced walk with the kept-upper bits below the hidden-upper ones: a wavefront takes
// G consecutive kept groups at a time and, for each hidden index, reads all G (adjacent units
// of one region of the state) before moving to the next hidden index, so each hidden row is
// read G x 1 KiB at a time instead of 1 KiB (G = 16). Launched with one block of G groups per
// wavefront, so all its stores follow its last load (CDNA retires vector memory operations
// in issue order: a load issued after a store would wait for the store). Each group still
// sums its hidden indices in ascending order, as k_marginal does, so the results are the
// same bits.
template <class R, class O, int G>
__global__ __launch_bounds__(kMargThreads) void k_marginal_inner(const Cx<R> *__restrict__ a, const MargArgs g,
                                                                 O *__restrict__ out) {
    constexpr int W = Unit<R>::W, WB = W == 2 ? 1 : 0;
    const u64 w = (u64)blockIdx.x * (kMargThreads / 64) + threadIdx.x / 64;
    if (w >= g.n_waves) return;
    const int lane = threadIdx.x & 63;
    const u64 hu_count = u64(1) << g.nh_u;
    const u64 groups = (u64(1) << g.c) >> g.nh_u;   // per wavefront, a multiple of G
    u64 grp = (w << g.c) >> g.nh_u;
    u64 ku_bits = deposit(grp, g.ku_mask);
    const u64 k_low = extract((u64)lane << WB, g.kl_mask);
    const bool writer = (lane & g.hl_lane) == 0;
    for (u64 gb = 0; gb < groups; gb += G) {
        u64 kb[G];
        kb[0] = ku_bits;
#pragma unroll
        for (int i = 1; i < G; ++i) kb[i] = next_sub(kb[i - 1], g.ku_mask);
        ku_bits = next_sub(kb[G - 1], g.ku_mask);
        double s0[G], s1[G];
#pragma unroll
        for (int i = 0; i < G; ++i) s0[i] = s1[i] = 0.;
        u64 hb = 0;
#pragma unroll 2
        for (u64 h = 0; h < hu_count; ++h) {
            Amps<R> v[G];
#pragma unroll
            for (int i = 0; i < G; ++i) v[i] = load_unit(a, ((kb[i] | hb) >> WB) + (u64)lane);
#pragma unroll
            for (int i = 0; i < G; ++i) {
                s0[i] += abs2d(v[i].v[0]);
                if (W == 2) s1[i] += abs2d(v[i].v[W - 1]);
            }
            hb = next_sub(hb, g.hu_mask);
        }
#pragma unroll
        for (int i = 0; i < G; ++i) {
            double t0 = s0[i], t1 = s1[i];
            if (W == 2 && g.h_in_unit) t0 = t0 + t1;
#pragma unroll
            for (int bit = 0; bit < 6; ++bit) {
                if ((g.hl_lane >> bit) & 1) {
                    t0 += __shfl_xor(t0, 1 << bit, 64);
                    if (W == 2 && !g.h_in_unit) t1 += __shfl_xor(t1, 1 << bit, 64);
                }
            }
            if (writer) {
                const u64 k = ((grp + i) << g.nk_low) | k_low;
                if (W == 2 && !g.h_in_unit)
                    store_pair(out, k, t0, t1);
                else
                    out[k] = (O)t0;
            }
        }
        grp += G;
    }
}

// out[k] = sum over slices s of partial[s * n_out + k], s ascending (few slices: one thread
// per output), or lanes over s then a butterfly (many slices: one wavefront per output).
template <class O>
__global__ __launch_bounds__(256) void k_marginal_slices_thread(const double *__restrict__ partial, u64 n_out,
                                                                u64 n_slices, O *__restrict__ out) {
    const u64 k = (u64)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_out) return;
    double s = partial[k];
    for (u64 i = 1; i < n_slices; ++i) s += partial[i * n_out + k];
    out[k] = (O)s;
}

template <class O>
__global__ __launch_bounds__(256) void k_marginal_slices_wave(const double *__restrict__ partial, u64 n_out,
                                                              u64 n_slices, O *__restrict__ out) {
    const u64 k = (u64)blockIdx.x * 4 + threadIdx.x / 64;
    if (k >= n_out) return;
    const int lane = threadIdx.x & 63;
    double s = 0.;
    for (u64 i = lane; i < n_slices; i += 64) s += partial[i * n_out + k];
    s = wave_total(s);
    if (lane == 0) out[k] = (O)s;
}

// Tiny vectors (fewer units than one wavefront's batch): one thread per output.
template <class R, class O>
__global__ __launch_bounds__(256) void k_marginal_small(const Cx<R> *__restrict__ a, u64 k_mask, u64 h_mask,
                                                        u64 n_out, O *__restrict__ out) {
    const u64 k = (u64)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_out) return;
    const u64 base = deposit(k, k_mask);
    u64 hb = 0;
    double s = 0.;
    do {
        s += abs2d(a[base | hb]);
        hb = next_sub(hb, h_mask);
    } while (hb != 0);
    out[k] = (O)s;
}

constexpr int kMaxCombine = 64;

struct CombineArgs {
    const void *src[kMaxCombine];
    int amp[kMaxCombine];   // 1: amplitudes (|a|^2 taken here), 0: an FP64 marginal
    int n_src, n_bytes;
};

// p[x] = prod_q f_q(index_q(x)), index_q(x) = OR over bytes b of table[(q n_bytes + b) 256 + byte_b(x)].
template <class R, class O>
__global__ __launch_bounds__(256) void k_combine(O *__restrict__ out, const CombineArgs g,
                                                 const u64 *__restrict__ tables, u64 n_out, u64 first) {
    const u64 x = first + (u64)blockIdx.x * 256 + threadIdx.x;
    if (x >= n_out) return;
    double v = 1.;
    for (int q = 0; q < g.n_src; ++q) {
        u64 idx = 0;
        const u64 *t = tables + (u64)q * g.n_bytes * 256;
        for (int b = 0; b < g.n_bytes; ++b) idx |= t[b * 256 + ((x >> (8 * b)) & 255)];
        if (g.amp[q])
            v *= abs2d(static_cast<const Cx<R> *>(g.src[q])[idx]);
        else
            v *= static_cast<const double *>(g.src[q])[idx];
    }
    out[x] = (O)v;
}

inline int popcount64(u64 v) { return __builtin_popcountll(v); }

// m[k] = sum over the hidden bits h_mask of |a|^2 at deposit(k, ~h_mask) | h, for the n-lane
// vector a, k over the 2^(n - popcount(h_mask)) kept combinations in local order.
template <class R, class O>
void launch_marginal(const Cx<R> *a, int n, u64 h_mask, O *out) {
    constexpr int W = Unit<R>::W, WB = W == 2 ? 1 : 0, L = 6 + WB;
    hipStream_t s = stream();
    const u64 all = n >= 64 ? ~u64(0) : (u64(1) << n) - 1;
    const u64 k_mask = all & ~h_mask;
    const int nk = popcount64(k_mask);
    const u64 n_out = u64(1) << nk;
    if (n < L + 2) {
        hipLaunchKernelGGL((k_marginal_small<R, O>), grid_for(n_out, 256), 256, 0, s, a, k_mask, h_mask, n_out, out);
        QG_HIP(hipGetLastError());
        return;
    }
    const u64 low = (u64(1) << L) - 1;
    MargArgs g;
    g.ku_mask = k_mask & ~low;
    g.hu_mask = h_mask & ~low;
    g.kl_mask = k_mask & low;
    g.hl_lane = (int)((h_mask >> WB) & 63);
    g.h_in_unit = (W == 2 && (h_mask & 1)) ? 1 : 0;
    g.nh_u = popcount64(g.hu_mask);
    g.nk_low = popcount64(g.kl_mask);
    g.n_out = n_out;
    const int upper = n - L;   // items = 2^upper over all wavefronts
    // items per wavefront: enough wavefronts to fill the device (2^17 on large states)
    int c = std::min(10, std::max(2, upper - 17));
    if (g.nh_u > c) {
        if (g.nh_u <= 10 && upper - g.nh_u >= 12)
            c = g.nh_u;                                          // a whole hidden range each
        else
            c = std::max(c, std::min(10, upper - 12));           // long slices, >= 2^12 wavefronts
    }
    c = std::max(2, std::min(c, upper));
    g.c = c;
    g.sliced = g.nh_u > c ? 1 : 0;
    g.n_waves = u64(1) << (upper - c);
    const unsigned blocks = (unsigned)((g.n_waves + kMargThreads / 64 - 1) / (kMargThreads / 64));
    // kept-upper bits below the hidden-upper ones, with at least 2^12 blocks of 16 kept groups:
    // one block of 16 adjacent groups per wavefront, each hidden row read 16 KiB at a time
    constexpr int kInnerBits = 4;
    if (g.nh_u >= 1 && g.nh_u <= 10 && g.ku_mask != 0 &&
        (g.ku_mask & (~g.ku_mask + 1)) < (g.hu_mask & (~g.hu_mask + 1)) &&
        upper - g.nh_u - kInnerBits >= 12) {
        MargArgs gi = g;
        gi.c = g.nh_u + kInnerBits;
        gi.sliced = 0;
        gi.n_waves = u64(1) << (upper - gi.c);
        const unsigned iblocks = (unsigned)((gi.n_waves + kMargThreads / 64 - 1) / (kMargThreads / 64));
        hipLaunchKernelGGL((k_marginal_inner<R, O, 1 << kInnerBits>), iblocks, kMargThreads, 0, s, a, gi, out);
        QG_HIP(hipGetLastError());
        return;
    }
    if (!g.sliced) {
        hipLaunchKernelGGL((k_marginal<R, O>), blocks, kMargThreads, 0, s, a, g, out, nullptr);
        QG_HIP(hipGetLastError());
        return;
    }
    const u64 n_slices = u64(1) << (g.nh_u - c);
    const size_t bytes = sizeof(double) * (size_t)(n_slices * n_out);
    double *partial = static_cast<double *>(dev_alloc(bytes));
    hipLaunchKernelGGL((k_marginal<R, O>), blocks, kMargThreads, 0, s, a, g, out, partial);
    if (n_slices <= 32)
        hipLaunchKernelGGL((k_marginal_slices_thread<O>), grid_for(n_out, 256), 256, 0, s, partial, n_out, n_slices,
                           out);
    else
        hipLaunchKernelGGL((k_marginal_slices_wave<O>), grid_for(n_out, 4), 256, 0, s, partial, n_out, n_slices, out);
    hipError_t e = hipGetLastError();
    dev_free(partial, bytes);   // stream-ordered reuse
    QG_HIP(e);
}

}  // namespace

template <class R>
void launch_prob_array(void *d_out, bool out_double, const std::vector<StateView> &qs, int n_lanes,
                       int n_hidden) {
    QG_REQUIRE(!qs.empty(), "no state vector");
    QG_REQUIRE((int)qs.size() <= kMaxCombine, "too many state vectors (%zu)", qs.size());
    QG_REQUIRE(n_lanes >= 0 && n_hidden >= 0 && n_lanes + n_hidden <= 62, "lane counts out of range");
    hipStream_t s = stream();
    const u64 n_out = u64(1) << n_lanes;
    struct Src {
        u64 h_mask = 0;
        std::vector<int> kept_ext;   // external bit of each kept lane, local order
        std::vector<int> kept_pos;   // its index bit in the source array
    };
    std::vector<Src> src(qs.size());
    for (size_t q = 0; q < qs.size(); ++q) {
        const StateView &v = qs[q];
        QG_REQUIRE((int)v.local_to_ext.size() == v.n_lanes, "lane table size mismatch");
        int rank = 0;
        for (int l = 0; l < v.n_lanes; ++l) {
            const int e = v.local_to_ext[l];
            QG_REQUIRE(0 <= e && e < n_lanes + n_hidden, "external lane %d out of range", e);
            if (e >= n_lanes) {
                src[q].h_mask |= u64(1) << l;
            } else {
                src[q].kept_ext.push_back(e);
                src[q].kept_pos.push_back(rank++);
            }
        }
    }
    // one vector whose kept lanes are external bits 0, 1, ... in local order: p = its marginal
    if (qs.size() == 1) {
        bool direct = (int)src[0].kept_ext.size() == n_lanes;
        for (size_t i = 0; direct && i < src[0].kept_ext.size(); ++i) direct = src[0].kept_ext[i] == (int)i;
        if (direct) {
            const Cx<R> *a = static_cast<const Cx<R> *>(qs[0].ptr);
            if (out_double)
                launch_marginal<R, double>(a, qs[0].n_lanes, src[0].h_mask, static_cast<double *>(d_out));
            else
                launch_marginal<R, R>(a, qs[0].n_lanes, src[0].h_mask, static_cast<R *>(d_out));
            return;
        }
    }
    // marginals of the vectors with hidden lanes, then the product over vectors
    std::vector<std::pair<void *, size_t>> temps;
    CombineArgs g;
    g.n_src = (int)qs.size();
    g.n_bytes = std::max(1, (n_lanes + 7) / 8);
    std::vector<u64> tables((size_t)g.n_src * g.n_bytes * 256, 0);
    u64 *d_tables = nullptr;
    const size_t tbytes = sizeof(u64) * tables.size();
    try {
        for (int q = 0; q < g.n_src; ++q) {
            const StateView &v = qs[q];
            const Src &sq = src[q];
            const bool amp = sq.h_mask == 0;
            g.amp[q] = amp ? 1 : 0;
            if (amp) {
                g.src[q] = v.ptr;
            } else {
                const size_t bytes = sizeof(double) << sq.kept_ext.size();
                void *m = dev_alloc(bytes);
                temps.emplace_back(m, bytes);
                launch_marginal<R, double>(static_cast<const Cx<R> *>(v.ptr), v.n_lanes, sq.h_mask,
                                           static_cast<double *>(m));
                g.src[q] = m;
            }
            // index bit of each external bit of x in this source: local lane (amplitudes) or
            // rank among the kept lanes (marginal)
            std::vector<int> pos_of_ext(64, -1);
            int rank = 0;
            for (int l = 0; l < v.n_lanes; ++l) {
                const int e = v.local_to_ext[l];
                if (e < n_lanes) pos_of_ext[e] = amp ? l : rank++;
            }
            for (int b = 0; b < g.n_bytes; ++b) {
                u64 *t = tables.data() + ((size_t)q * g.n_bytes + b) * 256;
                for (int val = 0; val < 256; ++val) {
                    u64 idx = 0;
                    for (int i = 0; i < 8; ++i) {
                        const int e = 8 * b + i;
                        if (((val >> i) & 1) && e < n_lanes && pos_of_ext[e] >= 0) idx |= u64(1) << pos_of_ext[e];
                    }
                    t[val] = idx;
                }
            }
        }
        d_tables = static_cast<u64 *>(dev_alloc(tbytes));
        QG_HIP(hipMemcpyAsync(d_tables, tables.data(), tbytes, hipMemcpyHostToDevice, s));
        for_each_piece(n_out, [&](u64 first, u64 count) {
            if (out_double)
                hipLaunchKernelGGL((k_combine<R, double>), grid_for(count, 256), 256, 0, s,
                                   static_cast<double *>(d_out), g, d_tables, n_out, first);
            else
                hipLaunchKernelGGL((k_combine<R, R>), grid_for(count, 256), 256, 0, s, static_cast<R *>(d_out), g,
                                   d_tables, n_out, first);
        });
        QG_HIP(hipGetLastError());
        // the host tables must outlive the copy
        QG_HIP(hipStreamSynchronize(s));
    } catch (...) {
        hipStreamSynchronize(s);
        for (auto &t : temps) dev_free(t.first, t.second);
        if (d_tables) dev_free(d_tables, tbytes);
        throw;
    }
    for (auto &t : temps) dev_free(t.first, t.second);
    dev_free(d_tables, tbytes);
}

template void launch_prob_array<float>(void *, bool, const std::vector<StateView> &, int, int);
template void launch_prob_array<double>(void *, bool, const std::vector<StateView> &, int, int);

}  // namespace qg
