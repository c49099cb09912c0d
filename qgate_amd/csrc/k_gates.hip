// Gate kernels: 2x2 (controlled) gates and dense k-qubit unitaries on one state vector.
//
// Reference semantics: qgate/simulator/src/CPUQubitProcessor.cpp:307-362 (applyGate,
// applyControlledGate) and DeviceProcPrimitives.cu:198-249: for every pair
// (i0, i1 = i0 | 2^t) with all control bits of i0 set,
//   a[i0] <- m00 a[i0] + m01 a[i1],  a[i1] <- m10 a[i0] + m11 a[i1].
//
// Layout: one interleaved complex array of 2^n amplitudes in HBM. Each work item moves
// 16 bytes per side (2 complex64 or 1 complex128 amplitudes), so a wavefront touches 1 KiB
// contiguous per side whenever the lowest inserted bit leaves room for it. The pair index
// is scattered into the state index in closed form (insert zero bits), replacing the
// reference's uploaded permutation tables.
#include <algorithm>
#include <type_traits>

#include "kernels.hpp"
#include "qg_runtime.hpp"

namespace qg {

namespace {

constexpr int kBlock = 256;

// NI > 0: compile-time count of inserted bits; NI == 0: count taken from b.n.
template <int NI>
__device__ __forceinline__ u64 scatter(u64 k, const BitInsert &b) {
    if constexpr (NI == 0) {
        return insert_zero_bits(k, b);
    } else {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const u64 low = (u64(1) << b.pos[i]) - 1;
            k = ((k & ~low) << 1) | (k & low);
        }
        return k;
    }
}

template <class R>
struct PairArgs {
    Mat2<R> m;
    BitInsert ins;   // sorted control and target positions
    u64 ctrl;        // control mask, OR-ed into every pair base
    u64 tbit;        // target bit
    u64 n_items;     // work items (W amplitudes per side each)
};

template <class R>
__device__ __forceinline__ void mix(const Mat2<R> &m, Cx<R> &x0, Cx<R> &x1) {
    const Cx<R> y0 = cadd(cmul(m.m00, x0), cmul(m.m01, x1));
    const Cx<R> y1 = cadd(cmul(m.m10, x0), cmul(m.m11, x1));
    x0 = y0;
    x1 = y1;
}

// Pairs (i0, i0 | tbit); W consecutive amplitudes per side per item, U items per thread
// with all loads issued before the first store.
template <class R, int W, int NI, int U>
__global__ __launch_bounds__(kBlock) void k_pairs(Cx<R> *__restrict__ a, const PairArgs<R> g) {
    typedef Vec<R, W> V;
    const u64 first = (u64)blockIdx.x * (kBlock * U) + threadIdx.x;
    V x0[U], x1[U];
    u64 i0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 w = first + (u64)u * kBlock;
        if (w < g.n_items) {
            i0[u] = scatter<NI>(w * W, g.ins) | g.ctrl;
            x0[u] = *reinterpret_cast<const V *>(a + i0[u]);
            x1[u] = *reinterpret_cast<const V *>(a + (i0[u] | g.tbit));
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 w = first + (u64)u * kBlock;
        if (w < g.n_items) {
#pragma unroll
            for (int j = 0; j < W; ++j) mix(g.m, x0[u].v[j], x1[u].v[j]);
            *reinterpret_cast<V *>(a + i0[u]) = x0[u];
            *reinterpret_cast<V *>(a + (i0[u] | g.tbit)) = x1[u];
        }
    }
}

// Target bit 0 of a complex64 vector: the pair (i0, i0 + 1) is one 16-byte access.
template <int NI, int U>
__global__ __launch_bounds__(kBlock) void k_pairs_bit0(Cx<float> *__restrict__ a,
                                                       const PairArgs<float> g) {
    typedef Vec<float, 2> V;
    const u64 first = (u64)blockIdx.x * (kBlock * U) + threadIdx.x;
    V x[U];
    u64 i0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 w = first + (u64)u * kBlock;
        if (w < g.n_items) {
            i0[u] = scatter<NI>(w, g.ins) | g.ctrl;
            x[u] = *reinterpret_cast<const V *>(a + i0[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 w = first + (u64)u * kBlock;
        if (w < g.n_items) {
            mix(g.m, x[u].v[0], x[u].v[1]);
            *reinterpret_cast<V *>(a + i0[u]) = x[u];
        }
    }
}

template <class R>
struct PhaseArgs {
    Cx<R> d;         // phase applied where every bit of `set` is 1
    BitInsert ins;
    u64 set;
    u64 n_items;
};

// Diagonal gate with m00 == 1 (U1, S, T, Z, CZ, cu1, ...): only amplitudes whose target and
// control bits are all 1 change, so half (or less) of the vector is read and written.
template <class R, int W, int NI, int U>
__global__ __launch_bounds__(kBlock) void k_phase(Cx<R> *__restrict__ a, const PhaseArgs<R> g) {
    typedef Vec<R, W> V;
    const u64 first = (u64)blockIdx.x * (kBlock * U) + threadIdx.x;
    V x[U];
    u64 idx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 w = first + (u64)u * kBlock;
        if (w < g.n_items) {
            idx[u] = scatter<NI>(w * W, g.ins) | g.set;
            x[u] = *reinterpret_cast<const V *>(a + idx[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 w = first + (u64)u * kBlock;
        if (w < g.n_items) {
#pragma unroll
            for (int j = 0; j < W; ++j) x[u].v[j] = cmul(g.d, x[u].v[j]);
            *reinterpret_cast<V *>(a + idx[u]) = x[u];
        }
    }
}

template <int U>
inline unsigned grid_for(u64 n_items) {
    const u64 per_block = (u64)kBlock * U;
    return (unsigned)((n_items + per_block - 1) / per_block);
}

// Calls f(std::integral_constant<int, NI>) with NI = n for n <= 4, else 0 (dynamic).
template <class F>
void with_ni(int n, F &&f) {
    switch (n) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    default: f(std::integral_constant<int, 0>()); break;
    }
}

}  // namespace

template <class R>
void launch_gate(Cx<R> *a, int n_lanes, const GateMatrix &gm, const std::vector<int> &ctrl,
                 int target) {
    constexpr int WV = Unit<R>::W;  // amplitudes per 16-byte access
    constexpr int U = 4;
    QG_REQUIRE(0 <= target && target < n_lanes, "target lane %d out of range", target);
    std::vector<int> pos(ctrl);
    pos.push_back(target);
    std::sort(pos.begin(), pos.end());
    for (size_t i = 0; i < pos.size(); ++i) {
        QG_REQUIRE(0 <= pos[i] && pos[i] < n_lanes, "lane %d out of range", pos[i]);
        QG_REQUIRE(i == 0 || pos[i] != pos[i - 1], "control and target lanes overlap");
    }
    BitInsert ins;
    ins.n = (int)pos.size();
    for (int i = 0; i < ins.n; ++i) ins.pos[i] = pos[i];
    u64 ctrl_mask = 0;
    for (int c : ctrl) ctrl_mask |= u64(1) << c;
    const u64 tbit = u64(1) << target;
    const int n_free = n_lanes - ins.n;
    hipStream_t s = stream();

    const auto &m = gm.m;
    const bool offdiag_zero = m[0][1] == 0. && m[1][0] == 0.;
    if (offdiag_zero && m[0][0] == 1.) {
        if (m[1][1] == 1.) return;  // identity on the touched subspace
        PhaseArgs<R> g;
        g.d = cx<R>((R)m[1][1].real(), (R)m[1][1].imag());
        g.ins = ins;
        g.set = ctrl_mask | tbit;
        const bool vec = WV > 1 && pos[0] >= 1;
        g.n_items = (u64(1) << n_free) / (vec ? WV : 1);
        with_ni(ins.n, [&](auto ni) {
            constexpr int NI = decltype(ni)::value;
            if (vec)
                hipLaunchKernelGGL((k_phase<R, WV, NI, U>), grid_for<U>(g.n_items), kBlock, 0, s, a, g);
            else
                hipLaunchKernelGGL((k_phase<R, 1, NI, U>), grid_for<U>(g.n_items), kBlock, 0, s, a, g);
        });
        QG_HIP(hipGetLastError());
        return;
    }

    PairArgs<R> g;
    g.m = round_matrix<R>(gm);
    g.ins = ins;
    g.ctrl = ctrl_mask;
    g.tbit = tbit;
    const u64 n_pairs = u64(1) << n_free;
    with_ni(ins.n, [&](auto ni) {
        constexpr int NI = decltype(ni)::value;
        if constexpr (std::is_same<R, float>::value) {
            if (target == 0) {
                g.n_items = n_pairs;
                hipLaunchKernelGGL((k_pairs_bit0<NI, U>), grid_for<U>(g.n_items), kBlock, 0, s, a, g);
                return;
            }
        }
        if (WV > 1 && pos[0] >= 1) {
            g.n_items = n_pairs / WV;
            hipLaunchKernelGGL((k_pairs<R, WV, NI, U>), grid_for<U>(g.n_items), kBlock, 0, s, a, g);
        } else {
            g.n_items = n_pairs;
            hipLaunchKernelGGL((k_pairs<R, 1, NI, U>), grid_for<U>(g.n_items), kBlock, 0, s, a, g);
        }
    });
    QG_HIP(hipGetLastError());
}

template void launch_gate<float>(Cx<float> *, int, const GateMatrix &, const std::vector<int> &, int);
template void launch_gate<double>(Cx<double> *, int, const GateMatrix &, const std::vector<int> &, int);

// ---- dense k-qubit unitary -----------------------------------------------------------------

namespace {

constexpr int kMaxFuse = 6;

template <class R>
struct UnitaryArgs {
    BitInsert ins;          // sorted fused lanes
    u64 member[1 << kMaxFuse];  // state-index offset of matrix index c
    int k;
    u64 n_groups;
};

// One thread per group of 2^K amplitudes: y = U x with U staged in LDS.
template <class R, int K>
__global__ __launch_bounds__(kBlock) void k_unitary(Cx<R> *__restrict__ a,
                                                    const Cx<R> *__restrict__ umat,
                                                    const UnitaryArgs<R> g) {
    constexpr int D = 1 << K;
    __shared__ Cx<R> su[D * D];
    for (int i = threadIdx.x; i < D * D; i += kBlock) su[i] = umat[i];
    __syncthreads();
    const u64 grp = (u64)blockIdx.x * kBlock + threadIdx.x;
    if (grp >= g.n_groups) return;
    const u64 base = insert_zero_bits(grp, g.ins);
    Cx<R> x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = a[base | g.member[c]];
#pragma unroll 1
    for (int r = 0; r < D; ++r) {
        Cx<R> acc = cx<R>(0, 0);
#pragma unroll
        for (int c = 0; c < D; ++c) acc = cadd(acc, cmul(su[r * D + c], x[c]));
        a[base | g.member[r]] = acc;
    }
}

}  // namespace

template <class R>
void launch_unitary(Cx<R> *a, int n_lanes, const std::vector<int> &lanes,
                    const std::vector<std::complex<double>> &mat) {
    const int k = (int)lanes.size();
    QG_REQUIRE(1 <= k && k <= kMaxFuse, "unitary width %d out of range", k);
    QG_REQUIRE(k <= n_lanes, "unitary wider than the state");
    const int D = 1 << k;
    QG_REQUIRE((int)mat.size() == D * D, "unitary matrix has wrong size");
    UnitaryArgs<R> g;
    std::vector<int> sorted(lanes);
    std::sort(sorted.begin(), sorted.end());
    g.ins.n = k;
    for (int i = 0; i < k; ++i) {
        QG_REQUIRE(0 <= sorted[i] && sorted[i] < n_lanes, "lane %d out of range", sorted[i]);
        QG_REQUIRE(i == 0 || sorted[i] != sorted[i - 1], "duplicate lane in unitary");
        g.ins.pos[i] = sorted[i];
    }
    for (int c = 0; c < D; ++c) {
        u64 off = 0;
        for (int j = 0; j < k; ++j)
            if (c & (1 << j)) off |= u64(1) << lanes[j];
        g.member[c] = off;
    }
    g.k = k;
    g.n_groups = u64(1) << (n_lanes - k);
    std::vector<Cx<R>> hm(D * D);
    for (int i = 0; i < D * D; ++i) hm[i] = cx<R>((R)mat[i].real(), (R)mat[i].imag());
    Cx<R> *dm = static_cast<Cx<R> *>(dev_alloc(sizeof(Cx<R>) * D * D));
    hipStream_t s = stream();
    QG_HIP(hipMemcpyAsync(dm, hm.data(), sizeof(Cx<R>) * D * D, hipMemcpyHostToDevice, s));
    const unsigned grid = (unsigned)((g.n_groups + kBlock - 1) / kBlock);
    switch (k) {
    case 1: hipLaunchKernelGGL((k_unitary<R, 1>), grid, kBlock, 0, s, a, dm, g); break;
    case 2: hipLaunchKernelGGL((k_unitary<R, 2>), grid, kBlock, 0, s, a, dm, g); break;
    case 3: hipLaunchKernelGGL((k_unitary<R, 3>), grid, kBlock, 0, s, a, dm, g); break;
    case 4: hipLaunchKernelGGL((k_unitary<R, 4>), grid, kBlock, 0, s, a, dm, g); break;
    case 5: hipLaunchKernelGGL((k_unitary<R, 5>), grid, kBlock, 0, s, a, dm, g); break;
    case 6: hipLaunchKernelGGL((k_unitary<R, 6>), grid, kBlock, 0, s, a, dm, g); break;
    }
    QG_HIP(hipGetLastError());
    QG_HIP(hipStreamSynchronize(s));   // hm is a host temporary
    dev_free(dm, sizeof(Cx<R>) * D * D);
}

template void launch_unitary<float>(Cx<float> *, int, const std::vector<int> &,
                                    const std::vector<std::complex<double>> &);
template void launch_unitary<double>(Cx<double> *, int, const std::vector<int> &,
                                     const std::vector<std::complex<double>> &);

}  // namespace qg
