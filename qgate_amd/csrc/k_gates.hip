// Gate kernels: 2x2 (controlled) gates and dense k-qubit unitaries on one state vector.
//
// Reference semantics: qgate/simulator/src/CPUQubitProcessor.cpp:307-362 (applyGate,
// applyControlledGate) and DeviceProcPrimitives.cu:198-249: for every pair
// (i0, i1 = i0 | 2^t) with all control bits of i0 set,
//   a[i0] <- m00 a[i0] + m01 a[i1],  a[i1] <- m10 a[i0] + m11 a[i1].
//
// Layout: one interleaved complex array of 2^n amplitudes in HBM, moved in 16-byte units
// (W = 2 complex64 or 1 complex128 amplitudes) with non-temporal loads and stores. By where
// the partner of a pair lies, a gate takes one of three streaming kernels:
//   * inside the unit (complex64, target bit 0)            k_unit_pair: one load per pair;
//   * inside the wavefront (partner 1..32 units away)      k_wave_pair: every lane loads its
//     own unit (a fully coalesced 1 KiB per wavefront) and fetches the partner with a
//     cross-lane exchange;
//   * further away                                         k_strided_pair: one unit from each
//     half, a wavefront covering 1 KiB contiguous on each side.
// Diagonal gates with m00 == 1 (U1, S, T, Z, CZ, cu1) take k_phase and touch only the
// amplitudes whose target and control bits are all 1.
// Control bits above the unit are folded into the work-item index (zero bits inserted in
// closed form, replacing the reference's uploaded permutation tables); control bits inside
// a wavefront chunk or a unit are checked per lane or per amplitude.
//   * partner 64 or 128 units away                         k_block_pair: a workgroup takes
//     a contiguous chunk of twice the distance and exchanges partners through LDS.
// Launch shapes (one wavefront per k_strided_pair block, two for partners 64-128 units apart;
// the spread block order for partners 256-1024 units and 16 MiB or more apart) come from
// per-target-bit bandwidth sweeps at 30 qubits.
#include <algorithm>
#include <complex>
#include <cstring>
#include <type_traits>
#include <vector>

#include "kernels.hpp"
#include "tile_common.hpp"
#include "qg_runtime.hpp"

namespace qg {

namespace {

// NI >= 0: compile-time count of inserted bits; NI == -1: count taken from b.n.
template <int NI>
__device__ __forceinline__ u64 scatter(u64 k, const BitInsert &b) {
    if constexpr (NI < 0) {
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
struct GateArgs {
    Mat2<R> m;
    BitInsert ins;      // zero bits inserted into the work-item index
    u64 item_set;       // bits OR-ed into the scattered index (controls; target for k_phase)
    u64 n_items;
    u64 dist;           // partner distance in units (strided) or lanes (wave)
    u64 lane_ctrl;      // control bits of the unit index below the wavefront chunk (k_wave_pair)
    u64 unit_chk;       // control bits of the unit index tested per work item (whole lines, below)
    unsigned amp_ctrl;  // control bits inside a unit (complex64 bit 0)
};

// Controls (and k_phase targets) on the lowest unit-index bits are not folded into the
// work-item index: folded, they would leave touched runs shorter than a 128-byte line
// (8 units), and the half-used lines, read and partly written, measured about twice the time
// of a full pass at 30 qubits (CX with its control on amplitude bit 1 or 2: 5.2 ms against
// 2.65 ms). Below kWholeLineBits such bits are tested per work item instead, over every unit,
// and units that fail the test are written back unchanged, so every line is read and
// written whole. k_wave_pair reads whole lines anyway; it writes them whole (lanes that
// fail still store) for bits below kWaveWholeLineBits.
constexpr int kWholeLineBits = 3;
constexpr int kWaveWholeLineBits = 2;

template <class R>
__device__ __forceinline__ void mix(const Mat2<R> &m, Cx<R> &x0, Cx<R> &x1) {
    const Cx<R> y0 = cadd(cmul(m.m00, x0), cmul(m.m01, x1));
    const Cx<R> y1 = cadd(cmul(m.m10, x0), cmul(m.m11, x1));
    x0 = y0;
    x1 = y1;
}

template <class R>
__device__ __forceinline__ bool amp_on(int j, unsigned amp_ctrl) {
    return ((unsigned)j & amp_ctrl) == amp_ctrl;
}

// Partner in the other half, `dist` units away.
// SPREAD: even blocks take the first half of the launch's blocks and odd blocks the second
// (a bijection for any grid size), so the blocks in flight read four runs of the state
// instead of two. With partners 16 MiB or more apart the two runs of a plain launch cover
// fewer HBM channels: at 30 qubits (complex64) targets 21..23 ran at 5.8-5.9 TB/s against
// 6.5 TB/s for targets up to 20; spread, targets 21..28 gain 2-4% (targets up to 20 lose).
template <class R, int NI, int B, bool SPREAD = false>
__global__ __launch_bounds__(B) void k_strided_pair(Cx<R> *__restrict__ a, const GateArgs<R> g, u64 first) {
    u64 blk = blockIdx.x;
    if constexpr (SPREAD) {
        const u64 half = ((u64)gridDim.x + 1) / 2;
        blk = (blk & 1) ? half + blk / 2 : blk / 2;
    }
    const u64 w = first + blk * B + threadIdx.x;
    if (w >= g.n_items) return;
    const u64 u0 = scatter<NI>(w, g.ins) | g.item_set, u1 = u0 | g.dist;
    Amps<R> x0 = load_unit(a, u0);
    Amps<R> x1 = load_unit(a, u1);
    if ((u0 & g.unit_chk) == g.unit_chk) {
#pragma unroll
        for (int j = 0; j < Unit<R>::W; ++j)
            if (amp_on<R>(j, g.amp_ctrl)) mix(g.m, x0.v[j], x1.v[j]);
    }
    store_unit(a, u0, x0);
    store_unit(a, u1, x1);
}

template <class R>
__device__ __forceinline__ Cx<R> shfl_xor_cx(const Cx<R> &v, int d) {
    return cx<R>(__shfl_xor(v.re, d, 64), __shfl_xor(v.im, d, 64));
}

// Partner in the same wavefront, `dist` (1..32) lanes away. One 64-unit chunk per wavefront.
template <class R, int NI>
__global__ __launch_bounds__(64) void k_wave_pair(Cx<R> *__restrict__ a, const GateArgs<R> g, u64 first) {
    const u64 w = first + blockIdx.x;   // one block per 64-unit chunk
    const int lane = threadIdx.x;
    const u64 unit = ((scatter<NI>(w, g.ins) | g.item_set) << 6) | (u64)lane;
    // a lane whose control bits in the chunk are not all set neither loads nor stores: its
    // partner (dist away on a bit that is no control) fails the same test
    const bool act = ((u64)lane & g.lane_ctrl) == g.lane_ctrl;
    Amps<R> x = {};
    if (act) x = load_unit(a, unit);
    Amps<R> p;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) p.v[j] = shfl_xor_cx(x.v[j], (int)g.dist);
    if (!act) return;
    const bool on = ((u64)lane & g.unit_chk) == g.unit_chk;   // off: written back unchanged
    const bool hi = ((u64)lane & g.dist) != 0;
    Amps<R> y = x;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) {
        if (!on || !amp_on<R>(j, g.amp_ctrl)) continue;
        Cx<R> lo_v = hi ? p.v[j] : x.v[j];
        Cx<R> hi_v = hi ? x.v[j] : p.v[j];
        mix(g.m, lo_v, hi_v);
        y.v[j] = hi ? hi_v : lo_v;
    }
    store_unit(a, unit, y);
}

// Partner 64 or 128 units away (unit bit tu = CB - 1): a workgroup of 2^CB threads takes a
// contiguous chunk of 2^CB units, one per thread, and exchanges partners through LDS, so the
// chunk is read and written as one run (as k_wave_pair's 1 KiB runs) instead of two runs of
// 1 KiB per wavefront. 30-qubit complex64 targets 7 and 8: 6.80 TB/s against 6.15-6.19 for
// k_strided_pair with two-wavefront blocks; for partners 256 units away or more, chunks of
// 2 x dist units measured no faster than k_strided_pair (1024-unit chunks 6.37 TB/s against
// 6.60 for the spread order) and chunks of several units per thread were slower.
// Control bits of the unit index are split as in k_wave_pair (above the chunk: folded into
// the chunk index; lane_ctrl: the lane sits out; unit_chk: tested, written back unchanged).
template <class R, int NI, int CB>
__global__ __launch_bounds__(1 << CB) void k_block_pair(Cx<R> *__restrict__ a, const GateArgs<R> g, u64 first) {
    typedef typename UnitVec<R>::T V;
    __shared__ V sx[1 << CB];
    const u64 w = first + blockIdx.x;   // one block per chunk
    const int lane = threadIdx.x;
    const u64 unit = ((scatter<NI>(w, g.ins) | g.item_set) << CB) | (u64)lane;
    const bool act = ((u64)lane & g.lane_ctrl) == g.lane_ctrl;   // partner fails alike
    Amps<R> x = {};
    if (act) x = load_unit(a, unit);
    V v;
    __builtin_memcpy(&v, &x, sizeof(V));
    sx[lane] = v;
    __syncthreads();
    if (!act) return;
    Amps<R> p;
    v = sx[lane ^ (int)g.dist];
    __builtin_memcpy(&p, &v, sizeof(V));
    const bool on = ((u64)lane & g.unit_chk) == g.unit_chk;   // off: written back unchanged
    const bool hi = ((u64)lane & g.dist) != 0;
    Amps<R> y = x;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) {
        if (!on || !amp_on<R>(j, g.amp_ctrl)) continue;
        Cx<R> lo_v = hi ? p.v[j] : x.v[j];
        Cx<R> hi_v = hi ? x.v[j] : p.v[j];
        mix(g.m, lo_v, hi_v);
        y.v[j] = hi ? hi_v : lo_v;
    }
    store_unit(a, unit, y);
}

// Target bit 0 of a complex64 vector: the pair is one unit.
template <int NI, int B>
__global__ __launch_bounds__(B) void k_unit_pair(Cx<float> *__restrict__ a, const GateArgs<float> g, u64 first) {
    const u64 w = first + (u64)blockIdx.x * B + threadIdx.x;
    if (w >= g.n_items) return;
    const u64 u = scatter<NI>(w, g.ins) | g.item_set;
    Amps<float> x = load_unit(a, u);
    if ((u & g.unit_chk) == g.unit_chk) mix(g.m, x.v[0], x.v[1]);
    store_unit(a, u, x);
}

// Diagonal with m00 == 1: multiply the amplitudes whose target and control bits are 1.
template <class R, int NI, int B>
__global__ __launch_bounds__(B) void k_phase(Cx<R> *__restrict__ a, const GateArgs<R> g, u64 first) {
    const u64 w = first + (u64)blockIdx.x * B + threadIdx.x;
    if (w >= g.n_items) return;
    const u64 u = scatter<NI>(w, g.ins) | g.item_set;
    Amps<R> x = load_unit(a, u);
    if ((u & g.unit_chk) == g.unit_chk) {
#pragma unroll
        for (int j = 0; j < Unit<R>::W; ++j)
            if (amp_on<R>(j, g.amp_ctrl)) x.v[j] = cmul(g.m.m11, x.v[j]);
    }
    store_unit(a, u, x);
}

// Calls f(std::integral_constant<int, NI>) with NI = n for n <= 4, else -1 (dynamic).
template <class F>
void with_ni(int n, F &&f) {
    switch (n) {
    case 0: f(std::integral_constant<int, 0>()); break;
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    default: f(std::integral_constant<int, -1>()); break;
    }
}

int log2_of(int w) { return w == 2 ? 1 : 0; }

BitInsert make_insert(std::vector<int> pos) {
    std::sort(pos.begin(), pos.end());
    BitInsert b;
    b.n = (int)pos.size();
    for (int i = 0; i < b.n; ++i) b.pos[i] = pos[i];
    return b;
}

}  // namespace

template <class R>
void launch_gate(Cx<R> *a, int n_lanes, const GateMatrix &gm, const std::vector<int> &ctrl,
                 int target) {
    constexpr int W = Unit<R>::W;
    const int lw = log2_of(W);                 // amplitude bits inside a unit
    QG_REQUIRE(0 <= target && target < n_lanes, "target lane %d out of range", target);
    std::vector<int> all(ctrl);
    all.push_back(target);
    std::sort(all.begin(), all.end());
    for (size_t i = 0; i < all.size(); ++i) {
        QG_REQUIRE(0 <= all[i] && all[i] < n_lanes, "lane %d out of range", all[i]);
        QG_REQUIRE(i == 0 || all[i] != all[i - 1], "control and target lanes overlap");
    }
    const int unit_bits = n_lanes - lw;        // bits of the unit index
    hipStream_t s = stream();

    GateArgs<R> g;
    g.m = round_matrix<R>(gm);
    g.amp_ctrl = 0;
    g.lane_ctrl = 0;
    g.unit_chk = 0;
    g.dist = 0;
    g.item_set = 0;
    // positions of the unit index below kWholeLineBits: tested per work item, not folded
    auto split_low = [&](std::vector<int> &pos) {
        std::vector<int> folded;
        for (int p : pos) {
            if (p < kWholeLineBits)
                g.unit_chk |= u64(1) << p;
            else
                folded.push_back(p);
        }
        pos.swap(folded);
    };
    // controls inside a unit (complex64 bit 0)
    std::vector<int> unit_ctrl;                // control positions in the unit index
    for (int c : ctrl) {
        if (c < lw)
            g.amp_ctrl |= 1u << c;
        else
            unit_ctrl.push_back(c - lw);
    }

    const auto &m = gm.m;
    const bool phase = m[0][1] == 0. && m[1][0] == 0. && m[0][0] == 1.;
    if (phase) {
        if (m[1][1] == 1.) return;                             // identity on the touched subspace
        std::vector<int> pos(unit_ctrl);
        if (target < lw)
            g.amp_ctrl |= 1u << target;
        else
            pos.push_back(target - lw);
        split_low(pos);
        g.ins = make_insert(pos);
        for (int p : pos) g.item_set |= u64(1) << p;
        g.n_items = u64(1) << (unit_bits - (int)pos.size());
        with_ni((int)pos.size(), [&](auto ni) {
            constexpr int NI = decltype(ni)::value;
            for_each_piece(g.n_items, [&](u64 first, u64 count) {
                hipLaunchKernelGGL((k_phase<R, NI, 64>), grid_for(count, 64), 64, 0, s, a, g, first);
            });
        });
        QG_HIP(hipGetLastError());
        return;
    }

    if constexpr (std::is_same<R, float>::value) {
        if (target == 0) {                                     // pair inside the unit
            split_low(unit_ctrl);
            g.ins = make_insert(unit_ctrl);
            for (int p : unit_ctrl) g.item_set |= u64(1) << p;
            g.n_items = u64(1) << (unit_bits - (int)unit_ctrl.size());
            with_ni((int)unit_ctrl.size(), [&](auto ni) {
                constexpr int NI = decltype(ni)::value;
                for_each_piece(g.n_items, [&](u64 first, u64 count) {
                    hipLaunchKernelGGL((k_unit_pair<NI, 64>), grid_for(count, 64), 64, 0, s, a, g, first);
                });
            });
            QG_HIP(hipGetLastError());
            return;
        }
    }

    const int tu = target - lw;                                // target bit of the unit index
    if (tu < 6 && unit_bits >= 6) {                            // partner in the wavefront
        std::vector<int> chunk_ctrl;
        for (int p : unit_ctrl) {
            if (p >= 6)
                chunk_ctrl.push_back(p - 6);
            else if (p < kWaveWholeLineBits)
                g.unit_chk |= u64(1) << p;
            else
                g.lane_ctrl |= u64(1) << p;
        }
        g.ins = make_insert(chunk_ctrl);
        for (int p : chunk_ctrl) g.item_set |= u64(1) << p;
        g.dist = u64(1) << tu;
        g.n_items = u64(1) << (unit_bits - 6 - (int)chunk_ctrl.size());
        with_ni((int)chunk_ctrl.size(), [&](auto ni) {
            constexpr int NI = decltype(ni)::value;
            // pieces in chunks: 64 work items each
            for_each_piece(g.n_items << 6, [&](u64 first, u64 count) {
                hipLaunchKernelGGL((k_wave_pair<R, NI>), (unsigned)(count >> 6), 64, 0, s, a, g, first >> 6);
            });
        });
        QG_HIP(hipGetLastError());
        return;
    }

    // partner in a chunk of 2 dist units, unless a control inside the chunk lies on a whole
    // line (unit bit 3 or more): folded into k_strided_pair's work-item index, such a control
    // halves the lines moved, where the chunk would leave half its lanes idle (30-qubit
    // complex64 CX, target 7, control 4..6: 5.9-6.0 TB/s strided against 4.9 in chunks)
    bool chunk = (tu == 6 || tu == 7) && unit_bits > tu;
    for (int p : unit_ctrl)
        if (p >= kWholeLineBits && p <= tu) chunk = false;
    if (chunk) {
        const int cb = tu + 1;
        std::vector<int> chunk_ctrl;
        for (int p : unit_ctrl) {
            if (p >= cb)
                chunk_ctrl.push_back(p - cb);
            else if (p < kWaveWholeLineBits)
                g.unit_chk |= u64(1) << p;
            else
                g.lane_ctrl |= u64(1) << p;
        }
        g.ins = make_insert(chunk_ctrl);
        for (int p : chunk_ctrl) g.item_set |= u64(1) << p;
        g.dist = u64(1) << tu;
        g.n_items = u64(1) << (unit_bits - cb - (int)chunk_ctrl.size());
        with_ni((int)chunk_ctrl.size(), [&](auto ni) {
            constexpr int NI = decltype(ni)::value;
            // pieces in chunks of 2^cb work items (pieces are powers of two >= 2^12)
            for_each_piece(g.n_items << cb, [&](u64 first, u64 count) {
                if (cb == 7)
                    hipLaunchKernelGGL((k_block_pair<R, NI, 7>), (unsigned)(count >> 7), 128, 0, s, a, g, first >> 7);
                else
                    hipLaunchKernelGGL((k_block_pair<R, NI, 8>), (unsigned)(count >> 8), 256, 0, s, a, g, first >> 8);
            });
        });
        QG_HIP(hipGetLastError());
        return;
    }

    split_low(unit_ctrl);
    std::vector<int> pos(unit_ctrl);
    pos.push_back(tu);
    g.ins = make_insert(pos);
    for (int p : unit_ctrl) g.item_set |= u64(1) << p;
    g.dist = u64(1) << tu;
    g.n_items = u64(1) << (unit_bits - (int)pos.size());
    // two wavefronts a block for partners 64-128 units apart (controlled gates that do not take
    // k_block_pair); blocks in the spread order (the two halves of the grid interleaved) for
    // partners 256 to 1024 units apart (target 9: 6.60 TB/s against 6.27) and 16 MiB or more
    const bool two_waves = tu == 6 || tu == 7;
    const bool spread = (tu >= 8 && tu <= 10) || (g.dist * sizeof(Cx<R>) * Unit<R>::W) >= (u64(16) << 20);
    with_ni((int)pos.size(), [&](auto ni) {
        constexpr int NI = decltype(ni)::value;
        for_each_piece(g.n_items, [&](u64 first, u64 count) {
            if (two_waves)
                hipLaunchKernelGGL((k_strided_pair<R, NI, 128>), grid_for(count, 128), 128, 0, s, a, g, first);
            else if (spread)
                hipLaunchKernelGGL((k_strided_pair<R, NI, 64, true>), grid_for(count, 64), 64, 0, s, a, g, first);
            else
                hipLaunchKernelGGL((k_strided_pair<R, NI, 64>), grid_for(count, 64), 64, 0, s, a, g, first);
        });
    });
    QG_HIP(hipGetLastError());
}

template void launch_gate<float>(Cx<float> *, int, const GateMatrix &, const std::vector<int> &, int);
template void launch_gate<double>(Cx<double> *, int, const GateMatrix &, const std::vector<int> &, int);

// ---- dense k-qubit unitary -----------------------------------------------------------------

namespace {

constexpr int kMaxFuse = 6;
constexpr int kBlock = 256;

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
                                                    const UnitaryArgs<R> g, u64 first) {
    constexpr int D = 1 << K;
    __shared__ Cx<R> su[D * D];
    for (int i = threadIdx.x; i < D * D; i += kBlock) su[i] = umat[i];
    __syncthreads();
    const u64 grp = first + (u64)blockIdx.x * kBlock + threadIdx.x;
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

// Dense block staged through LDS: one workgroup per tile of 2^M amplitudes (32 KiB) made of
// the block's lanes and the lowest other lanes, so the tile is read and written as rows of
// 2^low contiguous amplitudes in 16-byte units. Every group of 2^K amplitudes of the tile is
// multiplied by U in LDS (row sums from zero in column order, as k_unitary). With fewer
// groups than threads, T / G threads share a group and take D / (T / G) rows each, after a
// barrier that lets every thread read its inputs first.
template <class R>
struct DenseTileCfg;
template <>
struct DenseTileCfg<float> {
    static constexpr int M = 12, MaxK = 4;   // complex64 blocks of 5-6 lanes take the matrix cores
};
template <>
struct DenseTileCfg<double> {
    static constexpr int M = 11, MaxK = 4;   // 5 and 6 lanes take the FP64 matrix cores
};

struct DenseTileArgs {
    alignas(4) signed char pos[kTilePosMax];   // state bit of each local tile bit, ascending
    int low;                                   // local bits 0 .. low-1 are state bits 0 .. low-1
    unsigned rest;                             // local bits outside the block
    unsigned fmem[1 << kMaxFuse];              // local offset of matrix index c
};

template <class R, int K>
__global__ __launch_bounds__(kBlock) void k_unitary_tile(Cx<R> *__restrict__ a, const Cx<R> *__restrict__ umat,
                                                         const DenseTileArgs g, u64 first) {
    constexpr int M = DenseTileCfg<R>::M, D = 1 << K, G = 1 << (M - K), T = kBlock;
    constexpr int lw = Unit<R>::W == 2 ? 1 : 0;
    constexpr int S = G >= T ? 1 : T / G;   // threads per group
    constexpr int RPT = D / S;               // rows per thread
    __shared__ Cx<R> t[1 << M];
    __shared__ Cx<R> su[D * D];
    for (int i = threadIdx.x; i < D * D; i += T) su[i] = umat[i];
    const u64 base = tile_base(first + blockIdx.x, g.pos, M);
    tile_to_lds<R, T, (1 << (M - lw)) / T>(t, a, base, g.pos, M, g.low);
    __syncthreads();
    for (int q0 = 0; q0 < G; q0 += T / S) {
        const int q = q0 + (int)threadIdx.x / S, part = (int)threadIdx.x % S;
        // local index of the group: q deposited at the local bits outside the block
        unsigned lb = 0, qq = (unsigned)q;
        for (unsigned m = g.rest; m != 0; m &= m - 1) {
            if (qq & 1u) lb |= m & (~m + 1);
            qq >>= 1;
        }
        Cx<R> x[D];
#pragma unroll
        for (int c = 0; c < D; ++c) x[c] = t[lb | g.fmem[c]];
        if constexpr (S > 1) __syncthreads();
#pragma unroll 1
        for (int rr = 0; rr < RPT; ++rr) {
            const int r = part * RPT + rr;
            Cx<R> acc = cx<R>(0, 0);
#pragma unroll
            for (int c = 0; c < D; ++c) acc = cadd(acc, cmul(su[r * D + c], x[c]));
            t[lb | g.fmem[r]] = acc;
        }
    }
    __syncthreads();
    lds_to_tile<R, T, (1 << (M - lw)) / T>(a, t, base, g.pos, M, g.low);
}

template <class R>
bool use_dense_tile(int k, int n_lanes) {
    return k >= 2 && k <= DenseTileCfg<R>::MaxK && n_lanes >= DenseTileCfg<R>::M;
}

// Tile layout of a block on `lanes` (matrix index bit j <-> lanes[j]): the block's lanes and
// the lowest other lanes up to M tile bits.
template <class R>
DenseTileArgs dense_tile_args(int n_lanes, const std::vector<int> &lanes, int M = DenseTileCfg<R>::M) {
    std::vector<int> bits(lanes);
    for (int b = 0; (int)bits.size() < M && b < n_lanes; ++b)
        if (std::find(bits.begin(), bits.end(), b) == bits.end()) bits.push_back(b);
    std::sort(bits.begin(), bits.end());
    QG_REQUIRE((int)bits.size() == M, "dense tile needs %d lanes", M);
    DenseTileArgs g;
    std::memset(&g, 0, sizeof(g));
    for (int j = 0; j < M; ++j) g.pos[j] = (signed char)bits[j];
    g.low = 0;
    while (g.low < M && bits[g.low] == g.low) ++g.low;
    auto local = [&](int lane) { return (int)(std::find(bits.begin(), bits.end(), lane) - bits.begin()); };
    unsigned fmask = 0;
    for (int l : lanes) fmask |= 1u << local(l);
    g.rest = ((1u << M) - 1) & ~fmask;
    const int k = (int)lanes.size();
    for (int c = 0; c < (1 << k); ++c) {
        unsigned off = 0;
        for (int j = 0; j < k; ++j)
            if (c & (1 << j)) off |= 1u << local(lanes[j]);
        g.fmem[c] = off;
    }
    return g;
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// 16 x 16 x 4 matrix-core step in the precision of R: acc += A (16 x 4) B (4 x 16).
template <class R>
struct Mfma;
// Both take A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k. Result element v of
// lane l is D[row][l % 16] with row = 4 (l / 16) + v for FP32 and row = l / 16 + 4 v for FP64
// (measured on gfx950: one step on known A and B, every lane's four results identified).
template <>
struct Mfma<float> {
    typedef v4f Acc;
    static __device__ __forceinline__ Acc step(float a, float b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static constexpr int Pad = 16;        // column stride M + 16 floats: the four fragments on distinct banks
    static constexpr bool RowStep4 = false;   // element v: row 4 (l / 16) + v
};
template <>
struct Mfma<double> {
    typedef v4d Acc;
    static __device__ __forceinline__ Acc step(double a, double b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static constexpr int Pad = 8;         // 64-byte column offset, as for FP32
    static constexpr bool RowStep4 = true;    // element v: row l / 16 + 4 v
};

// Dense 2^K x 2^K unitary on matrix cores (K = 5, 6; FP32 for complex64, FP64 for
// complex128), where the product's 8 x 2^K flops per amplitude pass the vector units' rate.
// y = U x is the real product [yr; yi] = W [xr; xi], W = [[Ur, -Ui], [Ui, Ur]] (M x M,
// M = 2^(K+1)), with one column per group of 2^K amplitudes. A wavefront takes 16 groups at a
// time and accumulates the M / 16 output blocks of 16 rows x 16 groups with 16 x 16 x 4
// matrix-core steps, four input rows per step:
//   A operand, lane l: W[16 t + l % 16][4 kk + l / 16] from LDS (W stored column-major with
//     a padded column stride, so the four 16-lane fragments of a read fall on distinct banks);
//   B operand, lane l: row 4 kk + l / 16 of group l % 16, from registers (lane l loads the
//     amplitudes c = l / 16 (mod 4) of its group);
//   result element v of lane l: row 16 t + 4 (l / 16) + v (FP32) or 16 t + l / 16 + 4 v
//     (FP64) of group l % 16; blocks t and t + M / 32 hold the real and imaginary part of the
//     same amplitude.
// Persistent workgroups stride over the tasks, so W is staged into LDS once per workgroup.
// (complex128 blocks of 4 and 5 qubits on states of 11 / 10 qubits or more take
// k_unitary_tile_mfma.)
template <class R, int K>
__global__ __launch_bounds__(kBlock) void k_unitary_mfma(Cx<R> *__restrict__ a, const R *__restrict__ wcol,
                                                         const UnitaryArgs<R> g, u64 n_tasks) {
    typedef typename Mfma<R>::Acc Acc;
    constexpr int D = 1 << K, M = 2 * D, NT = M / 16, MS = M + Mfma<R>::Pad;
    __shared__ R w[M * MS];   // w[c * MS + r] = W[r][c]
    for (int i = threadIdx.x; i < M * M; i += kBlock) w[(i / M) * MS + i % M] = wcol[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, p = lane >> 4;
    const u64 stride = (u64)gridDim.x * (kBlock / 64);
    // member offsets are ORs of per-bit offsets: member[4 q + p] = member[4 q] | member[p], and
    // an output row is member[16 t + v] | member[4 p] (FP32) or member[16 t + 4 v] | member[p]
    // (FP64), so each lane holds its own part (two selects, no branches) and ORs in a uniform
    // part per access
    const u64 lane_in = ((p & 1) ? g.member[1] : 0) | ((p & 2) ? g.member[2] : 0);
    const u64 lane_out = Mfma<R>::RowStep4 ? lane_in : (((p & 1) ? g.member[4] : 0) | ((p & 2) ? g.member[8] : 0));
    // x_c (c = 4 q + p) of group j of a task
    auto load = [&](u64 task, R (&xr)[D / 4], R (&xi)[D / 4]) __attribute__((always_inline)) {
        const u64 base = insert_zero_bits(task * 16 + j, g.ins) | lane_in;
#pragma unroll
        for (int q = 0; q < D / 4; ++q) {
            const Cx<R> v = a[base | g.member[4 * q]];
            xr[q] = v.re;
            xi[q] = v.im;
        }
    };
    // products and stores of one task, from its loaded inputs
    auto run = [&](u64 task, const R (&xr)[D / 4], const R (&xi)[D / 4]) __attribute__((always_inline)) {
        const u64 base = insert_zero_bits(task * 16 + j, g.ins) | lane_out;
        Acc acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[t][v] = R(0);
#pragma unroll
        for (int kk = 0; kk < M / 4; ++kk) {
            const R b = kk < D / 4 ? xr[kk] : xi[kk - D / 4];
            const R *col = w + (4 * kk + p) * MS + j;
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = Mfma<R>::step(col[16 * t], b, acc[t]);
            // one step's matrix reads in flight at a time (hoisting them all would take every
            // register of the wave)
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int t = 0; t < NT / 2; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                a[base | g.member[16 * t + (Mfma<R>::RowStep4 ? 4 * v : v)]] = cx<R>(acc[t][v], acc[t + NT / 2][v]);
            }
    };
    // Two input buffers used in turn, the next task's loads issued before this task's products
    // and stores, unconditionally (past the last task they read this task again): CDNA retires
    // vector memory operations in issue order, and a copy between buffers or a load inside a
    // branch would make the compiler wait for every access in flight, this task's stores too.
    u64 task = (u64)blockIdx.x * (kBlock / 64) + wave;
    R ar[D / 4], ai[D / 4], br[D / 4], bi[D / 4];
    if (task < n_tasks) load(task, ar, ai);
    while (task < n_tasks) {
        const u64 t1 = task + stride;
        load(t1 < n_tasks ? t1 : task, br, bi);
        run(task, ar, ai);
        if (t1 >= n_tasks) break;
        const u64 t2 = t1 + stride;
        load(t2 < n_tasks ? t2 : t1, ar, ai);
        run(t1, br, bi);
        task = t2;
    }
}

// Dense 4- or 5-qubit complex128 block on the FP64 matrix cores, over tiles: a tile of 2^M
// amplitudes (the block's lanes and the lowest other lanes, loaded and stored as rows of
// 16-byte units, as k_unitary_tile) goes to LDS, and each wavefront takes tasks of 16 of its
// groups: the B operands come from the tile, the steps are those of k_unitary_mfma (two 16-row
// output blocks for K = 4, four for K = 5), and the results go back into the tile in place (a
// wavefront reads all of a task's amplitudes before it writes them, and tasks share no group).
// Persistent workgroups stride over the tiles, staging W (10 / 36 KiB) once. Measured on a
// 30-qubit state: K = 5 6.80 ms against 7.78 ms for the per-group gathers of k_unitary_mfma,
// whose 16 runs of 256 B per task cap them below the streaming rate; K = 4 6.26 ms against
// 6.82 ms for k_unitary_tile on the vector units (tile sizes and workgroup counts below).
// Loading the next tile into registers during the products was slower (K = 5: 7.4-8.1 ms;
// its loads then mix with the stores). The tile's loads are issued together in straight code:
// through the generic tile_to_lds / lds_to_tile the compiler split them around the stores and
// waited inside the task loop (7.8 ms). complex64 blocks gain nothing from tiles (K = 5:
// 3.6-3.7 ms against 3.6 ms on any tile size) and keep the gathers.
template <class R, int K, int M>
__global__ __launch_bounds__(kBlock) void k_unitary_tile_mfma(Cx<R> *__restrict__ a, const R *__restrict__ wcol,
                                                              const DenseTileArgs g, u64 n_tiles) {
    typedef typename Mfma<R>::Acc Acc;
    constexpr int D = 1 << K, MR = 2 * D, NT = MR / 16, MS = MR + Mfma<R>::Pad;
    constexpr int G = 1 << (M - K), TASKS = G / 16, T = kBlock, NW = T / 64;
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0, U = (1 << (M - lw)) / T;
    static_assert(G % 16 == 0, "a tile holds whole tasks of 16 groups");
    __shared__ Cx<R> t[1 << M];
    __shared__ R w[MR * MS];   // w[c * MS + r] = W[r][c]
    for (int i = threadIdx.x; i < MR * MR; i += T) w[(i / MR) * MS + i % MR] = wcol[i];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, p = lane >> 4;
    // member offsets are ORs of per-bit offsets (fmem[4 q + p] = fmem[4 q] | fmem[p]); an output
    // row is member 16 t + v + 4 p (FP32) or 16 t + 4 v + p (FP64)
    const unsigned lane_in = g.fmem[p];
    const unsigned lane_out = Mfma<R>::RowStep4 ? g.fmem[p] : g.fmem[4 * p];
    // unit c T + tid of a tile (local index (c T + tid) << lw) sits at state offset
    // tile_base | deposit(tid << lw) | deposit(c T << lw); all U loads of a thread are issued
    // before its first LDS write
    const u64 dt = deposit((unsigned)threadIdx.x << lw, g.pos, M, g.low);
    for (u64 tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const u64 base = tile_base(tile, g.pos, M) | dt;
        {
            Amps<R> x[U];
#pragma unroll
            for (int c = 0; c < U; ++c)
                x[c] = load_unit(a, (base | deposit((unsigned)(c * T) << lw, g.pos, M, g.low)) >> lw);
#pragma unroll
            for (int c = 0; c < U; ++c)
#pragma unroll
                for (int v = 0; v < W; ++v) t[((unsigned)(c * T + (int)threadIdx.x) << lw) + v] = x[c].v[v];
        }
        __syncthreads();   // (the first one also orders the staging of W)
        for (int task = wave; task < TASKS; task += NW) {
            // local index of group task * 16 + j: deposited at the tile bits outside the block
            unsigned gb = 0, qq = (unsigned)(task * 16 + j);
            for (unsigned m = g.rest; m != 0; m &= m - 1) {
                if (qq & 1u) gb |= m & (~m + 1);
                qq >>= 1;
            }
            R xr[D / 4], xi[D / 4];
#pragma unroll
            for (int q = 0; q < D / 4; ++q) {
                const Cx<R> v = t[gb | lane_in | g.fmem[4 * q]];
                xr[q] = v.re;
                xi[q] = v.im;
            }
            Acc acc[NT];
#pragma unroll
            for (int tt = 0; tt < NT; ++tt)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[tt][v] = R(0);
#pragma unroll
            for (int kk = 0; kk < MR / 4; ++kk) {
                const R b = kk < D / 4 ? xr[kk] : xi[kk - D / 4];
                const R *col = w + (4 * kk + p) * MS + j;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) acc[tt] = Mfma<R>::step(col[16 * tt], b, acc[tt]);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int tt = 0; tt < NT / 2; ++tt)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    t[gb | lane_out | g.fmem[16 * tt + (Mfma<R>::RowStep4 ? 4 * v : v)]] =
                        cx<R>(acc[tt][v], acc[tt + NT / 2][v]);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < U; ++c) {
            Amps<R> y;
#pragma unroll
            for (int v = 0; v < W; ++v) y.v[v] = t[((unsigned)(c * T + (int)threadIdx.x) << lw) + v];
            store_unit(a, (base | deposit((unsigned)(c * T) << lw, g.pos, M, g.low)) >> lw, y);
        }
        __syncthreads();   // every thread has read t for its stores before the next tile lands
    }
}

// Tile bits and resident workgroups per compute unit of k_unitary_tile_mfma<double, K>,
// measured on the 30-qubit state (ms; tile KiB x workgroups per CU):
//   K = 4: 32 x 2 6.26-6.38, 16 x 3..6 7.5-7.7, 32 x 3 7.19, 64 x 2 6.89, 32 x 1 8.66
//          (k_unitary_tile on the vector units: 6.82);
//   K = 5: 16 x 3 6.80-6.89, 32 x 2 6.98-7.07, 16 x 2 9.74, 64 x 1 9.5.
// complex128 3-qubit blocks took 7.2 ms at best on the same kernel (one 16-row block holding
// the real and imaginary rows) against 6.41 ms on the vector units, and stay there.
template <int K>
struct MfmaTileCfg;
template <>
struct MfmaTileCfg<4> {
    static constexpr int M = 11, PerCU = 2;
};
template <>
struct MfmaTileCfg<5> {
    static constexpr int M = 10, PerCU = 3;
};

// 4- and 5-qubit complex128 blocks on states of at least one tile take the tile kernel.
template <class R>
bool use_mfma_tile(int k, int n_lanes) {
    return std::is_same<R, double>::value && ((k == 4 && n_lanes >= MfmaTileCfg<4>::M) ||
                                              (k == 5 && n_lanes >= MfmaTileCfg<5>::M));
}

template <int K>
void launch_mfma_tile(Cx<double> *a, const double *w, int n_lanes, const std::vector<int> &lanes, hipStream_t s) {
    constexpr int M = MfmaTileCfg<K>::M;
    static const unsigned per_cu = [] {
        int n = 0;
        QG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_unitary_tile_mfma<double, K, M>, kBlock, 0));
        return (unsigned)std::max(1, std::min(n, MfmaTileCfg<K>::PerCU));
    }();
    const DenseTileArgs tg = dense_tile_args<double>(n_lanes, lanes, M);
    const u64 n_tiles = u64(1) << (n_lanes - M);
    const u64 grid = std::min<u64>(n_tiles, (u64)per_cu * (u64)std::max(1, runtime().n_cus));
    hipLaunchKernelGGL((k_unitary_tile_mfma<double, K, M>), (unsigned)grid, kBlock, 0, s, a, w, tg, n_tiles);
}

// Workgroups per compute unit of the matrix-core kernel (LDS of W: complex64 20 KiB for K = 5,
// 72 KiB for 6; complex128 36 KiB and 136 KiB).
template <class R>
constexpr int mfma_groups_per_cu(int k) {
    return sizeof(R) == 4 ? (k == 6 ? 2 : 8) : (k == 6 ? 1 : 4);
}

// The matrix-core kernel takes blocks of 5 or 6 qubits on states of at least 16 groups;
// everything else takes the vector kernels.
template <class R>
bool use_mfma(int k, int n_lanes) {
    return (k == 5 || k == 6) && n_lanes - k >= 4;
}

// Device matrix of a block: complex row-major for the vector kernel (D x D Cx<R>), the real
// form W column-major for the matrix-core kernel (M x M floats). Bytes of either.
template <class R>
bool real_form(int k, int n_lanes) {
    return use_mfma<R>(k, n_lanes) || use_mfma_tile<R>(k, n_lanes);
}

template <class R>
size_t unitary_bytes(int k, int n_lanes) {
    const size_t D = size_t(1) << k;
    return real_form<R>(k, n_lanes) ? sizeof(R) * 4 * D * D : sizeof(Cx<R>) * D * D;
}

template <class R>
void pack_unitary(int k, int n_lanes, const std::vector<std::complex<double>> &mat, char *out) {
    const int D = 1 << k;
    if (real_form<R>(k, n_lanes)) {
        const int M = 2 * D;
        R *w = reinterpret_cast<R *>(out);
        for (int r = 0; r < M; ++r)
            for (int c = 0; c < M; ++c) {
                const std::complex<double> u = mat[(size_t)(r % D) * D + (c % D)];
                const R re = (R)u.real(), im = (R)u.imag();
                w[(size_t)c * M + r] = (r < D) == (c < D) ? re : (r < D ? -im : im);
            }
        return;
    }
    Cx<R> *m = reinterpret_cast<Cx<R> *>(out);
    for (int i = 0; i < D * D; ++i) m[i] = cx<R>((R)mat[i].real(), (R)mat[i].imag());
}

template <class R>
UnitaryArgs<R> unitary_args(int n_lanes, const std::vector<int> &lanes) {
    const int k = (int)lanes.size();
    QG_REQUIRE(1 <= k && k <= kMaxFuse, "unitary width %d out of range", k);
    QG_REQUIRE(k <= n_lanes, "unitary wider than the state");
    UnitaryArgs<R> g;
    std::vector<int> sorted(lanes);
    std::sort(sorted.begin(), sorted.end());
    g.ins.n = k;
    for (int i = 0; i < k; ++i) {
        QG_REQUIRE(0 <= sorted[i] && sorted[i] < n_lanes, "lane %d out of range", sorted[i]);
        QG_REQUIRE(i == 0 || sorted[i] != sorted[i - 1], "duplicate lane in unitary");
        g.ins.pos[i] = sorted[i];
    }
    for (int c = 0; c < (1 << k); ++c) {
        u64 off = 0;
        for (int j = 0; j < k; ++j)
            if (c & (1 << j)) off |= u64(1) << lanes[j];
        g.member[c] = off;
    }
    for (int c = 1 << k; c < (1 << kMaxFuse); ++c) g.member[c] = 0;
    g.k = k;
    g.n_groups = u64(1) << (n_lanes - k);
    return g;
}

// Launch one block whose matrix is already on the device (packed by pack_unitary).
template <class R>
void launch_unitary_dev(Cx<R> *a, int n_lanes, const std::vector<int> &lanes, const void *dmat) {
    const UnitaryArgs<R> g = unitary_args<R>(n_lanes, lanes);
    const int k = g.k;
    hipStream_t s = stream();
    if constexpr (std::is_same<R, double>::value) {
        if (use_mfma_tile<R>(k, n_lanes)) {
            const R *w = static_cast<const R *>(dmat);
            if (k == 5)
                launch_mfma_tile<5>(a, w, n_lanes, lanes, s);
            else
                launch_mfma_tile<4>(a, w, n_lanes, lanes, s);
            return;
        }
    }
    if (use_mfma<R>(k, n_lanes)) {
        const u64 n_tasks = g.n_groups / 16;
        const int per_cu = mfma_groups_per_cu<R>(k);
        const unsigned grid =
            (unsigned)std::min<u64>((n_tasks + 3) / 4, (u64)std::max(1, runtime().n_cus * per_cu));
        const R *w = static_cast<const R *>(dmat);
        if (k == 5)
            hipLaunchKernelGGL((k_unitary_mfma<R, 5>), grid, kBlock, 0, s, a, w, g, n_tasks);
        else
            hipLaunchKernelGGL((k_unitary_mfma<R, 6>), grid, kBlock, 0, s, a, w, g, n_tasks);
        return;
    }
    const Cx<R> *dm = static_cast<const Cx<R> *>(dmat);
    if (use_dense_tile<R>(k, n_lanes)) {
        const DenseTileArgs tg = dense_tile_args<R>(n_lanes, lanes);
        const u64 n_tiles = u64(1) << (n_lanes - DenseTileCfg<R>::M);
        const u64 piece = launch_piece_items() / kBlock;
        for (u64 first = 0; first < n_tiles; first += piece) {
            const unsigned grid = (unsigned)std::min(piece, n_tiles - first);
            switch (k) {
            case 2: hipLaunchKernelGGL((k_unitary_tile<R, 2>), grid, kBlock, 0, s, a, dm, tg, first); break;
            case 3: hipLaunchKernelGGL((k_unitary_tile<R, 3>), grid, kBlock, 0, s, a, dm, tg, first); break;
            case 4: hipLaunchKernelGGL((k_unitary_tile<R, 4>), grid, kBlock, 0, s, a, dm, tg, first); break;
            }
        }
        return;
    }
    for_each_piece(g.n_groups, [&](u64 first, u64 count) {
        const unsigned grid = grid_for(count, kBlock);
        switch (k) {
        case 1: hipLaunchKernelGGL((k_unitary<R, 1>), grid, kBlock, 0, s, a, dm, g, first); break;
        case 2: hipLaunchKernelGGL((k_unitary<R, 2>), grid, kBlock, 0, s, a, dm, g, first); break;
        case 3: hipLaunchKernelGGL((k_unitary<R, 3>), grid, kBlock, 0, s, a, dm, g, first); break;
        case 4: hipLaunchKernelGGL((k_unitary<R, 4>), grid, kBlock, 0, s, a, dm, g, first); break;
        case 5: hipLaunchKernelGGL((k_unitary<R, 5>), grid, kBlock, 0, s, a, dm, g, first); break;
        case 6: hipLaunchKernelGGL((k_unitary<R, 6>), grid, kBlock, 0, s, a, dm, g, first); break;
        }
    });
}

// Unitary of a run of gates on `lanes` (matrix index bit j <-> lanes[j]), in FP64: each
// gate, in order, on every column of the running product (the pair update of
// CPUQubitProcessor.cpp:307-362 on a 2^k-amplitude register).
std::vector<std::complex<double>> block_unitary(const std::vector<int> &lanes,
                                                const std::vector<const QueuedGate *> &gates) {
    const int k = (int)lanes.size(), D = 1 << k;
    std::vector<std::complex<double>> u((size_t)D * D, 0.);
    for (int i = 0; i < D; ++i) u[(size_t)i * D + i] = 1.;
    auto bit_of = [&](int lane) {
        return (int)(std::find(lanes.begin(), lanes.end(), lane) - lanes.begin());
    };
    for (const QueuedGate *g : gates) {
        const int t = bit_of(g->target);
        int cmask = 0;
        for (int c : g->ctrl) cmask |= 1 << bit_of(c);
        const auto &m = g->m.m;
        for (int i0 = 0; i0 < D; ++i0) {
            if ((i0 >> t) & 1 || (i0 & cmask) != cmask) continue;
            const int i1 = i0 | (1 << t);
            for (int c = 0; c < D; ++c) {
                const std::complex<double> x0 = u[(size_t)i0 * D + c], x1 = u[(size_t)i1 * D + c];
                u[(size_t)i0 * D + c] = m[0][0] * x0 + m[0][1] * x1;
                u[(size_t)i1 * D + c] = m[1][0] * x0 + m[1][1] * x1;
            }
        }
    }
    return u;
}

}  // namespace

template <class R>
void launch_unitary(Cx<R> *a, int n_lanes, const std::vector<int> &lanes,
                    const std::vector<std::complex<double>> &mat) {
    const int k = (int)lanes.size();
    QG_REQUIRE(1 <= k && k <= kMaxFuse, "unitary width %d out of range", k);
    const int D = 1 << k;
    QG_REQUIRE((int)mat.size() == D * D, "unitary matrix has wrong size");
    unitary_args<R>(n_lanes, lanes);   // checks the lanes before anything is allocated
    const size_t bytes = unitary_bytes<R>(k, n_lanes);
    std::vector<char> hm(bytes);
    pack_unitary<R>(k, n_lanes, mat, hm.data());
    void *dm = dev_alloc(bytes);
    hipStream_t s = stream();
    hipError_t e = hipMemcpyAsync(dm, hm.data(), bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        launch_unitary_dev<R>(a, n_lanes, lanes, dm);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(s);   // hm is a host temporary
    dev_free(dm, bytes);
    QG_HIP(e);
    QG_HIP(e2);
}

template <class R>
void apply_dense_sequence(Cx<R> *a, int n_lanes, const std::vector<QueuedGate> &gates, int width) {
    QG_REQUIRE(2 <= width && width <= kMaxFuse, "dense block width %d out of range", width);
    // plan: each block scans the gates not yet applied in program order and takes every gate
    // that commutes with all the gates it left out before it (on every lane they share, both
    // act diagonally: a control, or the target of a diagonal gate) and whose lanes, controls
    // included, keep the block within `width` lanes. The gates of a block keep their program
    // order, and so do the ones left for later blocks. (Blocks of consecutive gates only: the
    // 2000-gate 30-qubit random circuit took 669 blocks at width 5, 541 at width 6; the scan
    // makes them about 360 and 290.) A gate wider than a block goes alone. The scan stops
    // after kLookahead gates left out or once every lane is held.
    struct Block {
        std::vector<int> lanes;
        std::vector<const QueuedGate *> gates;
    };
    constexpr int kLookahead = 256;
    const int n_gates = (int)gates.size();
    std::vector<char> done(n_gates, 0);
    std::vector<unsigned char> held(n_lanes);   // 0 free, 1 used diagonally, 2 otherwise
    std::vector<Block> blocks;
    for (int first = 0; first < n_gates;) {
        Block cur;
        std::fill(held.begin(), held.end(), (unsigned char)0);
        int left = 0, held_all = 0;
        for (int i = first; i < n_gates; ++i) {
            if (done[i]) continue;
            const QueuedGate &g = gates[i];
            const bool dg = is_diagonal(g.m);
            bool conflict = dg ? held[g.target] == 2 : held[g.target] != 0;
            for (int c : g.ctrl) conflict = conflict || held[c] == 2;
            if (!conflict) {
                std::vector<int> merged(cur.lanes);
                auto add = [&](int l) {
                    if (std::find(merged.begin(), merged.end(), l) == merged.end()) merged.push_back(l);
                };
                for (int c : g.ctrl) add(c);
                add(g.target);
                if ((int)merged.size() <= width) {
                    cur.lanes.swap(merged);
                    cur.gates.push_back(&g);
                    done[i] = 1;
                    continue;
                }
            }
            auto hold = [&](int lane, unsigned char v) {
                if (held[lane] < 2 && v == 2) ++held_all;
                held[lane] = std::max(held[lane], v);
            };
            hold(g.target, dg ? 1 : 2);
            for (int c : g.ctrl) hold(c, 1);
            if (++left >= kLookahead || held_all == n_lanes) break;
        }
        if (cur.gates.empty()) {   // the first gate is wider than a block: alone
            const QueuedGate &g = gates[first];
            cur.lanes = g.ctrl;
            cur.lanes.push_back(g.target);
            cur.gates.push_back(&g);
            done[first] = 1;
        }
        blocks.push_back(cur);
        while (first < n_gates && done[first]) ++first;
    }
    // matrices of the multi-gate blocks, uploaded together
    std::vector<size_t> offset(blocks.size(), 0);
    size_t total = 0;
    // (a block on one lane is a 2x2 gate, streamed by launch_gate below: no matrix upload)
    for (size_t b = 0; b < blocks.size(); ++b) {
        if (blocks[b].gates.size() < 2 || blocks[b].lanes.size() < 2) continue;
        offset[b] = total;
        total += (unitary_bytes<R>((int)blocks[b].lanes.size(), n_lanes) + 255) & ~size_t(255);
    }
    std::vector<char> host(std::max<size_t>(total, 1));
    for (size_t b = 0; b < blocks.size(); ++b)
        if (blocks[b].gates.size() >= 2 && blocks[b].lanes.size() >= 2)
            pack_unitary<R>((int)blocks[b].lanes.size(), n_lanes, block_unitary(blocks[b].lanes, blocks[b].gates),
                            host.data() + offset[b]);
    char *dev = total > 0 ? static_cast<char *>(dev_alloc(total)) : nullptr;
    hipStream_t s = stream();
    hipError_t e = total > 0 ? hipMemcpyAsync(dev, host.data(), total, hipMemcpyHostToDevice, s) : hipSuccess;
    try {
        for (size_t b = 0; b < blocks.size() && e == hipSuccess; ++b) {
            const Block &bl = blocks[b];
            if (bl.gates.size() == 1) {
                const QueuedGate &g = *bl.gates[0];
                launch_gate<R>(a, n_lanes, g.m, g.ctrl, g.target);
            } else if (bl.lanes.size() == 1) {   // the product of the block's gates, in FP64
                const std::vector<std::complex<double>> u = block_unitary(bl.lanes, bl.gates);
                GateMatrix gm;
                for (int r = 0; r < 2; ++r)
                    for (int c = 0; c < 2; ++c) gm.m[r][c] = u[(size_t)r * 2 + c];
                launch_gate<R>(a, n_lanes, gm, std::vector<int>(), bl.lanes[0]);
            } else {
                launch_unitary_dev<R>(a, n_lanes, bl.lanes, dev + offset[b]);
            }
            e = hipGetLastError();
        }
    } catch (...) {
        hipStreamSynchronize(s);
        if (dev) dev_free(dev, total);
        throw;
    }
    const hipError_t e2 = hipStreamSynchronize(s);   // the matrices are host temporaries
    if (dev) dev_free(dev, total);
    QG_HIP(e);
    QG_HIP(e2);
}

template void launch_unitary<float>(Cx<float> *, int, const std::vector<int> &,
                                    const std::vector<std::complex<double>> &);
template void launch_unitary<double>(Cx<double> *, int, const std::vector<int> &,
                                     const std::vector<std::complex<double>> &);
template void apply_dense_sequence<float>(Cx<float> *, int, const std::vector<QueuedGate> &, int);
template void apply_dense_sequence<double>(Cx<double> *, int, const std::vector<QueuedGate> &, int);

}  // namespace qg
