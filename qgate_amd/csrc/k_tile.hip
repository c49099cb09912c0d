// Tile passes: many gates per trip of the state vector through HBM.
//
// A pass picks a set T of m state bits (m <= 12 for complex64, 11 for complex128; 13 / 12 for
// register-resident passes over states of 28 or more qubits) that always
// holds the lowest `low` bits (so each tile row is a contiguous run of 2^low amplitudes) plus
// the targets of the gates it carries. The state splits into 2^(n-m) tiles of 2^m amplitudes
// (the bits outside T fixed per tile); each workgroup loads one tile into LDS, applies the
// pass's gates there in order and writes the tile back. Per gate:
//   * the target is a tile bit (or the gate is diagonal: a target outside the tile only picks
//     m00 or m11 for the whole tile);
//   * controls inside the tile are checked per amplitude; controls outside it are one
//     predicate per tile, so they never use a tile bit.
// Gates are applied in program order with the reference's pair update
// (CPUQubitProcessor.cpp:307-362): a[i0] <- m00 a[i0] + m01 a[i1], a[i1] <- m10 a[i0] + m11 a[i1].
// The gate list of a pass travels in the kernel arguments (at most 40 / 24 gates).
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "kernels.hpp"
#include "qg_runtime.hpp"
#include "tile_common.hpp"
#include "tile_reg.hpp"

namespace qg {

namespace {

constexpr int kTileThreads = 256;

template <class R>
struct TileOp {
    Mat2<R> m;
    u64 gctrl;          // control bits outside the tile (state positions)
    unsigned lctrl;     // control bits inside the tile (local positions)
    short target;       // local bit (pair, diag-local) or state bit (diag-outer)
    short kind;
};

template <class R>
struct TileProgram {
    int m;              // tile bits
    int low;            // pos[j] == j for j < low
    int n_ops;
    alignas(4) signed char pos[kTilePosMax];   // state bit of local bit j, ascending
    TileOp<R> ops[TileCfg<R>::MaxOps];
};

template <class R>
__device__ __forceinline__ void tile_mix(Cx<R> *t, const Mat2<R> &mm, int i0, int i1) {
    const Cx<R> x0 = t[i0], x1 = t[i1];
    t[i0] = cadd(cmul(mm.m00, x0), cmul(mm.m01, x1));
    t[i1] = cadd(cmul(mm.m10, x0), cmul(mm.m11, x1));
}

template <class R, int MB>
__global__ __launch_bounds__(kTileThreads) void k_tile_pass(Cx<R> *__restrict__ a, const TileProgram<R> prog,
                                                            u64 first) {
    __shared__ Cx<R> t[1 << MB];
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const int m = prog.m;
    const u64 base = tile_base(first + blockIdx.x, prog.pos, m);
    load_tile_lds<R, kTileThreads, MB>(t, a, base, prog.pos, m, prog.low);
    __syncthreads();
    const int n_amps = 1 << m;
    for (int k = 0; k < prog.n_ops; ++k) {
        const TileOp<R> &op = prog.ops[k];
        if ((base & op.gctrl) != op.gctrl) continue;   // same for the whole workgroup
        const unsigned lc = op.lctrl;
        if (op.kind == kOpPair) {
            const int tb = op.target;
            const unsigned lowmask = (1u << tb) - 1;
            for (int p = threadIdx.x; p < n_amps / 2; p += kTileThreads) {
                const unsigned i0 = (((unsigned)p & ~lowmask) << 1) | ((unsigned)p & lowmask);
                if ((i0 & lc) == lc) tile_mix(t, op.m, (int)i0, (int)(i0 | (1u << tb)));
            }
        }
        if (op.kind == kOpDiagLocal) {
            const int tb = op.target;
            for (int i = threadIdx.x; i < n_amps; i += kTileThreads) {
                if (((unsigned)i & lc) == lc) t[i] = cmul(((i >> tb) & 1) ? op.m.m11 : op.m.m00, t[i]);
            }
        }
        if (op.kind == kOpDiagOuter) {
            const Cx<R> f = ((base >> op.target) & 1) ? op.m.m11 : op.m.m00;
            for (int i = threadIdx.x; i < n_amps; i += kTileThreads) {
                if (((unsigned)i & lc) == lc) t[i] = cmul(f, t[i]);
            }
        }
        __syncthreads();
    }
    store_tile_lds<R, kTileThreads, MB>(a, t, base, prog.pos, m, prog.low);
}

// Pauli X (1), Y (2), Z (3); zero matrix otherwise.
void set_pauli(GateMatrix &g, int pauli) {
    const std::complex<double> J(0., 1.);
    g.m[0][0] = g.m[0][1] = g.m[1][0] = g.m[1][1] = 0.;
    if (pauli == 1) {
        g.m[0][1] = 1.;
        g.m[1][0] = 1.;
    } else if (pauli == 2) {
        g.m[0][1] = -J;
        g.m[1][0] = J;
    } else if (pauli == 3) {
        g.m[0][0] = 1.;
        g.m[1][1] = -1.;
    }
}

// Register tiles of the single-circuit passes on states of kWideLanes lanes or more: one bit
// wider than the LDS-pass tiles (64 KiB), 512-thread workgroups with 16 (complex64) / 8
// (complex128) amplitudes a thread, so the registers per thread are those of the 32 KiB tiles
// (two workgroups share a compute unit's LDS). A pass then carries the gates of one more
// target bit: the 2000-gate 30-qubit random circuit takes 123 passes instead of 144, and its
// step 499 ms instead of 507 (complex64), 991 ms instead of 1034 (complex128). On a
// 24-qubit state the narrower tiles are faster (an 864-gate circuit 3.74 ms against 4.07).
constexpr int kWideLanes = 28;
template <class R>
struct WideCfg;
template <>
struct WideCfg<float> {
    static constexpr int Bits = 13, Threads = 512;
};
template <>
struct WideCfg<double> {
    static constexpr int Bits = 12, Threads = 512;
};

// One pass being built: the tile bits and the gates it carries.
struct PassPlan {
    std::vector<int> bits;                 // state bits of the tile (unsorted while building)
    std::vector<const QueuedGate *> gates;
    bool has(int b) const { return std::find(bits.begin(), bits.end(), b) != bits.end(); }
};

template <class R>
void launch_pass(Cx<R> *a, int n_lanes, const PassPlan &plan, int low, int max_bits) {
    typedef TileCfg<R> C;
    if (plan.gates.size() == 1) {   // a lone gate streams faster than through LDS
        const QueuedGate &g = *plan.gates[0];
        launch_gate<R>(a, n_lanes, g.m, g.ctrl, g.target);
        return;
    }
    TileProgram<R> prog;
    std::vector<int> bits(plan.bits);
    // fill the tile up to its full size with the lowest free bits: bigger tiles mean fewer,
    // fuller workgroups and longer contiguous rows
    for (int b = 0; (int)bits.size() < max_bits && b < n_lanes; ++b)
        if (std::find(bits.begin(), bits.end(), b) == bits.end()) bits.push_back(b);
    std::sort(bits.begin(), bits.end());
    prog.m = (int)bits.size();
    prog.low = low;
    for (int j = 0; j < 16; ++j) prog.pos[j] = j < prog.m ? (signed char)bits[j] : (signed char)0;
    int local[64];
    for (int b = 0; b < 64; ++b) local[b] = -1;
    for (int j = 0; j < prog.m; ++j) local[bits[j]] = j;
    prog.n_ops = (int)plan.gates.size();
    QG_REQUIRE(prog.n_ops <= C::MaxOps, "tile pass holds too many gates");
    for (int k = 0; k < prog.n_ops; ++k) {
        const QueuedGate &g = *plan.gates[k];
        TileOp<R> &op = prog.ops[k];
        op.m = round_matrix<R>(g.m);
        op.gctrl = 0;
        op.lctrl = 0;
        for (int c : g.ctrl) {
            if (local[c] >= 0)
                op.lctrl |= 1u << local[c];
            else
                op.gctrl |= u64(1) << c;
        }
        if (local[g.target] >= 0) {
            op.target = (short)local[g.target];
            op.kind = (short)(is_diagonal(g.m) ? kOpDiagLocal : kOpPair);
        } else {
            QG_REQUIRE(is_diagonal(g.m), "tile pass: non-diagonal gate outside the tile");
            op.target = (short)g.target;
            op.kind = (short)kOpDiagOuter;
        }
    }
    const u64 n_tiles = u64(1) << (n_lanes - prog.m);
    hipStream_t s = stream();
    // one workgroup per tile; pieces keep every grid within launch_piece_items() threads
    const u64 piece = launch_piece_items() / kTileThreads;
    if ((prog.m == C::MaxBits || prog.m == WideCfg<R>::Bits) && use_register_tiles()) {
        RegProgram<R> rp;
        rp.n_ops = prog.n_ops;
        rp.low = prog.low;
        std::memcpy(rp.pos, prog.pos, sizeof(rp.pos));
        TileOp<R> tops[C::MaxOps];
        for (int k = 0; k < prog.n_ops; ++k) {
            tops[k] = prog.ops[k];
            TileOp<R> &op = tops[k];
            const Mat2<R> &mm = op.m;
            const R zero = R(0), one = R(1);
            if (op.kind == kOpPair) {
                const bool swap = mm.m00.re == zero && mm.m00.im == zero && mm.m11.re == zero && mm.m11.im == zero &&
                                  mm.m01.re == one && mm.m01.im == zero && mm.m10.re == one && mm.m10.im == zero;
                const bool real = mm.m00.im == zero && mm.m01.im == zero && mm.m10.im == zero && mm.m11.im == zero;
                op.kind = (short)(swap ? kOpPairSwap : real ? kOpPairReal : kOpPair);
            } else if (op.kind == kOpDiagLocal && mm.m00.re == one && mm.m00.im == zero) {
                op.kind = (short)kOpDiagPhase;
            }
        }
        if (prog.m == WideCfg<R>::Bits) {
            constexpr int WB = WideCfg<R>::Bits, WT = WideCfg<R>::Threads, A = (1 << WB) / WT;
            if (plan_phases<R>(rp, tops, prog.m, Log2<A>::v, Log2<WT / 64>::v)) {
                const u64 wpiece = launch_piece_items() / WT;
                for (u64 first = 0; first < n_tiles; first += wpiece) {
                    const u64 count = std::min(wpiece, n_tiles - first);
                    hipLaunchKernelGGL((k_tile_reg<R, WB, WT, false>), (unsigned)count, WT, 0, s, a, rp, first,
                                       (const Mat2<R> *)nullptr, 0, 0);
                }
                QG_HIP(hipGetLastError());
                return;
            }
        } else {
            constexpr int A = (1 << C::MaxBits) / kTileThreads;
            if (plan_phases<R>(rp, tops, prog.m, Log2<A>::v, Log2<kTileThreads / 64>::v)) {
                for (u64 first = 0; first < n_tiles; first += piece) {
                    const u64 count = std::min(piece, n_tiles - first);
                    hipLaunchKernelGGL((k_tile_reg<R, C::MaxBits, kTileThreads, false>), (unsigned)count, kTileThreads,
                                       0, s, a, rp, first, (const Mat2<R> *)nullptr, 0, 0);
                }
                QG_HIP(hipGetLastError());
                return;
            }
        }
    }
    if (prog.m > C::MaxBits) {
        // a wide tile whose gates need more register phases than a program holds
        for (u64 first = 0; first < n_tiles; first += piece) {
            const u64 count = std::min(piece, n_tiles - first);
            hipLaunchKernelGGL((k_tile_pass<R, WideCfg<R>::Bits>), (unsigned)count, kTileThreads, 0, s, a, prog, first);
        }
        QG_HIP(hipGetLastError());
        return;
    }
    for (u64 first = 0; first < n_tiles; first += piece) {
        const u64 count = std::min(piece, n_tiles - first);
        hipLaunchKernelGGL((k_tile_pass<R, C::MaxBits>), (unsigned)count, kTileThreads, 0, s, a, prog, first);
    }
    QG_HIP(hipGetLastError());
}

// ---- adjoint sweep ----------------------------------------------------------------------

// One pass of the adjoint backward sweep over a phi tile and a lambda tile: for each op in
// order, the gradient term Im<lambda| C G |phi> of the tile (FP64, written to
// partial[slot * n_tiles + tile]) and then phi, lambda <- U^dagger.
template <class R, int MB>
__global__ __launch_bounds__(kTileThreads) void k_tile_sweep(Cx<R> *__restrict__ phi, Cx<R> *__restrict__ lam,
                                                             const SweepPass pass,
                                                             const SweepOp<R> *__restrict__ ops,
                                                             double *__restrict__ partial, u64 n_tiles,
                                                             u64 first) {
    __shared__ Cx<R> tp[1 << MB];
    __shared__ Cx<R> tl[1 << MB];
    __shared__ double scratch[kTileThreads / 64];
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const int m = pass.m;
    const u64 tile = first + blockIdx.x;
    const u64 base = tile_base(tile, pass.pos, m);
    load_tile_lds<R, kTileThreads, MB>(tp, phi, base, pass.pos, m, pass.low);
    load_tile_lds<R, kTileThreads, MB>(tl, lam, base, pass.pos, m, pass.low);
    __syncthreads();
    const int n_amps = 1 << m;
    for (int k = 0; k < pass.n_ops; ++k) {
        const SweepOp<R> &op = ops[pass.op_begin + k];
        const bool on = (base & op.gctrl) == op.gctrl;
        const unsigned lc = op.lctrl;
        double s = 0.;
        if (on && op.kind == kOpPair) {
            const int tb = op.target;
            const unsigned lowmask = (1u << tb) - 1;
            for (int p = threadIdx.x; p < n_amps / 2; p += kTileThreads) {
                const unsigned i0 = (((unsigned)p & ~lowmask) << 1) | ((unsigned)p & lowmask);
                if ((i0 & lc) != lc) continue;
                const unsigned i1 = i0 | (1u << tb);
                const Cx<R> p0 = tp[i0], p1 = tp[i1], l0 = tl[i0], l1 = tl[i1];
                if (op.slot >= 0) {
                    const Cx<R> q0 = cadd(cmul(op.gen.m00, p0), cmul(op.gen.m01, p1));
                    const Cx<R> q1 = cadd(cmul(op.gen.m10, p0), cmul(op.gen.m11, p1));
                    s += (double)l0.re * (double)q0.im - (double)l0.im * (double)q0.re +
                         (double)l1.re * (double)q1.im - (double)l1.im * (double)q1.re;
                }
                tp[i0] = cadd(cmul(op.udag.m00, p0), cmul(op.udag.m01, p1));
                tp[i1] = cadd(cmul(op.udag.m10, p0), cmul(op.udag.m11, p1));
                tl[i0] = cadd(cmul(op.udag.m00, l0), cmul(op.udag.m01, l1));
                tl[i1] = cadd(cmul(op.udag.m10, l0), cmul(op.udag.m11, l1));
            }
        }
        if (on && op.kind != kOpPair) {   // diagonal: per amplitude, factor by the target bit
            const int tb = op.target;
            for (int i = threadIdx.x; i < n_amps; i += kTileThreads) {
                if (((unsigned)i & lc) != lc) continue;
                const bool one = op.kind == kOpDiagLocal ? ((i >> tb) & 1) != 0 : ((base >> tb) & 1) != 0;
                const Cx<R> x = tp[i], y = tl[i];
                if (op.slot >= 0) {
                    const Cx<R> q = cmul(one ? op.gen.m11 : op.gen.m00, x);
                    s += (double)y.re * (double)q.im - (double)y.im * (double)q.re;
                }
                const Cx<R> f = one ? op.udag.m11 : op.udag.m00;
                tp[i] = cmul(f, x);
                tl[i] = cmul(f, y);
            }
        }
        if (op.slot >= 0) {
            s = block_sum_tile<kTileThreads>(s, scratch);
            if (threadIdx.x == 0) partial[(u64)op.slot * n_tiles + tile] = s;
        }
        __syncthreads();
    }
    store_tile_lds<R, kTileThreads, MB>(phi, tp, base, pass.pos, m, pass.low);
    store_tile_lds<R, kTileThreads, MB>(lam, tl, base, pass.pos, m, pass.low);
}

// Sum of row[0..n) over a workgroup of T threads in a fixed order (valid in thread 0). Thread
// t adds row[t + k T] into 8 interleaved accumulators, so 8 independent loads are in flight,
// and the accumulators and then the threads are combined in a fixed tree.
template <int T>
__device__ __forceinline__ double row_sum(const double *__restrict__ row, u64 n, double *scratch) {
    double a0 = 0., a1 = 0., a2 = 0., a3 = 0., a4 = 0., a5 = 0., a6 = 0., a7 = 0., tail = 0.;
    u64 i = threadIdx.x;
    for (; i + 7 * (u64)T < n; i += 8 * (u64)T) {
        a0 += row[i];
        a1 += row[i + T];
        a2 += row[i + 2 * T];
        a3 += row[i + 3 * T];
        a4 += row[i + 4 * T];
        a5 += row[i + 5 * T];
        a6 += row[i + 6 * T];
        a7 += row[i + 7 * T];
    }
    for (; i < n; i += T) tail += row[i];
    const double s = (((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7))) + tail;
    return block_sum_tile<T>(s, scratch);
}

constexpr int kRowThreads = 1024;
// Device memory for the sweep's per-tile gradient partials of consecutive passes.
constexpr size_t kSweepPartialBytes = size_t(256) << 20;

// terms[slot_gate[r]] = sum over tiles of partial[r * n_tiles + tile], fixed order; one
// workgroup per slot row.
__global__ __launch_bounds__(kRowThreads) void k_slot_rows(const double *__restrict__ partial, u64 n_tiles,
                                                          const int *__restrict__ slot_gate,
                                                          double *__restrict__ terms) {
    __shared__ double scratch[kRowThreads / 64];
    const double s = row_sum<kRowThreads>(partial + (u64)blockIdx.x * n_tiles, n_tiles, scratch);
    if (threadIdx.x == 0) terms[slot_gate[blockIdx.x]] = s;
}

// terms[slot] = sum over tiles of partial[slot * n_tiles + tile], fixed order; one workgroup
// per slot.
__global__ __launch_bounds__(kRowThreads) void k_sum_slots(const double *__restrict__ partial, u64 n_tiles,
                                                          double *__restrict__ terms) {
    __shared__ double scratch[kRowThreads / 64];
    const double s = row_sum<kRowThreads>(partial + (u64)blockIdx.x * n_tiles, n_tiles, scratch);
    if (threadIdx.x == 0) terms[blockIdx.x] = s;
}

// ---- Pauli sums ---------------------------------------------------------------------------

template <class R>
struct PauliCfg;
template <>
struct PauliCfg<float> {
    static constexpr int MaxBits = 12;
    static constexpr int Low = 5;
};
template <>
struct PauliCfg<double> {
    static constexpr int MaxBits = 11;
    static constexpr int Low = 4;
};

struct PauliPass {
    int m, low;
    int g_begin, n_groups;
    int accumulate;     // add to lambda (later passes) instead of writing it
    int dot;            // last pass: partial[tile] = Re<psi|lambda> over the tile
    alignas(4) signed char pos[kTilePosMax];
};

// lambda (+)= sum over this pass's groups g of sum_{k in g} c_k (-1)^popcount(x & zm_k) psi[x ^ xm_g]
// for a tile holding every X bit of those groups; c_k carries the (-i)^(number of Y) factor.
template <class R, int MB>
__global__ __launch_bounds__(kTileThreads) void k_tile_pauli(const Cx<R> *__restrict__ psi, Cx<R> *__restrict__ lam,
                                                             const PauliPass pass,
                                                             const PauliGroup *__restrict__ groups,
                                                             const PauliTerm *__restrict__ terms,
                                                             double *__restrict__ partial, u64 first) {
    __shared__ Cx<R> tp[1 << MB];
    __shared__ double scratch[kTileThreads / 64];
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const int m = pass.m;
    const u64 tile = first + blockIdx.x;
    const u64 base = tile_base(tile, pass.pos, m);
    const int n_units = 1 << (m - lw);
    load_tile_lds<R, kTileThreads, MB>(tp, psi, base, pass.pos, m, pass.low);
    __syncthreads();
    double dot = 0.;
    const u64 dt = base | deposit((unsigned)threadIdx.x << lw, pass.pos, m, pass.low);
    for (int lu = threadIdx.x; lu < n_units; lu += kTileThreads) {
        const unsigned l = (unsigned)lu << lw;
        // state index of amplitude l: per-thread part | per-iteration part (disjoint bits)
        const u64 gl = dt | deposit((unsigned)(lu - (int)threadIdx.x) << lw, pass.pos, m, pass.low);
        const u64 u = gl >> lw;
        Amps<R> out;
        if (pass.accumulate) out = load_unit(lam, u);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const unsigned x = l + j;
            const u64 gx = gl | (u64)j;   // bit 0 is local bit 0 whenever W == 2
            double ar = pass.accumulate ? (double)out.v[j].re : 0., ai = pass.accumulate ? (double)out.v[j].im : 0.;
            for (int g = 0; g < pass.n_groups; ++g) {
                const PauliGroup gr = groups[pass.g_begin + g];
                double cr = 0., ci = 0.;
                for (int k = gr.first; k < gr.first + gr.count; ++k) {
                    const PauliTerm t = terms[k];
                    const double sgn = (__popcll(gx & t.zmask) & 1) ? -1. : 1.;
                    cr += sgn * t.cre;
                    ci += sgn * t.cim;
                }
                const Cx<R> v = tp[x ^ gr.xlocal];
                ar += cr * (double)v.re - ci * (double)v.im;
                ai += cr * (double)v.im + ci * (double)v.re;
            }
            out.v[j] = cx<R>((R)ar, (R)ai);
            if (pass.dot) dot += (double)tp[x].re * (double)out.v[j].re + (double)tp[x].im * (double)out.v[j].im;
        }
        store_unit(lam, u, out);
    }
    if (pass.dot) {
        dot = block_sum_tile<kTileThreads>(dot, scratch);
        if (threadIdx.x == 0) partial[tile] = dot;
    }
}

template <class R>
struct SweepCfg;
// Sweep tiles: register passes (RegBits, RegThreads) and LDS passes for smaller tiles or with
// register tiles off (LdsBits), both holding a phi and a lambda tile. 12-bit complex64 tiles
// (512 threads, 64 KiB of LDS) cut the 24-qubit ansatz's backward from 38 passes to 33 but
// measured no faster: the sweep is bound by its per-gate instruction work, not by passes.
// 128 threads on 11-bit tiles (16 amplitudes per state and thread, 203 VGPRs, two waves per
// SIMD) made that backward slower, 7.4 -> 8.6 ms.
template <>
struct SweepCfg<float> {
    static constexpr int RegBits = 11, RegThreads = 256;   // 8 amplitudes per state and thread
    static constexpr int LdsBits = 11;                     // two 2^11 x 8 B tiles = 32 KiB
    static constexpr int Low = 5;
};
template <>
struct SweepCfg<double> {
    static constexpr int RegBits = 10, RegThreads = 256;   // 4 amplitudes per state and thread
    static constexpr int LdsBits = 10;                     // two 2^10 x 16 B tiles = 32 KiB
    static constexpr int Low = 4;
};

}  // namespace

// Relative register-pass cost of a 2x2 by its form: diagonal or swap (moves, or one factor
// per thread when the bit is not a register bit), real, general.
static int form_cost(const GateMatrix &g) {
    if (is_diagonal(g)) return 1;
    if (g.m[0][0] == 0. && g.m[1][1] == 0. && g.m[0][1] == 1. && g.m[1][0] == 1.) return 1;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
            if (g.m[r][c].imag() != 0.) return 4;
    return 2;
}

// Runs of uncontrolled gates on one lane, as one gate each: a gate without controls joins the
// open run of its lane (product in FP64, rounded once where the pass takes it) when the
// product's form costs no more than the two apart (a real and a diagonal gate make a general
// one, so they stay apart), and a gate with controls closes the runs of every lane it touches
// before it is placed. A run's later gates move ahead of gates on other lanes only, which
// commute with them.
static std::vector<QueuedGate> merge_lane_runs(const std::vector<QueuedGate> &gates, int n_lanes) {
    std::vector<QueuedGate> out;
    out.reserve(gates.size());
    std::vector<int> open(n_lanes, -1);   // index in out of the lane's open run
    for (const QueuedGate &g : gates) {
        if (g.ctrl.empty()) {
            const int k = open[g.target];
            if (k >= 0) {
                const GateMatrix &m = out[k].m;
                GateMatrix p;
                for (int r = 0; r < 2; ++r)
                    for (int c = 0; c < 2; ++c) p.m[r][c] = g.m.m[r][0] * m.m[0][c] + g.m.m[r][1] * m.m[1][c];
                if (form_cost(p) <= form_cost(m) + form_cost(g.m)) {
                    out[k].m = p;
                    continue;
                }
            }
            open[g.target] = (int)out.size();
            out.push_back(g);
            continue;
        }
        open[g.target] = -1;
        for (int c : g.ctrl) open[c] = -1;
        out.push_back(g);
    }
    return out;
}

template <class R>
void apply_gate_sequence(Cx<R> *a, int n_lanes, const std::vector<QueuedGate> &gates_in, int max_bits,
                         int low_bits) {
    typedef TileCfg<R> C;
    // wide tiles by default from kWideLanes lanes up, on any state when asked for by width
    const bool wide = use_register_tiles() && (n_lanes >= kWideLanes || max_bits > C::MaxBits);
    const int cap = wide ? WideCfg<R>::Bits : C::MaxBits;
    const int mb = std::min(max_bits > 0 ? std::min(max_bits, cap) : cap, n_lanes);
    const int low = std::min(low_bits >= 0 ? low_bits : C::Low, mb);
    QG_REQUIRE(Unit<R>::W == 1 || low >= 1, "complex64 tiles need bit 0");
    const std::vector<QueuedGate> gates = merge_lane_runs(gates_in, n_lanes);
    // Each pass scans the gates not yet applied in program order and takes every gate that
    // commutes with all the gates it left out before it and whose target is a tile bit (or
    // can become one, or the gate is diagonal). Two gates commute when on every lane they
    // share both act diagonally there: a control, or the target of a diagonal gate. The
    // gates a pass takes keep their program order; the ones it leaves keep theirs for the
    // next passes. (Closing a pass at the first gate that does not fit, a 2000-gate random
    // circuit on 30 qubits took 120 passes of about 14 ops.) The scan stops after kLookahead
    // gates left out, or once every lane is held by one.
    constexpr int kLookahead = 256;
    const int n_gates = (int)gates.size();
    std::vector<char> done(n_gates, 0);
    std::vector<unsigned char> held(n_lanes);   // 0 free, 1 used diagonally, 2 otherwise
    PassPlan plan;
    for (int first = 0; first < n_gates;) {
        plan.bits.clear();
        plan.gates.clear();
        for (int b = 0; b < low; ++b) plan.bits.push_back(b);
        std::fill(held.begin(), held.end(), (unsigned char)0);
        int left = 0, held_all = 0;
        for (int i = first; i < n_gates && (int)plan.gates.size() < C::MaxOps; ++i) {
            if (done[i]) continue;
            const QueuedGate &g = gates[i];
            const bool dg = is_diagonal(g.m);
            bool conflict = dg ? held[g.target] == 2 : held[g.target] != 0;
            for (int c : g.ctrl) conflict = conflict || held[c] == 2;
            bool take = false;
            if (!conflict) {
                if (dg || plan.has(g.target)) {
                    take = true;
                } else if ((int)plan.bits.size() < mb) {
                    plan.bits.push_back(g.target);
                    take = true;
                }
            }
            if (take) {
                plan.gates.push_back(&g);
                done[i] = 1;
                continue;
            }
            auto hold = [&](int lane, unsigned char v) {
                if (held[lane] < 2 && v == 2) ++held_all;
                held[lane] = std::max(held[lane], v);
            };
            hold(g.target, dg ? 1 : 2);
            for (int c : g.ctrl) hold(c, 1);
            if (++left >= kLookahead || held_all == n_lanes) break;
        }
        if (plan.gates.empty()) {   // no tile bit left for the first gate (tile bits == low bits)
            plan.gates.push_back(&gates[first]);   // a lone gate takes the streaming kernels
            done[first] = 1;
        }
        launch_pass<R>(a, n_lanes, plan, low, mb);
        while (first < n_gates && done[first]) ++first;
    }
}

template void apply_gate_sequence<float>(Cx<float> *, int, const std::vector<QueuedGate> &, int, int);
template void apply_gate_sequence<double>(Cx<double> *, int, const std::vector<QueuedGate> &, int, int);

template <class R>
void adjoint_sweep(Cx<R> *phi, Cx<R> *lam, int n_lanes, const std::vector<SweepGate> &gates, double *d_terms,
                   int max_bits, int low_bits) {
    typedef SweepCfg<R> C;
    const bool regs = use_register_tiles();
    const int cap = regs ? C::RegBits : C::LdsBits;
    const int mb = std::min(max_bits > 0 ? std::min(max_bits, cap) : cap, n_lanes);
    const int low = std::min(low_bits >= 0 ? low_bits : C::Low, mb);
    // plan the passes as apply_gate_sequence does (each pass takes every gate not yet swept
    // that commutes with the gates it left out before it and fits the tile), with at most
    // kMaxSweepSlots rotations a pass. A gate counts as diagonal only if it is for every
    // parameter value (a fixed gate, or generator Z), so a reordered rotation has the same
    // derivative. `order` lists the gates pass by pass; a pass is a range [begin, end) of it.
    struct Plan {
        std::vector<int> bits;
        int begin, end, slots;
    };
    constexpr int kLookahead = 256;
    const int n_gates = (int)gates.size();
    std::vector<int> order;
    order.reserve(n_gates);
    std::vector<char> done(n_gates, 0);
    std::vector<unsigned char> held(n_lanes);   // 0 free, 1 used diagonally, 2 otherwise
    std::vector<Plan> plans;
    for (int first = 0; first < n_gates;) {
        Plan cur;
        for (int b = 0; b < low; ++b) cur.bits.push_back(b);
        cur.begin = (int)order.size();
        cur.slots = 0;
        auto has = [&](int b) { return std::find(cur.bits.begin(), cur.bits.end(), b) != cur.bits.end(); };
        std::fill(held.begin(), held.end(), (unsigned char)0);
        int left = 0, held_all = 0;
        for (int i = first; i < n_gates; ++i) {
            if (done[i]) continue;
            const SweepGate &g = gates[i];
            const bool diag = is_diagonal(g.udag) && (g.pauli == 0 || g.pauli == 3);
            bool conflict = diag ? held[g.target] == 2 : held[g.target] != 0;
            for (int c : g.ctrl) conflict = conflict || held[c] == 2;
            const bool slot_ok = g.pauli == 0 || cur.slots < kMaxSweepSlots;
            if (!conflict && slot_ok && (diag || has(g.target) || (int)cur.bits.size() < mb)) {
                if (!(diag || has(g.target))) cur.bits.push_back(g.target);
                if (g.pauli != 0) ++cur.slots;
                order.push_back(i);
                done[i] = 1;
                continue;
            }
            auto hold = [&](int lane, unsigned char v) {
                if (held[lane] < 2 && v == 2) ++held_all;
                held[lane] = std::max(held[lane], v);
            };
            hold(g.target, diag ? 1 : 2);
            for (int c : g.ctrl) hold(c, 1);
            if (++left >= kLookahead || held_all == n_lanes) break;
        }
        if ((int)order.size() == cur.begin) {   // no tile bit left (tile bits == low bits): as before,
            const SweepGate &g = gates[first];     // the first gate widens the tile by its target
            cur.bits.push_back(g.target);
            if (g.pauli != 0) ++cur.slots;
            order.push_back(first);
            done[first] = 1;
        }
        cur.end = (int)order.size();
        plans.push_back(cur);
        while (first < n_gates && done[first]) ++first;
    }

    constexpr int A = (1 << C::RegBits) / C::RegThreads;
    // device tables: ops (pass by pass), phases of register passes, slot -> gate index
    std::vector<SweepOp<R>> hops(std::max<size_t>(1, gates.size()));
    std::vector<SweepPhase> hphases;
    std::vector<int> slot_gate;
    std::vector<SweepPass> passes;
    std::vector<bool> pass_regs;
    int max_slots = 0;
    for (const Plan &pl : plans) {
        std::vector<int> bits(pl.bits);
        for (int b = 0; (int)bits.size() < mb && b < n_lanes; ++b)
            if (std::find(bits.begin(), bits.end(), b) == bits.end()) bits.push_back(b);
        std::sort(bits.begin(), bits.end());
        SweepPass ps;
        ps.m = (int)bits.size();
        ps.low = low;
        ps.op_begin = pl.begin;
        ps.n_ops = pl.end - pl.begin;
        ps.phase_begin = (int)hphases.size();
        ps.n_phases = 0;
        for (int j = 0; j < 16; ++j) ps.pos[j] = j < ps.m ? (signed char)bits[j] : (signed char)0;
        int local[64];
        for (int b = 0; b < 64; ++b) local[b] = -1;
        for (int j = 0; j < ps.m; ++j) local[bits[j]] = j;
        if (regs && ps.m == C::RegBits)   // diagonal gates where none of their lanes is a register bit
            place_diagonals(
                order, pl.begin, pl.end, local, ps.m, Log2<A>::v, Log2<C::RegThreads / 64>::v,
                [&](int gi) { return is_diagonal(gates[gi].udag) && (gates[gi].pauli == 0 || gates[gi].pauli == 3); },
                [&](int gi) { return gates[gi].target; },
                [&](int gi) {
                    std::vector<int> l(gates[gi].ctrl);
                    l.push_back(gates[gi].target);
                    return l;
                });
        int slots = 0;
        for (int i = pl.begin; i < pl.end; ++i) {
            const SweepGate &g = gates[order[i]];
            SweepOp<R> &op = hops[i];
            op.udag = round_matrix<R>(g.udag);
            GateMatrix gen;
            set_pauli(gen, g.pauli);
            op.gen = round_matrix<R>(gen);
            op.gctrl = 0;
            op.lctrl = 0;
            for (int c : g.ctrl) {
                if (local[c] >= 0)
                    op.lctrl |= 1u << local[c];
                else
                    op.gctrl |= u64(1) << c;
            }
            const bool diag = is_diagonal(g.udag) && (g.pauli == 0 || g.pauli == 3);
            if (local[g.target] >= 0) {
                op.target = (short)local[g.target];
                op.kind = (short)(diag ? kOpDiagLocal : kOpPair);
            } else {
                QG_REQUIRE(diag, "adjoint sweep: non-diagonal gate outside the tile");
                op.target = (short)g.target;
                op.kind = (short)kOpDiagOuter;
            }
            op.form = sweep_form(g.udag, g.pauli, op.kind != kOpPair);
            op.slot = g.pauli != 0 ? slots++ : -1;
            if (g.pauli != 0) slot_gate.push_back(order[i]);
        }
        ps.n_slots = slots;
        const bool use_regs = regs && ps.m == C::RegBits;
        QG_REQUIRE(use_regs || ps.m <= C::LdsBits, "adjoint sweep: tile of %d bits", ps.m);
        if (use_regs) {
            std::vector<int> pair(ps.n_ops);
            for (int k = 0; k < ps.n_ops; ++k)
                pair[k] = hops[pl.begin + k].kind == kOpPair ? hops[pl.begin + k].target : -1;
            const std::vector<PhaseLayout> L = plan_phase_layouts(pair, ps.m, Log2<A>::v, Log2<C::RegThreads / 64>::v);
            for (const PhaseLayout &P : L) {
                for (int k = P.begin; k < P.end; ++k) {
                    SweepOp<R> &op = hops[pl.begin + k];
                    const RegSplit r = split_for_phase(P, op.kind, op.target, op.lctrl);
                    op.rkind = r.kind;
                    op.rtarget = r.target;
                    op.creg = r.creg;
                    op.cthr = r.cthr;
                }
                hphases.push_back(phase_bits(P));
            }
            ps.n_phases = (int)L.size();
            QG_REQUIRE(ps.n_phases > 0, "adjoint sweep: register pass without phases");
        }
        pass_regs.push_back(use_regs);
        max_slots = std::max(max_slots, slots);
        passes.push_back(ps);
    }
    if (passes.empty()) return;
    if (hphases.empty()) hphases.push_back(SweepPhase());
    if (slot_gate.empty()) slot_gate.push_back(0);

    hipStream_t s = stream();
    // Every pass has the same tile count (tiles are filled up to mb bits). The partial rows of
    // consecutive passes accumulate in one buffer (up to kSweepPartialBytes) and are summed
    // by one launch when it is full or the sweep ends, not by a small launch after each pass.
    const u64 n_tiles = u64(1) << (n_lanes - passes[0].m);
    for (const SweepPass &ps : passes) QG_REQUIRE(ps.m == passes[0].m, "adjoint sweep: passes of unequal tiles");
    int total_slots = 0;
    for (const SweepPass &ps : passes) total_slots += ps.n_slots;
    const size_t row_bytes = sizeof(double) * (size_t)n_tiles;
    const int cap_slots = (int)std::max<size_t>(
        (size_t)max_slots, std::min<size_t>((size_t)total_slots, kSweepPartialBytes / row_bytes));
    const size_t part_bytes = std::max<size_t>(8, row_bytes * (size_t)cap_slots);
    const size_t ops_bytes = sizeof(SweepOp<R>) * hops.size();
    const size_t ph_bytes = sizeof(SweepPhase) * hphases.size();
    const size_t sg_bytes = sizeof(int) * slot_gate.size();
    const size_t o_ph = (ops_bytes + 255) & ~size_t(255), o_sg = (o_ph + ph_bytes + 255) & ~size_t(255);
    const size_t o_part = (o_sg + sg_bytes + 255) & ~size_t(255);
    const size_t total = o_part + part_bytes;
    char *buf = static_cast<char *>(dev_alloc(total));
    SweepOp<R> *d_ops = reinterpret_cast<SweepOp<R> *>(buf);
    SweepPhase *d_ph = reinterpret_cast<SweepPhase *>(buf + o_ph);
    int *d_sg = reinterpret_cast<int *>(buf + o_sg);
    double *d_part = reinterpret_cast<double *>(buf + o_part);
    hipError_t e = hipMemcpyAsync(d_ops, hops.data(), ops_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ph, hphases.data(), ph_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sg, slot_gate.data(), sg_bytes, hipMemcpyHostToDevice, s);
    int pending = 0, pending_begin = 0;   // rows in the buffer, and the slot_gate index of the first
    auto sum_rows = [&]() {                // terms of the pending rows land at d_terms[gate index]
        if (pending > 0)
            hipLaunchKernelGGL(k_slot_rows, (unsigned)pending, kRowThreads, 0, s, d_part, n_tiles,
                               d_sg + pending_begin, d_terms);
        pending_begin += pending;
        pending = 0;
    };
    for (size_t pi = 0; pi < passes.size() && e == hipSuccess; ++pi) {
        const SweepPass &ps = passes[pi];
        if (pending + ps.n_slots > cap_slots) sum_rows();
        double *part = d_part + (u64)pending * n_tiles;
        const int threads = pass_regs[pi] ? C::RegThreads : kTileThreads;
        const u64 piece = launch_piece_items() / (u64)threads;
        for (u64 first = 0; first < n_tiles; first += piece) {
            const u64 count = std::min(piece, n_tiles - first);
            if (pass_regs[pi])
                hipLaunchKernelGGL((k_tile_sweep_reg<R, C::RegBits, C::RegThreads, false>), (unsigned)count,
                                   C::RegThreads, 0, s, phi, lam, ps, d_ops, d_ph, part, n_tiles, first,
                                   (const Mat2<R> *)nullptr, 0, 0);
            else
                hipLaunchKernelGGL((k_tile_sweep<R, C::LdsBits>), (unsigned)count, kTileThreads, 0, s, phi, lam, ps,
                                   d_ops, part, n_tiles, first);
        }
        pending += ps.n_slots;
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        sum_rows();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // the tables are host temporaries
    dev_free(buf, total);
    QG_HIP(e);
}

template void adjoint_sweep<float>(Cx<float> *, Cx<float> *, int, const std::vector<SweepGate> &, double *, int,
                                   int);
template void adjoint_sweep<double>(Cx<double> *, Cx<double> *, int, const std::vector<SweepGate> &, double *,
                                    int, int);

template <class R>
double pauli_expectation(const Cx<R> *psi, Cx<R> *lam, int n_lanes, const std::vector<PauliGroup> &groups_in,
                         const std::vector<PauliTerm> &terms) {
    typedef PauliCfg<R> C;
    const int mb = std::min(C::MaxBits, n_lanes);
    const int low = std::min(C::Low, mb);
    // passes: consecutive groups whose X bits fit one tile together
    std::vector<PauliGroup> groups(groups_in);
    struct Plan {
        std::vector<int> bits;
        int begin, end;
    };
    std::vector<Plan> plans;
    Plan cur;
    auto open = [&](int at) {
        cur.bits.clear();
        for (int b = 0; b < low; ++b) cur.bits.push_back(b);
        cur.begin = cur.end = at;
    };
    auto missing = [&](u64 xm) {
        std::vector<int> need;
        for (int b = 0; b < n_lanes; ++b)
            if (((xm >> b) & 1) && std::find(cur.bits.begin(), cur.bits.end(), b) == cur.bits.end())
                need.push_back(b);
        return need;
    };
    const u64 always = Unit<R>::W == 2 ? 1 : 0;   // complex64 tiles always hold bit 0
    open(0);
    for (int i = 0; i < (int)groups.size(); ++i) {
        QG_REQUIRE(__builtin_popcountll(groups[i].xmask | always) <= mb,
                   "Pauli term with %d X/Y factors does not fit a %d-bit tile", __builtin_popcountll(groups[i].xmask), mb);
        std::vector<int> need = missing(groups[i].xmask);
        if ((int)(cur.bits.size() + need.size()) > mb) {
            if (cur.end > cur.begin) plans.push_back(cur);
            open(i);
            need = missing(groups[i].xmask);
            if ((int)(cur.bits.size() + need.size()) > mb) {   // a wide term: give up the low rows
                cur.bits.clear();
                need = missing(groups[i].xmask);
            }
        }
        for (int b : need) cur.bits.push_back(b);
        cur.end = i + 1;
    }
    plans.push_back(cur);

    std::vector<PauliPass> passes;
    for (const Plan &pl : plans) {
        std::vector<int> bits(pl.bits);
        // complex64 tiles move 2-amplitude units: bit 0 must be a tile bit
        if (Unit<R>::W == 2 && std::find(bits.begin(), bits.end(), 0) == bits.end()) bits.push_back(0);
        for (int b = 0; (int)bits.size() < mb && b < n_lanes; ++b)
            if (std::find(bits.begin(), bits.end(), b) == bits.end()) bits.push_back(b);
        QG_REQUIRE((int)bits.size() <= mb, "Pauli term does not fit a tile");
        std::sort(bits.begin(), bits.end());
        PauliPass ps;
        ps.m = (int)bits.size();
        int lowrun = 0;
        while (lowrun < ps.m && bits[lowrun] == lowrun) ++lowrun;
        ps.low = lowrun;
        ps.g_begin = pl.begin;
        ps.n_groups = pl.end - pl.begin;
        ps.accumulate = passes.empty() ? 0 : 1;
        ps.dot = 0;
        for (int j = 0; j < 16; ++j) ps.pos[j] = j < ps.m ? (signed char)bits[j] : (signed char)0;
        int local[64];
        for (int b = 0; b < 64; ++b) local[b] = -1;
        for (int j = 0; j < ps.m; ++j) local[bits[j]] = j;
        for (int i = pl.begin; i < pl.end; ++i) {
            unsigned xl = 0;
            for (int b = 0; b < n_lanes; ++b)
                if ((groups[i].xmask >> b) & 1) {
                    QG_REQUIRE(local[b] >= 0, "Pauli group outside its tile");
                    xl |= 1u << local[b];
                }
            groups[i].xlocal = xl;
        }
        passes.push_back(ps);
    }
    passes.back().dot = 1;

    hipStream_t s = stream();
    const size_t gbytes = sizeof(PauliGroup) * std::max<size_t>(1, groups.size());
    const size_t tbytes = sizeof(PauliTerm) * std::max<size_t>(1, terms.size());
    const u64 n_tiles_last = u64(1) << (n_lanes - passes.back().m);
    const size_t pbytes = sizeof(double) * (size_t)(n_tiles_last + 1);
    char *buf = static_cast<char *>(dev_alloc(gbytes + tbytes + pbytes));
    PauliGroup *d_groups = reinterpret_cast<PauliGroup *>(buf);
    PauliTerm *d_terms = reinterpret_cast<PauliTerm *>(buf + gbytes);
    double *d_part = reinterpret_cast<double *>(buf + gbytes + tbytes);
    Runtime &rt = runtime();
    hipError_t e = hipMemcpyAsync(d_groups, groups.data(), sizeof(PauliGroup) * groups.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_terms, terms.data(), sizeof(PauliTerm) * terms.size(), hipMemcpyHostToDevice, s);
    for (size_t pi = 0; pi < passes.size() && e == hipSuccess; ++pi) {
        const PauliPass &ps = passes[pi];
        const u64 n_tiles = u64(1) << (n_lanes - ps.m);
        const u64 piece = launch_piece_items() / kTileThreads;
        for (u64 first = 0; first < n_tiles; first += piece) {
            const u64 count = std::min(piece, n_tiles - first);
            hipLaunchKernelGGL((k_tile_pauli<R, C::MaxBits>), (unsigned)count, kTileThreads, 0, s, psi, lam, ps,
                               d_groups, d_terms, d_part, first);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_slots, 1, kRowThreads, 0, s, d_part, n_tiles_last, d_part + n_tiles_last);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(rt.h_pinned, d_part + n_tiles_last, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // group and term tables are host temporaries
    dev_free(buf, gbytes + tbytes + pbytes);
    QG_HIP(e);
    return rt.h_pinned[0];
}

template double pauli_expectation<float>(const Cx<float> *, Cx<float> *, int, const std::vector<PauliGroup> &,
                                         const std::vector<PauliTerm> &);
template double pauli_expectation<double>(const Cx<double> *, Cx<double> *, int, const std::vector<PauliGroup> &,
                                          const std::vector<PauliTerm> &);

}  // namespace qg
