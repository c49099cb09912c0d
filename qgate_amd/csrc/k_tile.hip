// Tile passes: many gates per trip of the state vector through HBM.
//
// A pass picks a set T of m state bits (m <= 12 for complex64, 11 for complex128) that always
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

namespace qg {

namespace {

constexpr int kTileThreads = 256;

template <class R>
struct TileCfg;
template <>
struct TileCfg<float> {
    static constexpr int MaxBits = 12;   // 2^12 x 8 B = 32 KiB of LDS
    static constexpr int MaxOps = 40;
    static constexpr int Low = 5;        // 32 contiguous amplitudes = 256 B per row
};
template <>
struct TileCfg<double> {
    static constexpr int MaxBits = 11;   // 2^11 x 16 B = 32 KiB
    static constexpr int MaxOps = 24;
    static constexpr int Low = 4;        // 16 x 16 B = 256 B per row
};

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

bool is_diagonal(const GateMatrix &g) { return g.m[0][1] == 0. && g.m[1][0] == 0.; }

// ---- register-resident tile passes ------------------------------------------------------
//
// The full-size pass (m == MaxBits) keeps the tile in registers: each of the T threads holds
// A = 2^m / T amplitudes. A pass is split into phases; in each phase RB = log2(A) tile bits
// are register bits (the targets of the phase's non-diagonal gates), 6 are lane bits and the
// rest wave bits, so the phase's pair updates are register-to-register with no LDS traffic
// or barriers. Between phases the tile goes through LDS once (write, barrier, read in the
// next phase's layout). A pass of g gates costs about g / RB LDS round trips instead of g.

constexpr int kMaxPhases = 40;

// Bit layout of one register phase, packed in dwords so that a phase read at a runtime
// index takes scalar loads (a byte array read at a runtime index would be a vector load,
// whose wait also waits for HBM traffic in flight). Byte k of `bits` is the local bit of
// register bit k (k < 5), of lane bit k - 5 (5 <= k < 11) or of wave bit k - 11 (k < 14).
struct PhaseBits {
    unsigned bits[4];
    int op_begin, n_ops;
    __host__ __device__ int byte(int k) const { return (int)((bits[k >> 2] >> ((k & 3) * 8)) & 0xffu); }
    __host__ __device__ int rb(int k) const { return byte(k); }
    __host__ __device__ int lb(int k) const { return byte(5 + k); }
    __host__ __device__ int wb(int k) const { return byte(11 + k); }
    void set(int k, int v) {
        const int sh = (k & 3) * 8;
        bits[k >> 2] = (bits[k >> 2] & ~(0xffu << sh)) | ((unsigned)v << sh);
    }
};
typedef PhaseBits RegPhase;

// Local-index bits of register index j in phase P (uniform: a scalar value).
template <int RB>
__device__ __forceinline__ unsigned reg_bits(const PhaseBits &P, int j) {
    unsigned l = 0;
#pragma unroll
    for (int k = 0; k < RB; ++k) l |= (((unsigned)j >> k) & 1u) << P.rb(k);
    return l;
}

// Register-pass op kinds. Pair updates on a register position (general, swap, real); a
// diagonal whose target is a register position (every amplitude, or only the |1> half when
// m00 == 1); a diagonal whose target is a lane/wave bit or a bit outside the tile, which is
// one factor per thread.
enum : int {
    kRegPair = 0, kRegPairSwap = 1, kRegPairReal = 2,
    kRegDiag = 3, kRegDiagPhase = 4, kRegDiagThread = 5, kRegDiagOuter = 6
};

// A gate in a register phase. Its controls are split by where their bits live: outside the
// tile (one test per tile), on lane/wave bits (one test per thread) and on register bits
// (register positions: tests on the compile-time register indices, made by the scalar unit).
template <class R>
struct RegOp {
    Mat2<R> m;
    u64 gctrl;         // controls outside the tile (state bits)
    unsigned cthr;     // controls on lane/wave bits (local bits)
    int creg;          // controls on register bits (mask of register positions)
    int target;        // register position (pairs, kRegDiag*), local bit (kRegDiagThread), state bit (kRegDiagOuter)
    int kind;
};

template <class R>
struct RegProgram {
    int n_ops, n_phases, low;
    alignas(4) signed char pos[kTilePosMax];
    RegPhase phases[kMaxPhases];
    RegOp<R> ops[TileCfg<R>::MaxOps];
};

template <int A>
struct Log2 {
    static constexpr int v = A <= 1 ? 0 : 1 + Log2<A / 2>::v;
};
template <>
struct Log2<1> {
    static constexpr int v = 0;
};

// f(std::integral_constant<int, J>) for J = B .. E-1: register arrays are only ever indexed
// with compile-time constants, so they stay in registers.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>());
        static_for<B + 1, E>(f);
    }
}

// Ends a register-pair update on register bit P. The markers differ per P, so the optimiser
// cannot merge the updates of different bits (dispatched on a runtime bit) into one block
// that picks the register elements through pointers, which would move those elements to
// private (scratch) memory.
template <int P>
__device__ __forceinline__ void pair_marker() {
    asm volatile("; register pair %0" ::"n"(P));
}

// Im(conj(a) b)
template <class R>
__device__ __forceinline__ R im_conj_mul(const Cx<R> &a, const Cx<R> &b) {
    return a.re * b.im - a.im * b.re;
}

// c ? a : b, component by component: a select of whole structs can be lowered to a select
// of their addresses, which would keep register arrays out of registers.
template <class R>
__device__ __forceinline__ Cx<R> csel(bool c, const Cx<R> &a, const Cx<R> &b) {
    return cx<R>(c ? a.re : b.re, c ? a.im : b.im);
}

// (y0, y1) = M (x0, x1) for a matrix of form F: 0 general, 1 swap, 2 real, 3 real diagonal
// and imaginary off-diagonal (RX-like: m00, m11 real; m01, m10 imaginary).
template <class R, int F>
__device__ __forceinline__ void apply_form(const Mat2<R> &mm, const Cx<R> &x0, const Cx<R> &x1, Cx<R> &y0, Cx<R> &y1) {
    if constexpr (F == 1) {
        y0 = x1;
        y1 = x0;
    } else if constexpr (F == 2) {
        const R a = mm.m00.re, b = mm.m01.re, c = mm.m10.re, d = mm.m11.re;
        y0 = cx<R>(a * x0.re + b * x1.re, a * x0.im + b * x1.im);
        y1 = cx<R>(c * x0.re + d * x1.re, c * x0.im + d * x1.im);
    } else if constexpr (F == 3) {
        const R a = mm.m00.re, b = mm.m01.im, c = mm.m10.im, d = mm.m11.re;
        y0 = cx<R>(a * x0.re - b * x1.im, a * x0.im + b * x1.re);
        y1 = cx<R>(d * x1.re - c * x0.im, d * x1.im + c * x0.re);
    } else {
        y0 = cadd(cmul(mm.m00, x0), cmul(mm.m01, x1));
        y1 = cadd(cmul(mm.m10, x0), cmul(mm.m11, x1));
    }
}

// Pair update on register bit P (form F, see apply_form) of the register indices j whose
// register-position controls creg are set.
template <class R, int A, int P, int F>
__device__ __forceinline__ void reg_pair(Cx<R> (&v)[A], const Mat2<R> &mm, int creg) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if constexpr (((j >> P) & 1) == 0) {
            constexpr int j1 = j + (1 << P);
            if ((j & creg) == creg) {
                const Cx<R> x0 = v[j], x1 = v[j1];
                apply_form<R, F>(mm, x0, x1, v[j], v[j1]);
            }
            __builtin_amdgcn_sched_barrier(0);   // one pair at a time: bounded register use
        }
    });
    pair_marker<P * 4 + F>();
}

// Pair update with controls: only register indices with (j & creg) == creg (scalar tests)
// in threads with ok set change; written as selects, so no masked region needs copies of
// the registers.
template <class R, int A, int P, int F>
__device__ __forceinline__ void reg_pair_sel(Cx<R> (&v)[A], const Mat2<R> &mm, int creg, bool ok) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if constexpr (((j >> P) & 1) == 0) {
            constexpr int j1 = j + (1 << P);
            const bool c = ok && (j & creg) == creg;
            const Cx<R> x0 = v[j], x1 = v[j1];
            Cx<R> y0, y1;
            apply_form<R, F>(mm, x0, x1, y0, y1);
            v[j] = csel(c, y0, x0);
            v[j1] = csel(c, y1, x1);
            __builtin_amdgcn_sched_barrier(0);
        }
    });
    pair_marker<32 + P * 4 + F>();
}

// Dispatch on the register position p (< RB, guaranteed by the planner) as one switch with
// an unreachable default: every path ends at one join. A chain of one-armed ifs would give
// each not-taken arm a copy of the whole register array.
template <class R, int A, int F>
__device__ __forceinline__ void reg_pair_sel_at(Cx<R> (&v)[A], const Mat2<R> &mm, int creg, bool ok, int p) {
    constexpr int RB = Log2<A>::v;
    switch (p) {
        case 0:
            if constexpr (RB > 0) {
                reg_pair_sel<R, A, 0, F>(v, mm, creg, ok);
                break;
            }
            [[fallthrough]];
        case 1:
            if constexpr (RB > 1) {
                reg_pair_sel<R, A, 1, F>(v, mm, creg, ok);
                break;
            }
            [[fallthrough]];
        case 2:
            if constexpr (RB > 2) {
                reg_pair_sel<R, A, 2, F>(v, mm, creg, ok);
                break;
            }
            [[fallthrough]];
        case 3:
            if constexpr (RB > 3) {
                reg_pair_sel<R, A, 3, F>(v, mm, creg, ok);
                break;
            }
            [[fallthrough]];
        default: __builtin_unreachable();
    }
}

template <class R, int A, int F>
__device__ __forceinline__ void reg_pair_if(Cx<R> (&v)[A], const Mat2<R> &mm, int p) {
    constexpr int RB = Log2<A>::v;
    if constexpr (RB > 0) {
        if (p == 0) reg_pair<R, A, 0, F>(v, mm, 0);
    }
    if constexpr (RB > 1) {
        if (p == 1) reg_pair<R, A, 1, F>(v, mm, 0);
    }
    if constexpr (RB > 2) {
        if (p == 2) reg_pair<R, A, 2, F>(v, mm, 0);
    }
    if constexpr (RB > 3) {
        if (p == 3) reg_pair<R, A, 3, F>(v, mm, 0);
    }
}

template <class R, int A, int F>
__device__ __forceinline__ void reg_pair_at(Cx<R> (&v)[A], const Mat2<R> &mm, int p) {
    constexpr int RB = Log2<A>::v;
    switch (p) {
        case 0:
            if constexpr (RB > 0) {
                reg_pair<R, A, 0, F>(v, mm, 0);
                break;
            }
            [[fallthrough]];
        case 1:
            if constexpr (RB > 1) {
                reg_pair<R, A, 1, F>(v, mm, 0);
                break;
            }
            [[fallthrough]];
        case 2:
            if constexpr (RB > 2) {
                reg_pair<R, A, 2, F>(v, mm, 0);
                break;
            }
            [[fallthrough]];
        case 3:
            if constexpr (RB > 3) {
                reg_pair<R, A, 3, F>(v, mm, 0);
                break;
            }
            [[fallthrough]];
        default: __builtin_unreachable();
    }
}

__host__ __device__ constexpr bool is_pair_kind(int k) { return k == kOpPair || k == kOpPairSwap || k == kOpPairReal; }

template <class R, int M, int T>
__global__ __launch_bounds__(T) void k_tile_reg(Cx<R> *__restrict__ a, const RegProgram<R> prog, u64 first) {
    constexpr int A = (1 << M) / T, RB = Log2<A>::v, WB = Log2<T / 64>::v;
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    __shared__ Cx<R> t[1 << M];
    const u64 base = tile_base(first + blockIdx.x, prog.pos, M);
    tile_to_lds<R, T, (1 << (M - lw)) / T>(t, a, base, prog.pos, M, prog.low);
    __syncthreads();
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Cx<R> v[A];
    for (int ph = 0; ph < prog.n_phases; ++ph) {
        const RegPhase P = prog.phases[ph];
        unsigned off = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) off |= ((lane >> k) & 1u) << P.lb(k);
#pragma unroll
        for (int k = 0; k < WB; ++k) off |= ((wave >> k) & 1u) << P.wb(k);
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            v[j] = t[off | reg_bits<RB>(P, j)];
        });
        // Diagonal gates with no register bit among their bits multiply all of a thread's
        // amplitudes by one factor. They are collected in tf and applied once at the end of
        // the phase: tf commutes with every other gate of the phase, which mixes or scales
        // amplitudes of one thread only.
        Cx<R> tf = cx<R>(R(1), R(0));
        bool pending = false;
        for (int k = 0; k < P.n_ops; ++k) {
            // plain locals: the lambdas below must not capture the kernel argument by address
            const int oi = P.op_begin + k;
            const u64 gctrl = prog.ops[oi].gctrl;
            if ((base & gctrl) != gctrl) continue;   // same for the whole workgroup
            const Mat2<R> mm = prog.ops[oi].m;
            const unsigned cthr = prog.ops[oi].cthr;
            const int creg = prog.ops[oi].creg, tgt = prog.ops[oi].target, kind = prog.ops[oi].kind;
            // no controls on lane, wave or register bits: straight-line updates, no tests
            const bool plain = cthr == 0 && creg == 0;
            const bool ok = (off & cthr) == cthr;
            if (kind <= kRegPairReal) {
                // one-armed ifs here (not one switch): they cost register copies per gate but
                // keep the kernel near 110 VGPRs, i.e. 4 workgroups per CU, which measured faster
                if (plain) {
                    if (kind == kRegPair) reg_pair_if<R, A, 0>(v, mm, tgt);
                    if (kind == kRegPairSwap) reg_pair_if<R, A, 1>(v, mm, tgt);
                    if (kind == kRegPairReal) reg_pair_if<R, A, 2>(v, mm, tgt);
                } else {
                    if (kind == kRegPair) reg_pair_sel_at<R, A, 0>(v, mm, creg, ok, tgt);
                    if (kind == kRegPairSwap) reg_pair_sel_at<R, A, 1>(v, mm, creg, ok, tgt);
                    if (kind == kRegPairReal) reg_pair_sel_at<R, A, 2>(v, mm, creg, ok, tgt);
                }
            } else if (kind <= kRegDiagPhase) {
                // target on register position tgt: the factor of index j is a scalar choice
                if (plain) {
                    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        v[j] = cmul(csel(((j >> tgt) & 1) != 0, mm.m11, mm.m00), v[j]);
                        __builtin_amdgcn_sched_barrier(0);
                    });
                } else {
                    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        const Cx<R> y = cmul(csel(((j >> tgt) & 1) != 0, mm.m11, mm.m00), v[j]);
                        v[j] = csel(ok && (j & creg) == creg, y, v[j]);
                        __builtin_amdgcn_sched_barrier(0);
                    });
                }
            } else {
                const u64 src = kind == kRegDiagThread ? (u64)off : base;
                Cx<R> f = csel(((src >> tgt) & 1) != 0, mm.m11, mm.m00);
                f = csel(cthr == 0 || ok, f, cx<R>(R(1), R(0)));
                if (creg == 0) {
                    tf = cmul(tf, f);
                    pending = true;
                } else {
                    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        if ((j & creg) == creg) v[j] = cmul(f, v[j]);
                        __builtin_amdgcn_sched_barrier(0);
                    });
                }
            }
        }
        if (pending) {
            static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                v[j] = cmul(tf, v[j]);
            });
        }
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            t[off | reg_bits<RB>(P, j)] = v[j];
        });
        __syncthreads();
    }
    lds_to_tile<R, T, (1 << (M - lw)) / T>(a, t, base, prog.pos, M, prog.low);
}

// Phases of a pass: consecutive gates whose pair targets (pair[k] >= 0: local bit of a
// non-diagonal gate) fit rb_count register bits. Each phase also fixes its lane bits (the
// lowest 6 other tile bits) and wave bits (the rest).
struct PhaseLayout {
    std::vector<int> regs, lanes, waves;
    int begin, end;
};

std::vector<PhaseLayout> plan_phase_layouts(const std::vector<int> &pair, int m, int rb_count, int wb_count) {
    std::vector<PhaseLayout> out;
    int k = 0;
    const int n = (int)pair.size();
    while (k < n) {
        PhaseLayout P;
        P.begin = k;
        for (; k < n; ++k) {
            if (pair[k] < 0) continue;
            if (std::find(P.regs.begin(), P.regs.end(), pair[k]) != P.regs.end()) continue;
            if ((int)P.regs.size() == rb_count) break;
            P.regs.push_back(pair[k]);
        }
        P.end = k;
        // pad with the highest free bits, so that the lane bits are the lowest ones: a
        // wavefront then reads and writes LDS at consecutive addresses (no bank conflicts)
        for (int b = m - 1; (int)P.regs.size() < rb_count && b >= 0; --b)
            if (std::find(P.regs.begin(), P.regs.end(), b) == P.regs.end()) P.regs.push_back(b);
        std::sort(P.regs.begin(), P.regs.end());
        for (int b = 0; b < m; ++b)
            if (std::find(P.regs.begin(), P.regs.end(), b) == P.regs.end()) {
                if ((int)P.lanes.size() < 6)
                    P.lanes.push_back(b);
                else
                    P.waves.push_back(b);
            }
        QG_REQUIRE((int)P.waves.size() == wb_count, "tile layout does not match the workgroup");
        out.push_back(P);
    }
    return out;
}

int reg_position(const PhaseLayout &P, int bit) {
    return (int)(std::find(P.regs.begin(), P.regs.end(), bit) - P.regs.begin());
}

// Phase layouts of a register program: ops[k].target holds local bits on entry; pair gates
// get their register position. Returns false if the pass needs more phases than it holds.
// Register-pass encoding of a gate in phase P: kind (tile-pass kind, local bits: kOpPair*,
// kOpDiag*), target (local bit, or state bit for kOpDiagOuter) and local controls lctrl.
struct RegSplit {
    int kind, target, creg;
    unsigned cthr;
};

RegSplit split_for_phase(const PhaseLayout &P, int kind, int target, unsigned lctrl) {
    RegSplit r;
    r.creg = 0;
    r.cthr = 0;
    const int rb = (int)P.regs.size();
    for (int b = 0; b < 32; ++b) {
        if (((lctrl >> b) & 1u) == 0) continue;
        const int p = reg_position(P, b);
        if (p < rb)
            r.creg |= 1 << p;
        else
            r.cthr |= 1u << b;
    }
    if (kind == kOpPair || kind == kOpPairSwap || kind == kOpPairReal) {
        r.kind = kind == kOpPair ? kRegPair : kind == kOpPairSwap ? kRegPairSwap : kRegPairReal;
        r.target = reg_position(P, target);
        QG_REQUIRE(r.target < rb, "register phase: pair target is not a register bit");
    } else if (kind == kOpDiagOuter) {
        r.kind = kRegDiagOuter;
        r.target = target;
    } else {
        const int p = reg_position(P, target);
        if (p < rb) {
            r.kind = kind == kOpDiagPhase ? kRegDiagPhase : kRegDiag;
            r.target = p;
        } else {
            r.kind = kRegDiagThread;
            r.target = target;
        }
    }
    return r;
}

// Phase layouts of a register program built from tile-pass ops (local bits): false if the
// pass needs more phases than a program holds.
template <class R>
bool plan_phases(RegProgram<R> &prog, const TileOp<R> *ops, int m, int rb_count, int wb_count) {
    std::vector<int> pair(prog.n_ops);
    for (int k = 0; k < prog.n_ops; ++k) pair[k] = is_pair_kind(ops[k].kind) ? ops[k].target : -1;
    const std::vector<PhaseLayout> L = plan_phase_layouts(pair, m, rb_count, wb_count);
    if ((int)L.size() > kMaxPhases) return false;
    prog.n_phases = (int)L.size();
    for (int i = 0; i < prog.n_phases; ++i) {
        RegPhase &P = prog.phases[i];
        std::memset(&P, 0, sizeof(P));
        for (int b = 0; b < rb_count; ++b) P.set(b, L[i].regs[b]);
        for (int b = 0; b < 6; ++b) P.set(5 + b, L[i].lanes[b]);
        for (int b = 0; b < wb_count; ++b) P.set(11 + b, L[i].waves[b]);
        P.op_begin = L[i].begin;
        P.n_ops = L[i].end - L[i].begin;
        for (int k = L[i].begin; k < L[i].end; ++k) {
            const RegSplit r = split_for_phase(L[i], ops[k].kind, ops[k].target, ops[k].lctrl);
            RegOp<R> &op = prog.ops[k];
            op.m = ops[k].m;
            op.gctrl = ops[k].gctrl;
            op.cthr = r.cthr;
            op.creg = r.creg;
            op.target = r.target;
            op.kind = r.kind;
        }
    }
    return true;
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
    if (prog.m == C::MaxBits && use_register_tiles()) {
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
        constexpr int A = (1 << C::MaxBits) / kTileThreads;
        if (plan_phases<R>(rp, tops, prog.m, Log2<A>::v, Log2<kTileThreads / 64>::v)) {
            for (u64 first = 0; first < n_tiles; first += piece) {
                const u64 count = std::min(piece, n_tiles - first);
                hipLaunchKernelGGL((k_tile_reg<R, C::MaxBits, kTileThreads>), (unsigned)count, kTileThreads, 0, s, a,
                                   rp, first);
            }
            QG_HIP(hipGetLastError());
            return;
        }
    }
    for (u64 first = 0; first < n_tiles; first += piece) {
        const u64 count = std::min(piece, n_tiles - first);
        hipLaunchKernelGGL((k_tile_pass<R, C::MaxBits>), (unsigned)count, kTileThreads, 0, s, a, prog, first);
    }
    QG_HIP(hipGetLastError());
}

// ---- adjoint sweep ----------------------------------------------------------------------

template <class R>
struct SweepOp {
    Mat2<R> udag;       // applied to phi and lambda
    Mat2<R> gen;        // generator of a rotation (entries 0, +-1, +-i)
    u64 gctrl;
    unsigned lctrl;
    int target;         // local bit (pair, diag-local) or state bit (diag-outer)
    int kind;           // (dwords: a uniform dword field is one scalar load)
    int slot;           // gradient slot within the pass, -1 for fixed gates
    int form;           // register passes: 4 * matrix form + generator (sweep_form)
    int rkind, rtarget; // register passes: kind and target in the phase (RegOp)
    int creg;           // register passes: controls on register bits (register positions)
    unsigned cthr;      // register passes: controls on lane/wave bits (local bits)
};

struct SweepPass {
    int m, low;
    int op_begin, n_ops;
    int phase_begin, n_phases, n_slots;
    alignas(4) signed char pos[kTilePosMax];
};

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

// Register-resident sweep pass (m == SweepBits): phi and lambda tiles in registers, one LDS
// round trip per phase; each rotation's tile term is reduced over the wavefront at once and
// over the workgroup at the end of the pass.
constexpr int kMaxSweepSlots = 96;

typedef PhaseBits SweepPhase;   // op_begin relative to the pass

// One DPP step on a double: lane l receives v of the lane that control CTRL names, in the rows
// enabled by ROWS (other rows receive 0). DPP moves run on the vector ALU, with no LDS
// round trip.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_double(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum of v over the 64 lanes of a wavefront in a fixed order, valid in lane 63: butterflies
// within quads (quad_perm [1,0,3,2], [2,3,0,1]) and rows (row_ror 4, 8) give every lane its
// 16-lane row sum; row_bcast 15 then adds row 0 into row 1 and row 2 into row 3, and
// row_bcast 31 adds rows 0+1 into row 3.
__device__ __forceinline__ double wave_sum_lane63(double v) {
    v += dpp_double<0xb1, 0xf>(v);
    v += dpp_double<0x4e, 0xf>(v);
    v += dpp_double<0x124, 0xf>(v);
    v += dpp_double<0x128, 0xf>(v);
    v += dpp_double<0x142, 0xa>(v);
    v += dpp_double<0x143, 0xc>(v);
    return v;
}

// Register sweep pair update on register bit P: the gradient term Im<lambda|G|phi> over the
// pair for generator G (0 none, 1 X, 2 Y, 3 Z), accumulated in `acc` (state precision; the
// caller widens each thread's sum to FP64), then U^dagger of form F on phi and lambda.
// SEL: controls present; a pair changes (and contributes) only if ok && (j & creg) == creg,
// written as selects so that no masked region needs copies of the registers.
template <class R, int A, int P, int F, int G, bool SEL>
__device__ __forceinline__ void reg_sweep_pair(Cx<R> (&vp)[A], Cx<R> (&vl)[A], const Mat2<R> &ud, int creg, bool ok,
                                               R &acc) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if constexpr (((j >> P) & 1) == 0) {
            constexpr int j1 = j + (1 << P);
            const Cx<R> p0 = vp[j], p1 = vp[j1], l0 = vl[j], l1 = vl[j1];
            R t = R(0);
            if constexpr (G == 1) t = im_conj_mul(l0, p1) + im_conj_mul(l1, p0);
            if constexpr (G == 2) t = (l1.re * p0.re + l1.im * p0.im) - (l0.re * p1.re + l0.im * p1.im);
            if constexpr (G == 3) t = im_conj_mul(l0, p0) - im_conj_mul(l1, p1);
            Cx<R> q0, q1, m0, m1;
            apply_form<R, F>(ud, p0, p1, q0, q1);
            apply_form<R, F>(ud, l0, l1, m0, m1);
            if constexpr (SEL) {
                const bool c = ok && (j & creg) == creg;
                if constexpr (G != 0) acc += c ? t : R(0);
                vp[j] = csel(c, q0, p0);
                vp[j1] = csel(c, q1, p1);
                vl[j] = csel(c, m0, l0);
                vl[j1] = csel(c, m1, l1);
            } else {
                if constexpr (G != 0) acc += t;
                vp[j] = q0;
                vp[j1] = q1;
                vl[j] = m0;
                vl[j1] = m1;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    });
    pair_marker<64 + ((P * 4 + F) * 4 + G) * 2 + (SEL ? 1 : 0)>();
}

template <class R, int A, int F, int G, bool SEL>
__device__ __forceinline__ void reg_sweep_pair_pos(Cx<R> (&vp)[A], Cx<R> (&vl)[A], const Mat2<R> &ud, int creg, bool ok,
                                                   int p, R &acc) {
    constexpr int RB = Log2<A>::v;
    if constexpr (sizeof(R) == 8) {
        // complex128: one-armed ifs keep the kernel's register count (and occupancy) lower
        if constexpr (RB > 0) {
            if (p == 0) reg_sweep_pair<R, A, 0, F, G, SEL>(vp, vl, ud, creg, ok, acc);
        }
        if constexpr (RB > 1) {
            if (p == 1) reg_sweep_pair<R, A, 1, F, G, SEL>(vp, vl, ud, creg, ok, acc);
        }
        if constexpr (RB > 2) {
            if (p == 2) reg_sweep_pair<R, A, 2, F, G, SEL>(vp, vl, ud, creg, ok, acc);
        }
        if constexpr (RB > 3) {
            if (p == 3) reg_sweep_pair<R, A, 3, F, G, SEL>(vp, vl, ud, creg, ok, acc);
        }
        return;
    }
    switch (p) {
        case 0:
            if constexpr (RB > 0) {
                reg_sweep_pair<R, A, 0, F, G, SEL>(vp, vl, ud, creg, ok, acc);
                break;
            }
            [[fallthrough]];
        case 1:
            if constexpr (RB > 1) {
                reg_sweep_pair<R, A, 1, F, G, SEL>(vp, vl, ud, creg, ok, acc);
                break;
            }
            [[fallthrough]];
        case 2:
            if constexpr (RB > 2) {
                reg_sweep_pair<R, A, 2, F, G, SEL>(vp, vl, ud, creg, ok, acc);
                break;
            }
            [[fallthrough]];
        case 3:
            if constexpr (RB > 3) {
                reg_sweep_pair<R, A, 3, F, G, SEL>(vp, vl, ud, creg, ok, acc);
                break;
            }
            [[fallthrough]];
        default: __builtin_unreachable();
    }
}

// Pair op of form code 4 F + G (see sweep_form): fixed gates in the general, swap and real
// forms, RY (real, Y) and RX (form 3, X); any other combination takes the general matrix.
template <class R, int A>
__device__ __forceinline__ void reg_sweep_pair_at(Cx<R> (&vp)[A], Cx<R> (&vl)[A], const Mat2<R> &ud, int creg, bool ok,
                                                  bool plain, int p, int form, R &acc) {
    if (plain) {
        switch (form) {
            case 4 * 2 + 2: reg_sweep_pair_pos<R, A, 2, 2, false>(vp, vl, ud, 0, true, p, acc); break;
            case 4 * 3 + 1: reg_sweep_pair_pos<R, A, 3, 1, false>(vp, vl, ud, 0, true, p, acc); break;
            case 4 * 1 + 0: reg_sweep_pair_pos<R, A, 1, 0, false>(vp, vl, ud, 0, true, p, acc); break;
            case 4 * 2 + 0: reg_sweep_pair_pos<R, A, 2, 0, false>(vp, vl, ud, 0, true, p, acc); break;
            case 1: reg_sweep_pair_pos<R, A, 0, 1, false>(vp, vl, ud, 0, true, p, acc); break;
            case 2: reg_sweep_pair_pos<R, A, 0, 2, false>(vp, vl, ud, 0, true, p, acc); break;
            case 3: reg_sweep_pair_pos<R, A, 0, 3, false>(vp, vl, ud, 0, true, p, acc); break;
            default: reg_sweep_pair_pos<R, A, 0, 0, false>(vp, vl, ud, 0, true, p, acc); break;
        }
        return;
    }
    // controlled gates: general matrix, any generator
    switch (form & 3) {
        case 1: reg_sweep_pair_pos<R, A, 0, 1, true>(vp, vl, ud, creg, ok, p, acc); break;
        case 2: reg_sweep_pair_pos<R, A, 0, 2, true>(vp, vl, ud, creg, ok, p, acc); break;
        case 3: reg_sweep_pair_pos<R, A, 0, 3, true>(vp, vl, ud, creg, ok, p, acc); break;
        default: reg_sweep_pair_pos<R, A, 0, 0, true>(vp, vl, ud, creg, ok, p, acc); break;
    }
}

// Registers of one state in phase layout (off, P) <-> LDS tile t.
template <class R, int A, int RB>
__device__ __forceinline__ void regs_from_lds(Cx<R> (&v)[A], const Cx<R> *t, unsigned off, const PhaseBits &P) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        v[j] = t[off | reg_bits<RB>(P, j)];
    });
}

template <class R, int A, int RB>
__device__ __forceinline__ void regs_to_lds(Cx<R> *t, const Cx<R> (&v)[A], unsigned off, const PhaseBits &P) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        t[off | reg_bits<RB>(P, j)] = v[j];
    });
}

template <int WB>
__device__ __forceinline__ unsigned phase_offset(const PhaseBits &P, unsigned lane, unsigned wave) {
    unsigned off = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) off |= ((lane >> k) & 1u) << P.lb(k);
#pragma unroll
    for (int k = 0; k < WB; ++k) off |= ((wave >> k) & 1u) << P.wb(k);
    return off;
}

// phi and lambda tiles in registers, one LDS tile each for the layout changes. Every LDS
// location written at the end of a phase is the one the same thread read at its start, so
// only the hand-overs between layouts need barriers.
template <class R, int M, int T>
__global__ __launch_bounds__(T) void k_tile_sweep_reg(Cx<R> *__restrict__ phi, Cx<R> *__restrict__ lam,
                                                      const SweepPass pass, const SweepOp<R> *__restrict__ ops,
                                                      const SweepPhase *__restrict__ phases,
                                                      double *__restrict__ partial, u64 n_tiles, u64 first) {
    constexpr int A = (1 << M) / T, RB = Log2<A>::v, WB = Log2<T / 64>::v, NW = T / 64;
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    __shared__ Cx<R> tp[1 << M];
    __shared__ Cx<R> tl[1 << M];
    __shared__ double red[kMaxSweepSlots][NW];
    const u64 tile = first + blockIdx.x;
    const u64 base = tile_base(tile, pass.pos, M);
    constexpr int upt = (1 << (M - lw)) / T;   // units per thread
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Cx<R> vp[A], vl[A];
    SweepPhase P = phases[pass.phase_begin];
    unsigned off = phase_offset<WB>(P, lane, wave);
    tile_to_lds<R, T, upt>(tp, phi, base, pass.pos, M, pass.low);
    tile_to_lds<R, T, upt>(tl, lam, base, pass.pos, M, pass.low);
    __syncthreads();
    regs_from_lds<R, A, RB>(vp, tp, off, P);
    regs_from_lds<R, A, RB>(vl, tl, off, P);
    for (int ph = 0; ph < pass.n_phases; ++ph) {
        if (ph > 0) {   // registers from phase ph - 1's layout to phase ph's
            const SweepPhase Q = phases[pass.phase_begin + ph];
            const unsigned qoff = phase_offset<WB>(Q, lane, wave);
            regs_to_lds<R, A, RB>(tp, vp, off, P);
            regs_to_lds<R, A, RB>(tl, vl, off, P);
            __syncthreads();
            regs_from_lds<R, A, RB>(vp, tp, qoff, Q);
            regs_from_lds<R, A, RB>(vl, tl, qoff, Q);
            P = Q;
            off = qoff;
        }
        // As in the forward register pass, diagonal gates on no register bit are one factor per
        // thread (tf), applied to phi and lambda at the end of the phase. Diagonal gates
        // leave w_j = Im(conj(lambda_j) phi_j) unchanged (|factor| = 1), so the generator-Z
        // terms of a run of diagonal gates all use the same w: an RZ on a lane, wave or outer
        // bit contributes +-U with U = sum_j w_j, computed once per run (have_u).
        // Runs of gates that leave the registers alone (fused diagonals, gates whose controls
        // outside the tile are not all set here) go through an inner loop that carries only
        // tf, U and the slot terms: the register arrays cross a loop back-edge (where the
        // compiler inserts copies of them) once per register-changing gate, not once per gate.
        Cx<R> tf = cx<R>(R(1), R(0));
        bool pending = false, have_u = false;
        R u_sum = R(0);
        int k = 0;
        while (k < P.n_ops) {
            for (; k < P.n_ops; ++k) {
                const SweepOp<R> &op = ops[pass.op_begin + P.op_begin + k];
                const bool on = (base & op.gctrl) == op.gctrl;
                const int kind = op.rkind, creg = op.creg, form = op.form;
                if (on && (kind < kRegDiagThread || creg != 0)) break;   // changes the registers
                R acc = R(0);
                if (on) {
                    const unsigned cthr = op.cthr;
                    const bool ok = (off & cthr) == cthr;
                    const u64 src = kind == kRegDiagThread ? (u64)off : base;
                    const bool one = ((src >> op.rtarget) & 1) != 0;
                    Cx<R> f = csel(one, op.udag.m11, op.udag.m00);
                    f = csel(cthr == 0 || ok, f, cx<R>(R(1), R(0)));
                    if ((form & 3) == 3) {
                        if (!have_u) {
                            u_sum = R(0);
                            static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                                constexpr int j = decltype(jc)::value;
                                u_sum += im_conj_mul(vl[j], vp[j]);
                            });
                            have_u = true;
                        }
                        acc = ok ? (one ? -u_sum : u_sum) : R(0);
                    }
                    tf = cmul(tf, f);
                    pending = true;
                }
                if (op.slot >= 0) {
                    const double s = wave_sum_lane63((double)acc);
                    if (lane == 63) red[op.slot][wave] = s;
                }
            }
            if (k == P.n_ops) break;
            // one gate that changes the registers (it is on)
            const SweepOp<R> &op = ops[pass.op_begin + P.op_begin + k];
            ++k;
            const unsigned cthr = op.cthr;
            const int creg = op.creg, tgt = op.rtarget, kind = op.rkind, form = op.form;
            const bool ok = (off & cthr) == cthr;
            const bool plain = cthr == 0 && creg == 0;
            R acc = R(0);
            if (kind <= kRegPairReal) {
                // tf (pending diagonal factor) commutes with the pair update, and the pair's
                // gradient term is unchanged by a common unit-modulus factor of phi and lambda
                reg_sweep_pair_at<R, A>(vp, vl, op.udag, creg, ok, plain, tgt, form, acc);
                have_u = false;
            } else {
                const bool zgen = (form & 3) == 3;
                const Cx<R> f0 = op.udag.m00, f1 = op.udag.m11;
                if (kind <= kRegDiagPhase) {
                    // target on register position tgt: the factor of index j is a scalar choice
                    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        const bool one = ((j >> tgt) & 1) != 0;
                        const Cx<R> f = csel(one, f1, f0);
                        const Cx<R> x = vp[j], y = vl[j];
                        const R w = zgen ? im_conj_mul(y, x) : R(0);
                        const Cx<R> nx = cmul(f, x), ny = cmul(f, y);
                        if (plain) {
                            acc += one ? -w : w;
                            vp[j] = nx;
                            vl[j] = ny;
                        } else {
                            const bool c = ok && (j & creg) == creg;
                            acc += c ? (one ? -w : w) : R(0);
                            vp[j] = csel(c, nx, x);
                            vl[j] = csel(c, ny, y);
                        }
                    });
                } else {   // lane, wave or outer target with controls on register bits
                    const u64 src = kind == kRegDiagThread ? (u64)off : base;
                    const bool one = ((src >> tgt) & 1) != 0;
                    Cx<R> f = csel(one, f1, f0);
                    f = csel(cthr == 0 || ok, f, cx<R>(R(1), R(0)));
                    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        const bool c = (j & creg) == creg;   // scalar
                        const Cx<R> x = vp[j], y = vl[j];
                        if (zgen) {
                            const R w = im_conj_mul(y, x);
                            acc += (c && ok) ? (one ? -w : w) : R(0);
                        }
                        if (c) {
                            vp[j] = cmul(f, x);
                            vl[j] = cmul(f, y);
                        }
                    });
                }
            }
            if (op.slot >= 0) {
                const double s = wave_sum_lane63((double)acc);
                if (lane == 63) red[op.slot][wave] = s;
            }
        }
        if (pending) {
            static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                vp[j] = cmul(tf, vp[j]);
                vl[j] = cmul(tf, vl[j]);
            });
        }
    }
    regs_to_lds<R, A, RB>(tp, vp, off, P);
    regs_to_lds<R, A, RB>(tl, vl, off, P);
    __syncthreads();
    lds_to_tile<R, T, upt>(phi, tp, base, pass.pos, M, pass.low);
    lds_to_tile<R, T, upt>(lam, tl, base, pass.pos, M, pass.low);
    for (int k = threadIdx.x; k < pass.n_slots; k += T) {
        double t = 0.;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += red[k][w];
        partial[(u64)k * n_tiles + tile] = t;
    }
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

template <class R>
void apply_gate_sequence(Cx<R> *a, int n_lanes, const std::vector<QueuedGate> &gates, int max_bits,
                         int low_bits) {
    typedef TileCfg<R> C;
    const int mb = std::min(max_bits > 0 ? std::min(max_bits, C::MaxBits) : C::MaxBits, n_lanes);
    const int low = std::min(low_bits >= 0 ? low_bits : C::Low, mb);
    QG_REQUIRE(Unit<R>::W == 1 || low >= 1, "complex64 tiles need bit 0");
    PassPlan plan;
    auto open = [&]() {
        plan.bits.clear();
        plan.gates.clear();
        for (int b = 0; b < low; ++b) plan.bits.push_back(b);
    };
    open();
    for (const QueuedGate &g : gates) {
        const bool need = !(is_diagonal(g.m) || plan.has(g.target));
        if ((int)plan.bits.size() + (need ? 1 : 0) > mb || (int)plan.gates.size() >= C::MaxOps) {
            launch_pass<R>(a, n_lanes, plan, low, mb);
            open();
        }
        if (!is_diagonal(g.m) && !plan.has(g.target)) plan.bits.push_back(g.target);
        plan.gates.push_back(&g);
    }
    if (!plan.gates.empty()) launch_pass<R>(a, n_lanes, plan, low, mb);
}

template void apply_gate_sequence<float>(Cx<float> *, int, const std::vector<QueuedGate> &, int, int);
template void apply_gate_sequence<double>(Cx<double> *, int, const std::vector<QueuedGate> &, int, int);

// Form code of a sweep op for the register passes: 4 * matrix form + generator (0 none, 1 X,
// 2 Y, 3 Z). Pair ops: matrix form 1 swap, 2 real, 3 real diagonal with imaginary
// off-diagonal, 0 anything else. Diagonal ops: 1 if m00 == 1 (only the |1> half changes),
// else 0. The forms are exact properties of the entries, which rounding to the state
// precision keeps.
int sweep_form(const GateMatrix &u, int pauli, bool diagonal) {
    const std::complex<double> m00 = u.m[0][0], m01 = u.m[0][1], m10 = u.m[1][0], m11 = u.m[1][1];
    int f = 0;
    if (diagonal) {
        f = m00 == 1. ? 1 : 0;
    } else if (m00 == 0. && m11 == 0. && m01 == 1. && m10 == 1.) {
        f = 1;
    } else if (m00.imag() == 0. && m01.imag() == 0. && m10.imag() == 0. && m11.imag() == 0.) {
        f = 2;
    } else if (m00.imag() == 0. && m11.imag() == 0. && m01.real() == 0. && m10.real() == 0.) {
        f = 3;
    }
    return 4 * f + pauli;
}

template <class R>
void adjoint_sweep(Cx<R> *phi, Cx<R> *lam, int n_lanes, const std::vector<SweepGate> &gates, double *d_terms,
                   int max_bits, int low_bits) {
    typedef SweepCfg<R> C;
    const bool regs = use_register_tiles();
    const int cap = regs ? C::RegBits : C::LdsBits;
    const int mb = std::min(max_bits > 0 ? std::min(max_bits, cap) : cap, n_lanes);
    const int low = std::min(low_bits >= 0 ? low_bits : C::Low, mb);
    // plan the passes: like apply_gate_sequence, with at most kMaxSweepSlots rotations a pass
    struct Plan {
        std::vector<int> bits;
        int begin, end, slots;
    };
    std::vector<Plan> plans;
    Plan cur;
    auto open = [&](int at) {
        cur.bits.clear();
        for (int b = 0; b < low; ++b) cur.bits.push_back(b);
        cur.begin = cur.end = at;
        cur.slots = 0;
    };
    auto has = [&](int b) { return std::find(cur.bits.begin(), cur.bits.end(), b) != cur.bits.end(); };
    open(0);
    for (int i = 0; i < (int)gates.size(); ++i) {
        const SweepGate &g = gates[i];
        const bool diag = is_diagonal(g.udag) && (g.pauli == 0 || g.pauli == 3);
        const bool need = !(diag || has(g.target));
        if ((int)cur.bits.size() + (need ? 1 : 0) > mb || (g.pauli != 0 && cur.slots == kMaxSweepSlots)) {
            plans.push_back(cur);
            open(i);
        }
        if (need) cur.bits.push_back(g.target);
        if (g.pauli != 0) ++cur.slots;
        cur.end = i + 1;
    }
    if (cur.end > cur.begin) plans.push_back(cur);

    constexpr int A = (1 << C::RegBits) / C::RegThreads;
    // device tables: ops (pass by pass), phases of register passes, slot -> gate index
    std::vector<SweepOp<R>> hops(std::max<size_t>(1, gates.size()));
    std::vector<SweepPhase> hphases;
    std::vector<int> slot_gate;
    std::vector<SweepPass> passes;
    std::vector<int> pass_slot_begin;
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
        pass_slot_begin.push_back((int)slot_gate.size());
        int slots = 0;
        for (int i = pl.begin; i < pl.end; ++i) {
            const SweepGate &g = gates[i];
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
            if (g.pauli != 0) slot_gate.push_back(i);
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
                SweepPhase h;
                std::memset(&h, 0, sizeof(h));
                for (int b = 0; b < (int)P.regs.size(); ++b) h.set(b, P.regs[b]);
                for (int b = 0; b < 6; ++b) h.set(5 + b, P.lanes[b]);
                for (int b = 0; b < (int)P.waves.size(); ++b) h.set(11 + b, P.waves[b]);
                h.op_begin = P.begin;
                h.n_ops = P.end - P.begin;
                for (int k = P.begin; k < P.end; ++k) {
                    SweepOp<R> &op = hops[pl.begin + k];
                    const RegSplit r = split_for_phase(P, op.kind, op.target, op.lctrl);
                    op.rkind = r.kind;
                    op.rtarget = r.target;
                    op.creg = r.creg;
                    op.cthr = r.cthr;
                }
                hphases.push_back(h);
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
    size_t part_bytes = 8;
    for (const SweepPass &ps : passes)
        part_bytes = std::max(part_bytes, sizeof(double) * (size_t)max_slots * (size_t)(u64(1) << (n_lanes - ps.m)));
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
    for (size_t pi = 0; pi < passes.size() && e == hipSuccess; ++pi) {
        const SweepPass &ps = passes[pi];
        const u64 n_tiles = u64(1) << (n_lanes - ps.m);
        const int threads = pass_regs[pi] ? C::RegThreads : kTileThreads;
        const u64 piece = launch_piece_items() / (u64)threads;
        for (u64 first = 0; first < n_tiles; first += piece) {
            const u64 count = std::min(piece, n_tiles - first);
            if (pass_regs[pi])
                hipLaunchKernelGGL((k_tile_sweep_reg<R, C::RegBits, C::RegThreads>), (unsigned)count,
                                   C::RegThreads, 0, s, phi, lam, ps, d_ops, d_ph, d_part, n_tiles, first);
            else
                hipLaunchKernelGGL((k_tile_sweep<R, C::LdsBits>), (unsigned)count, kTileThreads, 0, s, phi, lam, ps,
                                   d_ops, d_part, n_tiles, first);
        }
        if (ps.n_slots > 0)   // terms of this pass's slots land at d_terms[gate index]
            hipLaunchKernelGGL(k_slot_rows, (unsigned)ps.n_slots, kRowThreads, 0, s, d_part, n_tiles, d_sg + pass_slot_begin[pi],
                               d_terms);
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
