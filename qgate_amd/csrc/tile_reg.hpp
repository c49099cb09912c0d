// Register-resident tile passes and their planning, shared by the single-circuit passes
// (k_tile.hip) and the batched passes (k_batch.hip).
#pragma once

#include <algorithm>
#include <complex>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gate_matrix.hpp"
#include "qg_runtime.hpp"
#include "tile_common.hpp"

namespace qg {

namespace {

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
    int gate;          // batched passes: index of the gate in its circuit (m unused)
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

template <class R>
__device__ __forceinline__ R fma_r(R a, R b, R c) {
    if constexpr (sizeof(R) == 4)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

// acc + Im(conj(a) b), as two FMAs into the accumulator
template <class R>
__device__ __forceinline__ R add_im_conj_mul(R acc, const Cx<R> &a, const Cx<R> &b) {
    return fma_r(-a.im, b.re, fma_r(a.re, b.im, acc));
}

// acc + the gradient term of a pair (p0, p1) / (l0, l1) for generator G (1 X, 2 Y, 3 Z):
// Im<l|G|p> over the pair, four FMAs into the accumulator
template <class R, int G>
__device__ __forceinline__ R add_pair_term(R acc, const Cx<R> &p0, const Cx<R> &p1, const Cx<R> &l0,
                                           const Cx<R> &l1) {
    if constexpr (G == 1) return add_im_conj_mul(add_im_conj_mul(acc, l0, p1), l1, p0);
    if constexpr (G == 2) {
        acc = fma_r(l1.re, p0.re, acc);
        acc = fma_r(l1.im, p0.im, acc);
        acc = fma_r(-l0.re, p1.re, acc);
        return fma_r(-l0.im, p1.im, acc);
    }
    if constexpr (G == 3) return add_im_conj_mul(add_im_conj_mul(acc, l0, p0), p1, l1);
    return acc;
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
        // one multiply and three FMAs per component (a sum of two complex products written
        // as such takes two multiplies, two FMAs and an add)
        y0 = cx<R>(fma_r(mm.m00.re, x0.re, fma_r(-mm.m00.im, x0.im, fma_r(mm.m01.re, x1.re, -mm.m01.im * x1.im))),
                   fma_r(mm.m00.re, x0.im, fma_r(mm.m00.im, x0.re, fma_r(mm.m01.re, x1.im, mm.m01.im * x1.re))));
        y1 = cx<R>(fma_r(mm.m10.re, x0.re, fma_r(-mm.m10.im, x0.im, fma_r(mm.m11.re, x1.re, -mm.m11.im * x1.im))),
                   fma_r(mm.m10.re, x0.im, fma_r(mm.m10.im, x0.re, fma_r(mm.m11.re, x1.im, mm.m11.im * x1.re))));
    }
}

// ---- register ops between two register arrays ----------------------------------------------
//
// The ops of a phase read one register array and write the other, which the next op reads;
// a uniform flag says which array holds the amplitudes. Every output of an op is a new
// value, so all dispatch arms of an op leave their results in the same registers and no join
// needs copies. (An in-place update of one loop-carried array made the compiler keep two
// copies of it anyway and move the whole array between them on most paths: 36 to 130 moves
// a gate, against 128 FMAs for a general 2x2.) Ops that change no amplitude (skipped for
// controls outside the tile, per-thread factors) leave the flag as it is and move nothing.

// Arm codes of an op: pairs on register bit P of form F (0 general, 1 swap, 2 real) at
// 4 F + P, with controls on lane, wave or register bits at 12 + 4 F + P; diagonals on
// register bit P at 24 + P, with m00 == 1 at 28 + P; a diagonal on a register bit with
// controls 32; a per-thread factor on the register indices with controls creg 33.
enum : int { kArmPairSel = 12, kArmDiag = 24, kArmDiagPhase = 28, kArmDiagCtrl = 32, kArmFactorReg = 33 };

// Ends an arm. The markers differ per arm and per array, so the optimiser merges no two arms
// into one block that picks register elements through pointers (which would move them to
// scratch memory) or through selects.
template <int H, int C>
__device__ __forceinline__ void arm_marker() {
    asm volatile("; register op %0" ::"n"(H * 64 + C));
}

// y = x with the pair update of form F on register bit P applied to every pair (plain), or
// to the register indices j with (j & creg) == creg in threads with ok set (SEL: selects).
template <class R, int A, int P, int F, bool SEL>
__device__ __forceinline__ void op_pair(const Cx<R> (&x)[A], Cx<R> (&y)[A], const Mat2<R> &mm, int creg, bool ok) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if constexpr (((j >> P) & 1) == 0) {
            constexpr int j1 = j + (1 << P);
            if constexpr (SEL) {
                Cx<R> y0, y1;
                apply_form<R, F>(mm, x[j], x[j1], y0, y1);
                const bool c = ok && (j & creg) == creg;
                y[j] = csel(c, y0, x[j]);
                y[j1] = csel(c, y1, x[j1]);
            } else {
                apply_form<R, F>(mm, x[j], x[j1], y[j], y[j1]);
            }
            __builtin_amdgcn_sched_barrier(0);   // one pair at a time: bounded register use
        }
    });
}

// y = x times m11 where bit P of the register index is set, else m00 (PHASE: m00 == 1, those
// indices copied), the choice made at compile time.
template <class R, int A, int P, bool PHASE>
__device__ __forceinline__ void op_diag(const Cx<R> (&x)[A], Cx<R> (&y)[A], const Mat2<R> &mm) {
    static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if constexpr (((j >> P) & 1) != 0)
            y[j] = cmul(mm.m11, x[j]);
        else if constexpr (PHASE)
            y[j] = x[j];
        else
            y[j] = cmul(mm.m00, x[j]);
        __builtin_amdgcn_sched_barrier(0);
    });
}

template <class R, int A, int H, int C>
__device__ __forceinline__ void arm_body(const Cx<R> (&x)[A], Cx<R> (&y)[A], const Mat2<R> &mm, int creg, bool ok,
                                         int tgt, const Cx<R> &f) {
    constexpr int RB = Log2<A>::v;
    // Complex64: each arm takes its own copy of the matrix entries (scalar registers, through
    // an empty asm statement). A product of an entry and an amplitude that two arms share
    // would otherwise be computed once above their branch and be live through either arm,
    // which raised the kernel above 128 VGPRs (3 workgroups per CU instead of 4).
    const Cx<R>(&xl)[A] = x;
    Mat2<R> ml = mm;
    if constexpr (sizeof(R) == 4)
        asm volatile("; register op matrix %8"
                     : "+s"(ml.m00.re), "+s"(ml.m00.im), "+s"(ml.m01.re), "+s"(ml.m01.im), "+s"(ml.m10.re),
                       "+s"(ml.m10.im), "+s"(ml.m11.re), "+s"(ml.m11.im)
                     : "n"(H * 64 + C));
    if constexpr (C < kArmDiag) {
        constexpr int F = (C % kArmPairSel) / 4, P = C % 4;
        if constexpr (P < RB)
            op_pair<R, A, P, F, (C >= kArmPairSel)>(xl, y, ml, creg, ok);
        else
            __builtin_unreachable();
    } else if constexpr (C < kArmDiagCtrl) {
        constexpr int P = C % 4;
        if constexpr (P < RB)
            op_diag<R, A, P, (C >= kArmDiagPhase)>(xl, y, ml);
        else
            __builtin_unreachable();
    } else if constexpr (C == kArmDiagCtrl) {
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            const Cx<R> u = cmul(csel(((j >> tgt) & 1) != 0, ml.m11, ml.m00), xl[j]);
            y[j] = csel(ok && (j & creg) == creg, u, xl[j]);
            __builtin_amdgcn_sched_barrier(0);
        });
    } else {
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            y[j] = csel((j & creg) == creg, cmul(f, xl[j]), xl[j]);
            __builtin_amdgcn_sched_barrier(0);
        });
    }
    arm_marker<H, C>();
}

// One op from x into y (H: 0 when x is the phase's first array, 1 for the second), one
// switch: every arm ends at one join.
template <class R, int A, int H>
__device__ __forceinline__ void reg_op(const Cx<R> (&x)[A], Cx<R> (&y)[A], const Mat2<R> &mm, int arm, int creg,
                                       bool ok, int tgt, const Cx<R> &f) {
#define QG_REG_ARM(c)                                          \
    case c:                                                    \
        arm_body<R, A, H, c>(x, y, mm, creg, ok, tgt, f);      \
        break;
    switch (arm) {
        QG_REG_ARM(0) QG_REG_ARM(1) QG_REG_ARM(2) QG_REG_ARM(3) QG_REG_ARM(4) QG_REG_ARM(5) QG_REG_ARM(6)
        QG_REG_ARM(7) QG_REG_ARM(8) QG_REG_ARM(9) QG_REG_ARM(10) QG_REG_ARM(11) QG_REG_ARM(12) QG_REG_ARM(13)
        QG_REG_ARM(14) QG_REG_ARM(15) QG_REG_ARM(16) QG_REG_ARM(17) QG_REG_ARM(18) QG_REG_ARM(19)
        QG_REG_ARM(20) QG_REG_ARM(21) QG_REG_ARM(22) QG_REG_ARM(23) QG_REG_ARM(24) QG_REG_ARM(25)
        QG_REG_ARM(26) QG_REG_ARM(27) QG_REG_ARM(28) QG_REG_ARM(29) QG_REG_ARM(30) QG_REG_ARM(31)
        QG_REG_ARM(32) QG_REG_ARM(33)
        default: __builtin_unreachable();
    }
#undef QG_REG_ARM
}

// The phase's amplitudes back to LDS, times the collected per-thread factor if there is one.
template <class R, int A, int H>
__device__ __forceinline__ void phase_store(Cx<R> *t, const Cx<R> (&v)[A], const PhaseBits &P, unsigned off,
                                            const Cx<R> &tf, bool pending) {
    constexpr int RB = Log2<A>::v;
    // the store addresses are formed here again (one OR each) from an opaque copy of off: the
    // compiler would otherwise keep the phase load's A addresses live through every op
    asm volatile("; phase store %1" : "+v"(off) : "n"(H));
    if (pending) {
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            t[off | reg_bits<RB>(P, j)] = cmul(tf, v[j]);
        });
    } else {
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            t[off | reg_bits<RB>(P, j)] = v[j];
        });
    }
    arm_marker<H, 63>();
}

__host__ __device__ constexpr bool is_pair_kind(int k) { return k == kOpPair || k == kOpPairSwap || k == kOpPairReal; }

// One op of a register phase, decoded for reg_op.
template <class R>
struct RegStep {
    Mat2<R> mm;
    Cx<R> f;          // kArmFactorReg: the factor
    int arm, creg, tgt;
    bool ok;          // the thread passes the op's controls on lane and wave bits
};

// The first op of phase P at index k or later that changes amplitudes (P.n_ops if none),
// decoded into d. Ops skipped on the way: those whose controls outside the tile fail (the
// same for the whole workgroup), and per-thread factors on every register index, which are
// multiplied into tf.
template <class R, bool B>
__device__ __forceinline__ int next_reg_op(const RegProgram<R> &prog, const RegPhase &P, int k, u64 base,
                                           unsigned off, const Mat2<R> *tm, Cx<R> &tf, bool &pending,
                                           RegStep<R> &d) {
    const Cx<R> one = cx<R>(R(1), R(0));
    for (; k < P.n_ops; ++k) {
        const int oi = P.op_begin + k;
        const u64 gctrl = prog.ops[oi].gctrl;
        if ((base & gctrl) != gctrl) continue;
        if constexpr (B)
            d.mm = tm[2 * prog.ops[oi].gate];
        else
            d.mm = prog.ops[oi].m;
        const unsigned cthr = prog.ops[oi].cthr;
        const int creg = prog.ops[oi].creg, tgt = prog.ops[oi].target, kind = prog.ops[oi].kind;
        // no controls on lane, wave or register bits: straight-line updates, no tests
        const bool plain = cthr == 0 && creg == 0;
        d.ok = (off & cthr) == cthr;
        d.creg = creg;
        d.tgt = tgt;
        d.f = one;
        if (kind <= kRegPairReal) {
            d.arm = (plain ? 0 : kArmPairSel) + 4 * kind + tgt;
        } else if (kind <= kRegDiagPhase) {
            d.arm = plain ? (kind == kRegDiag ? kArmDiag : kArmDiagPhase) + tgt : kArmDiagCtrl;
        } else {
            const u64 src = kind == kRegDiagThread ? (u64)off : base;
            Cx<R> f = csel(((src >> tgt) & 1) != 0, d.mm.m11, d.mm.m00);
            f = csel(cthr == 0 || d.ok, f, one);
            if (creg == 0) {
                tf = cmul(tf, f);
                pending = true;
                continue;
            }
            d.f = f;
            d.arm = kArmFactorReg;
        }
        return k;
    }
    return P.n_ops;
}

// Batched passes (B): the state vector holds many circuits of one gate structure, item i's
// amplitudes at [i << item_shift, (i + 1) << item_shift), and no tile spans two items. Gate g
// then takes its matrix from the tile item's table, imats[2 (item n_gates + g)] (tables hold U
// and U^dagger per gate; imats points at the one the pass applies), instead of from the op.
// One tile per workgroup: persistent workgroups that load the next tile into registers while
// they work on the current one need 32 more VGPRs (3 workgroups per CU instead of 4) and
// measured slower (30-qubit complex64 step 610 ms against 507 ms).
template <class R, int M, int T, bool B>
__global__ __launch_bounds__(T) void k_tile_reg(Cx<R> *__restrict__ a, const RegProgram<R> prog, u64 first,
                                                const Mat2<R> *__restrict__ imats, int n_gates, int item_shift) {
    constexpr int A = (1 << M) / T, RB = Log2<A>::v, WB = Log2<T / 64>::v;
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    __shared__ Cx<R> t[1 << M];
    const u64 base = tile_base(first + blockIdx.x, prog.pos, M);
    const Mat2<R> *tm = nullptr;
    if constexpr (B) tm = imats + 2 * (base >> item_shift) * (u64)n_gates;
    tile_to_lds<R, T, (1 << (M - lw)) / T>(t, a, base, prog.pos, M, prog.low);
    __syncthreads();
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Cx<R> one = cx<R>(R(1), R(0));
    Cx<R> va[A], vb[A];
    for (int ph = 0; ph < prog.n_phases; ++ph) {
        const RegPhase P = prog.phases[ph];
        unsigned off = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) off |= ((lane >> k) & 1u) << P.lb(k);
#pragma unroll
        for (int k = 0; k < WB; ++k) off |= ((wave >> k) & 1u) << P.wb(k);
        static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            va[j] = t[off | reg_bits<RB>(P, j)];
        });
        // Diagonal gates with no register bit among their bits multiply all of a thread's
        // amplitudes by one factor. They are collected in tf and applied once at the end of
        // the phase: tf commutes with every other gate of the phase, which mixes or scales
        // amplitudes of one thread only.
        Cx<R> tf = one;
        bool pending = false;
        // ops in twos, va -> vb then vb -> va: only va is carried round the loop
        int k = 0;
        for (;;) {
            RegStep<R> d;
            k = next_reg_op<R, B>(prog, P, k, base, off, tm, tf, pending, d);
            if (k >= P.n_ops) {
                phase_store<R, A, 0>(t, va, P, off, tf, pending);
                break;
            }
            reg_op<R, A, 0>(va, vb, d.mm, d.arm, d.creg, d.ok, d.tgt, d.f);
            k = next_reg_op<R, B>(prog, P, k + 1, base, off, tm, tf, pending, d);
            if (k >= P.n_ops) {
                phase_store<R, A, 1>(t, vb, P, off, tf, pending);
                break;
            }
            reg_op<R, A, 1>(vb, va, d.mm, d.arm, d.creg, d.ok, d.tgt, d.f);
            ++k;
        }
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

inline std::vector<PhaseLayout> plan_phase_layouts(const std::vector<int> &pair, int m, int rb_count, int wb_count) {
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

// A register pass keeps its pair updates in order and puts each diagonal gate at a position,
// between the pair updates on its lanes before and after it, whose phase has none of its tile
// lanes as a register bit: there it is a per-thread factor, not an update of the register
// arrays. (A pass that takes the gates of several layers otherwise meets a layer's RZ and CZ
// right after the RY that made their lanes register bits.) Diagonal gates commute with each
// other and with every pair update whose target is none of their lanes, as do the generators
// of diagonal rotations, so the states and gradient terms are unchanged. order[begin, end) are
// the pass's gates; is_diag(g), target_of(g) and lanes_of(g) (target and controls) describe
// gate g; local maps a state lane to its tile bit (-1 outside the tile).
template <class IsDiag, class TargetOf, class LanesOf>
inline void place_diagonals(std::vector<int> &order, int begin, int end, const int *local, int m, int rb_count,
                            int wb_count, IsDiag is_diag, TargetOf target_of, LanesOf lanes_of) {
    std::vector<int> pairs, pair_local;
    for (int i = begin; i < end; ++i)
        if (!is_diag(order[i])) {
            pairs.push_back(order[i]);
            pair_local.push_back(local[target_of(order[i])]);
        }
    const int np = (int)pairs.size();
    if (np == 0) return;
    const std::vector<PhaseLayout> L = plan_phase_layouts(pair_local, m, rb_count, wb_count);
    std::vector<int> phase_of(np, 0);
    for (int ph = 0; ph < (int)L.size(); ++ph)
        for (int k = L[ph].begin; k < L[ph].end; ++k) phase_of[k] = ph;
    // a gate after pair update gp - 1 (before the first: gp = 0) runs in its phase
    auto gap_phase = [&](int gp) { return gp >= 1 ? phase_of[gp - 1] : 0; };
    std::vector<std::vector<int>> at(np + 1);
    int seen = 0;
    for (int i = begin; i < end; ++i) {
        const int gi = order[i];
        if (!is_diag(gi)) {
            ++seen;
            continue;
        }
        const std::vector<int> lanes = lanes_of(gi);
        auto hits = [&](int k) { return std::find(lanes.begin(), lanes.end(), target_of(pairs[k])) != lanes.end(); };
        int lo = 0, hi = np;
        for (int k = seen - 1; k >= 0; --k)
            if (hits(k)) {
                lo = k + 1;
                break;
            }
        for (int k = seen; k < np; ++k)
            if (hits(k)) {
                hi = k;
                break;
            }
        auto bad = [&](int gp) {
            const std::vector<int> &rg = L[gap_phase(gp)].regs;
            for (int l : lanes)
                if (local[l] >= 0 && std::find(rg.begin(), rg.end(), local[l]) != rg.end()) return true;
            return false;
        };
        int best = seen;
        for (int d = 1; bad(best) && (seen - d >= lo || seen + d <= hi); ++d) {
            if (seen - d >= lo && !bad(seen - d))
                best = seen - d;
            else if (seen + d <= hi && !bad(seen + d))
                best = seen + d;
        }
        at[best].push_back(gi);
    }
    int w = begin;
    for (int gp = 0; gp <= np; ++gp) {
        for (int gi : at[gp]) order[w++] = gi;
        if (gp < np) order[w++] = pairs[gp];
    }
}

// Packed form of a phase layout (ops [begin, end) of its pass).
inline PhaseBits phase_bits(const PhaseLayout &P) {
    PhaseBits h;
    std::memset(&h, 0, sizeof(h));
    for (int b = 0; b < (int)P.regs.size(); ++b) h.set(b, P.regs[b]);
    for (int b = 0; b < (int)P.lanes.size(); ++b) h.set(5 + b, P.lanes[b]);
    for (int b = 0; b < (int)P.waves.size(); ++b) h.set(11 + b, P.waves[b]);
    h.op_begin = P.begin;
    h.n_ops = P.end - P.begin;
    return h;
}

inline int reg_position(const PhaseLayout &P, int bit) {
    return (int)(std::find(P.regs.begin(), P.regs.end(), bit) - P.regs.begin());
}

// Register-pass encoding of a gate in phase P: kind (tile-pass kind, local bits: kOpPair*,
// kOpDiag*), target (local bit, or state bit for kOpDiagOuter) and local controls lctrl.
struct RegSplit {
    int kind, target, creg;
    unsigned cthr;
};

inline RegSplit split_for_phase(const PhaseLayout &P, int kind, int target, unsigned lctrl) {
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

// Phase layouts of a register program built from tile-pass ops (local bits; an Op has m,
// gctrl, lctrl, target and kind as TileOp does): false if the pass needs more phases than a
// program holds.
template <class R, class Op>
bool plan_phases(RegProgram<R> &prog, const Op *ops, int m, int rb_count, int wb_count) {
    std::vector<int> pair(prog.n_ops);
    for (int k = 0; k < prog.n_ops; ++k) pair[k] = is_pair_kind(ops[k].kind) ? ops[k].target : -1;
    const std::vector<PhaseLayout> L = plan_phase_layouts(pair, m, rb_count, wb_count);
    if ((int)L.size() > kMaxPhases) return false;
    prog.n_phases = (int)L.size();
    for (int i = 0; i < prog.n_phases; ++i) {
        prog.phases[i] = phase_bits(L[i]);
        for (int k = L[i].begin; k < L[i].end; ++k) {
            const RegSplit r = split_for_phase(L[i], ops[k].kind, ops[k].target, ops[k].lctrl);
            RegOp<R> &op = prog.ops[k];
            op.m = ops[k].m;
            op.gctrl = ops[k].gctrl;
            op.cthr = r.cthr;
            op.creg = r.creg;
            op.target = r.target;
            op.kind = r.kind;
            op.gate = 0;
        }
    }
    return true;
}

// ---- register-resident adjoint sweep ------------------------------------------------------

// Form code of a sweep op for the register passes: 4 * matrix form + generator (0 none, 1 X,
// 2 Y, 3 Z). Pair ops: matrix form 1 swap, 2 real, 3 real diagonal with imaginary
// off-diagonal, 0 anything else. Diagonal ops: 1 if m00 == 1 (only the |1> half changes),
// else 0. The forms are exact properties of the entries, which rounding to the state
// precision keeps.
inline int sweep_form(const GateMatrix &u, int pauli, bool diagonal) {
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
    int gate;           // batched passes: index of the gate in its circuit (udag unused)
};

struct SweepPass {
    int m, low;
    int op_begin, n_ops;
    int phase_begin, n_phases, n_slots;
    alignas(4) signed char pos[kTilePosMax];
};

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
            Cx<R> q0, q1, m0, m1;
            apply_form<R, F>(ud, p0, p1, q0, q1);
            apply_form<R, F>(ud, l0, l1, m0, m1);
            if constexpr (SEL) {
                const bool c = ok && (j & creg) == creg;
                if constexpr (G != 0) acc += c ? add_pair_term<R, G>(R(0), p0, p1, l0, l1) : R(0);
                vp[j] = csel(c, q0, p0);
                vp[j1] = csel(c, q1, p1);
                vl[j] = csel(c, m0, l0);
                vl[j1] = csel(c, m1, l1);
            } else {
                acc = add_pair_term<R, G>(acc, p0, p1, l0, l1);
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
// forms, RY (real, Y) and RX (form 3, X); any other combination takes the general matrix with
// the op's own generator G (an RX whose matrix happens to be real, RX(0) = 1, keeps its X
// term).
template <class R, int A>
__device__ __forceinline__ void reg_sweep_pair_at(Cx<R> (&vp)[A], Cx<R> (&vl)[A], const Mat2<R> &ud, int creg, bool ok,
                                                  bool plain, int p, int form, R &acc) {
    if (plain) {
        switch (form) {
            case 4 * 2 + 2: reg_sweep_pair_pos<R, A, 2, 2, false>(vp, vl, ud, 0, true, p, acc); break;
            case 4 * 3 + 1: reg_sweep_pair_pos<R, A, 3, 1, false>(vp, vl, ud, 0, true, p, acc); break;
            case 4 * 1 + 0: reg_sweep_pair_pos<R, A, 1, 0, false>(vp, vl, ud, 0, true, p, acc); break;
            case 4 * 2 + 0: reg_sweep_pair_pos<R, A, 2, 0, false>(vp, vl, ud, 0, true, p, acc); break;
            default:
                switch (form & 3) {
                    case 1: reg_sweep_pair_pos<R, A, 0, 1, false>(vp, vl, ud, 0, true, p, acc); break;
                    case 2: reg_sweep_pair_pos<R, A, 0, 2, false>(vp, vl, ud, 0, true, p, acc); break;
                    case 3: reg_sweep_pair_pos<R, A, 0, 3, false>(vp, vl, ud, 0, true, p, acc); break;
                    default: reg_sweep_pair_pos<R, A, 0, 0, false>(vp, vl, ud, 0, true, p, acc); break;
                }
                break;
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
// B: batched passes, as in k_tile_reg (imats points at the U^dagger entries).
template <class R, int M, int T, bool B>
__global__ __launch_bounds__(T) void k_tile_sweep_reg(Cx<R> *__restrict__ phi, Cx<R> *__restrict__ lam,
                                                      const SweepPass pass, const SweepOp<R> *__restrict__ ops,
                                                      const SweepPhase *__restrict__ phases,
                                                      double *__restrict__ partial, u64 n_tiles, u64 first,
                                                      const Mat2<R> *__restrict__ imats, int n_gates,
                                                      int item_shift) {
    constexpr int A = (1 << M) / T, RB = Log2<A>::v, WB = Log2<T / 64>::v, NW = T / 64;
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    __shared__ Cx<R> tp[1 << M];
    __shared__ Cx<R> tl[1 << M];
    __shared__ double red[kMaxSweepSlots][NW];
    const u64 tile = first + blockIdx.x;
    const u64 base = tile_base(tile, pass.pos, M);
    const Mat2<R> *tm = nullptr;
    if constexpr (B) tm = imats + 2 * (base >> item_shift) * (u64)n_gates;
    // U^dagger of an op: from the op, or from the tile item's table
    auto udag_of = [&](const SweepOp<R> &op) __attribute__((always_inline)) -> Mat2<R> {
        if constexpr (B)
            return tm[2 * op.gate];
        else
            return op.udag;
    };
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
                    const Mat2<R> ud = udag_of(op);
                    Cx<R> f = csel(one, ud.m11, ud.m00);
                    f = csel(cthr == 0 || ok, f, cx<R>(R(1), R(0)));
                    if ((form & 3) == 3) {
                        if (!have_u) {
                            u_sum = R(0);
                            static_for<0, A>([&](auto jc) __attribute__((always_inline)) {
                                constexpr int j = decltype(jc)::value;
                                u_sum = add_im_conj_mul(u_sum, vl[j], vp[j]);
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
                reg_sweep_pair_at<R, A>(vp, vl, udag_of(op), creg, ok, plain, tgt, form, acc);
                have_u = false;
            } else {
                const bool zgen = (form & 3) == 3;
                const Mat2<R> ud = udag_of(op);
                const Cx<R> f0 = ud.m00, f1 = ud.m11;
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

}  // namespace

}  // namespace qg
