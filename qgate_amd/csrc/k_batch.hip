// Batches of parameterised circuits: <H> and the full adjoint gradient of many circuits of
// uneven qubit counts in one captured HIP graph. (No reference counterpart.)
//
// Circuits with the same gate structure, qubit count and observable terms form a group. The
// group's state vectors sit back to back in HBM, so item i's amplitude x is element
// (i << n) | x: the group is one state vector of n + log2(items) bits whose top bits select
// the circuit. Tiles never take those top bits, so every tile belongs to one circuit, and
// the per-circuit gate matrices (computed on the device from each circuit's parameters) and
// observable coefficients are picked by the tile's item index. A group then runs as:
//   k_batch_mats     U and U^dagger of every (item, gate), from the parameters (FP64 sincos)
//   k_batch_zero     |0...0> in every item
//   k_batch_pass     forward tile passes (gates applied in LDS, many per pass)
//   k_batch_pauli    lambda = H psi in tile passes, the last one with Re<psi|lambda> partials
//   k_batch_sweep    backward sweep passes over psi and lambda with gradient partials
//   k_rows_sum       partials -> values and per-gate gradient terms (fixed order)
//   k_batch_grads    per-gate terms -> per-parameter gradients (fixed order)
// The pass kernels are persistent: a fixed grid of workgroups strides over the tiles. Tile
// size, workgroup size and grid size per group are picked by timing candidates on the first
// run (the autotuner below); the whole sequence for all groups is then captured once into a
// HIP graph and replayed by later runs with new parameters.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/qgate_hip.h"
#include "circuit.hpp"
#include "kernels.hpp"
#include "qg_runtime.hpp"
#include "tile_common.hpp"
#include "tile_reg.hpp"

using namespace qg;

namespace {

struct BGate {          // one gate of a group's structure
    double mat[8];      // fixed gates: row-major (re, im) pairs
    int pauli;          // 0 fixed, 1 X, 2 Y, 3 Z rotation
    int param;
};

template <class R>
struct BMats {
    Mat2<R> fwd, bwd;
};

struct BOp {            // a gate inside a pass
    u64 gctrl;          // controls outside the tile (state bits of the item)
    unsigned lctrl;     // controls inside the tile (local bits)
    int target;         // local bit, or the item's state bit for kOpDiagOuter
    int kind;           // (dwords: a uniform dword field is one scalar load)
    int gate;           // index in the structure
    int pauli;          // generator of a rotation (sweep)
    int slot;           // gradient slot within the pass (sweep), -1 otherwise
};

struct BPass {
    int m, low;
    int op_begin, n_ops;    // ops (gate passes) or groups (Pauli passes)
    int n_slots;            // sweep: gradient slots; Pauli: 1 on the last pass
    int accumulate;         // Pauli: add to lambda
    alignas(4) signed char pos[kTilePosMax];
};

struct BTermMask {
    u64 zmask;
};

// ---- kernels ------------------------------------------------------------------------------

template <class R>
__global__ __launch_bounds__(256) void k_batch_mats(const BGate *__restrict__ gates, int n_gates,
                                                    const double *__restrict__ params, int n_params, u64 n_items,
                                                    BMats<R> *__restrict__ mats) {
    const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_items * (u64)n_gates) return;
    const u64 item = i / (u64)n_gates;
    const int g = (int)(i % (u64)n_gates);
    const BGate gt = gates[g];
    double m[8];
    for (int k = 0; k < 8; ++k) m[k] = gt.mat[k];
    if (gt.pauli != 0) {
        const double th = params[item * (u64)n_params + (u64)gt.param];
        const double c = cos(0.5 * th), s = sin(0.5 * th);
        const double isx = gt.pauli == 1 ? 1. : 0., isy = gt.pauli == 2 ? 1. : 0., isz = gt.pauli == 3 ? 1. : 0.;
        // exp(-i th G / 2) = c I - i s G (qgate's RX / RY / RZ, GateMatrix.cpp:96-124)
        m[0] = c;        m[1] = -s * isz;     // m00
        m[2] = -s * isy; m[3] = -s * isx;     // m01
        m[4] = s * isy;  m[5] = -s * isx;     // m10
        m[6] = c;        m[7] = s * isz;      // m11
    }
    BMats<R> out;
    out.fwd.m00 = cx<R>((R)m[0], (R)m[1]);
    out.fwd.m01 = cx<R>((R)m[2], (R)m[3]);
    out.fwd.m10 = cx<R>((R)m[4], (R)m[5]);
    out.fwd.m11 = cx<R>((R)m[6], (R)m[7]);
    out.bwd.m00 = cx<R>((R)m[0], (R)-m[1]);
    out.bwd.m01 = cx<R>((R)m[4], (R)-m[5]);
    out.bwd.m10 = cx<R>((R)m[2], (R)-m[3]);
    out.bwd.m11 = cx<R>((R)m[6], (R)-m[7]);
    mats[i] = out;
}

// every item to |0...0>
template <class R>
__global__ __launch_bounds__(256) void k_batch_zero(Cx<R> *a, u64 n_units, int unit_bits) {
    const u64 u = (u64)blockIdx.x * 256 + threadIdx.x;
    if (u >= n_units) return;
    Amps<R> v;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) v.v[j] = cx<R>(0, 0);
    if ((u & ((u64(1) << unit_bits) - 1)) == 0) v.v[0].re = R(1);
    store_unit(a, u, v);
}

template <class R, int T, int MB>
__device__ __forceinline__ void load_tile(Cx<R> *t, const Cx<R> *a, u64 base, const BPass &ps) {
    load_tile_lds<R, T, MB>(t, a, base, ps.pos, ps.m, ps.low);
}

template <class R, int T, int MB>
__device__ __forceinline__ void store_tile(Cx<R> *a, const Cx<R> *t, u64 base, const BPass &ps) {
    store_tile_lds<R, T, MB>(a, t, base, ps.pos, ps.m, ps.low);
}

// Forward pass: the pass's gates on every tile, each with its item's matrix.
template <class R, int MB, int T>
__global__ __launch_bounds__(T) void k_batch_pass(Cx<R> *__restrict__ a, const BPass ps, const BOp *__restrict__ ops,
                                                  const BMats<R> *__restrict__ mats, int n_gates, int n, u64 n_tiles) {
    __shared__ Cx<R> t[1 << MB];
    const int n_amps = 1 << ps.m;
    for (u64 w = blockIdx.x; w < n_tiles; w += gridDim.x) {
        const u64 base = tile_base(w, ps.pos, ps.m);
        const u64 item = base >> n;
        load_tile<R, T, MB>(t, a, base, ps);
        __syncthreads();
        for (int k = 0; k < ps.n_ops; ++k) {
            const BOp op = ops[ps.op_begin + k];
            if ((base & op.gctrl) != op.gctrl) continue;
            const Mat2<R> mm = mats[item * (u64)n_gates + (u64)op.gate].fwd;
            const unsigned lc = op.lctrl;
            if (op.kind == kOpPair) {
                const int tb = op.target;
                const unsigned lowmask = (1u << tb) - 1;
                for (int p = threadIdx.x; p < n_amps / 2; p += T) {
                    const unsigned i0 = (((unsigned)p & ~lowmask) << 1) | ((unsigned)p & lowmask);
                    if ((i0 & lc) != lc) continue;
                    const unsigned i1 = i0 | (1u << tb);
                    const Cx<R> x0 = t[i0], x1 = t[i1];
                    t[i0] = cadd(cmul(mm.m00, x0), cmul(mm.m01, x1));
                    t[i1] = cadd(cmul(mm.m10, x0), cmul(mm.m11, x1));
                }
            } else {
                const int tb = op.target;
                const bool outer = op.kind == kOpDiagOuter;
                const bool obit = ((base >> tb) & 1) != 0;
                for (int i = threadIdx.x; i < n_amps; i += T) {
                    if (((unsigned)i & lc) != lc) continue;
                    const bool one = outer ? obit : ((i >> tb) & 1) != 0;
                    t[i] = cmul(one ? mm.m11 : mm.m00, t[i]);
                }
            }
            __syncthreads();
        }
        store_tile<R, T, MB>(a, t, base, ps);
        __syncthreads();
    }
}

// lambda (+)= sum over the pass's groups of c_k(item) (-1)^popcount(x & zm_k) psi[x ^ xm]; the
// last pass also leaves partial[w] = Re<psi|lambda> over tile w.
template <class R, int MB, int T>
__global__ __launch_bounds__(T) void k_batch_pauli(const Cx<R> *__restrict__ psi, Cx<R> *__restrict__ lam,
                                                   const BPass ps, const PauliGroup *__restrict__ groups,
                                                   const BTermMask *__restrict__ tmask,
                                                   const double *__restrict__ coefs, int n_terms, int n,
                                                   double *__restrict__ partial, u64 n_tiles) {
    __shared__ Cx<R> tp[1 << MB];
    __shared__ double scratch[T / 64];
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const int n_units = 1 << (ps.m - lw);
    for (u64 w = blockIdx.x; w < n_tiles; w += gridDim.x) {
        const u64 base = tile_base(w, ps.pos, ps.m);
        const u64 item = base >> n;
        const double *c = coefs + item * (u64)n_terms * 2;
        load_tile<R, T, MB>(tp, psi, base, ps);
        __syncthreads();
        double dot = 0.;
        const u64 dt = base | deposit((unsigned)threadIdx.x << lw, ps.pos, ps.m, ps.low);
        for (int lu = threadIdx.x; lu < n_units; lu += T) {
            const unsigned l = (unsigned)lu << lw;
            // state index of amplitude l: per-thread part | per-iteration part (disjoint bits)
            const u64 gl = dt | deposit((unsigned)(lu - (int)threadIdx.x) << lw, ps.pos, ps.m, ps.low);
            const u64 u = gl >> lw;
            Amps<R> out;
            if (ps.accumulate) out = load_unit(lam, u);
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const unsigned x = l + j;
                const u64 gx = gl | (u64)j;   // bit 0 is local bit 0 whenever W == 2
                double ar = ps.accumulate ? (double)out.v[j].re : 0., ai = ps.accumulate ? (double)out.v[j].im : 0.;
                for (int g = 0; g < ps.n_ops; ++g) {
                    const PauliGroup gr = groups[ps.op_begin + g];
                    double cr = 0., ci = 0.;
                    for (int k = gr.first; k < gr.first + gr.count; ++k) {
                        const double sgn = (__popcll(gx & tmask[k].zmask) & 1) ? -1. : 1.;
                        cr += sgn * c[2 * k];
                        ci += sgn * c[2 * k + 1];
                    }
                    const Cx<R> v = tp[x ^ gr.xlocal];
                    ar += cr * (double)v.re - ci * (double)v.im;
                    ai += cr * (double)v.im + ci * (double)v.re;
                }
                out.v[j] = cx<R>((R)ar, (R)ai);
                dot += (double)tp[x].re * (double)out.v[j].re + (double)tp[x].im * (double)out.v[j].im;
            }
            store_unit(lam, u, out);
        }
        if (ps.n_slots) {
            dot = block_sum_tile<T>(dot, scratch);
            if (threadIdx.x == 0) partial[w] = dot;
        }
        __syncthreads();
    }
}

// Backward sweep pass over psi and lambda tiles: per gate, the gradient term of a rotation
// (partial[slot * n_tiles + w] for tile w = item * tiles_per_item + tile index in the item)
// and U^dagger on both.
template <class R, int MB, int T>
__global__ __launch_bounds__(T) void k_batch_sweep(Cx<R> *__restrict__ phi, Cx<R> *__restrict__ lam, const BPass ps,
                                                   const BOp *__restrict__ ops, const BMats<R> *__restrict__ mats,
                                                   int n_gates, int n, double *__restrict__ partial,
                                                   u64 n_tiles) {
    __shared__ Cx<R> tp[1 << MB];
    __shared__ Cx<R> tl[1 << MB];
    __shared__ double scratch[T / 64];
    const int n_amps = 1 << ps.m;
    for (u64 w = blockIdx.x; w < n_tiles; w += gridDim.x) {
        const u64 base = tile_base(w, ps.pos, ps.m);
        const u64 item = base >> n;
        load_tile<R, T, MB>(tp, phi, base, ps);
        load_tile<R, T, MB>(tl, lam, base, ps);
        __syncthreads();
        for (int k = 0; k < ps.n_ops; ++k) {
            const BOp op = ops[ps.op_begin + k];
            const bool on = (base & op.gctrl) == op.gctrl;
            const Mat2<R> ud = mats[item * (u64)n_gates + (u64)op.gate].bwd;
            // generator, branch-free: X (1), Y (2), Z (3)
            const R gx = op.pauli == 1 ? R(1) : R(0), gy = op.pauli == 2 ? R(1) : R(0), gz = op.pauli == 3 ? R(1) : R(0);
            Mat2<R> gen;
            gen.m00 = cx<R>(gz, R(0));
            gen.m01 = cx<R>(gx, -gy);
            gen.m10 = cx<R>(gx, gy);
            gen.m11 = cx<R>(-gz, R(0));
            const unsigned lc = op.lctrl;
            double s = 0.;
            if (on && op.kind == kOpPair) {
                const int tb = op.target;
                const unsigned lowmask = (1u << tb) - 1;
                for (int p = threadIdx.x; p < n_amps / 2; p += T) {
                    const unsigned i0 = (((unsigned)p & ~lowmask) << 1) | ((unsigned)p & lowmask);
                    if ((i0 & lc) != lc) continue;
                    const unsigned i1 = i0 | (1u << tb);
                    const Cx<R> p0 = tp[i0], p1 = tp[i1], l0 = tl[i0], l1 = tl[i1];
                    const Cx<R> q0 = cadd(cmul(gen.m00, p0), cmul(gen.m01, p1));
                    const Cx<R> q1 = cadd(cmul(gen.m10, p0), cmul(gen.m11, p1));
                    s += (double)l0.re * (double)q0.im - (double)l0.im * (double)q0.re +
                         (double)l1.re * (double)q1.im - (double)l1.im * (double)q1.re;
                    tp[i0] = cadd(cmul(ud.m00, p0), cmul(ud.m01, p1));
                    tp[i1] = cadd(cmul(ud.m10, p0), cmul(ud.m11, p1));
                    tl[i0] = cadd(cmul(ud.m00, l0), cmul(ud.m01, l1));
                    tl[i1] = cadd(cmul(ud.m10, l0), cmul(ud.m11, l1));
                }
            }
            if (on && op.kind != kOpPair) {
                const int tb = op.target;
                const bool outer = op.kind == kOpDiagOuter;
                const bool obit = ((base >> tb) & 1) != 0;
                for (int i = threadIdx.x; i < n_amps; i += T) {
                    if (((unsigned)i & lc) != lc) continue;
                    const bool one = outer ? obit : ((i >> tb) & 1) != 0;
                    const Cx<R> x = tp[i], y = tl[i];
                    const Cx<R> q = cmul(one ? gen.m11 : gen.m00, x);
                    s += (double)y.re * (double)q.im - (double)y.im * (double)q.re;
                    const Cx<R> f = one ? ud.m11 : ud.m00;
                    tp[i] = cmul(f, x);
                    tl[i] = cmul(f, y);
                }
            }
            if (op.slot >= 0) {
                s = block_sum_tile<T>(s, scratch);
                if (threadIdx.x == 0) partial[(u64)op.slot * n_tiles + w] = s;
            }
            __syncthreads();
        }
        store_tile<R, T, MB>(phi, tp, base, ps);
        store_tile<R, T, MB>(lam, tl, base, ps);
        __syncthreads();
    }
}

// Row r = slot * n_items + item (slot-major: the tiles of all items for one slot are
// consecutive): out[item * out_stride + slot_out[slot]] = sum of partial[r * len .. + len),
// fixed order; one wavefront per row.
__global__ __launch_bounds__(64) void k_rows_sum(const double *__restrict__ partial, u64 len, u64 n_rows, u64 n_items,
                                                 const int *__restrict__ slot_out, int out_stride,
                                                 double *__restrict__ out) {
    const u64 r = blockIdx.x;
    if (r >= n_rows) return;
    const double *row = partial + r * len;
    double s = 0.;
    for (u64 i = threadIdx.x; i < len; i += 64) s += row[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) out[(r % n_items) * (u64)out_stride + (u64)slot_out[r / n_items]] = s;
}

// grads[item * n_params + p] = sum of terms[item * n_gates + g] over the rotations g with
// parameter p, in sweep order (pgate[pptr[p] .. pptr[p + 1])).
__global__ __launch_bounds__(256) void k_batch_grads(const double *__restrict__ terms, int n_gates,
                                                     const int *__restrict__ pptr, const int *__restrict__ pgate,
                                                     int n_params, u64 n_items, double *__restrict__ grads) {
    const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_items * (u64)n_params) return;
    const u64 item = i / (u64)n_params;
    const int p = (int)(i % (u64)n_params);
    double s = 0.;
    for (int j = pptr[p]; j < pptr[p + 1]; ++j) s += terms[item * (u64)n_gates + (u64)pgate[j]];
    grads[i] = s;
}

// ---- host ---------------------------------------------------------------------------------

template <class R>
struct BCfg;
template <>
struct BCfg<float> {
    static constexpr int MaxBits = 12;       // forward / Pauli tiles: 32 KiB
    static constexpr int SweepBits = 11;     // sweep: two tiles, 32 KiB
    static constexpr int Low = 5;
};
template <>
struct BCfg<double> {
    static constexpr int MaxBits = 11;
    static constexpr int SweepBits = 10;
    static constexpr int Low = 4;
};
// Register-resident passes: 256 threads, 2^12 / 2^11 amplitudes per forward tile (16 / 8 per
// thread) and 2^11 / 2^10 per sweep tile (8 / 4 per thread and state), as in the
// single-circuit passes.
constexpr int kBatchRegThreads = 256;
static_assert(BCfg<float>::MaxBits == TileCfg<float>::MaxBits && BCfg<double>::MaxBits == TileCfg<double>::MaxBits,
              "forward tiles");
static_assert(sizeof(BMats<float>) == 2 * sizeof(Mat2<float>) && sizeof(BMats<double>) == 2 * sizeof(Mat2<double>),
              "BMats is (U, U^dagger) back to back");

bool diag_matrix(const GateMatrix &g) { return g.m[0][1] == 0. && g.m[1][0] == 0.; }

struct Layout {
    std::vector<int> bits;   // sorted
    int low;
    int begin, end;
};

// Passes over the gates of a circuit in a planned order: each pass takes every gate not yet
// planned that commutes with the gates it left out before it (on every lane they share both
// act diagonally: a control, or the target of a diagonal gate) and whose non-diagonal target
// fits the tile, up to max_ops gates and max_counted counted ones. full[i] holds gate i's
// lanes used non-diagonally (its target unless diagonal: the bits it needs in the tile),
// diag[i] the others. `order` receives the gates pass by pass; the layouts' [begin, end)
// index it. (Passes of consecutive gates only, as plan_layouts: the 4096-circuit batch of
// 8..20-qubit ansatz circuits took about twice as many passes.) The scan stops after
// kLookahead gates left out or once every lane is held.
std::vector<Layout> plan_layouts_ahead(const std::vector<u64> &full, const std::vector<u64> &diag, int n, int mb,
                                       int low, bool c64, int max_ops, const std::vector<char> *counted,
                                       int max_counted, std::vector<int> &order) {
    constexpr int kLookahead = 256;
    const int ne = (int)full.size();
    const u64 all = n >= 64 ? ~u64(0) : (u64(1) << n) - 1;
    order.clear();
    std::vector<char> done(ne, 0);
    std::vector<Layout> out;
    for (int first = 0; first < ne;) {
        Layout cur;
        cur.begin = (int)order.size();
        u64 bits = (u64(1) << low) - 1, h1 = 0, h2 = 0;   // h2: lanes held non-diagonally
        int n_counted = 0, left = 0;
        for (int i = first; i < ne && (int)order.size() - cur.begin < max_ops; ++i) {
            if (done[i]) continue;
            QG_REQUIRE(__builtin_popcountll(full[i] | (c64 ? 1 : 0)) <= mb, "gate does not fit a tile");
            const int c = counted && (*counted)[i] ? 1 : 0;
            const bool conflict = (full[i] & (h1 | h2)) != 0 || (diag[i] & h2) != 0;
            if (!conflict && n_counted + c <= max_counted && __builtin_popcountll(bits | full[i]) <= mb) {
                bits |= full[i];
                n_counted += c;
                order.push_back(i);
                done[i] = 1;
                continue;
            }
            h2 |= full[i];
            h1 |= diag[i];
            if (++left >= kLookahead || (h2 & all) == all) break;
        }
        if ((int)order.size() == cur.begin) {   // the first gate does not fit beside the low bits
            bits = full[first];
            order.push_back(first);
            done[first] = 1;
        }
        cur.end = (int)order.size();
        for (int b = 0; b < n; ++b)
            if ((bits >> b) & 1) cur.bits.push_back(b);
        out.push_back(cur);
        while (first < ne && done[first]) ++first;
    }
    if (out.empty()) {
        Layout e;
        e.begin = e.end = 0;
        for (int b = 0; b < low; ++b) e.bits.push_back(b);
        out.push_back(e);
    }
    for (Layout &l : out) {
        if (c64 && std::find(l.bits.begin(), l.bits.end(), 0) == l.bits.end()) l.bits.push_back(0);
        for (int b = 0; (int)l.bits.size() < mb && b < n; ++b)
            if (std::find(l.bits.begin(), l.bits.end(), b) == l.bits.end()) l.bits.push_back(b);
        std::sort(l.bits.begin(), l.bits.end());
        int run = 0;
        while (run < (int)l.bits.size() && l.bits[run] == run) ++run;
        l.low = run;
    }
    return out;
}

// Passes over consecutive entries: need_masks[i] holds the bits entry i requires in the tile.
// A pass also ends before it would hold more than max_ops entries, or more than max_counted
// entries with counted[i] set (rotations of a sweep pass: one gradient slot each).
std::vector<Layout> plan_layouts(const std::vector<u64> &need_masks, int n, int mb, int low, bool c64,
                                 int max_ops = 1 << 30, const std::vector<char> *counted = nullptr,
                                 int max_counted = 1 << 30) {
    std::vector<Layout> out;
    Layout cur;
    int n_counted = 0;
    auto open = [&](int at) {
        cur.bits.clear();
        for (int b = 0; b < low; ++b) cur.bits.push_back(b);
        cur.begin = cur.end = at;
        n_counted = 0;
    };
    auto missing = [&](u64 m) {
        std::vector<int> v;
        for (int b = 0; b < n; ++b)
            if (((m >> b) & 1) && std::find(cur.bits.begin(), cur.bits.end(), b) == cur.bits.end()) v.push_back(b);
        return v;
    };
    open(0);
    for (int i = 0; i < (int)need_masks.size(); ++i) {
        QG_REQUIRE(__builtin_popcountll(need_masks[i] | (c64 ? 1 : 0)) <= mb, "gate or term does not fit a tile");
        std::vector<int> need = missing(need_masks[i]);
        const int c = counted && (*counted)[i] ? 1 : 0;
        if ((int)(cur.bits.size() + need.size()) > mb || cur.end - cur.begin == max_ops ||
            n_counted + c > max_counted) {
            if (cur.end > cur.begin) out.push_back(cur);
            open(i);
            need = missing(need_masks[i]);
            if ((int)(cur.bits.size() + need.size()) > mb) {
                cur.bits.clear();
                need = missing(need_masks[i]);
            }
        }
        for (int b : need) cur.bits.push_back(b);
        cur.end = i + 1;
        n_counted += c;
    }
    if (cur.end > cur.begin || out.empty()) out.push_back(cur);
    for (Layout &l : out) {
        if (c64 && std::find(l.bits.begin(), l.bits.end(), 0) == l.bits.end()) l.bits.push_back(0);
        for (int b = 0; (int)l.bits.size() < mb && b < n; ++b)
            if (std::find(l.bits.begin(), l.bits.end(), b) == l.bits.end()) l.bits.push_back(b);
        std::sort(l.bits.begin(), l.bits.end());
        int run = 0;
        while (run < (int)l.bits.size() && l.bits[run] == run) ++run;
        l.low = run;
    }
    return out;
}

BPass make_pass(const Layout &l, int local[64]) {
    BPass p;
    p.m = (int)l.bits.size();
    p.low = l.low;
    p.op_begin = l.begin;
    p.n_ops = l.end - l.begin;
    p.n_slots = 0;
    p.accumulate = 0;
    for (int j = 0; j < 16; ++j) p.pos[j] = j < p.m ? (signed char)l.bits[j] : (signed char)0;
    for (int b = 0; b < 64; ++b) local[b] = -1;
    for (int j = 0; j < p.m; ++j) local[l.bits[j]] = j;
    return p;
}

struct Tuning {
    int tile_bits = 0;   // forward / Pauli tile bits (the sweep: at most BCfg::SweepBits)
    int threads = 256;   // LDS passes
    int grid_per_cu = 4; // LDS passes: persistent workgroups per compute unit
    int registers = 0;   // 1: forward and sweep passes keep their tiles in registers
    double ms = 0.;      // measured time of one run of the group
};

struct Group {
    int n = 0;
    const void *key = nullptr;
    std::vector<CircuitOp> ops;
    int n_params = 0;
    std::vector<u64> xmask, zmask;          // observable terms (sorted by X mask)
    std::vector<PauliGroup> pgroups;
    std::vector<int> items;                 // batch item index of each group item
    std::vector<double> params;             // n_items x n_params
    std::vector<double> coefs;              // n_items x n_terms x 2 (with the (-i)^#Y factor)
    Tuning tune;
    // device
    void *psi = nullptr, *lam = nullptr;
    size_t state_bytes = 0;
    char *block = nullptr;
    size_t block_bytes = 0;
    double *d_params = nullptr, *d_coefs = nullptr, *d_partial = nullptr, *d_terms = nullptr;
    double *d_values = nullptr, *d_grads = nullptr;
    void *d_mats = nullptr;
    BGate *d_gates = nullptr;
    BOp *d_fops = nullptr, *d_sops = nullptr;
    PauliGroup *d_pgroups = nullptr;
    BTermMask *d_tmask = nullptr;
    int *d_slot_out = nullptr, *d_pptr = nullptr, *d_pgate = nullptr, *d_zero = nullptr;
    std::vector<BPass> fpasses, ppasses, spasses;
    std::vector<int> spass_slot_begin;      // offset of each sweep pass's slot table in d_slot_out
    // register-resident passes (tune.registers): one program per forward pass (kernel
    // arguments), sweep ops and phases in device tables
    bool fregs = false, sregs = false;
    std::vector<RegProgram<float>> rprog32;
    std::vector<RegProgram<double>> rprog64;
    std::vector<SweepPass> rspasses;
    void *d_rsops = nullptr;                // SweepOp<R>[]
    SweepPhase *d_rphases = nullptr;
    size_t n_items() const { return items.size(); }
};

}  // namespace

struct qghip_batch {
    Prec prec;
    std::vector<std::pair<int, int>> items;   // (group, index in group)
    std::vector<Group> groups;
    bool built = false;
    bool tuned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int fixed_tile_bits = 0, fixed_threads = 0, fixed_grid = 0;   // > 0: skip tuning for that knob
    int fixed_registers = -1;                                      // >= 0: skip tuning for that knob
};

namespace {

template <class F>
int guarded_batch(F &&f) {
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

void require(bool ok, const char *msg) {
    if (!ok) throw std::invalid_argument(msg);
}

void free_group(Group &g) {
    if (g.psi) dev_free(g.psi, g.state_bytes);
    if (g.lam) dev_free(g.lam, g.state_bytes);
    if (g.block) dev_free(g.block, g.block_bytes);
    g.psi = g.lam = nullptr;
    g.block = nullptr;
}

void drop_graph(qghip_batch *b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

template <class R>
std::vector<RegProgram<R>> &reg_programs(Group &g);
template <>
std::vector<RegProgram<float>> &reg_programs<float>(Group &g) {
    return g.rprog32;
}
template <>
std::vector<RegProgram<double>> &reg_programs<double>(Group &g) {
    return g.rprog64;
}

// A gate's matrix as far as its form goes: fixed gates as they are, rotations at a generic
// angle (every angle gives the same form: RX general / imaginary off-diagonal, RY real, RZ
// diagonal with m00 != 1), so one classification holds for every item.
GateMatrix form_matrix(const CircuitOp &op) { return op.kind == 1 ? rotation(op.pauli, 1.) : op.mat; }

// Register-pass input of one forward gate (the fields plan_phases reads).
template <class R>
struct BRegInput {
    Mat2<R> m;
    u64 gctrl;
    unsigned lctrl;
    int target, kind;
};

// Tile-pass kind of a forward gate in a register pass: pair kinds split into swap / real /
// general, local diagonals with m00 == 1 into kOpDiagPhase (from the matrix rounded to R).
template <class R>
int reg_forward_kind(const CircuitOp &op, int kind) {
    const Mat2<R> mm = round_matrix<R>(form_matrix(op));
    const R zero = R(0), one = R(1);
    if (kind == kOpPair) {
        const bool swap = op.kind == 0 && mm.m00.re == zero && mm.m00.im == zero && mm.m11.re == zero &&
                          mm.m11.im == zero && mm.m01.re == one && mm.m01.im == zero && mm.m10.re == one &&
                          mm.m10.im == zero;
        const bool real = mm.m00.im == zero && mm.m01.im == zero && mm.m10.im == zero && mm.m11.im == zero;
        return swap ? kOpPairSwap : real ? kOpPairReal : kOpPair;
    }
    if (kind == kOpDiagLocal && op.kind == 0 && mm.m00.re == one && mm.m00.im == zero) return kOpDiagPhase;
    return kind;
}

// Plans and device tables of a group for a tuning (allocates; not during capture).
template <class R>
void build_group(Group &g, const Tuning &t) {
    typedef BCfg<R> C;
    const bool c64 = Unit<R>::W == 2;
    const int n = g.n;
    const int n_gates = (int)g.ops.size();
    const size_t ni = g.n_items();
    // complex64 tiles move two-amplitude units, so they hold bit 0 plus at least one target
    const int min_bits = std::min(n, c64 ? 2 : 1);
    const int mf = std::max(min_bits, std::min(t.tile_bits, n));
    // sweep tiles hold two states: at most SweepBits; a group that fits one tile sweeps it whole
    const int ms = std::max(min_bits, std::min(std::min(t.tile_bits, (int)C::SweepBits), n));
    const int lowf = std::min(C::Low, mf), lows = std::min(C::Low, ms);
    // register-resident passes need full-size tiles (the kernels' compile-time tile bits)
    typedef TileCfg<R> TC;
    const bool fregs = t.registers != 0 && mf == TC::MaxBits && use_register_tiles();
    const bool sregs = t.registers != 0 && ms == C::SweepBits && use_register_tiles();

    // forward passes
    // (a rotation counts as diagonal only with generator Z: diagonal for every item's angle)
    std::vector<u64> need(n_gates), dlanes(n_gates);
    for (int i = 0; i < n_gates; ++i) {
        const CircuitOp &op = g.ops[i];
        const bool diag = op.kind == 1 ? op.pauli == 3 : diag_matrix(op.mat);
        need[i] = diag ? 0 : (u64(1) << op.target);
        dlanes[i] = diag ? (u64(1) << op.target) : 0;
        for (int c : op.ctrl) dlanes[i] |= u64(1) << c;
    }
    auto lanes_of = [&](int gi) {
        std::vector<int> l(g.ops[gi].ctrl);
        l.push_back(g.ops[gi].target);
        return l;
    };
    std::vector<int> ford;   // forward gates pass by pass
    std::vector<Layout> fl =
        plan_layouts_ahead(need, dlanes, n, mf, lowf, c64, fregs ? TC::MaxOps : 1 << 30, nullptr, 1 << 30, ford);
    std::vector<BOp> fops(std::max(1, n_gates));
    g.fpasses.clear();
    std::vector<RegProgram<R>> &progs = reg_programs<R>(g);
    progs.clear();
    for (const Layout &l : fl) {
        int local[64];
        BPass p = make_pass(l, local);
        if (fregs) {   // diagonal gates where none of their lanes is a register bit
            constexpr int A = (1 << TC::MaxBits) / kBatchRegThreads;
            place_diagonals(
                ford, l.begin, l.end, local, p.m, Log2<A>::v, Log2<kBatchRegThreads / 64>::v,
                [&](int gi) { return need[gi] == 0; }, [&](int gi) { return g.ops[gi].target; }, lanes_of);
        }
        for (int i = l.begin; i < l.end; ++i) {
            const CircuitOp &op = g.ops[ford[i]];
            BOp &o = fops[i];
            o.gctrl = 0;
            o.lctrl = 0;
            for (int c : op.ctrl) {
                if (local[c] >= 0)
                    o.lctrl |= 1u << local[c];
                else
                    o.gctrl |= u64(1) << c;
            }
            const bool diag = need[ford[i]] == 0;
            o.kind = (short)(local[op.target] >= 0 ? (diag ? kOpDiagLocal : kOpPair) : kOpDiagOuter);
            o.target = (short)(local[op.target] >= 0 ? local[op.target] : op.target);
            o.gate = ford[i];
            o.pauli = op.kind == 1 ? op.pauli : 0;
            o.slot = -1;
        }
        g.fpasses.push_back(p);
        if (fregs) {
            RegProgram<R> rp;
            std::memset(&rp, 0, sizeof(rp));
            rp.n_ops = p.n_ops;
            rp.low = p.low;
            std::memcpy(rp.pos, p.pos, sizeof(rp.pos));
            std::vector<BRegInput<R>> in(std::max(1, p.n_ops));
            for (int k = 0; k < p.n_ops; ++k) {
                const BOp &o = fops[l.begin + k];
                std::memset(&in[k], 0, sizeof(in[k]));
                in[k].gctrl = o.gctrl;
                in[k].lctrl = o.lctrl;
                in[k].target = o.target;
                in[k].kind = reg_forward_kind<R>(g.ops[ford[l.begin + k]], o.kind);
            }
            constexpr int A = (1 << TC::MaxBits) / kBatchRegThreads;
            QG_REQUIRE(plan_phases<R>(rp, in.data(), p.m, Log2<A>::v, Log2<kBatchRegThreads / 64>::v),
                       "batch: register pass with too many phases");
            for (int k = 0; k < p.n_ops; ++k) rp.ops[k].gate = ford[l.begin + k];
            progs.push_back(rp);
        }
    }
    // sweep passes: the gates in reverse order
    std::vector<u64> sneed(n_gates), sdl(n_gates);
    for (int i = 0; i < n_gates; ++i) sneed[i] = need[n_gates - 1 - i];
    for (int i = 0; i < n_gates; ++i) sdl[i] = dlanes[n_gates - 1 - i];
    std::vector<char> srot(n_gates);
    for (int i = 0; i < n_gates; ++i) srot[i] = g.ops[n_gates - 1 - i].kind == 1 ? 1 : 0;
    std::vector<int> sord;   // sweep entries (entry e is gate n_gates - 1 - e) pass by pass
    std::vector<Layout> sl = plan_layouts_ahead(sneed, sdl, n, ms, lows, c64, 1 << 30, &srot, kMaxSweepSlots, sord);
    std::vector<BOp> sops(std::max(1, n_gates));
    std::vector<SweepOp<R>> rsops(std::max(1, n_gates));
    std::memset(rsops.data(), 0, sizeof(SweepOp<R>) * rsops.size());
    std::vector<SweepPhase> rphases;
    std::vector<int> slot_out;
    g.spasses.clear();
    g.rspasses.clear();
    g.spass_slot_begin.clear();
    int max_slots = 1;
    for (const Layout &l : sl) {
        int local[64];
        BPass p = make_pass(l, local);
        if (sregs) {   // diagonal gates where none of their lanes is a register bit
            constexpr int A = (1 << C::SweepBits) / kBatchRegThreads;
            place_diagonals(
                sord, l.begin, l.end, local, p.m, Log2<A>::v, Log2<kBatchRegThreads / 64>::v,
                [&](int e) { return sneed[e] == 0; }, [&](int e) { return g.ops[n_gates - 1 - e].target; },
                [&](int e) { return lanes_of(n_gates - 1 - e); });
        }
        g.spass_slot_begin.push_back((int)slot_out.size());
        int slots = 0;
        for (int i = l.begin; i < l.end; ++i) {
            const int gi = n_gates - 1 - sord[i];
            const CircuitOp &op = g.ops[gi];
            BOp &o = sops[i];
            o.gctrl = 0;
            o.lctrl = 0;
            for (int c : op.ctrl) {
                if (local[c] >= 0)
                    o.lctrl |= 1u << local[c];
                else
                    o.gctrl |= u64(1) << c;
            }
            const bool diag = sneed[sord[i]] == 0;
            o.kind = (short)(local[op.target] >= 0 ? (diag ? kOpDiagLocal : kOpPair) : kOpDiagOuter);
            o.target = (short)(local[op.target] >= 0 ? local[op.target] : op.target);
            o.gate = gi;
            o.pauli = op.kind == 1 ? op.pauli : 0;
            o.slot = op.kind == 1 ? slots++ : -1;
            if (op.kind == 1) slot_out.push_back(gi);
            SweepOp<R> &r = rsops[i];
            r.gctrl = o.gctrl;
            r.lctrl = o.lctrl;
            r.target = o.target;
            r.kind = o.kind;
            r.slot = o.slot;
            r.form = sweep_form(dagger(form_matrix(op)), o.pauli, o.kind != kOpPair);
            r.gate = gi;
        }
        p.n_slots = slots;
        max_slots = std::max(max_slots, slots);
        g.spasses.push_back(p);
        if (sregs) {
            SweepPass rp;
            std::memset(&rp, 0, sizeof(rp));
            rp.m = p.m;
            rp.low = p.low;
            rp.op_begin = l.begin;
            rp.n_ops = p.n_ops;
            rp.n_slots = slots;
            rp.phase_begin = (int)rphases.size();
            std::memcpy(rp.pos, p.pos, sizeof(rp.pos));
            std::vector<int> pair(p.n_ops);
            for (int k = 0; k < p.n_ops; ++k)
                pair[k] = rsops[l.begin + k].kind == kOpPair ? rsops[l.begin + k].target : -1;
            constexpr int A = (1 << C::SweepBits) / kBatchRegThreads;
            const std::vector<PhaseLayout> L =
                plan_phase_layouts(pair, p.m, Log2<A>::v, Log2<kBatchRegThreads / 64>::v);
            for (const PhaseLayout &P : L) {
                for (int k = P.begin; k < P.end; ++k) {
                    SweepOp<R> &r = rsops[l.begin + k];
                    const RegSplit sp = split_for_phase(P, r.kind, r.target, r.lctrl);
                    r.rkind = sp.kind;
                    r.rtarget = sp.target;
                    r.creg = sp.creg;
                    r.cthr = sp.cthr;
                }
                rphases.push_back(phase_bits(P));
            }
            rp.n_phases = (int)L.size();
            QG_REQUIRE(rp.n_ops == 0 || rp.n_phases > 0, "batch: register sweep pass without phases");
            g.rspasses.push_back(rp);
        }
    }
    if (rphases.empty()) rphases.push_back(SweepPhase());
    if (slot_out.empty()) slot_out.push_back(0);
    // Pauli passes
    std::vector<u64> pneed;
    for (const PauliGroup &pg : g.pgroups) pneed.push_back(pg.xmask);
    std::vector<Layout> pl = plan_layouts(pneed, n, mf, lowf, c64);
    std::vector<PauliGroup> pgroups(g.pgroups);
    g.ppasses.clear();
    for (size_t k = 0; k < pl.size(); ++k) {
        int local[64];
        BPass p = make_pass(pl[k], local);
        for (int i = pl[k].begin; i < pl[k].end; ++i) {
            unsigned xl = 0;
            for (int b = 0; b < n; ++b)
                if ((pgroups[i].xmask >> b) & 1) xl |= 1u << local[b];
            pgroups[i].xlocal = xl;
        }
        p.accumulate = k > 0 ? 1 : 0;
        p.n_slots = k + 1 == pl.size() ? 1 : 0;
        g.ppasses.push_back(p);
    }
    std::vector<BTermMask> tmask;
    for (u64 z : g.zmask) tmask.push_back(BTermMask{z});
    if (tmask.empty()) tmask.push_back(BTermMask{0});
    // parameter -> rotations (sweep order)
    std::vector<int> pptr(g.n_params + 1, 0), pgate;
    for (int p = 0; p < g.n_params; ++p) {
        pptr[p] = (int)pgate.size();
        for (int gi = n_gates - 1; gi >= 0; --gi)
            if (g.ops[gi].kind == 1 && g.ops[gi].param == p) pgate.push_back(gi);
    }
    pptr[g.n_params] = (int)pgate.size();
    if (pgate.empty()) pgate.push_back(0);
    std::vector<BGate> gates(std::max(1, n_gates));
    for (int i = 0; i < n_gates; ++i) {
        gate_matrix_to_array(g.ops[i].mat, gates[i].mat);
        gates[i].pauli = g.ops[i].kind == 1 ? g.ops[i].pauli : 0;
        gates[i].param = g.ops[i].param;
    }

    // device memory
    const size_t state_bytes = (amp_bytes(Unit<R>::W == 2 ? precFP32 : precFP64) << n) * ni;
    if (g.psi == nullptr || g.state_bytes != state_bytes) {
        free_group(g);
        g.state_bytes = state_bytes;
        g.psi = dev_alloc(g.state_bytes);
        g.lam = dev_alloc(g.state_bytes);
    } else if (g.block) {
        dev_free(g.block, g.block_bytes);
        g.block = nullptr;
    }
    u64 max_part = 0;
    for (const BPass &p : g.spasses)
        max_part = std::max<u64>(max_part, (u64)ni * (u64)max_slots * (u64(1) << (n - p.m)));
    max_part = std::max<u64>(max_part, (u64)ni << (n - g.ppasses.back().m));
    const size_t n_terms = g.zmask.size();
    std::vector<std::pair<size_t, size_t>> parts;   // (offset, bytes)
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        off = (off + 255) & ~size_t(255);
        const size_t at = off;
        off += std::max<size_t>(bytes, 8);
        return at;
    };
    const size_t o_params = carve(sizeof(double) * ni * std::max(1, g.n_params));
    const size_t o_coefs = carve(sizeof(double) * ni * std::max<size_t>(1, n_terms) * 2);
    const size_t o_part = carve(sizeof(double) * max_part);
    const size_t o_terms = carve(sizeof(double) * ni * std::max(1, n_gates));
    const size_t o_values = carve(sizeof(double) * ni);
    const size_t o_grads = carve(sizeof(double) * ni * std::max(1, g.n_params));
    const size_t o_mats = carve(sizeof(BMats<R>) * ni * std::max(1, n_gates));
    const size_t o_gates = carve(sizeof(BGate) * gates.size());
    const size_t o_fops = carve(sizeof(BOp) * fops.size());
    const size_t o_sops = carve(sizeof(BOp) * sops.size());
    const size_t o_rsops = carve(sizeof(SweepOp<R>) * rsops.size());
    const size_t o_rph = carve(sizeof(SweepPhase) * rphases.size());
    const size_t o_pg = carve(sizeof(PauliGroup) * std::max<size_t>(1, pgroups.size()));
    const size_t o_tm = carve(sizeof(BTermMask) * tmask.size());
    const size_t o_slot = carve(sizeof(int) * slot_out.size());
    const size_t o_pptr = carve(sizeof(int) * pptr.size());
    const size_t o_pgate = carve(sizeof(int) * pgate.size());
    const size_t o_zero = carve(sizeof(int));
    g.block_bytes = off;
    g.block = static_cast<char *>(dev_alloc(g.block_bytes));
    g.d_params = reinterpret_cast<double *>(g.block + o_params);
    g.d_coefs = reinterpret_cast<double *>(g.block + o_coefs);
    g.d_partial = reinterpret_cast<double *>(g.block + o_part);
    g.d_terms = reinterpret_cast<double *>(g.block + o_terms);
    g.d_values = reinterpret_cast<double *>(g.block + o_values);
    g.d_grads = reinterpret_cast<double *>(g.block + o_grads);
    g.d_mats = g.block + o_mats;
    g.d_gates = reinterpret_cast<BGate *>(g.block + o_gates);
    g.d_fops = reinterpret_cast<BOp *>(g.block + o_fops);
    g.d_sops = reinterpret_cast<BOp *>(g.block + o_sops);
    g.d_rsops = g.block + o_rsops;
    g.d_rphases = reinterpret_cast<SweepPhase *>(g.block + o_rph);
    g.d_pgroups = reinterpret_cast<PauliGroup *>(g.block + o_pg);
    g.d_tmask = reinterpret_cast<BTermMask *>(g.block + o_tm);
    g.d_slot_out = reinterpret_cast<int *>(g.block + o_slot);
    g.d_pptr = reinterpret_cast<int *>(g.block + o_pptr);
    g.d_pgate = reinterpret_cast<int *>(g.block + o_pgate);
    g.d_zero = reinterpret_cast<int *>(g.block + o_zero);
    hipStream_t s = stream();
    const int zero = 0;
    QG_HIP(hipMemcpyAsync(g.d_gates, gates.data(), sizeof(BGate) * gates.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_fops, fops.data(), sizeof(BOp) * fops.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_sops, sops.data(), sizeof(BOp) * sops.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_rsops, rsops.data(), sizeof(SweepOp<R>) * rsops.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_rphases, rphases.data(), sizeof(SweepPhase) * rphases.size(), hipMemcpyHostToDevice,
                          s));
    if (!pgroups.empty())
        QG_HIP(hipMemcpyAsync(g.d_pgroups, pgroups.data(), sizeof(PauliGroup) * pgroups.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_tmask, tmask.data(), sizeof(BTermMask) * tmask.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_slot_out, slot_out.data(), sizeof(int) * slot_out.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_pptr, pptr.data(), sizeof(int) * pptr.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_pgate, pgate.data(), sizeof(int) * pgate.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_zero, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    QG_HIP(hipMemcpyAsync(g.d_coefs, g.coefs.data(), sizeof(double) * g.coefs.size(), hipMemcpyHostToDevice, s));
    QG_HIP(hipStreamSynchronize(s));   // the tables above are host temporaries
    g.fregs = fregs;
    g.sregs = sregs;
    g.tune = t;
}

template <class R, int T>
void launch_group_t(Group &g) {
    typedef BCfg<R> C;
    constexpr int W = Unit<R>::W;
    hipStream_t s = stream();
    const int n = g.n;
    const int n_gates = (int)g.ops.size();
    const u64 ni = g.n_items();
    const unsigned grid_cap = (unsigned)std::max(1, runtime().n_cus * g.tune.grid_per_cu);
    BMats<R> *mats = static_cast<BMats<R> *>(g.d_mats);
    Cx<R> *psi = static_cast<Cx<R> *>(g.psi), *lam = static_cast<Cx<R> *>(g.lam);
    const u64 nm = ni * (u64)std::max(1, n_gates);
    if (n_gates > 0)
        hipLaunchKernelGGL((k_batch_mats<R>), (unsigned)((nm + 255) / 256), 256, 0, s, g.d_gates, n_gates, g.d_params,
                           g.n_params, ni, mats);
    const u64 n_units = (ni << n) / W;
    const int unit_bits = n - (W == 2 ? 1 : 0);
    hipLaunchKernelGGL((k_batch_zero<R>), (unsigned)((n_units + 255) / 256), 256, 0, s, psi, n_units, unit_bits);
    const Mat2<R> *umats = reinterpret_cast<const Mat2<R> *>(mats);   // U of (item, gate): umats[2 (item n_gates + gate)]
    const u64 piece = launch_piece_items() / kBatchRegThreads;         // register passes: one tile per workgroup
    for (size_t k = 0; k < g.fpasses.size(); ++k) {
        const BPass &p = g.fpasses[k];
        if (p.n_ops == 0) continue;
        const u64 n_tiles = ni << (n - p.m);
        if (g.fregs) {
            const RegProgram<R> &rp = reg_programs<R>(g)[k];
            for (u64 first = 0; first < n_tiles; first += piece)
                hipLaunchKernelGGL((k_tile_reg<R, TileCfg<R>::MaxBits, kBatchRegThreads, true>),
                                   (unsigned)std::min(piece, n_tiles - first), kBatchRegThreads, 0, s, psi, rp, first,
                                   umats, n_gates, n);
            continue;
        }
        const unsigned grid = (unsigned)std::min<u64>(n_tiles, grid_cap);
        hipLaunchKernelGGL((k_batch_pass<R, C::MaxBits, T>), grid, T, 0, s, psi, p, g.d_fops, mats, n_gates, n, n_tiles);
    }
    for (const BPass &p : g.ppasses) {
        const u64 n_tiles = ni << (n - p.m);
        const unsigned grid = (unsigned)std::min<u64>(n_tiles, grid_cap);
        hipLaunchKernelGGL((k_batch_pauli<R, C::MaxBits, T>), grid, T, 0, s, psi, lam, p, g.d_pgroups, g.d_tmask,
                           g.d_coefs, (int)g.zmask.size(), n, g.d_partial, n_tiles);
    }
    {
        const BPass &p = g.ppasses.back();
        const u64 len = u64(1) << (n - p.m);
        hipLaunchKernelGGL(k_rows_sum, (unsigned)ni, 64, 0, s, g.d_partial, len, ni, ni, g.d_zero, 1, g.d_values);
    }
    for (size_t k = 0; k < g.spasses.size(); ++k) {
        const BPass &p = g.spasses[k];
        if (p.n_ops == 0) continue;
        const u64 tpi = u64(1) << (n - p.m);
        const u64 n_tiles = ni * tpi;
        if (g.sregs) {
            const SweepPass &rp = g.rspasses[k];
            const Mat2<R> *udmats = umats + 1;   // U^dagger entries
            for (u64 first = 0; first < n_tiles; first += piece)
                hipLaunchKernelGGL((k_tile_sweep_reg<R, C::SweepBits, kBatchRegThreads, true>),
                                   (unsigned)std::min(piece, n_tiles - first), kBatchRegThreads, 0, s, psi, lam, rp,
                                   static_cast<const SweepOp<R> *>(g.d_rsops), g.d_rphases, g.d_partial, n_tiles,
                                   first, udmats, n_gates, n);
        } else {
            const unsigned grid = (unsigned)std::min<u64>(n_tiles, grid_cap);
            hipLaunchKernelGGL((k_batch_sweep<R, C::SweepBits, T>), grid, T, 0, s, psi, lam, p, g.d_sops, mats,
                               n_gates, n, g.d_partial, n_tiles);
        }
        if (p.n_slots > 0)
            hipLaunchKernelGGL(k_rows_sum, (unsigned)(ni * (u64)p.n_slots), 64, 0, s, g.d_partial, tpi,
                               ni * (u64)p.n_slots, ni, g.d_slot_out + g.spass_slot_begin[k], n_gates, g.d_terms);
    }
    if (g.n_params > 0) {
        const u64 nq = ni * (u64)g.n_params;
        hipLaunchKernelGGL(k_batch_grads, (unsigned)((nq + 255) / 256), 256, 0, s, g.d_terms, n_gates, g.d_pptr,
                           g.d_pgate, g.n_params, ni, g.d_grads);
    }
}

template <class R>
void launch_group(Group &g) {
    if (g.tune.threads == 128)
        launch_group_t<R, 128>(g);
    else
        launch_group_t<R, 256>(g);
    QG_HIP(hipGetLastError());
}

void upload_params(Group &g) {
    if (g.params.empty()) return;
    QG_HIP(hipMemcpyAsync(g.d_params, g.params.data(), sizeof(double) * g.params.size(), hipMemcpyHostToDevice,
                          stream()));
}

template <class R>
double time_group(Group &g, int reps) {
    hipStream_t s = stream();
    upload_params(g);
    launch_group<R>(g);   // warm-up
    hipEvent_t e0, e1;
    QG_HIP(hipEventCreate(&e0));
    QG_HIP(hipEventCreate(&e1));
    QG_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) launch_group<R>(g);
    QG_HIP(hipEventRecord(e1, s));
    QG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    QG_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return ms / reps;
}

// Autotuner: candidates are timed on the group's own structure and a sample of its items, in
// two stages. First the tile shape: register-resident passes (full-size tiles; groups of at
// least the sweep tile's qubit count) and LDS passes of up to three tile sizes, at 256
// threads and 4 workgroups per compute unit. Then, for the fastest shape, the workgroup size
// and persistent grid of the LDS passes (every group has Pauli passes in LDS). Knobs fixed
// by qghip_batch_set_tuning are not searched.
template <class R>
void tune_group(qghip_batch *b, Group &g) {
    typedef BCfg<R> C;
    std::vector<std::pair<int, int>> shapes;   // (tile bits, registers)
    auto add_shape = [&](int tb, int rg) {
        if (std::find(shapes.begin(), shapes.end(), std::make_pair(tb, rg)) == shapes.end())
            shapes.push_back(std::make_pair(tb, rg));
    };
    const bool reg_ok = use_register_tiles() && g.n >= C::SweepBits && b->fixed_registers != 0 &&
                        (b->fixed_tile_bits == 0 || b->fixed_tile_bits == C::MaxBits);
    if (reg_ok) add_shape(C::MaxBits, 1);
    if (!reg_ok || b->fixed_registers != 1) {
        if (b->fixed_tile_bits > 0) {
            add_shape(std::min(b->fixed_tile_bits, C::MaxBits), 0);
        } else {
            for (int d = 0; d < 3; ++d) {
                const int v = std::min(C::MaxBits - d, g.n);
                if (v >= 1) add_shape(v, 0);
            }
        }
    }
    const std::vector<int> threads = b->fixed_threads > 0 ? std::vector<int>{b->fixed_threads} : std::vector<int>{256, 128};
    const std::vector<int> grids = b->fixed_grid > 0 ? std::vector<int>{b->fixed_grid} : std::vector<int>{4, 8, 2};
    // the sample: the structure is the same for every item, so the ranking carries over; it
    // holds at most 2^24 amplitudes (and 64 items), which still spans several workgroups per
    // compute unit
    Group sample;
    sample.n = g.n;
    sample.key = g.key;
    sample.ops = g.ops;
    sample.n_params = g.n_params;
    sample.xmask = g.xmask;
    sample.zmask = g.zmask;
    sample.pgroups = g.pgroups;
    const size_t cap = std::max<size_t>(8, (size_t(1) << 24) >> std::min(g.n, 24));
    const size_t keep = std::min<size_t>(g.n_items(), std::min<size_t>(64, cap));
    sample.items.assign(g.items.begin(), g.items.begin() + keep);
    sample.params.assign(g.params.begin(), g.params.begin() + keep * (size_t)g.n_params);
    sample.coefs.assign(g.coefs.begin(), g.coefs.begin() + keep * g.zmask.size() * 2);
    Tuning best;
    best.ms = -1.;
    auto try_tuning = [&](int tb, int rg, int th, int gr) {
        Tuning t;
        t.tile_bits = tb;
        t.registers = rg;
        t.threads = th;
        t.grid_per_cu = gr;
        build_group<R>(sample, t);
        t.ms = time_group<R>(sample, 2);
        if (best.ms < 0. || t.ms < best.ms) best = t;
    };
    try {
        for (const auto &sh : shapes) try_tuning(sh.first, sh.second, threads[0], grids[0]);
        const Tuning shape = best;
        for (int th : threads)
            for (int gr : grids)
                if (th != threads[0] || gr != grids[0]) try_tuning(shape.tile_bits, shape.registers, th, gr);
    } catch (...) {
        free_group(sample);
        throw;
    }
    free_group(sample);
    build_group<R>(g, best);
    g.tune = best;
}

template <class R>
void prepare(qghip_batch *b) {
    if (b->built) return;
    drop_graph(b);
    for (Group &g : b->groups) {
        if (!b->tuned || g.tune.tile_bits == 0)   // new settings, or a group added since tuning
            tune_group<R>(b, g);
        else
            build_group<R>(g, g.tune);
    }
    b->tuned = true;
    // capture every group's launch sequence into one graph
    hipStream_t s = stream();
    QG_HIP(hipStreamSynchronize(s));
    QG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    try {
        for (Group &g : b->groups) launch_group<R>(g);
    } catch (...) {
        hipGraph_t tmp = nullptr;
        hipStreamEndCapture(s, &tmp);
        if (tmp) hipGraphDestroy(tmp);
        throw;
    }
    QG_HIP(hipStreamEndCapture(s, &b->graph));
    QG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->built = true;
}

// Fold (-i)^(number of Y) into each coefficient; sort terms by X mask into groups.
void set_observable(Group &g, int n_terms, const double *coefs, const u64 *xm, const u64 *zm, bool first,
                    std::vector<double> &folded) {
    std::vector<int> order(n_terms);
    for (int i = 0; i < n_terms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return xm[a] < xm[c]; });
    if (first) {
        g.xmask.clear();
        g.zmask.clear();
        g.pgroups.clear();
        for (int idx : order) {
            g.xmask.push_back(xm[idx]);
            g.zmask.push_back(zm[idx]);
            if (g.pgroups.empty() || g.pgroups.back().xmask != xm[idx])
                g.pgroups.push_back(PauliGroup{xm[idx], (int)g.xmask.size() - 1, 0, 0u});
            g.pgroups.back().count += 1;
        }
    }
    static const double pr[4] = {1., 0., -1., 0.}, pi[4] = {0., -1., 0., 1.};
    folded.clear();
    for (int idx : order) {
        const int ny = __builtin_popcountll(xm[idx] & zm[idx]) & 3;
        folded.push_back(coefs[idx] * pr[ny]);
        folded.push_back(coefs[idx] * pi[ny]);
    }
}

bool same_observable(const Group &g, int n_terms, const u64 *xm, const u64 *zm) {
    if ((int)g.xmask.size() != n_terms) return false;
    std::vector<int> order(n_terms);
    for (int i = 0; i < n_terms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return xm[a] < xm[c]; });
    for (int i = 0; i < n_terms; ++i)
        if (g.xmask[i] != xm[order[i]] || g.zmask[i] != zm[order[i]]) return false;
    return true;
}

}  // namespace

extern "C" {

int qghip_batch_new(int prec, qghip_batch **out) {
    return guarded_batch([&] {
        require(out != nullptr, "null argument");
        require(prec == QGHIP_FP64 || prec == QGHIP_FP32, "bad precision");
        qghip_batch *b = new qghip_batch;
        b->prec = prec == QGHIP_FP64 ? precFP64 : precFP32;
        *out = b;
    });
}

int qghip_batch_delete(qghip_batch *b) {
    return guarded_batch([&] {
        if (b == nullptr) return;
        drop_graph(b);
        for (Group &g : b->groups) free_group(g);
        delete b;
    });
}

int qghip_batch_add(qghip_batch *b, const qghip_circuit *c, const double *params, int n_params, int n_terms,
                    const double *coefs, const uint64_t *xmask, const uint64_t *zmask, int *item) {
    return guarded_batch([&] {
        require(b != nullptr && c != nullptr && item != nullptr, "null argument");
        require(c->prec == b->prec, "circuit precision does not match the batch");
        require(n_params >= c->n_params && (c->n_params == 0 || params != nullptr), "too few parameters");
        require(n_terms >= 1 && coefs && xmask && zmask, "empty observable");
        const u64 lim = (u64(1) << c->n_qubits) - 1;
        for (int i = 0; i < n_terms; ++i)
            require((xmask[i] & ~lim) == 0 && (zmask[i] & ~lim) == 0, "Pauli term outside the register");
        require(c->n_qubits <= 30, "batched circuits are limited to 30 qubits");
        int gi = -1;
        for (int k = 0; k < (int)b->groups.size(); ++k) {
            const Group &g = b->groups[k];
            if (g.key == c && g.n == c->n_qubits && (int)g.ops.size() == (int)c->ops.size() &&
                same_observable(g, n_terms, xmask, zmask)) {
                gi = k;
                break;
            }
        }
        if (gi < 0) {
            Group g;
            g.n = c->n_qubits;
            g.key = c;
            g.ops = c->ops;
            g.n_params = c->n_params;
            std::vector<double> tmp;
            set_observable(g, n_terms, coefs, xmask, zmask, true, tmp);
            b->groups.push_back(g);
            gi = (int)b->groups.size() - 1;
        }
        Group &g = b->groups[gi];
        std::vector<double> folded;
        set_observable(g, n_terms, coefs, xmask, zmask, false, folded);
        g.coefs.insert(g.coefs.end(), folded.begin(), folded.end());
        g.params.insert(g.params.end(), params, params + g.n_params);
        g.items.push_back((int)b->items.size());
        b->items.push_back(std::make_pair(gi, (int)g.items.size() - 1));
        b->built = false;
        *item = (int)b->items.size() - 1;
    });
}

int qghip_batch_set_params(qghip_batch *b, int item, const double *params, int n_params) {
    return guarded_batch([&] {
        require(b != nullptr && params != nullptr, "null argument");
        require(0 <= item && item < (int)b->items.size(), "item out of range");
        Group &g = b->groups[b->items[item].first];
        require(n_params >= g.n_params, "too few parameters");
        std::copy(params, params + g.n_params, g.params.begin() + (size_t)b->items[item].second * g.n_params);
    });
}

int qghip_batch_set_tuning(qghip_batch *b, int tile_bits, int threads, int grid_per_cu, int registers) {
    return guarded_batch([&] {
        require(b != nullptr, "null argument");
        require(registers >= -1 && registers <= 1, "registers must be -1, 0 or 1");
        require(tile_bits == 0 || (1 <= tile_bits && tile_bits <= 12), "tile_bits out of range");
        require(threads == 0 || threads == 128 || threads == 256, "threads must be 128 or 256");
        require(grid_per_cu >= 0 && grid_per_cu <= 64, "grid_per_cu out of range");
        b->fixed_tile_bits = tile_bits;
        b->fixed_threads = threads;
        b->fixed_grid = grid_per_cu;
        b->fixed_registers = registers;
        b->tuned = false;
        b->built = false;
    });
}

int qghip_batch_run(qghip_batch *b, int use_graph, double *values, double *grads, int64_t grads_size) {
    return guarded_batch([&] {
        require(b != nullptr && values != nullptr, "null argument");
        int64_t total = 0;
        for (const Group &g : b->groups) total += (int64_t)g.n_items() * g.n_params;
        require(grads == nullptr || grads_size >= total, "gradient array too small");
        if (b->prec == precFP32)
            prepare<float>(b);
        else
            prepare<double>(b);
        hipStream_t s = stream();
        for (Group &g : b->groups) upload_params(g);
        if (use_graph) {
            QG_HIP(hipGraphLaunch(b->exec, s));
        } else {
            for (Group &g : b->groups) {
                if (b->prec == precFP32)
                    launch_group<float>(g);
                else
                    launch_group<double>(g);
            }
        }
        // results, item by item in batch order
        std::vector<std::vector<double>> gv(b->groups.size()), gg(b->groups.size());
        for (size_t k = 0; k < b->groups.size(); ++k) {
            Group &g = b->groups[k];
            gv[k].resize(g.n_items());
            gg[k].resize(g.n_items() * (size_t)g.n_params);
            QG_HIP(hipMemcpyAsync(gv[k].data(), g.d_values, sizeof(double) * gv[k].size(), hipMemcpyDeviceToHost, s));
            if (!gg[k].empty())
                QG_HIP(hipMemcpyAsync(gg[k].data(), g.d_grads, sizeof(double) * gg[k].size(), hipMemcpyDeviceToHost, s));
        }
        QG_HIP(hipStreamSynchronize(s));
        // grads are laid out per item in batch order, each item with its circuit's n_params entries
        int64_t off = 0;
        for (size_t i = 0; i < b->items.size(); ++i) {
            const int gi = b->items[i].first, k = b->items[i].second;
            const Group &g = b->groups[gi];
            values[i] = gv[gi][k];
            if (grads) {
                std::copy(gg[gi].begin() + (size_t)k * g.n_params, gg[gi].begin() + (size_t)(k + 1) * g.n_params,
                          grads + off);
            }
            off += g.n_params;
        }
    });
}

int qghip_batch_describe(const qghip_batch *b, char *buf, int size) {
    return guarded_batch([&] {
        require(b != nullptr && buf != nullptr && size > 0, "null argument");
        std::string out = "[";
        for (size_t k = 0; k < b->groups.size(); ++k) {
            const Group &g = b->groups[k];
            char line[512];
            snprintf(line, sizeof(line),
                     "%s{\"qubits\": %d, \"items\": %zu, \"gates\": %zu, \"tile_bits\": %d, \"threads\": %d, "
                     "\"grid_per_cu\": %d, \"registers\": %d, \"forward_in_registers\": %d, "
                     "\"sweep_in_registers\": %d, \"tuned_ms\": %.4f, \"forward_passes\": %zu, "
                     "\"sweep_passes\": %zu}",
                     k ? ", " : "", g.n, g.n_items(), g.ops.size(), g.tune.tile_bits, g.tune.threads,
                     g.tune.grid_per_cu, g.tune.registers, g.fregs ? 1 : 0, g.sregs ? 1 : 0, g.tune.ms,
                     g.fpasses.size(), g.spasses.size());
            out += line;
        }
        out += "]";
        require((int)out.size() < size, "buffer too small");
        std::memcpy(buf, out.c_str(), out.size() + 1);
    });
}

}  // extern "C"
