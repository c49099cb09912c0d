// State kernels: |0...0> init, P(bit = 0), collapse, collapse-and-split, reset, join.
//
// Reference semantics (qgate/simulator/src/CPUQubitProcessor.cpp):
//   resetQubitStates      :61-73    a = |0...0>
//   calcProbability       :75-125   sum of |a_i|^2 over i with bit `lane` = 0
//   join                  :129-213  Kronecker product, last source in the lowest bits
//   decohere              :217-246  kept half *= 1/sqrt(p) (or 1/sqrt(1-p)), other half = 0
//   decohereAndSeparate   :248-284  kept half * norm -> new (n-1)-lane vector
//   applyReset            :286-305  a[i0] = a[i1], a[i1] = 0
// Every pass streams 16-byte units (2 complex64 or 1 complex128 amplitudes) with
// non-temporal hints. Probability sums accumulate in FP64 whatever the state precision, in a
// fixed order (per-block partials, then one block over the partials), so repeated calc_prob
// calls on the same state return the same bits. decohere_prob (a recorded collapse fused with
// the next probability) sums per-wavefront partials instead: the same terms in another fixed
// order, so its result may differ from a separate calc_prob in the last bits.
#include <algorithm>

#include "kernels.hpp"
#include "qg_runtime.hpp"

namespace qg {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

// a[i] = (i == 0)
template <class R>
__global__ __launch_bounds__(kBlock) void k_zero_state(Cx<R> *a, u64 n_units, u64 first) {
    const u64 u = first + (u64)blockIdx.x * kBlock + threadIdx.x;
    if (u >= n_units) return;
    Amps<R> v;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) v.v[j] = cx<R>(0, 0);
    if (u == 0) v.v[0].re = R(1);
    store_unit(a, u, v);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of `v` in a fixed order; valid in thread 0.
template <int B>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[B / 64];
    v = wave_sum(v);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double s = 0.;
    if (threadIdx.x == 0)
        for (int i = 0; i < B / 64; ++i) s += part[i];
    return s;
}

constexpr int kProbUnitsPerThread = 16;

// Partial sums of |a|^2 over the units of the half with bit `lane` = 0 (lane at unit level:
// a zero bit inserted into the unit index), or over every unit counting only the amplitudes
// with in-unit bit 0 clear (amp_mask = 1, complex64 lane 0).
template <class R>
__global__ __launch_bounds__(kBlock) void k_prob_partial(const Cx<R> *__restrict__ a, BitInsert ins,
                                                         int amp_mask, u64 n_units, double *partial) {
    double s = 0.;
    const u64 base = (u64)blockIdx.x * (kBlock * kProbUnitsPerThread) + threadIdx.x;
#pragma unroll 4
    for (int r = 0; r < kProbUnitsPerThread; ++r) {
        const u64 w = base + (u64)r * kBlock;
        if (w < n_units) {
            const Amps<R> v = load_unit(a, insert_zero_bits(w, ins));
#pragma unroll
            for (int j = 0; j < Unit<R>::W; ++j)
                if ((j & amp_mask) == 0) s += abs2d(v.v[j]);
        }
    }
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[k] = sum of partial[k * kSumRun ...] over one run of kSumRun partials, in a fixed order.
constexpr int kSumRun = 8192;

__global__ __launch_bounds__(1024) void k_sum_runs(const double *partial, u64 n, double *out) {
    const u64 base = (u64)blockIdx.x * kSumRun;
    double s = 0.;
#pragma unroll
    for (int r = 0; r < kSumRun / 1024; ++r) {
        const u64 i = base + (u64)r * 1024 + threadIdx.x;
        if (i < n) s += partial[i];
    }
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_sum_partials(const double *partial, u64 n, double *out) {
    double s = 0.;
    for (u64 i = threadIdx.x; i < n; i += 1024) s += partial[i];
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) *out = s;
}

// Collapse in place. Lane at unit level (dist = unit distance of the lane bit): the kept unit
// is scaled, its partner zeroed. Lane inside the unit (dist == 0): per amplitude.
template <class R>
__global__ __launch_bounds__(kWave) void k_decohere(Cx<R> *__restrict__ a, BitInsert ins, u64 dist,
                                                    int value, R norm, u64 n_items, u64 first) {
    const u64 w = first + (u64)blockIdx.x * kWave + threadIdx.x;
    if (w >= n_items) return;
    if (dist == 0) {   // complex64 lane 0: bit 0 of the amplitude inside the unit
        Amps<R> v = load_unit(a, w);
#pragma unroll
        for (int j = 0; j < Unit<R>::W; ++j)
            v.v[j] = (j == value) ? cx<R>(v.v[j].re * norm, v.v[j].im * norm) : cx<R>(0, 0);
        store_unit(a, w, v);
        return;
    }
    if (ins.n == 0) {
        // unit bit `dist` below the wavefront's 64 units: thread w owns unit w and scales or
        // clears it, so each store instruction writes whole lines (a keep/drop pair per
        // thread wrote half lines twice, at about twice the time)
        Amps<R> v;
        if (((w & dist) != 0) == (value != 0)) {
            v = load_unit(a, w);
#pragma unroll
            for (int j = 0; j < Unit<R>::W; ++j) v.v[j] = cx<R>(v.v[j].re * norm, v.v[j].im * norm);
        } else {
#pragma unroll
            for (int j = 0; j < Unit<R>::W; ++j) v.v[j] = cx<R>(0, 0);
        }
        store_unit(a, w, v);
        return;
    }
    const u64 u0 = insert_zero_bits(w, ins);
    const u64 keep = value ? (u0 | dist) : u0, drop = value ? u0 : (u0 | dist);
    Amps<R> v = load_unit(a, keep), z;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) {
        v.v[j] = cx<R>(v.v[j].re * norm, v.v[j].im * norm);
        z.v[j] = cx<R>(0, 0);
    }
    store_unit(a, keep, v);
    store_unit(a, drop, z);
}

// k_decohere's collapse (same three forms, one item per thread) and, on the collapsed
// amplitudes as stored, the sum per wavefront of |a|^2 over the amplitudes whose unit index has
// no bit of `p_unit` set and whose in-unit index has no bit of `p_amp` set: the probability of
// the next lane, in the same pass. One item per thread in single-wavefront blocks with no
// barrier keeps the collapse's own speed (a 4 GiB collapse: 1.02 ms, against 1.18 ms with 8
// items per thread in 256-thread blocks; the FP64 sum itself costs nothing measurable).
template <class R>
__global__ __launch_bounds__(kWave) void k_decohere_prob(Cx<R> *__restrict__ a, BitInsert ins, u64 dist,
                                                         int value, R norm, u64 n_items, u64 first, u64 p_unit,
                                                         int p_amp, double *partial) {
    const u64 w = first + (u64)blockIdx.x * kWave + threadIdx.x;
    double s = 0.;
    if (w < n_items) {   // no early return: the whole wavefront takes part in the sum
        Amps<R> v;
        u64 unit = w;
        bool kept = true;
        if (dist != 0 && ins.n != 0) {   // the pair (u0, u0 | dist)
            const u64 u0 = insert_zero_bits(w, ins);
            unit = value ? (u0 | dist) : u0;
            Amps<R> z;
#pragma unroll
            for (int j = 0; j < Unit<R>::W; ++j) z.v[j] = cx<R>(0, 0);
            store_unit(a, value ? u0 : (u0 | dist), z);
        } else if (dist != 0) {          // unit w, kept or cleared
            kept = ((w & dist) != 0) == (value != 0);
        }
        if (kept) {
            v = load_unit(a, unit);
#pragma unroll
            for (int j = 0; j < Unit<R>::W; ++j) {
                const bool on = dist != 0 || j == value;
                v.v[j] = on ? cx<R>(v.v[j].re * norm, v.v[j].im * norm) : cx<R>(0, 0);
            }
            if ((unit & p_unit) == 0) {
#pragma unroll
                for (int j = 0; j < Unit<R>::W; ++j)
                    if ((j & p_amp) == 0) s += abs2d(v.v[j]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < Unit<R>::W; ++j) v.v[j] = cx<R>(0, 0);
        }
        store_unit(a, unit, v);
    }
    s = wave_sum(s);
    if (threadIdx.x == 0) partial[first / kWave + blockIdx.x] = s;
}

// dst[k] = norm * src[k with a `value` bit inserted at `lane`].
template <class R>
__global__ __launch_bounds__(kWave) void k_separate(Cx<R> *__restrict__ dst, const Cx<R> *__restrict__ src,
                                                    BitInsert ins, u64 vbit, int value, R norm,
                                                    u64 n_units, u64 first) {
    const u64 u = first + (u64)blockIdx.x * kWave + threadIdx.x;
    if (u >= n_units) return;
    Amps<R> v;
    bool gathered = false;
    if constexpr (Unit<R>::W == 2) {
        if (ins.n == 0) {   // complex64 lane 0: dst unit u takes amplitude `value` of src units 2u, 2u+1
            const Amps<R> s0 = load_unit(src, 2 * u), s1 = load_unit(src, 2 * u + 1);
            const Cx<R> a0 = s0.v[value], a1 = s1.v[value];
            v.v[0] = cx<R>(norm * a0.re, norm * a0.im);
            v.v[1] = cx<R>(norm * a1.re, norm * a1.im);
            gathered = true;
        }
    }
    if (!gathered) {
        v = load_unit(src, insert_zero_bits(u, ins) | vbit);
#pragma unroll
        for (int j = 0; j < Unit<R>::W; ++j) v.v[j] = cx<R>(norm * v.v[j].re, norm * v.v[j].im);
    }
    store_unit(dst, u, v);
}

template <class R>
__global__ __launch_bounds__(kWave) void k_reset(Cx<R> *__restrict__ a, BitInsert ins, u64 dist,
                                                 u64 n_items, u64 first) {
    const u64 w = first + (u64)blockIdx.x * kWave + threadIdx.x;
    if (w >= n_items) return;
    if (dist == 0) {   // complex64 lane 0
        Amps<R> v = load_unit(a, w);
        v.v[0] = v.v[Unit<R>::W - 1];
        v.v[Unit<R>::W - 1] = cx<R>(0, 0);
        store_unit(a, w, v);
        return;
    }
    const u64 u0 = insert_zero_bits(w, ins);
    const Amps<R> v = load_unit(a, u0 | dist);
    Amps<R> z;
#pragma unroll
    for (int j = 0; j < Unit<R>::W; ++j) z.v[j] = cx<R>(0, 0);
    store_unit(a, u0, v);
    store_unit(a, u0 | dist, z);
}

// A Pauli string P = (x) sigma_k over the lanes of x | z (X: x bit, Z: z bit, Y: both) acts
// on a basis state as P|y> = c (-1)^|y & z| |y ^ x>, c = i^#Y, so
//   (P psi)_y = c (-1)^|(y ^ x) & z| psi_(y ^ x).
// The kernels below visit each pair (y0, y1 = y0 ^ x) once (y0 with bit h = top bit of x
// clear) and use that P is Hermitian: ((1 + P) psi)_y0 and ((1 + P) psi)_y1 have equal moduli.
//   mode 0: x == 0 (Z string): diagonal, every unit, each amplitude on its own;
//   mode 1: complex64 with x == 1: the pair is the two amplitudes of one unit;
//   mode 2: the pair spans two units u0, u1 = u0 ^ xu, amplitude j of u0 with j ^ xj of u1.
struct PauliArgs {
    u64 z;
    u64 xu;           // x at unit level
    BitInsert ins;    // zero bit at unit level h - lw (mode 2)
    int xj;           // x & (W - 1)
    int mode;
    double cr, ci;    // c = cr + i ci
};

__device__ __forceinline__ double parity_sign(u64 y, u64 z) { return (__popcll(y & z) & 1) ? -1. : 1.; }

// |a0 + c (-1)^|y1 & z| a1|^2 in FP64
template <class R>
__device__ __forceinline__ double pauli_pair_weight(const Cx<R> &a0, const Cx<R> &a1, u64 y1, const PauliArgs &P) {
    const double sg = parity_sign(y1, P.z);
    const double br = (double)a1.re, bi = (double)a1.im;
    const double tr = (double)a0.re + sg * (P.cr * br - P.ci * bi);
    const double ti = (double)a0.im + sg * (P.cr * bi + P.ci * br);
    return tr * tr + ti * ti;
}

// Per-block sums of mode 0: |psi_y|^2 over even |y & z|; modes 1, 2: the pair weights above
// (the host halves them).
template <class R>
__global__ __launch_bounds__(kBlock) void k_pauli_prob_partial(const Cx<R> *__restrict__ a, const PauliArgs P,
                                                               u64 n_items, double *partial) {
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    double s = 0.;
    const u64 base = (u64)blockIdx.x * (kBlock * kProbUnitsPerThread) + threadIdx.x;
#pragma unroll 4
    for (int r = 0; r < kProbUnitsPerThread; ++r) {
        const u64 w = base + (u64)r * kBlock;
        if (w >= n_items) continue;
        if (P.mode == 0) {
            const Amps<R> v = load_unit(a, w);
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (parity_sign((w << lw) | j, P.z) > 0.) s += abs2d(v.v[j]);
        } else if (P.mode == 1) {
            const Amps<R> v = load_unit(a, w);
            s += pauli_pair_weight(v.v[0], v.v[W - 1], (w << 1) | 1, P);
        } else {
            const u64 u0 = insert_zero_bits(w, P.ins), u1 = u0 ^ P.xu;
            const Amps<R> v0 = load_unit(a, u0), v1 = load_unit(a, u1);
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const int jj = j ^ P.xj;
                s += pauli_pair_weight(v0.v[j], jj ? v1.v[W - 1] : v1.v[0], (u1 << lw) | (u64)jj, P);
            }
        }
    }
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// (cr + i ci) v for cr, ci in {-1, 0, 1}: exact
template <class R>
__device__ __forceinline__ Cx<R> unit_phase(double cr, double ci, const Cx<R> &v) {
    const R r = (R)cr, i = (R)ci;
    return cx<R>(r * v.re - i * v.im, r * v.im + i * v.re);
}

// Collapse onto the eigenspace P = (-1)^value: psi <- (1 + sgn P) psi / (2 sqrt(p)), with
// half_norm = 1 / (2 sqrt(p)) and sgn = (-1)^value; mode 0 keeps amplitudes with
// |y & z| = value (mod 2), scaled by norm = 1 / sqrt(p), and zeroes the rest.
template <class R>
__global__ __launch_bounds__(kWave) void k_pauli_decohere(Cx<R> *__restrict__ a, const PauliArgs P, int value,
                                                          R norm, R half_norm, u64 n_items, u64 first) {
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const u64 w = first + (u64)blockIdx.x * kWave + threadIdx.x;
    if (w >= n_items) return;
    const double sgn = value ? -1. : 1.;
    if (P.mode == 0) {
        Amps<R> v = load_unit(a, w);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const bool keep = (int)(__popcll(((w << lw) | j) & P.z) & 1) == value;
            v.v[j] = keep ? cx<R>(v.v[j].re * norm, v.v[j].im * norm) : cx<R>(0, 0);
        }
        store_unit(a, w, v);
        return;
    }
    if (P.mode == 1) {
        Amps<R> v = load_unit(a, w);
        const Cx<R> a0 = v.v[0], a1 = v.v[W - 1];
        const double s1 = sgn * parity_sign((w << 1) | 1, P.z), s0 = sgn * parity_sign(w << 1, P.z);
        const Cx<R> t0 = cadd(a0, unit_phase(s1 * P.cr, s1 * P.ci, a1));
        const Cx<R> t1 = cadd(a1, unit_phase(s0 * P.cr, s0 * P.ci, a0));
        v.v[0] = cx<R>(t0.re * half_norm, t0.im * half_norm);
        v.v[W - 1] = cx<R>(t1.re * half_norm, t1.im * half_norm);
        store_unit(a, w, v);
        return;
    }
    const u64 u0 = insert_zero_bits(w, P.ins), u1 = u0 ^ P.xu;
    const Amps<R> v0 = load_unit(a, u0), v1 = load_unit(a, u1);
    Amps<R> n0, n1;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const int jj = j ^ P.xj;
        const Cx<R> b0 = v0.v[j], b1 = jj ? v1.v[W - 1] : v1.v[0];
        const double s1 = sgn * parity_sign((u1 << lw) | (u64)jj, P.z);
        const double s0 = sgn * parity_sign((u0 << lw) | (u64)j, P.z);
        const Cx<R> t0 = cadd(b0, unit_phase(s1 * P.cr, s1 * P.ci, b1));
        const Cx<R> t1 = cadd(b1, unit_phase(s0 * P.cr, s0 * P.ci, b0));
        n0.v[j] = cx<R>(t0.re * half_norm, t0.im * half_norm);
        if (jj) n1.v[W - 1] = cx<R>(t1.re * half_norm, t1.im * half_norm);
        else n1.v[0] = cx<R>(t1.re * half_norm, t1.im * half_norm);
    }
    store_unit(a, u0, n0);
    store_unit(a, u1, n1);
}

template <class R>
__global__ void k_set_basis1(Cx<R> *a, int value) {
    a[0] = value ? cx<R>(0, 0) : cx<R>(1, 0);
    a[1] = value ? cx<R>(1, 0) : cx<R>(0, 0);
}

constexpr int kMaxJoin = 64;

template <class R>
struct JoinArgs {
    const Cx<R> *src[kMaxJoin];
    int shift[kMaxJoin];     // bit offset of source k in the product index
    u64 mask[kMaxJoin];      // 2^lanes_k - 1
    int n_src;
    u64 product;             // product size; dst beyond it is zero
    u64 n_dst;
};

// dst[i] = src[m-2][i_{m-2}] * src[m-1][i_{m-1}], then src[k][i_k] * (...) for k = m-3..0;
// the same association as the reference's kron + in-place kron passes.
template <class R>
__global__ __launch_bounds__(kBlock) void k_join(Cx<R> *__restrict__ dst, const JoinArgs<R> g, u64 first) {
    const u64 i = first + (u64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.n_dst) return;
    if (i >= g.product) {
        dst[i] = cx<R>(0, 0);
        return;
    }
    const int m = g.n_src;
    if (m == 1) {
        dst[i] = g.src[0][i & g.mask[0]];
        return;
    }
    Cx<R> v = cmul(g.src[m - 2][(i >> g.shift[m - 2]) & g.mask[m - 2]],
                   g.src[m - 1][(i >> g.shift[m - 1]) & g.mask[m - 1]]);
    for (int k = m - 3; k >= 0; --k) v = cmul(g.src[k][(i >> g.shift[k]) & g.mask[k]], v);
    dst[i] = v;
}

BitInsert one_bit(int pos) {
    BitInsert b;
    b.n = 1;
    b.pos[0] = pos;
    return b;
}

BitInsert no_bits() {
    BitInsert b;
    b.n = 0;
    return b;
}

constexpr int lw_of(int w) { return w == 2 ? 1 : 0; }

}  // namespace

template <class R>
void launch_set_zero_state(Cx<R> *a, int n_lanes) {
    constexpr int W = Unit<R>::W;
    hipStream_t s = stream();
    if (n_lanes >= lw_of(W)) {
        const u64 n_units = (u64(1) << n_lanes) / W;
        for_each_piece(n_units, [&](u64 first, u64 count) {
            hipLaunchKernelGGL((k_zero_state<R>), grid_for(count, kBlock), kBlock, 0, s, a, n_units, first);
        });
    }
    QG_HIP(hipGetLastError());
}

namespace {

u64 prob_blocks(u64 n_items) {
    const u64 per_block = (u64)kBlock * kProbUnitsPerThread;
    return (n_items + per_block - 1) / per_block;
}

// Doubles the summation of n_partials partial sums needs: the partials, one sum per run of
// kSumRun (when there is more than one run's worth) and the total.
u64 sum_buffer_len(u64 n_partials) {
    const u64 n_runs = n_partials > (u64)kSumRun ? (n_partials + kSumRun - 1) / kSumRun : 0;
    return n_partials + n_runs + 1;
}

// launch(partial) writes n_partials partial sums; they are summed in a fixed order on the
// device (runs of kSumRun first when there are more than one run's worth) and the total is
// returned (the stream is synchronised). The buffer is taken from the allocator, or is
// `buf` (sum_buffer_len(n_partials) doubles, every one written before it is read) when given.
template <class F>
double sum_prob_partials(u64 n_partials, F &&launch, double *buf = nullptr) {
    Runtime &rt = runtime();
    hipStream_t s = stream();
    const u64 n_runs = n_partials > (u64)kSumRun ? (n_partials + kSumRun - 1) / kSumRun : 0;
    const size_t bytes = sizeof(double) * (size_t)sum_buffer_len(n_partials);
    double *partial = buf != nullptr ? buf : static_cast<double *>(dev_alloc(bytes));
    launch(partial);
    const double *last = partial;
    u64 n_last = n_partials;
    if (n_runs > 0) {
        hipLaunchKernelGGL(k_sum_runs, (unsigned)n_runs, 1024, 0, s, partial, n_partials, partial + n_partials);
        last = partial + n_partials;
        n_last = n_runs;
    }
    double *total = partial + n_partials + n_runs;
    hipLaunchKernelGGL(k_sum_partials, 1, 1024, 0, s, last, n_last, total);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(rt.h_pinned, total, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (buf == nullptr) dev_free(partial, bytes);
    QG_HIP(e);
    return rt.h_pinned[0];
}

// Units k_prob_partial sums for P(lane = 0): the half with the lane bit clear (lane at unit
// level), or every unit counting even amplitudes only (complex64 lane 0).
template <class R>
struct ProbRange {
    BitInsert ins;
    int amp_mask;
    u64 n_units;
    ProbRange(int n_lanes, int lane) {
        constexpr int W = Unit<R>::W, lw = lw_of(W);
        const u64 n_units_total = (u64(1) << n_lanes) / W;
        if (lane < lw) {
            ins = no_bits();
            amp_mask = 1;
            n_units = n_units_total;
        } else {
            ins = one_bit(lane - lw);
            amp_mask = 0;
            n_units = n_units_total / 2;
        }
    }
};

template <class R>
double calc_prob_with(const Cx<R> *a, int n_lanes, int lane, double *buf) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    hipStream_t s = stream();
    const ProbRange<R> r(n_lanes, lane);
    const u64 n_blocks = prob_blocks(r.n_units);
    return sum_prob_partials(n_blocks, [&](double *partial) {
        hipLaunchKernelGGL((k_prob_partial<R>), (unsigned)n_blocks, kBlock, 0, s, a, r.ins, r.amp_mask, r.n_units,
                           partial);
    }, buf);
}

}  // namespace

template <class R>
double calc_prob(const Cx<R> *a, int n_lanes, int lane) {
    return calc_prob_with<R>(a, n_lanes, lane, nullptr);
}

template <class R>
u64 calc_prob_buffer_len(int n_lanes, int lane) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    return sum_buffer_len(prob_blocks(ProbRange<R>(n_lanes, lane).n_units));
}

template <class R>
double calc_prob_in(const Cx<R> *a, int n_lanes, int lane, double *buf, u64 buf_len) {
    QG_REQUIRE(buf != nullptr, "null partial-sum buffer");
    QG_REQUIRE(calc_prob_buffer_len<R>(n_lanes, lane) <= buf_len, "partial-sum buffer too small");
    return calc_prob_with<R>(a, n_lanes, lane, buf);
}

template <class R>
double decohere_prob(Cx<R> *a, int n_lanes, int lane, int value, R norm, int prob_lane) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    QG_REQUIRE(0 <= prob_lane && prob_lane < n_lanes, "lane %d out of range", prob_lane);
    constexpr int W = Unit<R>::W, lw = lw_of(W);
    const u64 n_units = (u64(1) << n_lanes) / W;
    const u64 dist = lane < lw ? 0 : u64(1) << (lane - lw);
    const bool per_unit = dist < (u64)kWave;   // as launch_decohere
    const BitInsert ins = per_unit ? no_bits() : one_bit(lane - lw);
    const u64 n_items = per_unit ? n_units : n_units / 2;
    const u64 p_unit = prob_lane < lw ? 0 : u64(1) << (prob_lane - lw);
    const int p_amp = prob_lane < lw ? 1 << prob_lane : 0;
    // one partial per wavefront; pieces are powers of two of at least 2^12 items, so each
    // piece starts on a wavefront boundary
    return sum_prob_partials((n_items + kWave - 1) / kWave, [&](double *partial) {
        for_each_piece(n_items, [&](u64 first, u64 count) {
            hipLaunchKernelGGL((k_decohere_prob<R>), grid_for(count, kWave), kWave, 0, stream(), a, ins, dist,
                               value, norm, n_items, first, p_unit, p_amp, partial);
        });
    });
}

template <class R>
void launch_decohere(Cx<R> *a, int n_lanes, int lane, int value, R norm) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    constexpr int W = Unit<R>::W, lw = lw_of(W);
    const u64 n_units = (u64(1) << n_lanes) / W;
    hipStream_t s = stream();
    const u64 dist = lane < lw ? 0 : u64(1) << (lane - lw);
    // one unit per thread for a lane inside the unit or inside a wavefront's units, else one
    // (kept, dropped) pair per thread
    const bool per_unit = dist < (u64)kWave;
    const BitInsert ins = per_unit ? no_bits() : one_bit(lane - lw);
    const u64 n_items = per_unit ? n_units : n_units / 2;
    for_each_piece(n_items, [&](u64 first, u64 count) {
        hipLaunchKernelGGL((k_decohere<R>), grid_for(count, kWave), kWave, 0, s, a, ins, dist, value, norm,
                           n_items, first);
    });
    QG_HIP(hipGetLastError());
}

template <class R>
void launch_decohere_and_separate(Cx<R> *dst, const Cx<R> *src, int n_lanes, int lane, int value,
                                  R norm) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    QG_REQUIRE(n_lanes >= 2, "separating needs at least two lanes");
    constexpr int W = Unit<R>::W, lw = lw_of(W);
    const u64 n_dst_units = (u64(1) << (n_lanes - 1)) / W;
    hipStream_t s = stream();
    const BitInsert ins = lane < lw ? no_bits() : one_bit(lane - lw);
    const u64 vbit = (lane < lw || !value) ? 0 : (u64(1) << (lane - lw));
    for_each_piece(n_dst_units, [&](u64 first, u64 count) {
        hipLaunchKernelGGL((k_separate<R>), grid_for(count, kWave), kWave, 0, s, dst, src, ins, vbit, value,
                           norm, n_dst_units, first);
    });
    QG_HIP(hipGetLastError());
}

template <class R>
void launch_reset(Cx<R> *a, int n_lanes, int lane) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    constexpr int W = Unit<R>::W, lw = lw_of(W);
    const u64 n_units = (u64(1) << n_lanes) / W;
    hipStream_t s = stream();
    const BitInsert ins = lane < lw ? no_bits() : one_bit(lane - lw);
    const u64 dist = lane < lw ? 0 : u64(1) << (lane - lw);
    const u64 n_items = lane < lw ? n_units : n_units / 2;
    for_each_piece(n_items, [&](u64 first, u64 count) {
        hipLaunchKernelGGL((k_reset<R>), grid_for(count, kWave), kWave, 0, s, a, ins, dist, n_items, first);
    });
    QG_HIP(hipGetLastError());
}

template <class R>
void launch_set_basis1(Cx<R> *a, int value) {
    hipLaunchKernelGGL((k_set_basis1<R>), 1, 1, 0, stream(), a, value);
    QG_HIP(hipGetLastError());
}

template <class R>
void launch_join(Cx<R> *dst, int dst_lanes, const std::vector<const Cx<R> *> &srcs,
                 const std::vector<int> &src_lanes) {
    const int m = (int)srcs.size();
    QG_REQUIRE(1 <= m && m <= kMaxJoin, "join of %d state vectors is not supported", m);
    JoinArgs<R> g;
    g.n_src = m;
    int shift = 0;
    for (int k = m - 1; k >= 0; --k) {
        g.src[k] = srcs[k];
        g.shift[k] = shift;
        g.mask[k] = (u64(1) << src_lanes[k]) - 1;
        shift += src_lanes[k];
    }
    QG_REQUIRE(shift <= dst_lanes, "joined lanes exceed the destination");
    g.product = u64(1) << shift;
    g.n_dst = u64(1) << dst_lanes;
    for_each_piece(g.n_dst, [&](u64 first, u64 count) {
        hipLaunchKernelGGL((k_join<R>), grid_for(count, kBlock), kBlock, 0, stream(), dst, g, first);
    });
    QG_HIP(hipGetLastError());
}

namespace {

template <class R>
PauliArgs pauli_args(int n_lanes, u64 x, u64 z, u64 *n_items) {
    constexpr int W = Unit<R>::W, lw = lw_of(W);
    QG_REQUIRE(n_lanes >= 1 && n_lanes < 64, "lane count %d out of range", n_lanes);
    const u64 lanes = (u64(1) << n_lanes) - 1;
    QG_REQUIRE(((x | z) & ~lanes) == 0, "Pauli lanes outside the state vector");
    QG_REQUIRE((x | z) != 0, "the Pauli string is the identity");
    PauliArgs P;
    P.z = z;
    P.ins = no_bits();
    P.xu = 0;
    P.xj = 0;
    switch (__builtin_popcountll(x & z) & 3) {   // c = i^#Y
        case 0: P.cr = 1.; P.ci = 0.; break;
        case 1: P.cr = 0.; P.ci = 1.; break;
        case 2: P.cr = -1.; P.ci = 0.; break;
        default: P.cr = 0.; P.ci = -1.; break;
    }
    const u64 n_units = (u64(1) << n_lanes) / W;
    if (x == 0) {
        P.mode = 0;
        *n_items = n_units;
    } else if (W == 2 && x == 1) {
        P.mode = 1;
        *n_items = n_units;
    } else {
        const int h = 63 - __builtin_clzll(x);
        P.mode = 2;
        P.ins = one_bit(h - lw);
        P.xu = x >> lw;
        P.xj = (int)(x & (W - 1));
        *n_items = n_units / 2;
    }
    return P;
}

}  // namespace

template <class R>
double calc_pauli_prob(const Cx<R> *a, int n_lanes, u64 x, u64 z) {
    Runtime &rt = runtime();
    hipStream_t s = stream();
    u64 n_items;
    const PauliArgs P = pauli_args<R>(n_lanes, x, z, &n_items);
    const u64 per_block = (u64)kBlock * kProbUnitsPerThread;
    const u64 n_blocks = (n_items + per_block - 1) / per_block;
    const size_t bytes = sizeof(double) * (size_t)(n_blocks + 1);
    double *partial = static_cast<double *>(dev_alloc(bytes));
    hipLaunchKernelGGL((k_pauli_prob_partial<R>), (unsigned)n_blocks, kBlock, 0, s, a, P, n_items, partial);
    hipLaunchKernelGGL(k_sum_partials, 1, 1024, 0, s, partial, n_blocks, partial + n_blocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(rt.h_pinned, partial + n_blocks, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(partial, bytes);
    QG_HIP(e);
    // pair modes summed |((1 + P) psi)_y0|^2 once per pair: p = sum / 2 (exact)
    return P.mode == 0 ? rt.h_pinned[0] : 0.5 * rt.h_pinned[0];
}

template <class R>
void launch_pauli_decohere(Cx<R> *a, int n_lanes, u64 x, u64 z, int value, double prob) {
    QG_REQUIRE(value == 0 || value == 1, "value must be 0 or 1");
    u64 n_items;
    const PauliArgs P = pauli_args<R>(n_lanes, x, z, &n_items);
    const double norm = value == 0 ? 1. / std::sqrt(prob) : 1. / std::sqrt(1. - prob);
    const R rnorm = (R)norm, half = (R)(0.5 * norm);
    hipStream_t s = stream();
    for_each_piece(n_items, [&](u64 first, u64 count) {
        hipLaunchKernelGGL((k_pauli_decohere<R>), grid_for(count, kWave), kWave, 0, s, a, P, value, rnorm, half,
                           n_items, first);
    });
    QG_HIP(hipGetLastError());
}

namespace {

// dst = src over n_units 16-byte units, one unit per thread, one wavefront per block (the
// shape of the fastest gate kernels), non-temporal both ways: the streaming ceiling a state
// pass can reach.
__global__ __launch_bounds__(kWave) void k_stream_copy(f32x4 *__restrict__ dst, const f32x4 *__restrict__ src,
                                                       u64 n_units, u64 first) {
    const u64 u = first + (u64)blockIdx.x * kWave + threadIdx.x;
    if (u < n_units) __builtin_nontemporal_store(__builtin_nontemporal_load(src + u), dst + u);
}

}  // namespace

void launch_stream_copy(void *dst, const void *src, u64 bytes) {
    QG_REQUIRE(bytes % 16 == 0, "copy size must be a multiple of 16 bytes");
    const u64 n_units = bytes / 16;
    hipStream_t s = stream();
    for_each_piece(n_units, [&](u64 first, u64 count) {
        hipLaunchKernelGGL(k_stream_copy, grid_for(count, kWave), kWave, 0, s, static_cast<f32x4 *>(dst),
                           static_cast<const f32x4 *>(src), n_units, first);
    });
    QG_HIP(hipGetLastError());
}

#define QG_INST(R)                                                                              \
    template double calc_pauli_prob<R>(const Cx<R> *, int, u64, u64);                         \
    template void launch_pauli_decohere<R>(Cx<R> *, int, u64, u64, int, double);              \
    template void launch_set_zero_state<R>(Cx<R> *, int);                                     \
    template double calc_prob<R>(const Cx<R> *, int, int);                                    \
    template u64 calc_prob_buffer_len<R>(int, int);                                           \
    template double calc_prob_in<R>(const Cx<R> *, int, int, double *, u64);                  \
    template double decohere_prob<R>(Cx<R> *, int, int, int, R, int);                         \
    template void launch_decohere<R>(Cx<R> *, int, int, int, R);                              \
    template void launch_decohere_and_separate<R>(Cx<R> *, const Cx<R> *, int, int, int, R);  \
    template void launch_reset<R>(Cx<R> *, int, int);                                         \
    template void launch_set_basis1<R>(Cx<R> *, int);                                         \
    template void launch_join<R>(Cx<R> *, int, const std::vector<const Cx<R> *> &,            \
                                 const std::vector<int> &);
QG_INST(float)
QG_INST(double)
#undef QG_INST

}  // namespace qg
