// State kernels: |0...0> init, P(bit = 0), collapse, collapse-and-split, reset, join.
//
// Reference semantics (qgate/simulator/src/CPUQubitProcessor.cpp):
//   resetQubitStates      :61-73    a = |0...0>
//   calcProbability       :75-125   sum of |a_i|^2 over i with bit `lane` = 0
//   join                  :129-213  Kronecker product, last source in the lowest bits
//   decohere              :217-246  kept half *= 1/sqrt(p) (or 1/sqrt(1-p)), other half = 0
//   decohereAndSeparate   :248-284  kept half * norm -> new (n-1)-lane vector
//   applyReset            :286-305  a[i0] = a[i1], a[i1] = 0
// Probability sums are accumulated in FP64 whatever the state precision.
#include "kernels.hpp"
#include "qg_runtime.hpp"

namespace qg {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(u64 n, int per_thread = 1) {
    const u64 per_block = (u64)kBlock * per_thread;
    return (unsigned)((n + per_block - 1) / per_block);
}

// a[i] = (i == 0), written as 16-byte units.
template <class R, int W>
__global__ __launch_bounds__(kBlock) void k_zero_state(Cx<R> *a, u64 n_units) {
    const u64 u = (u64)blockIdx.x * kBlock + threadIdx.x;
    if (u >= n_units) return;
    Vec<R, W> v;
#pragma unroll
    for (int j = 0; j < W; ++j) v.v[j] = cx<R>(0, 0);
    if (u == 0) v.v[0].re = R(1);
    reinterpret_cast<Vec<R, W> *>(a)[u] = v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of `v` (fixed order); valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kBlock / 64];
    v = wave_sum(v);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double s = 0.;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) s += part[i];
    return s;
}

// Partial sums of |a_i|^2 over the half with bit `lane` = 0; grid-stride over W-wide units.
template <class R, int W>
__global__ __launch_bounds__(kBlock) void k_prob_partial(const Cx<R> *__restrict__ a, BitInsert ins,
                                                         u64 n_units, double *partial) {
    double s = 0.;
    const u64 stride = (u64)gridDim.x * kBlock;
    for (u64 u = (u64)blockIdx.x * kBlock + threadIdx.x; u < n_units; u += stride) {
        const u64 i = insert_zero_bits(u * W, ins);
        const Vec<R, W> v = *reinterpret_cast<const Vec<R, W> *>(a + i);
#pragma unroll
        for (int j = 0; j < W; ++j) s += abs2d(v.v[j]);
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *partial, int n, double *out) {
    double s = 0.;
    for (int i = threadIdx.x; i < n; i += kBlock) s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

template <class R, int W>
__global__ __launch_bounds__(kBlock) void k_decohere(Cx<R> *__restrict__ a, BitInsert ins, u64 bit,
                                                     int value, R norm, u64 n_units) {
    const u64 u = (u64)blockIdx.x * kBlock + threadIdx.x;
    if (u >= n_units) return;
    typedef Vec<R, W> V;
    const u64 i0 = insert_zero_bits(u * W, ins);
    const u64 keep = value ? (i0 | bit) : i0, drop = value ? i0 : (i0 | bit);
    V v = *reinterpret_cast<const V *>(a + keep);
    V z;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        v.v[j] = cx<R>(v.v[j].re * norm, v.v[j].im * norm);
        z.v[j] = cx<R>(0, 0);
    }
    *reinterpret_cast<V *>(a + keep) = v;
    *reinterpret_cast<V *>(a + drop) = z;
}

template <class R, int W>
__global__ __launch_bounds__(kBlock) void k_separate(Cx<R> *__restrict__ dst, const Cx<R> *__restrict__ src,
                                                     BitInsert ins, u64 vbit, R norm, u64 n_units) {
    const u64 u = (u64)blockIdx.x * kBlock + threadIdx.x;
    if (u >= n_units) return;
    typedef Vec<R, W> V;
    const u64 i = insert_zero_bits(u * W, ins) | vbit;
    V v = *reinterpret_cast<const V *>(src + i);
#pragma unroll
    for (int j = 0; j < W; ++j) v.v[j] = cx<R>(norm * v.v[j].re, norm * v.v[j].im);
    reinterpret_cast<V *>(dst)[u] = v;
}

template <class R, int W>
__global__ __launch_bounds__(kBlock) void k_reset(Cx<R> *__restrict__ a, BitInsert ins, u64 bit,
                                                  u64 n_units) {
    const u64 u = (u64)blockIdx.x * kBlock + threadIdx.x;
    if (u >= n_units) return;
    typedef Vec<R, W> V;
    const u64 i0 = insert_zero_bits(u * W, ins);
    const V v = *reinterpret_cast<const V *>(a + (i0 | bit));
    V z;
#pragma unroll
    for (int j = 0; j < W; ++j) z.v[j] = cx<R>(0, 0);
    *reinterpret_cast<V *>(a + i0) = v;
    *reinterpret_cast<V *>(a + (i0 | bit)) = z;
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
__global__ __launch_bounds__(kBlock) void k_join(Cx<R> *__restrict__ dst, const JoinArgs<R> g) {
    const u64 i = (u64)blockIdx.x * kBlock + threadIdx.x;
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

BitInsert one_bit(int lane) {
    BitInsert b;
    b.n = 1;
    b.pos[0] = lane;
    return b;
}

}  // namespace

template <class R>
void launch_set_zero_state(Cx<R> *a, int n_lanes) {
    const u64 n = u64(1) << n_lanes;
    hipStream_t s = stream();
    if (Unit<R>::W > 1 && n >= (u64)Unit<R>::W) {
        constexpr int W = Unit<R>::W;
        hipLaunchKernelGGL((k_zero_state<R, W>), blocks_for(n / W), kBlock, 0, s, a, n / W);
    } else {
        hipLaunchKernelGGL((k_zero_state<R, 1>), blocks_for(n), kBlock, 0, s, a, n);
    }
    QG_HIP(hipGetLastError());
}

template <class R>
double calc_prob(const Cx<R> *a, int n_lanes, int lane) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    Runtime &rt = runtime();
    hipStream_t s = stream();
    const BitInsert ins = one_bit(lane);
    const u64 half = u64(1) << (n_lanes - 1);
    constexpr int W = Unit<R>::W;
    const bool vec = W > 1 && lane >= 1;
    const u64 n_units = vec ? half / W : half;
    const int grid = (int)std::min<u64>((u64)reduce_grid(), (n_units + kBlock - 1) / kBlock);
    if (vec)
        hipLaunchKernelGGL((k_prob_partial<R, W>), grid, kBlock, 0, s, a, ins, n_units, rt.d_partials);
    else
        hipLaunchKernelGGL((k_prob_partial<R, 1>), grid, kBlock, 0, s, a, ins, n_units, rt.d_partials);
    QG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, 1, kBlock, 0, s, rt.d_partials, grid,
                       rt.d_partials + reduce_grid());
    QG_HIP(hipGetLastError());
    QG_HIP(hipMemcpyAsync(rt.h_pinned, rt.d_partials + reduce_grid(), sizeof(double),
                          hipMemcpyDeviceToHost, s));
    QG_HIP(hipStreamSynchronize(s));
    return rt.h_pinned[0];
}

template <class R>
void launch_decohere(Cx<R> *a, int n_lanes, int lane, int value, R norm) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    const u64 half = u64(1) << (n_lanes - 1);
    constexpr int W = Unit<R>::W;
    hipStream_t s = stream();
    if (W > 1 && lane >= 1)
        hipLaunchKernelGGL((k_decohere<R, W>), blocks_for(half / W), kBlock, 0, s, a, one_bit(lane),
                           u64(1) << lane, value, norm, half / W);
    else
        hipLaunchKernelGGL((k_decohere<R, 1>), blocks_for(half), kBlock, 0, s, a, one_bit(lane),
                           u64(1) << lane, value, norm, half);
    QG_HIP(hipGetLastError());
}

template <class R>
void launch_decohere_and_separate(Cx<R> *dst, const Cx<R> *src, int n_lanes, int lane, int value,
                                  R norm) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    const u64 half = u64(1) << (n_lanes - 1);
    const u64 vbit = value ? (u64(1) << lane) : 0;
    constexpr int W = Unit<R>::W;
    hipStream_t s = stream();
    if (W > 1 && lane >= 1)
        hipLaunchKernelGGL((k_separate<R, W>), blocks_for(half / W), kBlock, 0, s, dst, src,
                           one_bit(lane), vbit, norm, half / W);
    else
        hipLaunchKernelGGL((k_separate<R, 1>), blocks_for(half), kBlock, 0, s, dst, src,
                           one_bit(lane), vbit, norm, half);
    QG_HIP(hipGetLastError());
}

template <class R>
void launch_reset(Cx<R> *a, int n_lanes, int lane) {
    QG_REQUIRE(0 <= lane && lane < n_lanes, "lane %d out of range", lane);
    const u64 half = u64(1) << (n_lanes - 1);
    constexpr int W = Unit<R>::W;
    hipStream_t s = stream();
    if (W > 1 && lane >= 1)
        hipLaunchKernelGGL((k_reset<R, W>), blocks_for(half / W), kBlock, 0, s, a, one_bit(lane),
                           u64(1) << lane, half / W);
    else
        hipLaunchKernelGGL((k_reset<R, 1>), blocks_for(half), kBlock, 0, s, a, one_bit(lane),
                           u64(1) << lane, half);
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
    hipLaunchKernelGGL((k_join<R>), blocks_for(g.n_dst), kBlock, 0, stream(), dst, g);
    QG_HIP(hipGetLastError());
}

#define QG_INST(R)                                                                              \
    template void launch_set_zero_state<R>(Cx<R> *, int);                                     \
    template double calc_prob<R>(const Cx<R> *, int, int);                                    \
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
