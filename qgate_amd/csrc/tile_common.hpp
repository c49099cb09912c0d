// Device helpers shared by the tile-pass kernels (k_tile.hip, k_batch.hip).
#pragma once

#include "qg_common.hpp"

namespace qg {

// How a gate acts inside a tile: on a pair of tile amplitudes (target is a tile bit), as a
// diagonal on a tile bit, or as a diagonal whose target lies outside the tile (one factor for
// the whole tile).
enum : int { kOpPair = 0, kOpDiagLocal = 1, kOpDiagOuter = 2 };

// State index of the first amplitude of tile `tile`: zero bits inserted at the tile bits
// pos[0..m) (ascending).
__device__ __forceinline__ u64 tile_base(u64 tile, const signed char *pos, int m) {
    for (int j = 0; j < m; ++j) {
        const u64 low = (u64(1) << pos[j]) - 1;
        tile = ((tile & ~low) << 1) | (tile & low);
    }
    return tile;
}

// State offset of local tile index l (pos[j] == j for j < low).
__device__ __forceinline__ u64 deposit(unsigned l, const signed char *pos, int m, int low) {
    u64 g = l & ((1u << low) - 1);
    for (int j = low; j < m; ++j) g |= u64((l >> j) & 1u) << pos[j];
    return g;
}

// Sum of v over a workgroup of B threads in a fixed order; valid in thread 0. scratch holds
// B / 64 doubles.
template <int B>
__device__ __forceinline__ double block_sum_tile(double v, double *scratch) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (threadIdx.x % 64 == 0) scratch[threadIdx.x / 64] = v;
    __syncthreads();
    double s = 0.;
    if (threadIdx.x == 0)
        for (int i = 0; i < B / 64; ++i) s += scratch[i];
    return s;
}

}  // namespace qg
