// Device helpers shared by the tile-pass kernels (k_tile.hip, k_batch.hip).
#pragma once

#include "qg_common.hpp"

namespace qg {

// How a gate acts inside a tile: on a pair of tile amplitudes (target is a tile bit), as a
// diagonal on a tile bit, or as a diagonal whose target lies outside the tile (one factor for
// the whole tile).
enum : int { kOpPair = 0, kOpDiagLocal = 1, kOpDiagOuter = 2 };
// Cheaper forms of the same updates (register-resident passes): a pair gate that only swaps
// (X, CX, CCX targets), a pair gate with a real matrix (H, RY), and a diagonal on a tile bit
// with m00 == 1 (Z, S, T, U1, CZ, cu1: only the |1> half changes).
enum : int { kOpPairSwap = 3, kOpPairReal = 4, kOpDiagPhase = 5 };

// Tile layouts hold at most kTilePosMax tile bits; pos[] arrays have that many entries.
constexpr int kTilePosMax = 16;

// tile_base / deposit read pos[] only at compile-time indices (loops fully unrolled, entries
// beyond m skipped by a uniform test). A kernel-argument byte read at a runtime index is a
// vector load, and its wait (vmcnt) would also wait for every tile load and store in flight,
// serialising a workgroup's HBM traffic; at constant indices the reads are scalar loads.

// pos[j] read as part of its aligned dword (pos arrays are declared alignas(4)), so that
// bytes at any offset take the same scalar loads.
__device__ __forceinline__ int pos_at(const signed char *pos, int j) {
    const unsigned w = reinterpret_cast<const unsigned *>(pos)[j >> 2];
    return (int)(signed char)(unsigned char)(w >> ((j & 3) * 8));
}

// State index of the first amplitude of tile `tile`: zero bits inserted at the tile bits
// pos[0..m) (ascending).
__device__ __forceinline__ u64 tile_base(u64 tile, const signed char *pos, int m) {
#pragma unroll
    for (int j = 0; j < kTilePosMax; ++j) {
        if (j < m) {
            const u64 low = (u64(1) << pos_at(pos, j)) - 1;
            tile = ((tile & ~low) << 1) | (tile & low);
        }
    }
    return tile;
}

// State offset of local tile index l (pos[j] == j for j < low).
__device__ __forceinline__ u64 deposit(unsigned l, const signed char *pos, int m, int low) {
    u64 g = l & ((1u << low) - 1);
#pragma unroll
    for (int j = 0; j < kTilePosMax; ++j)
        if (j >= low && j < m) g |= u64((l >> j) & 1u) << pos_at(pos, j);
    return g;
}

// Tile <-> LDS transfers of a workgroup of T threads: every thread issues up to CHUNK 16-byte
// loads (or stores) before touching LDS, so each workgroup keeps CHUNK x T accesses in flight
// instead of waiting on one load at a time. Unit lu = c T + tid has local index
// (c T + tid) << lw, whose two parts occupy disjoint bits, so its state offset is
// deposit(tid << lw) | deposit(c T << lw): one deposit per thread and one per chunk (the
// same for the whole workgroup) instead of one per access. All addresses of a round are
// formed before its first access.
template <class R, int T, int CHUNK = 8>
__device__ __forceinline__ void tile_to_lds(Cx<R> *t, const Cx<R> *a, u64 base, const signed char *pos, int m,
                                            int low) {
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const int n_units = 1 << (m - lw);
    const u64 dt = base | deposit((unsigned)threadIdx.x << lw, pos, m, low);
    for (int b0 = 0; b0 < n_units; b0 += T * CHUNK) {
        u64 ad[CHUNK];
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) ad[c] = (dt | deposit((unsigned)(b0 + c * T) << lw, pos, m, low)) >> lw;
        Amps<R> x[CHUNK];
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) {
            const int lu = b0 + c * T + (int)threadIdx.x;
            if (lu < n_units) x[c] = load_unit(a, ad[c]);
        }
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) {
            const int lu = b0 + c * T + (int)threadIdx.x;
            if (lu < n_units) {
#pragma unroll
                for (int j = 0; j < W; ++j) t[((unsigned)lu << lw) + j] = x[c].v[j];
            }
        }
    }
}

template <class R, int T, int CHUNK = 8>
__device__ __forceinline__ void lds_to_tile(Cx<R> *a, const Cx<R> *t, u64 base, const signed char *pos, int m,
                                            int low) {
    constexpr int W = Unit<R>::W, lw = W == 2 ? 1 : 0;
    const int n_units = 1 << (m - lw);
    const u64 dt = base | deposit((unsigned)threadIdx.x << lw, pos, m, low);
    for (int b0 = 0; b0 < n_units; b0 += T * CHUNK) {
        u64 ad[CHUNK];
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) ad[c] = (dt | deposit((unsigned)(b0 + c * T) << lw, pos, m, low)) >> lw;
        Amps<R> x[CHUNK];
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) {
            const int lu = b0 + c * T + (int)threadIdx.x;
            if (lu < n_units) {
#pragma unroll
                for (int j = 0; j < W; ++j) x[c].v[j] = t[((unsigned)lu << lw) + j];
            }
        }
#pragma unroll
        for (int c = 0; c < CHUNK; ++c) {
            const int lu = b0 + c * T + (int)threadIdx.x;
            if (lu < n_units) store_unit(a, ad[c], x[c]);
        }
    }
}

// Tile transfers for a tile of m <= MB bits: the full-size tile (m == MB, the common case)
// takes the compile-time-sized transfer (one round, no bounds tests), whose loads and stores
// the compiler issues back to back; smaller tiles take the generic loop.
template <class R, int T, int MB>
__device__ __forceinline__ void load_tile_lds(Cx<R> *t, const Cx<R> *a, u64 base, const signed char *pos, int m,
                                              int low) {
    constexpr int lw = Unit<R>::W == 2 ? 1 : 0, full = (1 << (MB - lw)) / T;
    if (m == MB)
        tile_to_lds<R, T, full>(t, a, base, pos, MB, low);
    else
        tile_to_lds<R, T, 8>(t, a, base, pos, m, low);
}

template <class R, int T, int MB>
__device__ __forceinline__ void store_tile_lds(Cx<R> *a, const Cx<R> *t, u64 base, const signed char *pos, int m,
                                               int low) {
    constexpr int lw = Unit<R>::W == 2 ? 1 : 0, full = (1 << (MB - lw)) / T;
    if (m == MB)
        lds_to_tile<R, T, full>(a, t, base, pos, MB, low);
    else
        lds_to_tile<R, T, 8>(a, t, base, pos, m, low);
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
