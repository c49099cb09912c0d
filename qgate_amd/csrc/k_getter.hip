// Readout kernels: amplitudes / probabilities in the external qreg ordering, marginal
// probability arrays, and the FP64 sampling pool (prefix sum + batched search).
//
// Reference semantics:
//   get_states        qgate/simulator/src/CPUQubitsStatesGetter.cpp:34-157, glue.cpp:447-506
//   prepareProbArray  (k_marginal.hip)
//   SamplingPool      CPUSamplingPool.cpp:8-81: inclusive prefix sum, normalised by 1/sum,
//                     observation = upper_bound(cdf, u), then zero bits inserted at the
//                     empty-lane positions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"
#include "qg_runtime.hpp"

namespace qg {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxQs = 64;

inline unsigned blocks_for(u64 n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct ViewArgs {
    const void *ptr[kMaxQs];
    int lane_begin[kMaxQs + 1];
    int ext_pos[64];   // external bit of each lane, lanes of qstates q at [lane_begin[q], lane_begin[q+1])
    int n_qs;
};

__device__ __forceinline__ u64 local_index(u64 ext, const ViewArgs &g, int q) {
    u64 local = 0;
    const int b = g.lane_begin[q], e = g.lane_begin[q + 1];
    for (int l = b; l < e; ++l) local |= ((ext >> g.ext_pos[l]) & 1) << (l - b);
    return local;
}

template <class R>
__device__ __forceinline__ R abs2r(const Cx<R> &c) {
    return c.re * c.re + c.im * c.im;
}

// mathop 0: product of amplitudes; mathop 1: product of |amplitude|^2 (in R, as the
// reference getter does).
template <class R, int MATHOP>
__global__ __launch_bounds__(kBlock) void k_get_states(void *out, const ViewArgs g, u64 empty_mask,
                                                       i64 start, i64 step, u64 count) {
    const u64 j = (u64)blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const u64 ext = (u64)(start + (i64)j * step);
    if (MATHOP == 0) {
        Cx<R> v = cx<R>(1, 0);
        if (ext & empty_mask) {
            v = cx<R>(0, 0);
        } else {
            for (int q = 0; q < g.n_qs; ++q)
                v = cmul(v, static_cast<const Cx<R> *>(g.ptr[q])[local_index(ext, g, q)]);
        }
        static_cast<Cx<R> *>(out)[j] = v;
    } else {
        R v = R(1);
        if (ext & empty_mask) {
            v = R(0);
        } else {
            for (int q = 0; q < g.n_qs; ++q)
                v *= abs2r(static_cast<const Cx<R> *>(g.ptr[q])[local_index(ext, g, q)]);
        }
        static_cast<R *>(out)[j] = v;
    }
}

ViewArgs make_view(const std::vector<StateView> &qs) {
    QG_REQUIRE((int)qs.size() <= kMaxQs, "too many state vectors (%zu)", qs.size());
    ViewArgs g;
    g.n_qs = (int)qs.size();
    int l = 0;
    for (int q = 0; q < g.n_qs; ++q) {
        g.ptr[q] = qs[q].ptr;
        g.lane_begin[q] = l;
        QG_REQUIRE((int)qs[q].local_to_ext.size() == qs[q].n_lanes, "lane table size mismatch");
        for (int e : qs[q].local_to_ext) {
            QG_REQUIRE(l < 64, "more than 63 lanes");
            QG_REQUIRE(0 <= e && e < 63, "external lane %d out of range", e);
            g.ext_pos[l++] = e;
        }
    }
    g.lane_begin[g.n_qs] = l;
    return g;
}

bool is_identity_view(const std::vector<StateView> &qs) {
    if (qs.size() != 1) return false;
    for (int l = 0; l < qs[0].n_lanes; ++l)
        if (qs[0].local_to_ext[l] != l) return false;
    return true;
}

}  // namespace

template <class R>
void get_states(void *host_out, int mathop, const std::vector<StateView> &qs, u64 empty_mask,
                i64 n_states, i64 start, i64 step) {
    QG_REQUIRE(mathop == 0 || mathop == 1, "Unknown math op %d.", mathop);
    if (n_states <= 0) return;
    hipStream_t s = stream();
    const size_t item = mathop == 0 ? sizeof(Cx<R>) : sizeof(R);
    if (mathop == 0 && step == 1 && empty_mask == 0 && is_identity_view(qs)) {
        // external order == local order: the state itself goes through the pinned staging
        const Cx<R> *a = static_cast<const Cx<R> *>(qs[0].ptr);
        copy_to_host(host_out, a + start, item * (size_t)n_states);
        return;
    }
    // gathered in chunks of one staging buffer into two device buffers used in turn: the
    // kernel for chunk k + 1 runs while chunk k is copied out on the host
    const ViewArgs g = make_view(qs);
    const i64 chunk = (i64)(kStageBytes / item);
    const i64 staged = std::min<i64>(chunk, n_states);
    const size_t dbytes = item * (size_t)staged;
    char *d_out = static_cast<char *>(dev_alloc(2 * dbytes));
    try {
        HostStager stager;
        int k = 0;
        for (i64 done = 0; done < n_states; done += chunk, k ^= 1) {
            const i64 cnt = std::min<i64>(chunk, n_states - done);
            const i64 st = start + done * step;
            char *d = d_out + k * dbytes;
            if (mathop == 0)
                hipLaunchKernelGGL((k_get_states<R, 0>), blocks_for(cnt), kBlock, 0, s, d, g,
                                   empty_mask, st, step, (u64)cnt);
            else
                hipLaunchKernelGGL((k_get_states<R, 1>), blocks_for(cnt), kBlock, 0, s, d, g,
                                   empty_mask, st, step, (u64)cnt);
            QG_HIP(hipGetLastError());
            stager.push(static_cast<char *>(host_out) + item * done, d, item * (size_t)cnt);
        }
        stager.finish();
    } catch (...) {
        hipStreamSynchronize(s);
        dev_free(d_out, 2 * dbytes);
        throw;
    }
    dev_free(d_out, 2 * dbytes);
}

template void get_states<float>(void *, int, const std::vector<StateView> &, u64, i64, i64, i64);
template void get_states<double>(void *, int, const std::vector<StateView> &, u64, i64, i64, i64);

// ---- sampling pool -----------------------------------------------------------------------
//
// CDF definition (FP64, deterministic, monotone). Bins are grouped into chunks of 16,
// tiles of 256 chunks and super-tiles of 1024 tiles. With P the running sum inside a chunk,
// Es / Ms / Hs the inclusive prefix of the previous chunk in the tile / tile in the
// super-tile / super-tile in the pool (0 for the first), each computed serially:
//     cdf(bin) = Hs + (Ms + (Es + P))          (each + rounded in FP64)
// P at the end of a chunk is the chunk's serial sum, and the stored chunk prefixes are
// scanned from those sums; the tile total stored for the upper levels is the tile's last
// chunk prefix, and each upper level stores its own serial prefix, so the value at the end
// of a tile or super-tile equals the stored inclusive prefix exactly. Inside a chunk the
// association of the sums differs from the stored prefix by rounding, so a draw in that
// ulp-wide gap past the chunk's last bin returns the chunk's last bin with p > 0. A draw u
// returns the first bin with cdf * (1 / total) > u (upper_bound semantics), found level by
// level. A bounded pool stores no p: a draw recomputes |a|^2 of the chunk's 16 bins from the
// state in FP64, the same values the build summed, so its CDF is the same.

namespace {

constexpr int kChunk = 16;
constexpr int kTileChunks = 256;
constexpr int kTile = kChunk * kTileChunks;   // 4096 bins
constexpr int kSuperTiles = 1024;             // 2^22 bins per super-tile
// Threads per tile block: 8 bins each, 512 threads (the first 256 sum the chunks); the
// persistent grid has two blocks per CU (pool_build).
constexpr int kPoolThreads = 512;
constexpr int kPoolRows = kTile / kPoolThreads;
static_assert(kPoolThreads >= kTileChunks, "one thread per chunk for the chunk sums");

// One DPP move of a double: lane l receives v of the lane that control CTRL names, in the
// rows enabled by ROWS; lanes with no source lane and disabled rows receive 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

struct StateSrc {};
struct ProbSrc {};

// Tile values loaded as 16-byte units: two complex64 amplitudes, one complex128 amplitude or
// two FP64 probabilities per unit.
template <class R, bool FROM_STATE>
struct PoolUnit {
    static constexpr int V = FROM_STATE ? Unit<R>::W : 2;           // values per unit
    static constexpr int U = kTile / V / kPoolThreads;              // units per thread
    typedef typename std::conditional<FROM_STATE, typename UnitVec<R>::T, f64x2>::type T;
};

// Persistent workgroups over the tiles (stride gridDim.x); each thread issues the next tile's
// loads before working on the current one, so HBM reads stay in flight through the scan. Per
// tile: p = |a|^2 (from the state) or p given; chunk sums; fixed-order scan of the 256 chunk
// sums; the tile total; the last bin with p > 0.
template <class R, bool FROM_STATE, bool STORE_P>
__global__ __launch_bounds__(kPoolThreads) void k_pool_tile(const Cx<R> *__restrict__ a, i64 n_bins,
                                                      double *__restrict__ p, double *__restrict__ e_inc,
                                                      double *__restrict__ t_tot, i64 *__restrict__ t_last,
                                                      i64 n_tiles) {
    typedef PoolUnit<R, FROM_STATE> PU;
    constexpr int V = PU::V, U = PU::U;
    typedef typename PU::T T;
    __shared__ double sp[kTile + kTile / kChunk];   // one pad slot per chunk
    __shared__ double sc[kTileChunks];
    __shared__ int sl[kTileChunks];
    const T *src = FROM_STATE ? reinterpret_cast<const T *>(a) : reinterpret_cast<const T *>(p);
    const i64 n_units = n_bins / V;   // n_bins is a power of two >= 2
    // Loads are unconditional (units past the end read the last unit, and their values are
    // replaced by zeros where they are used): a load inside a branch leaves the compiler's
    // count of loads in flight unknown, and it then waits for all of them before the scan.
    auto load = [&](i64 tile, T (&raw)[U]) __attribute__((always_inline)) {
        const i64 ubase = tile * (kTile / V);
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const i64 u = ubase + r * kPoolThreads + threadIdx.x;
            raw[r] = __builtin_nontemporal_load(src + (u < n_units ? u : n_units - 1));
        }
    };
    // the per-tile work on loaded values `cur`
    auto process = [&](i64 tile, const T (&cur)[U]) __attribute__((always_inline)) {
        const i64 base = tile * kTile;
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const bool in = tile * (kTile / V) + r * kPoolThreads + (i64)threadIdx.x < n_units;
            double v[V];
            if constexpr (FROM_STATE) {
                Amps<R> am;
                __builtin_memcpy(&am, &cur[r], sizeof(T));
#pragma unroll
                for (int j = 0; j < V; ++j) v[j] = in ? abs2d(am.v[j]) : 0.;
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) v[j] = in ? (double)cur[r][j] : 0.;
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int li = (r * kPoolThreads + threadIdx.x) * V + j;
                const i64 bin = base + li;
                // probabilities from the state, and zeros for the padding up to whole tiles
                if (STORE_P && (FROM_STATE || bin >= n_bins)) __builtin_nontemporal_store(v[j], p + bin);
                sp[li + li / kChunk] = v[j];
            }
        }
        __syncthreads();
        // serial sum over this thread's chunk (threads below kTileChunks) and the local index of
        // its last bin with p > 0 (-1: none)
        const int c = threadIdx.x;
        if (c < kTileChunks) {
            double s = 0.;
            int last = -1;
            for (int k = 0; k < kChunk; ++k) {
                const int li = c * kChunk + k;
                const double x = sp[li + li / kChunk];
                s = (k == 0) ? x : s + x;
                if (x > 0.) last = li;
            }
            sc[c] = s;
            sl[c] = last;
        }
        __syncthreads();
        // One wavefront scans the chunk totals in a fixed order: each lane adds up four
        // consecutive chunks, the lanes' totals are scanned across the wavefront with DPP
        // moves (within rows of 16 at distances 1, 2, 4, 8, then row 0's total into row 1, row
        // 2's into row 3, rows 0-1's into rows 2-3), and each lane adds the total of the lanes
        // before it (the scan shifted by one lane) to its four partial sums. The same wavefront
        // stores the chunk prefixes, the tile total (its last chunk prefix, so the levels above
        // agree with it exactly) and the tile's last bin with p > 0 (the highest lane holding
        // one, by ballot). The next tile's chunk sums rewrite sc and sl only after the barrier
        // that this wavefront reaches once it is done here.
        static_assert(kTileChunks == 4 * 64, "one wavefront scans the chunk totals, four per lane");
        if (threadIdx.x < 64) {
            const int l = threadIdx.x;
            const double q0 = sc[4 * l], q1 = q0 + sc[4 * l + 1], q2 = q1 + sc[4 * l + 2], q3 = q2 + sc[4 * l + 3];
            double incl = q3;
            incl += dpp_f64<0x111, 0xf>(incl);   // row_shr:1
            incl += dpp_f64<0x112, 0xf>(incl);   // row_shr:2
            incl += dpp_f64<0x114, 0xf>(incl);   // row_shr:4
            incl += dpp_f64<0x118, 0xf>(incl);   // row_shr:8
            incl += dpp_f64<0x142, 0xa>(incl);   // row_bcast:15 into rows 1 and 3
            incl += dpp_f64<0x143, 0xc>(incl);   // row_bcast:31 into rows 2 and 3
            const double before = dpp_f64<0x138, 0xf>(incl);   // wave_shr:1, lane 0: 0
            // the prefixes go back to sc (this wavefront has read it; its LDS accesses complete
            // in order) and out 64 consecutive chunks a store, non-temporal: full 512-byte runs
            // (four 8-byte stores a lane at a 32-byte stride, the tile scan took 1.79 ms at 30
            // qubits, against 1.60 ms)
            sc[4 * l] = before + q0;
            sc[4 * l + 1] = before + q1;
            sc[4 * l + 2] = before + q2;
            sc[4 * l + 3] = before + q3;
            double *eo = e_inc + tile * kTileChunks;
#pragma unroll
            for (int i = 0; i < 4; ++i) __builtin_nontemporal_store(sc[64 * i + l], eo + 64 * i + l);
            if (l == 63) t_tot[tile] = before + q3;
            const int m = max(max(sl[4 * l], sl[4 * l + 1]), max(sl[4 * l + 2], sl[4 * l + 3]));
            const unsigned long long has = __ballot(m >= 0);
            const int hl = has ? 63 - __builtin_clzll(has) : 0;   // highest lane with a nonzero bin
            const int ml = __builtin_amdgcn_readlane(m, hl);
            if (l == 0) t_last[tile] = has ? base + ml : (i64)-1;
        }
    };
    // two register buffers used in turn (no copies between them, whose reads would wait for
    // the loads in flight): the next tile's loads are issued before this tile's stores, and
    // CDNA retires vector memory operations in issue order
    // (the prefetch is unconditional too: past the last tile it reads the current one again)
    const i64 stride = gridDim.x;
    T bufa[U], bufb[U];
    i64 tile = blockIdx.x;
    if (tile < n_tiles) load(tile, bufa);
    while (tile < n_tiles) {
        const i64 t1 = tile + stride;
        load(t1 < n_tiles ? t1 : tile, bufb);
        process(tile, bufa);
        if (t1 >= n_tiles) break;
        const i64 t2 = t1 + stride;
        load(t2 < n_tiles ? t2 : t1, bufa);
        process(t1, bufb);
        tile = t2;
    }
}

// One wave per super-tile: serial inclusive scan of the tile totals in place; super total out.
__global__ __launch_bounds__(64) void k_pool_super(double *__restrict__ m_inc, i64 n_tiles,
                                                   double *__restrict__ s_tot, const i64 *__restrict__ t_last,
                                                   i64 *__restrict__ s_last) {
    __shared__ double st[kSuperTiles];
    const i64 base = (i64)blockIdx.x * kSuperTiles;
    for (int i = threadIdx.x; i < kSuperTiles; i += 64)
        st[i] = (base + i < n_tiles) ? m_inc[base + i] : 0.;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.;
        for (int i = 0; i < kSuperTiles; ++i) {
            acc = (i == 0) ? st[0] : acc + st[i];
            st[i] = acc;
        }
        s_tot[blockIdx.x] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSuperTiles; i += 64)
        if (base + i < n_tiles) m_inc[base + i] = st[i];
    // the super-tile's last nonzero bin
    i64 last = -1;
    for (int i = threadIdx.x; i < kSuperTiles; i += 64)
        if (base + i < n_tiles) last = max(last, t_last[base + i]);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) last = max(last, (i64)__shfl_xor((long long)last, d, 64));
    if (threadIdx.x == 0) s_last[blockIdx.x] = last;
}

// Serial inclusive scan of the super-tile totals in place; the last value is the total.
__global__ void k_pool_top(double *h_inc, i64 n_supers, double *total, const i64 *s_last, i64 *last_nonzero) {
    double acc = 0.;
    i64 last = -1;
    for (i64 i = 0; i < n_supers; ++i) {
        acc = (i == 0) ? h_inc[0] : acc + h_inc[i];
        h_inc[i] = acc;
        last = max(last, s_last[i]);
    }
    *total = acc;
    *last_nonzero = last + 1;   // 0: no nonzero bin
}

struct SampleArgs {
    const double *p, *e_inc, *m_inc, *h_inc;
    const void *amps;                  // bounded pool: bins read as |a|^2 of the state
    int amp_kind;                      // 0: p, 1: complex64 amplitudes, 2: complex128
    const i64 *t_last, *s_last;        // last nonzero bin per tile / super-tile
    i64 n, n_chunks, n_tiles, n_supers;
    double norm;
    i64 last_nonzero;
    BitInsert empty;
};

__device__ __forceinline__ double bin_prob(const SampleArgs &g, i64 bin) {
    if (g.amp_kind == 0) return g.p[bin];
    if (bin >= g.n) return 0.;   // padding up to whole tiles
    if (g.amp_kind == 1) return abs2d(static_cast<const Cx<float> *>(g.amps)[bin]);
    return abs2d(static_cast<const Cx<double> *>(g.amps)[bin]);
}

// first index in [lo, hi) with f(index) true, for f monotone false..true; hi if none
template <class F>
__device__ __forceinline__ i64 first_true(i64 lo, i64 hi, F f) {
    while (lo < hi) {
        const i64 mid = lo + (hi - lo) / 2;
        if (f(mid))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_pool_sample(const SampleArgs g, const double *__restrict__ u,
                                                        i64 *__restrict__ obs, i64 n) {
    const i64 i = (i64)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double x = u[i];
    const double norm = g.norm;
    i64 bin;
    const i64 gsup = first_true(0, g.n_supers, [&](i64 s) { return g.h_inc[s] * norm > x; });
    if (gsup == g.n_supers) {
        bin = g.last_nonzero;   // u beyond the normalised total: clamp to the last nonzero bin
    } else {
        const double hs = gsup > 0 ? g.h_inc[gsup - 1] : 0.;
        const i64 t0 = gsup * kSuperTiles;
        const i64 t1 = (t0 + kSuperTiles < g.n_tiles) ? t0 + kSuperTiles : g.n_tiles;
        const i64 t = first_true(t0, t1, [&](i64 k) { return (hs + g.m_inc[k]) * norm > x; });
        if (t == t1) {
            // not reached: the super-tile's prefix is its last tile's (exactly equal sums)
            bin = g.s_last[gsup];
        } else {
            const double ms = t > t0 ? g.m_inc[t - 1] : 0.;
            const i64 c0 = t * kTileChunks, c1 = c0 + kTileChunks;
            const i64 c = first_true(c0, c1, [&](i64 k) { return (hs + (ms + g.e_inc[k])) * norm > x; });
            if (c == c1) {
                // not reached: the tile's prefix is its last chunk's (exactly equal sums)
                bin = g.t_last[t];
            } else {
                const double es = c > c0 ? g.e_inc[c - 1] : 0.;
                double run = 0.;
                i64 last = -1;
                bin = -1;
                for (int k = 0; k < kChunk; ++k) {
                    const double pk = bin_prob(g, c * kChunk + k);
                    run = (k == 0) ? pk : run + pk;
                    if (pk > 0.) last = c * kChunk + k;
                    if ((hs + (ms + (es + run))) * norm > x) {
                        bin = c * kChunk + k;
                        break;
                    }
                }
                // u in the rounding gap past the chunk's serial sum: its last bin with p > 0
                if (bin < 0) bin = last;
            }
        }
        if (bin < 0) bin = g.last_nonzero;
    }
    obs[i] = (i64)insert_zero_bits((u64)bin, g.empty);
}

}  // namespace

size_t pool_bytes(int n_lanes, bool with_p) {
    const i64 n = i64(1) << n_lanes;
    const i64 n_tiles = (n + kTile - 1) / kTile;
    const i64 n_chunks = n_tiles * kTileChunks;
    const i64 n_supers = (n_tiles + kSuperTiles - 1) / kSuperTiles;
    const size_t n_p = with_p ? (size_t)n_tiles * kTile : 0;
    const size_t n_m = (size_t)n_supers * kSuperTiles;
    return sizeof(double) * (n_p + n_chunks + n_m + n_supers + 8) + sizeof(i64) * (n_tiles + n_supers);
}

void pool_alloc(PoolData &pd, int n_lanes, bool with_p) {
    pd.n = i64(1) << n_lanes;
    pd.n_tiles = (pd.n + kTile - 1) / kTile;
    pd.n_chunks = pd.n_tiles * kTileChunks;
    pd.n_supers = (pd.n_tiles + kSuperTiles - 1) / kSuperTiles;
    const size_t n_p = with_p ? (size_t)pd.n_tiles * kTile : 0;
    const size_t n_m = (size_t)pd.n_supers * kSuperTiles;
    pd.bytes = pool_bytes(n_lanes, with_p);
    pd.block = dev_alloc(pd.bytes);
    double *d = static_cast<double *>(pd.block);
    pd.p = with_p ? d : nullptr;
    pd.amps = nullptr;
    pd.amp_kind = 0;
    pd.e_inc = d + n_p;
    pd.m_inc = pd.e_inc + pd.n_chunks;
    pd.h_inc = pd.m_inc + n_m;
    pd.d_last_nonzero = reinterpret_cast<i64 *>(pd.h_inc + pd.n_supers);
    pd.t_last = reinterpret_cast<i64 *>(pd.h_inc + pd.n_supers + 8);
    pd.s_last = pd.t_last + pd.n_tiles;
}

void pool_free(PoolData &pd) {
    if (pd.block) dev_free(pd.block, pd.bytes);
    pd.block = nullptr;
}

namespace {

template <class R, bool FROM_STATE, bool STORE_P>
void pool_build(PoolData &pd, const Cx<R> *a) {
    hipStream_t s = stream();
    double *d_total = pd.h_inc + pd.n_supers + 1;
    // two workgroups per compute unit (four fit the tile's 37 KB of LDS, and measured slower:
    // a 30-qubit complex64 scan 1.60 ms with two, 1.71 ms with four)
    const i64 grid = std::min<i64>(pd.n_tiles, (i64)runtime().n_cus * 2);
    hipLaunchKernelGGL((k_pool_tile<R, FROM_STATE, STORE_P>), (unsigned)grid, kPoolThreads, 0, s, a, pd.n,
                       pd.p, pd.e_inc, pd.m_inc, pd.t_last, pd.n_tiles);
    QG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pool_super, (unsigned)pd.n_supers, 64, 0, s, pd.m_inc, pd.n_tiles, pd.h_inc, pd.t_last,
                       pd.s_last);
    QG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pool_top, 1, 1, 0, s, pd.h_inc, pd.n_supers, d_total, pd.s_last, pd.d_last_nonzero);
    QG_HIP(hipGetLastError());
    Runtime &rt = runtime();
    QG_HIP(hipMemcpyAsync(rt.h_pinned, d_total, sizeof(double), hipMemcpyDeviceToHost, s));
    QG_HIP(hipMemcpyAsync(rt.h_pinned + 1, pd.d_last_nonzero, sizeof(i64), hipMemcpyDeviceToHost, s));
    QG_HIP(hipStreamSynchronize(s));
    pd.total = rt.h_pinned[0];
    std::memcpy(&pd.last_nonzero, rt.h_pinned + 1, sizeof(i64));
    pd.last_nonzero -= 1;
    QG_REQUIRE(!(0.05 < std::abs(pd.total - 1.)),
               "error in probability sum is beyond 0.05., %g.", pd.total);
    pd.norm = 1. / pd.total;
}

}  // namespace

template <class R>
void pool_build_from_state(PoolData &pd, const Cx<R> *a, int n_lanes, bool bounded) {
    pool_alloc(pd, n_lanes, !bounded);
    try {
        if (bounded) {
            pool_build<R, true, false>(pd, a);
            pd.amps = a;
            pd.amp_kind = sizeof(R) == 4 ? 1 : 2;
        } else {
            pool_build<R, true, true>(pd, a);
        }
    } catch (...) {
        pool_free(pd);
        throw;
    }
}

template void pool_build_from_state<float>(PoolData &, const Cx<float> *, int, bool);
template void pool_build_from_state<double>(PoolData &, const Cx<double> *, int, bool);

void pool_build_from_probs(PoolData &pd, int n_lanes) {
    (void)n_lanes;
    QG_REQUIRE(pd.p != nullptr, "pool without a probability array");
    pool_build<double, false, true>(pd, nullptr);
}

void pool_sample(const PoolData &pd, i64 *host_obs, i64 n_samples, const double *host_u,
                 const std::vector<int> &empty_lanes) {
    if (n_samples <= 0) return;
    SampleArgs g;
    g.p = pd.p;
    g.e_inc = pd.e_inc;
    g.m_inc = pd.m_inc;
    g.h_inc = pd.h_inc;
    g.amps = pd.amps;
    g.amp_kind = pd.amp_kind;
    g.t_last = pd.t_last;
    g.s_last = pd.s_last;
    g.n = pd.n;
    g.n_chunks = pd.n_chunks;
    g.n_tiles = pd.n_tiles;
    g.n_supers = pd.n_supers;
    g.norm = pd.norm;
    g.last_nonzero = pd.last_nonzero < 0 ? 0 : pd.last_nonzero;
    std::vector<int> empty(empty_lanes);
    std::sort(empty.begin(), empty.end());
    QG_REQUIRE(empty.size() < 63, "too many empty lanes");
    g.empty.n = (int)empty.size();
    for (size_t i = 0; i < empty.size(); ++i) g.empty.pos[i] = empty[i];
    hipStream_t s = stream();
    const size_t bytes = sizeof(double) * (size_t)n_samples;
    char *buf = static_cast<char *>(dev_alloc(2 * bytes));
    double *d_u = reinterpret_cast<double *>(buf);
    i64 *d_obs = reinterpret_cast<i64 *>(buf + bytes);
    try {
        QG_HIP(hipMemcpyAsync(d_u, host_u, bytes, hipMemcpyHostToDevice, s));
        for_each_piece((u64)n_samples, [&](u64 first, u64 count) {
            hipLaunchKernelGGL(k_pool_sample, blocks_for(count), kBlock, 0, s, g, d_u + first, d_obs + first,
                               (i64)count);
        });
        QG_HIP(hipGetLastError());
        QG_HIP(hipMemcpyAsync(host_obs, d_obs, bytes, hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
    } catch (...) {
        dev_free(buf, 2 * bytes);
        throw;
    }
    dev_free(buf, 2 * bytes);
}

}  // namespace qg
