// Shared types for the qgate HIP runtime: complex amplitudes, bit insertion, errors.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace qg {

typedef uint64_t u64;
typedef int64_t i64;

// Interleaved complex amplitude: the state vector is an array of these (re, im).
template <class R>
struct __align__(2 * sizeof(R)) Cx {
    R re, im;
};

// W consecutive amplitudes moved as one 16-byte (c64, W = 2) or 16-byte (c128, W = 1) access.
template <class R, int W>
struct __align__(W * 2 * sizeof(R)) Vec {
    Cx<R> v[W];
};

// Amplitudes per 16-byte access: 2 for complex64, 1 for complex128.
template <class R>
struct Unit {
    static constexpr int W = 16 / (2 * (int)sizeof(R));
};

// 16-byte access unit as a vector type (the form the non-temporal builtins take).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <class R>
struct UnitVec;
template <>
struct UnitVec<float> {
    typedef f32x4 T;
};
template <>
struct UnitVec<double> {
    typedef f64x2 T;
};

// One 16-byte unit of amplitudes (W = Unit<R>::W of them), loaded / stored with streaming
// (non-temporal) hints: state-vector passes touch every byte once, so they bypass caching.
template <class R>
struct Amps {
    Cx<R> v[Unit<R>::W];
};

template <class R>
__device__ __forceinline__ Amps<R> load_unit(const Cx<R> *a, u64 unit) {
    typedef typename UnitVec<R>::T T;
    T t = __builtin_nontemporal_load(reinterpret_cast<const T *>(a) + unit);
    Amps<R> r;
    __builtin_memcpy(&r, &t, sizeof(T));
    return r;
}

template <class R>
__device__ __forceinline__ void store_unit(Cx<R> *a, u64 unit, const Amps<R> &r) {
    typedef typename UnitVec<R>::T T;
    T t;
    __builtin_memcpy(&t, &r, sizeof(T));
    __builtin_nontemporal_store(t, reinterpret_cast<T *>(a) + unit);
}

template <class R>
struct Mat2 {
    Cx<R> m00, m01, m10, m11;
};

template <class R>
__host__ __device__ __forceinline__ Cx<R> cx(R re, R im) {
    Cx<R> c;
    c.re = re;
    c.im = im;
    return c;
}

template <class R>
__device__ __forceinline__ Cx<R> cmul(const Cx<R> &a, const Cx<R> &b) {
    return cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

template <class R>
__device__ __forceinline__ Cx<R> cadd(const Cx<R> &a, const Cx<R> &b) {
    return cx<R>(a.re + b.re, a.im + b.im);
}

template <class R>
__device__ __forceinline__ Cx<R> cscale(const Cx<R> &a, R s) {
    return cx<R>(a.re * s, a.im * s);
}

template <class R>
__device__ __forceinline__ double abs2d(const Cx<R> &a) {
    double re = (double)a.re, im = (double)a.im;
    return re * re + im * im;
}

// Compact index -> state index with zero bits inserted at ascending positions `pos`.
// This is the closed form of the reference's 8-bit-chunk permutation tables
// (qgate/simulator/src/BitPermTable.h:32-70, BitPermTable.cpp:5-33).
struct BitInsert {
    int n;
    int pos[63];
};

__host__ __device__ __forceinline__ u64 insert_zero_bits(u64 k, const BitInsert &b) {
    for (int i = 0; i < b.n; ++i) {
        const u64 low = (u64(1) << b.pos[i]) - 1;
        k = ((k & ~low) << 1) | (k & low);
    }
    return k;
}

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void raise_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define QG_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t qg_e_ = (expr);                                                            \
        if (qg_e_ != hipSuccess)                                                              \
            ::qg::raise_error("%s:%d %s failed: %s", __FILE__, __LINE__, #expr,               \
                              hipGetErrorString(qg_e_));                                      \
    } while (0)

#define QG_REQUIRE(cond, ...)                                                                 \
    do {                                                                                      \
        if (!(cond)) ::qg::raise_error(__VA_ARGS__);                                          \
    } while (0)

enum Prec { precFP64 = 0, precFP32 = 1 };

inline size_t amp_bytes(Prec p) { return p == precFP64 ? 16 : 8; }

}  // namespace qg
