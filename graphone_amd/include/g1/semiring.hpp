// Semiring policies of the SpMV/SpMM kernels.
//
// A policy supplies the reference's SpMV functor contract {initialize, combine(nonzero, x), reduce}
// (include/GraphOne/archi/blas/spmv/def.hpp:25-45, engine/blas_gc.h:113-128) split into an
// accumulator so that a compensated sum can carry (sum, compensation) pairs across threads and tiles:
//   identity()      the accumulator of an empty fold (initialize())
//   combine(w, x)   the per-nonzero product
//   add(acc, p)     acc = reduce(acc, p), p to the right of everything already in acc
//   merge(a, b)     the fold of a followed by b (a holds the lower nonzero positions)
//   finalize(acc)   the y value
// merge(identity(), b) returns b bit for bit, so prepending an empty fold never changes a result.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"

namespace g1 {

// SSSP / BFS semiring over uint32 (apps/sssp_blas.cu:55-73): combine saturates at INF, reduce = min.
// min is exact and order free, so any fold order gives the same bits.
struct MinPlusU32 {
    using value_t = uint32_t;
    using x_t = uint32_t;
    using y_t = uint32_t;
    struct acc_t { uint32_t v; };
    static constexpr bool kOrderFree = true;

    __device__ __forceinline__ static acc_t identity() { return acc_t{kInf}; }
    __device__ __forceinline__ static uint32_t combine(uint32_t w, uint32_t x) {
        return x == kInf ? kInf : w + x;
    }
    __device__ __forceinline__ static void add(acc_t& a, uint32_t p) { a.v = p < a.v ? p : a.v; }
    __device__ __forceinline__ static acc_t merge(acc_t a, acc_t b) { return acc_t{b.v < a.v ? b.v : a.v}; }
    __device__ __forceinline__ static uint32_t finalize(acc_t a) { return a.v; }
    __device__ __forceinline__ static acc_t shfl_down(acc_t a, int d, int width) {
        return acc_t{static_cast<uint32_t>(__shfl_down(static_cast<int>(a.v), d, width))};
    }
};

// Min-plus over a one-byte x (the delta-stepping engine's heavy pull step): x[u] is a bucket member's
// distance minus the bucket's start (0..253), 254 for a vertex settled in an earlier bucket and 255 for
// every other vertex, so the gathered vector is a quarter of the u32 one and stays in the L2 caches; y =
// min over the in-edges of x + w for x < 254, the bucket start added back by the apply. Order free, like
// MinPlusU32.
struct MinPlusU8X {
    using value_t = uint32_t;
    using x_t = uint8_t;
    using y_t = uint32_t;
    struct acc_t { uint32_t v; };
    static constexpr bool kOrderFree = true;

    __device__ __forceinline__ static acc_t identity() { return acc_t{kInf}; }
    __device__ __forceinline__ static uint32_t combine(uint32_t w, uint8_t x) {
        return (x >= 254u || w >= kInf - 255u) ? kInf : w + x;
    }
    __device__ __forceinline__ static void add(acc_t& a, uint32_t p) { a.v = p < a.v ? p : a.v; }
    __device__ __forceinline__ static acc_t merge(acc_t a, acc_t b) { return acc_t{b.v < a.v ? b.v : a.v}; }
    __device__ __forceinline__ static uint32_t finalize(acc_t a) { return a.v; }
    __device__ __forceinline__ static acc_t shfl_down(acc_t a, int d, int width) {
        return acc_t{static_cast<uint32_t>(__shfl_down(static_cast<int>(a.v), d, width))};
    }
};

// PageRank semiring over float (apps/pagerank_blas.cu:82-111: combine = info * weight, reduce = +),
// accumulated with Kahan-Babuska (Neumaier) compensation: s holds the running sum, c the rounding
// error collected so far; the value is s + c.
struct PlusTimesKahanF32 {
    using value_t = float;
    using x_t = float;
    using y_t = float;
    struct acc_t { float s, c; };
    static constexpr bool kOrderFree = false;

    __device__ __forceinline__ static acc_t identity() { return acc_t{0.0f, 0.0f}; }
    __device__ __forceinline__ static float combine(float w, float x) { return x * w; }
    __device__ __forceinline__ static void add(acc_t& a, float p) {
        const float t = a.s + p;
        const float e = (fabsf(a.s) >= fabsf(p)) ? ((a.s - t) + p) : ((p - t) + a.s);
        a.c += e;
        a.s = t;
    }
    __device__ __forceinline__ static acc_t merge(acc_t a, acc_t b) {
        const float t = a.s + b.s;
        const float e = (fabsf(a.s) >= fabsf(b.s)) ? ((a.s - t) + b.s) : ((b.s - t) + a.s);
        return acc_t{t, (a.c + b.c) + e};
    }
    __device__ __forceinline__ static float finalize(acc_t a) { return a.s + a.c; }
    __device__ __forceinline__ static acc_t shfl_down(acc_t a, int d, int width) {
        return acc_t{__shfl_down(a.s, d, width), __shfl_down(a.c, d, width)};
    }
};

// Adapter from a reference-style SpMV functor F {initialize, combine, reduce} (the BlasGcEngine
// wrapper of engine/blas_gc.h:113-128) to an accumulator policy. The fold is the naive one of
// archi/blas/spmv/cpu/naive.hpp:5-20: sum = reduce(sum, combine(A, x)), left to right.
template <typename F, typename W, typename X, typename Y>
struct FunctorSemiring {
    using value_t = W;
    using x_t = X;
    using y_t = Y;
    struct acc_t { Y v; };
    static constexpr bool kOrderFree = false;

    __device__ __forceinline__ static acc_t identity() { return acc_t{F::initialize()}; }
    __device__ __forceinline__ static Y combine(const W& w, const X& x) { return F::combine(w, x); }
    __device__ __forceinline__ static void add(acc_t& a, const Y& p) { a.v = F::reduce(a.v, p); }
    __device__ __forceinline__ static acc_t merge(acc_t a, acc_t b) { return acc_t{F::reduce(a.v, b.v)}; }
    __device__ __forceinline__ static Y finalize(acc_t a) { return a.v; }
    __device__ __forceinline__ static acc_t shfl_down(acc_t a, int d, int width) {
        return acc_t{__shfl_down(a.v, d, width)};
    }
};

}  // namespace g1
