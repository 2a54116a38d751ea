// Platform layer of the graph_one C++ API on MI355X: function qualifiers, memory (through the C-ABI of
// libgraphone_hip.so), launch helpers, device-wide fill / transform / transform_reduce, atomics and
// the counting iterator / reduction functors the apps use.
//
// Reference counterparts: macros include/GraphOne/archi/macro/{macro.h,cuda.cuh}; memory
// include/GraphOne/archi/mem/cuda.cuh:11-59; LaunchKernel/LaunchSync include/GraphOne/archi/kernel/cuda.cuh;
// Thrust wrappers include/GraphOne/archi/thrust/functions.h:17-104; AtomicMin/UpperBound
// include/GraphOne/archi/only/cuda.cuh:37-70. Here every device-wide algorithm is a small HIP kernel of
// this header (no Thrust), and every runtime call goes through the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "GraphOne/type.hpp"
#include "GraphOne/utils.h"
#include "graphone_hip.h"

#define __ONE_DEV__ __device__
#define __ONE_DEV_INL__ __device__ __forceinline__
#define __ONE_ARCH__ __host__ __device__
#define __ONE_ARCH_INL__ __host__ __device__ __forceinline__
#define __ONE_KERNEL__ __global__

namespace graph_one {

// Counting iterator and reduction functors (the apps' uses of thrust::counting_iterator,
// thrust::make_counting_iterator, thrust::maximum and thrust::plus).
template <typename T>
struct counting_iterator {
    T value;
    __host__ __device__ counting_iterator() : value(0) {}
    __host__ __device__ explicit counting_iterator(T v) : value(v) {}
    __host__ __device__ T operator[](size_t i) const { return static_cast<T>(value + i); }
    __host__ __device__ T operator*() const { return value; }
    __host__ __device__ counting_iterator operator+(size_t i) const { return counting_iterator(static_cast<T>(value + i)); }
    __host__ __device__ ptrdiff_t operator-(const counting_iterator& o) const {
        return static_cast<ptrdiff_t>(value) - static_cast<ptrdiff_t>(o.value);
    }
};
template <typename T>
__host__ __device__ counting_iterator<T> make_counting_iterator(T v) {
    return counting_iterator<T>(v);
}
template <typename T>
struct maximum {
    __host__ __device__ T operator()(const T& a, const T& b) const { return a < b ? b : a; }
};
template <typename T>
struct minimum {
    __host__ __device__ T operator()(const T& a, const T& b) const { return b < a ? b : a; }
};
template <typename T>
struct plus {
    __host__ __device__ T operator()(const T& a, const T& b) const { return a + b; }
};

namespace archi {

inline void check(g1_status s, const char* where) {
    if (s != G1_OK) LOG_ERROR(where, " failed: ", g1_last_error_message());
}

// ---------------------------------------------------------------- memory
template <arch_t arch, typename T>
T* memalloc(size_t count) {
    if (count == 0) return nullptr;
    if constexpr (arch == arch_t::cpu) {
        return new T[count]();
    } else {
        void* p = nullptr;
        check(g1_malloc(&p, count * sizeof(T)), "g1_malloc");
        return static_cast<T*>(p);
    }
}

template <arch_t arch, typename T>
void memfree(T* p) {
    if (!p) return;
    if constexpr (arch == arch_t::cpu) delete[] p;
    else check(g1_free(p), "g1_free");
}

template <arch_t to, arch_t from, typename T>
void memcpy(T* dst, const T* src, size_t count) {
    if (count == 0) return;
    const size_t bytes = count * sizeof(T);
    if constexpr (to == arch_t::cpu && from == arch_t::cpu) std::memcpy(dst, src, bytes);
    else if constexpr (to == arch_t::hip && from == arch_t::cpu) check(g1_memcpy_h2d(dst, src, bytes, nullptr), "h2d");
    else if constexpr (to == arch_t::cpu && from == arch_t::hip) check(g1_memcpy_d2h(dst, src, bytes, nullptr), "d2h");
    else {
        check(g1_memcpy_d2d(dst, src, bytes, nullptr), "d2d");
        check(g1_stream_synchronize(nullptr), "sync");
    }
}

template <typename T>
__global__ void fill_value_kernel(T* __restrict__ dst, T value, size_t count) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = value;
}

inline unsigned grid_blocks(size_t items, unsigned cap = 4096) {
    size_t b = (items + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return static_cast<unsigned>(b);
}

inline void launch_check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) LOG_ERROR(what, " launch failed: ", hipGetErrorString(e));
}

template <arch_t arch>
void LaunchSync() {
    if constexpr (arch != arch_t::cpu) check(g1_stream_synchronize(nullptr), "sync");
}

template <arch_t arch, typename T>
void memset_zero(T* p, size_t count) {
    if (count == 0) return;
    if constexpr (arch == arch_t::cpu) {
        std::memset(static_cast<void*>(p), 0, count * sizeof(T));
    } else {
        fill_value_kernel<T><<<grid_blocks(count), 256>>>(p, T{}, count);
        launch_check("fill");
        LaunchSync<arch>();
    }
}

// ---------------------------------------------------------------- device-wide algorithms
template <arch_t arch, typename It, typename T>
void fill(It first, It last, const T& value) {
    const size_t n = static_cast<size_t>(last - first);
    if (n == 0) return;
    using V = std::remove_reference_t<decltype(*first)>;
    if constexpr (arch == arch_t::cpu) {
        std::fill(first, last, static_cast<V>(value));
    } else {
        fill_value_kernel<V><<<grid_blocks(n), 256>>>(&*first, static_cast<V>(value), n);
        launch_check("fill");
        LaunchSync<arch>();
    }
}

template <typename InIt, typename OutIt, typename Op>
__global__ void transform_kernel(InIt in, size_t n, OutIt out, Op op) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = op(in[i]);
}

template <typename InIt1, typename InIt2, typename OutIt, typename Op>
__global__ void transform2_kernel(InIt1 a, InIt2 b, size_t n, OutIt out, Op op) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = op(a[i], b[i]);
}

template <arch_t arch, typename InIt, typename OutIt, typename Op>
OutIt transform(InIt first, InIt last, OutIt out, Op op) {
    const size_t n = static_cast<size_t>(last - first);
    if (n == 0) return out;
    if constexpr (arch == arch_t::cpu) {
        for (size_t i = 0; i < n; ++i) out[i] = op(first[i]);
    } else {
        transform_kernel<<<grid_blocks(n), 256>>>(first, n, out, op);
        launch_check("transform");
        LaunchSync<arch>();
    }
    return out + n;
}

template <arch_t arch, typename InIt1, typename InIt2, typename OutIt, typename Op>
OutIt transform(InIt1 first1, InIt1 last1, InIt2 first2, OutIt out, Op op) {
    const size_t n = static_cast<size_t>(last1 - first1);
    if (n == 0) return out;
    if constexpr (arch == arch_t::cpu) {
        for (size_t i = 0; i < n; ++i) out[i] = op(first1[i], first2[i]);
    } else {
        transform2_kernel<<<grid_blocks(n), 256>>>(first1, first2, n, out, op);
        launch_check("transform");
        LaunchSync<arch>();
    }
    return out + n;
}

// Per-block folds (each block folds a contiguous run of the input left to right, then its threads'
// results in order), so the result does not depend on scheduling; the host folds the block results.
template <typename InIt, typename T, typename Unary, typename Binary>
__global__ void transform_reduce_kernel(InIt in, size_t n, Unary uop, T init, Binary bop, T* partial) {
    __shared__ T s[256];
    const size_t per_block = (n + gridDim.x - 1) / gridDim.x;
    const size_t b0 = blockIdx.x * per_block;
    const size_t b1 = b0 + per_block < n ? b0 + per_block : n;
    const size_t per_thread = (per_block + 255) / 256;
    const size_t t0 = b0 + threadIdx.x * per_thread;
    T acc = init;
    bool any = false;
    for (size_t i = t0; i < t0 + per_thread && i < b1; ++i) {
        acc = any ? bop(acc, uop(in[i])) : bop(init, uop(in[i]));
        any = true;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = s[0];
        for (int i = 1; i < 256; ++i) t = bop(t, s[i]);
        partial[blockIdx.x] = t;
    }
}

template <arch_t arch, typename InIt, typename Unary, typename T, typename Binary>
T transform_reduce(InIt first, InIt last, Unary uop, T init, Binary bop) {
    const size_t n = static_cast<size_t>(last - first);
    if (n == 0) return init;
    if constexpr (arch == arch_t::cpu) {
        T acc = init;
        for (size_t i = 0; i < n; ++i) acc = bop(acc, uop(first[i]));
        return acc;
    } else {
        const unsigned blocks = grid_blocks(n, 1024);
        T* partial = memalloc<arch_t::hip, T>(blocks);
        transform_reduce_kernel<<<blocks, 256>>>(first, n, uop, init, bop, partial);
        launch_check("transform_reduce");
        std::vector<T> h(blocks);
        memcpy<arch_t::cpu, arch_t::hip, T>(h.data(), partial, blocks);
        memfree<arch_t::hip, T>(partial);
        T acc = init;
        for (unsigned b = 0; b < blocks; ++b) acc = bop(acc, h[b]);
        return acc;
    }
}

template <typename T, typename Idx, typename Out>
__global__ void scatter_kernel(const T* __restrict__ src, size_t n, const Idx* __restrict__ map, Out out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[map[i]] = src[i];
}

template <arch_t arch, typename T, typename Idx, typename Out>
void scatter(const T* first, const T* last, const Idx* map, Out out) {
    const size_t n = static_cast<size_t>(last - first);
    if (n == 0) return;
    if constexpr (arch == arch_t::cpu) {
        for (size_t i = 0; i < n; ++i) out[map[i]] = first[i];
    } else {
        scatter_kernel<<<grid_blocks(n), 256>>>(first, n, map, out);
        launch_check("scatter");
        LaunchSync<arch>();
    }
}

// ---------------------------------------------------------------- device helpers
// Atomic minimum returning the previous value (reference: archi::cuda::AtomicMin, only/cuda.cuh:52-70).
template <typename T>
__device__ __forceinline__ T AtomicMin(T* address, T value) {
    return ::atomicMin(address, value);
}
template <>
__device__ __forceinline__ float AtomicMin<float>(float* address, float value) {
    unsigned int* a = reinterpret_cast<unsigned int*>(address);
    unsigned int old = *a;
    for (;;) {
        const float cur = __uint_as_float(old);
        if (!(value < cur)) return cur;
        const unsigned int prev = ::atomicCAS(a, old, __float_as_uint(value));
        if (prev == old) return cur;
        old = prev;
    }
}
template <>
__device__ __forceinline__ double AtomicMin<double>(double* address, double value) {
    unsigned long long* a = reinterpret_cast<unsigned long long*>(address);
    unsigned long long old = *a;
    for (;;) {
        const double cur = __longlong_as_double(static_cast<long long>(old));
        if (!(value < cur)) return cur;
        const unsigned long long prev = ::atomicCAS(a, old, static_cast<unsigned long long>(__double_as_longlong(value)));
        if (prev == old) return cur;
        old = prev;
    }
}

// Index of the first element greater than target in [begin, end) (reference: only/cuda.cuh:37-50).
template <typename It, typename T>
__host__ __device__ __forceinline__ It UpperBound(It begin, It end, const T& target) {
    while (begin < end) {
        It mid = begin + ((end - begin) >> 1);
        if (*mid <= target) begin = mid + 1; else end = mid;
    }
    return begin;
}

}  // namespace archi
}  // namespace graph_one

