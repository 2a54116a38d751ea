// Device helpers shared by every gfx950 kernel of the library: wave64 ballots and scans, block scans,
// and the hardware constants the launch shapes are derived from.
//
// Written for CDNA wave64 directly: a wavefront is 64 lanes, __ballot returns a 64-bit mask and
// lane prefix counts come from mbcnt. Blocks are multiples of 64 lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace g1 {

constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `mask` at lanes below the calling lane.
__device__ __forceinline__ uint32_t mask_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t ballot(bool pred) { return __ballot(pred); }

// Inclusive prefix sum across the 64 lanes of a wave.
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T t = __shfl_up(v, d);
        if (lane >= static_cast<uint32_t>(d)) v += t;
    }
    return v;
}

// Sum across the 64 lanes of a wave (every lane gets the total).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Exclusive prefix sum over a block of BLOCK threads (BLOCK a multiple of 64). `lds` must hold
// BLOCK/64 elements. Returns the exclusive prefix of the calling thread and writes the block total.
// Contains two __syncthreads(); every thread of the block must call it.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T& total, T* lds) {
    constexpr int kWaves = BLOCK / kWave;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    T incl = wave_inclusive_sum(v);
    if (lane == kWave - 1) lds[wave] = incl;
    __syncthreads();
    T wave_off = 0;
    T tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        T s = lds[w];
        if (static_cast<uint32_t>(w) < wave) wave_off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wave_off + incl - v;
}

// Block-wide count of threads with `pred` set (every thread gets the count).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_count(bool pred, uint32_t* lds) {
    constexpr int kWaves = BLOCK / kWave;
    const uint64_t m = ballot(pred);
    if (lane_id() == 0) lds[threadIdx.x / kWave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) c += lds[w];
    __syncthreads();
    return c;
}

// Block-wide counts of threads with `a` set and with `b` set, packed as count(a) | count(b) << 16 (every
// thread gets them). BLOCK <= 65535.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_count_pair(bool a, bool b, uint32_t* lds) {
    constexpr int kWaves = BLOCK / kWave;
    const uint32_t ca = static_cast<uint32_t>(__popcll(ballot(a)));
    const uint32_t cb = static_cast<uint32_t>(__popcll(ballot(b)));
    if (lane_id() == 0) lds[threadIdx.x / kWave] = ca | (cb << 16);
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) c += lds[w];
    __syncthreads();
    return c;
}

// Index of the last element of the ascending array a[0..len) that is <= key, found by one wave with
// a 64-way search (upper_bound(key) - 1). Requires len >= 1 and a[0] <= key. Every lane of the wave
// must call it; every lane gets the result.
__device__ __forceinline__ uint32_t wave_search_last_le(const uint32_t* __restrict__ a, uint32_t len,
                                                        uint32_t key) {
    const uint32_t lane = lane_id();
    uint32_t lo = 0, hi = len;  // answer in [lo, hi)
    while (hi - lo > static_cast<uint32_t>(kWave)) {
        const uint32_t step = (hi - lo + kWave - 1) / kWave;
        const uint32_t pos = lo + lane * step;
        const bool le = (pos < hi) && (a[pos] <= key);
        const uint32_t k = static_cast<uint32_t>(__popcll(ballot(le)));  // >= 1 since a[lo] <= key
        const uint32_t nlo = lo + (k - 1) * step;
        const uint32_t nhi = nlo + step < hi ? nlo + step : hi;
        lo = nlo;
        hi = nhi;
    }
    const uint32_t pos = lo + lane;
    const bool le = (pos < hi) && (a[pos] <= key);
    const uint32_t k = static_cast<uint32_t>(__popcll(ballot(le)));
    return lo + k - 1;
}

// upper_bound(key) - 1 over an ascending array in LDS (or any memory): index of the last element <= key.
__device__ __forceinline__ uint32_t search_last_le(const uint32_t* a, uint32_t len, uint32_t key) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

}  // namespace g1
