// On-device CSR conversions for gfx950: the in-edge CSR (transpose) of an out-edge CSR, and the hub-first
// relabelling of plain PageRank. Replaces the reference's ToCsc (include/GraphOne/mat/convert.h:62-88),
// which expands row offsets to row indices (OffsetsToIndices, :16-45), sorts the edges by column with a
// stable key sort, then turns the sorted columns back into offsets (IndicesToOffsets, :47-59).
//
// Here the same three steps are three kinds of kernels:
//   expand    csr_expand_rows_kernel: each thread takes 16 consecutive edges, finds the row of the first by
//             binary search over the row offsets and walks forward (no atomics, every edge written once);
//   sort      an LSD radix sort by destination, 8 bits per pass, stable, carrying the source and the weight.
//             Per pass: radix_hist_kernel counts digits per 4096-edge tile (LDS), the digit-major counts are
//             exclusive-scanned, and radix_scatter_kernel writes every edge to its slot. Stability inside a
//             tile comes from wave64 ballots: a lane's peers (lanes with its digit) are the AND of eight
//             ballots, its rank is the number of peers below it (mbcnt), and waves and 256-edge rounds are
//             folded in order through LDS counters;
//   offsets   row_starts_kernel: t_row_off[v] = number of sorted destinations below v (lower_bound).
// The input CSR lists edges by ascending source, so after the stable sort every in-edge row holds its
// sources in ascending order and weights follow their edges: the transpose is a function of the CSR alone,
// byte for byte the same whatever the launch shape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"

namespace g1 {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortRounds;  // edges per radix tile
constexpr uint32_t kSortBins = 256;                          // 8-bit digits
constexpr uint32_t kExpandPer = 16;                           // edges per thread of the expand / copy kernels
constexpr uint32_t kScanTile = 4096;                          // items per workgroup of the scan kernels

// Last index r in [0, len) with a[r] <= key (a ascending, a[0] <= key).
__device__ __forceinline__ uint32_t last_le_u32(const uint32_t* __restrict__ a, uint32_t len, uint32_t key) {
    uint32_t lo = 0, hi = len;  // a[lo] <= key, answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= key) lo = mid; else hi = mid;
    }
    return lo;
}

// Number of elements of the ascending array a[0..len) below key.
__device__ __forceinline__ uint32_t count_below_u32(const uint32_t* __restrict__ a, uint32_t len, uint32_t key) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// src[e] = the row of edge e (row_off[n] = m).
static __global__ void __launch_bounds__(256) csr_expand_rows_kernel(uint32_t n, uint32_t m,
                                                                     const uint32_t* __restrict__ row_off,
                                                                     uint32_t* __restrict__ src) {
    const uint64_t e0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kExpandPer;
    if (e0 >= m) return;
    const uint32_t e_end = static_cast<uint32_t>(e0 + kExpandPer < m ? e0 + kExpandPer : m);
    uint32_t r = last_le_u32(row_off, n + 1, static_cast<uint32_t>(e0));
    for (uint32_t e = static_cast<uint32_t>(e0); e < e_end; ++e) {
        while (row_off[r + 1] <= e) ++r;
        src[e] = r;
    }
}

// Digit counts of one 4096-key tile: hist[d * nb + tile] (digit-major, so one exclusive scan of the whole
// array gives every (digit, tile) its first output slot).
static __global__ void __launch_bounds__(kSortThreads) radix_hist_kernel(const uint32_t* __restrict__ keys,
                                                                         uint32_t count, uint32_t shift,
                                                                         uint32_t nb, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[kSortBins];
    s_h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < count) atomicAdd(&s_h[(keys[i] >> shift) & (kSortBins - 1)], 1u);
    }
    __syncthreads();
    hist[static_cast<size_t>(threadIdx.x) * nb + blockIdx.x] = s_h[threadIdx.x];
}

// One stable pass: key i goes to slot hist_scanned[d * nb + tile] + (keys of its tile with digit d before it).
// The tile is read in 16 rounds of 256 consecutive keys (thread t takes key round * 256 + t, so lane and
// wave order is key order). kW: a weight array travels too.
template <bool kW>
static __global__ void __launch_bounds__(kSortThreads)
radix_scatter_kernel(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ uin, const uint32_t* __restrict__ win,
                     uint32_t count, uint32_t shift, uint32_t nb, const uint32_t* __restrict__ hist,
                     uint32_t* __restrict__ kout, uint32_t* __restrict__ uout, uint32_t* __restrict__ wout) {
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ uint32_t s_base[kSortBins];         // next free slot of each digit for this tile
    __shared__ uint32_t s_wcnt[kWaves][kSortBins];  // keys per (wave, digit) of the current round
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t / kWave;
    s_base[t] = hist[static_cast<size_t>(t) * nb + blockIdx.x];
    const uint32_t tile0 = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s_wcnt[w][t] = 0u;  // column t: only thread t touches it outside (B)-(C)
        __syncthreads();                                       // (A)
        const uint32_t i = tile0 + r * kSortThreads + t;
        const bool valid = i < count;
        uint32_t k = 0, u = 0, wv = 0, d = 0;
        if (valid) {
            k = kin[i];
            u = uin[i];
            if (kW) wv = win[i];
            d = (k >> shift) & (kSortBins - 1);
        }
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = ((d >> b) & 1u) != 0;
            const uint64_t bb = ballot(valid && bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = mask_prefix(peers);
        if (valid && rank == 0) s_wcnt[wave][d] = static_cast<uint32_t>(__popcll(peers));
        __syncthreads();  // (B)
        if (valid) {
            uint32_t off = s_base[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) off += s_wcnt[w][d];
            kout[off] = k;
            uout[off] = u;
            if (kW) wout[off] = wv;
        }
        __syncthreads();  // (C)
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) add += s_wcnt[w][t];
        s_base[t] += add;
    }
}

// t_row_off[v] = number of sorted keys below v, for v in [0, n].
static __global__ void __launch_bounds__(256) row_starts_kernel(uint32_t n, uint32_t m,
                                                                const uint32_t* __restrict__ sorted_keys,
                                                                uint32_t* __restrict__ t_row_off) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += stride)
        t_row_off[v] = (v == n) ? m : count_below_u32(sorted_keys, m, v);
}

// ---- exclusive scan of u32 (totals below 2^32): partial sums per 4096-item tile, one workgroup scans the
// partials, then every tile rescans its items from its partial.

static __global__ void __launch_bounds__(256) scan_partials_kernel(const uint32_t* __restrict__ in, uint32_t count,
                                                                   uint32_t* __restrict__ partials) {
    __shared__ uint32_t s_red[256 / kWave];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t acc = 0;
    for (uint32_t k = threadIdx.x; k < kScanTile; k += 256) {
        const uint32_t i = base + k;
        if (i < count) acc += in[i];
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < 256 / kWave; ++w) tot += s_red[w];
        partials[blockIdx.x] = tot;
    }
}

// One workgroup: exclusive scan of partials[0..count) in place, 1024 at a time with a running carry.
static __global__ void __launch_bounds__(1024) scan_carry_kernel(uint32_t* __restrict__ partials, uint32_t count) {
    __shared__ uint32_t s_lds[1024 / kWave];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += 1024) {  // uniform
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < count ? partials[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_sum<1024>(v, total, s_lds);
        if (i < count) partials[i] = carry + ex;
        carry += total;
    }
}

// out[i] = partials[tile] + exclusive sum of in[tile start .. i); in and out may alias. Thread t owns 16
// consecutive items of the tile.
static __global__ void __launch_bounds__(256) scan_apply_kernel(const uint32_t* in, uint32_t count,
                                                                const uint32_t* __restrict__ partials, uint32_t* out) {
    __shared__ uint32_t s_lds[256 / kWave];
    constexpr uint32_t kPer = kScanTile / 256;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        v[k] = (base + k < count) ? in[base + k] : 0u;
        local += v[k];
    }
    uint32_t total;
    uint32_t run = partials[blockIdx.x] + block_exclusive_sum<256>(local, total, s_lds);
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        if (base + k < count) out[base + k] = run;
        run += v[k];
    }
}

// u32 -> f32 (edge weights of the plus-times engines); exact for weights below 2^24.
static __global__ void __launch_bounds__(256) u32_to_f32_kernel(size_t count, const uint32_t* __restrict__ in,
                                                                float* __restrict__ out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        out[i] = static_cast<float>(in[i]);
}

// ---- hub-first relabelling (plain PageRank): keys[v] = maxd - outdeg(v), ids[v] = v; after a stable sort by
// key, perm = ids (new id i is old vertex perm[i]: out-degree descending, ties by id).
static __global__ void __launch_bounds__(256) degree_keys_kernel(uint32_t n, uint32_t maxd,
                                                                 const uint32_t* __restrict__ row_off,
                                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        keys[v] = maxd - (row_off[v + 1] - row_off[v]);
        ids[v] = v;
    }
}

// inv[perm[i]] = i; lens[i] = in-degree of perm[i], lens[n] = 0 (scanned into the new row offsets);
// nosum[i] = osum[perm[i]].
static __global__ void __launch_bounds__(256) relabel_rows_kernel(uint32_t n, const uint32_t* __restrict__ perm,
                                                                  const uint32_t* __restrict__ t_row_off,
                                                                  const float* __restrict__ osum, uint32_t* __restrict__ inv,
                                                                  uint32_t* __restrict__ lens, float* __restrict__ nosum) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        if (i == n) {
            lens[n] = 0u;
            continue;
        }
        const uint32_t v = perm[i];
        inv[v] = i;
        lens[i] = t_row_off[v + 1] - t_row_off[v];
        nosum[i] = osum[v];
    }
}

// keys[v] = the number of leading zero bits of out-degree(v) + 1 (0..31: the larger the degree group, the
// smaller the key), ids[v] = v.
static __global__ void __launch_bounds__(256) degree_group_keys_kernel(uint32_t n, const uint32_t* __restrict__ row_off,
                                                                       uint32_t* __restrict__ keys,
                                                                       uint32_t* __restrict__ ids) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        keys[v] = static_cast<uint32_t>(__clz(static_cast<int>(row_off[v + 1] - row_off[v] + 1u)));
        ids[v] = v;
    }
}

// inv[perm[i]] = i; lens[i] = out-degree of perm[i], lens[n] = 0 (scanned into the new row offsets).
static __global__ void __launch_bounds__(256) hub_rows_kernel(uint32_t n, const uint32_t* __restrict__ perm,
                                                              const uint32_t* __restrict__ row_off,
                                                              uint32_t* __restrict__ inv, uint32_t* __restrict__ lens) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        if (i == n) {
            lens[n] = 0u;
            continue;
        }
        const uint32_t v = perm[i];
        inv[v] = i;
        lens[i] = row_off[v + 1] - row_off[v];
    }
}

// New row i is old row perm[i] of (t_row_off, t_col, t_wf) with its entries in their order and every
// column renamed: ncol[j] = inv[t_col[t_row_off[perm[i]] + (j - nro[i])]] (the weight follows unchanged).
// Each thread takes 16 consecutive slots j. Used for the in-edge CSR (float weights, PageRank) and for the
// out-edge CSR (u32 weights, the push engines' copy).
template <typename W>
static __global__ void __launch_bounds__(256) relabel_copy_kernel(uint32_t n, uint32_t m, const uint32_t* __restrict__ nro,
                                                                  const uint32_t* __restrict__ perm,
                                                                  const uint32_t* __restrict__ inv,
                                                                  const uint32_t* __restrict__ t_row_off,
                                                                  const uint32_t* __restrict__ t_col,
                                                                  const W* __restrict__ t_wf,
                                                                  uint32_t* __restrict__ ncol, W* __restrict__ nwf) {
    const uint64_t j0 = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kExpandPer;
    if (j0 >= m) return;
    const uint32_t j_end = static_cast<uint32_t>(j0 + kExpandPer < m ? j0 + kExpandPer : m);
    uint32_t i = last_le_u32(nro, n + 1, static_cast<uint32_t>(j0));
    for (uint32_t j = static_cast<uint32_t>(j0); j < j_end; ++j) {
        while (nro[i + 1] <= j) ++i;
        const uint32_t k = t_row_off[perm[i]] + (j - nro[i]);
        ncol[j] = inv[t_col[k]];
        if (nwf) nwf[j] = t_wf[k];
    }
}

}  // namespace g1
