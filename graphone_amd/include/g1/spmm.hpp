// Batched multi-source semiring SpMM (k sources as k dense columns, row-major n x k) for gfx950.
//
// Same merge-path tiles as the SpMV of g1/spmv_merge.hpp (same coordinates, kMergeThreads "virtual
// threads" of kMergeItems merge steps per tile); here each virtual thread is executed by a group of K
// lanes, lane j owning column j, so the K lanes walk the same steps in lockstep and load the K values of
// a source row with one coalesced access. A workgroup runs its tile's virtual threads in rounds of
// 256/K; after each round the lanes of column j (threads 0..K-1) fold, left to right, the round's heads
// and tails into the open row exactly as the SpMV's closing threads do. The fold order of every column
// is therefore the SpMV's fold order, bit for bit: a batched plus-times column equals the single-source
// SpMV of that column.
//
// Rows are finished by a policy (row_done) called together by the K lanes of a group, with the row's
// K column values; the policy fuses the engine's apply step. Rows that cross tiles are finished by the
// fixup kernel with the same policy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/semiring.hpp"
#include "g1/spmv_merge.hpp"

namespace g1 {

constexpr int kSpmmMaxK = 64;

// Per-tile record of the SpMM: K carries and K deferred heads, then flags.
template <typename S, int K>
struct SpmmTileRec {
    typename S::acc_t carry[K];
    typename S::acc_t head[K];
    uint32_t flags;
};

template <typename S, int K>
inline size_t spmm_records_bytes(uint32_t n_rows, uint32_t nnz) {
    return static_cast<size_t>(merge_num_tiles(n_rows, nnz)) * sizeof(SpmmTileRec<S, K>);
}

// Lanes of the calling thread's K-lane group as a mask of the group's ballot bits.
template <int K>
__device__ __forceinline__ uint64_t group_bits(uint64_t wave_ballot) {
    if constexpr (K == 64) return wave_ballot;
    const uint32_t shift = (lane_id() / K) * K;
    return (wave_ballot >> shift) & ((1ull << K) - 1ull);
}

template <typename S, int K, typename P>
__global__ void __launch_bounds__(256)
spmm_tile_kernel(uint32_t n_rows, uint32_t nnz, const uint32_t* __restrict__ row_offsets,
                 const uint32_t* __restrict__ col, const typename S::value_t* __restrict__ vals,
                 const MergeCoord* __restrict__ coords, SpmmTileRec<S, K>* __restrict__ recs, P pol,
                 uint32_t* __restrict__ tile_counts) {
    using acc_t = typename S::acc_t;
    using value_t = typename S::value_t;
    constexpr int G = 256 / K;  // virtual threads per round
    __shared__ uint32_t s_rowend[kMergeTile + 1];
    __shared__ uint32_t s_col[kMergeTile];
    __shared__ value_t s_w[kMergeTile];
    __shared__ acc_t s_head[G][K];
    __shared__ acc_t s_tail[G][K];
    __shared__ uint32_t s_hasend[G];
    __shared__ uint32_t s_first[G];
    __shared__ uint32_t s_red[4];

    const uint32_t tile = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const MergeCoord c0 = coords[tile];
    const MergeCoord c1 = coords[tile + 1];
    const uint32_t rows = c1.x - c0.x;
    const uint32_t nzs = c1.y - c0.y;
    const uint32_t total = rows + nzs;
    {
        uint32_t re[kMergeItems], ci[kMergeItems];
        value_t wv[kMergeItems];
#pragma unroll
        for (int r = 0; r < kMergeItems; ++r) {
            const uint32_t k = r * 256 + tid;
            if (k < rows) re[r] = row_offsets[c0.x + 1 + k];
            if (k < nzs) {
                ci[r] = col[c0.y + k];
                wv[r] = vals ? vals[c0.y + k] : value_t(1);
            }
        }
#pragma unroll
        for (int r = 0; r < kMergeItems; ++r) {
            const uint32_t k = r * 256 + tid;
            if (k < rows) s_rowend[k] = re[r];
            if (k < nzs) {
                s_col[k] = ci[r];
                s_w[k] = wv[r];
            }
        }
        if (tid == 0) s_rowend[rows] = kInf;
    }
    __syncthreads();

    const uint32_t j = tid % K;  // column of this lane
    const uint32_t g = tid / K;  // virtual thread slot within a round
    acc_t carry = S::identity();  // threads 0..K-1: open-row fold of column j
    bool seen_end = false;
    uint32_t counted = 0;  // rows this thread's policy calls reported
    // The first closed row is deferred to the fixup when it began in an earlier tile.
    const bool first_started_before = (rows > 0) && (row_offsets[c0.x] < c0.y);

    for (int round = 0; round < kMergeThreads / G; ++round) {
        const uint32_t vt = round * G + g;
        uint32_t d = vt * kMergeItems;
        if (d > total) d = total;
        uint32_t d_end = d + kMergeItems;
        if (d_end > total) d_end = total;
        uint32_t lo = d > nzs ? d - nzs : 0u;
        uint32_t hi = d < rows ? d : rows;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (static_cast<uint64_t>(s_rowend[mid]) + 1 <= static_cast<uint64_t>(c0.y) + (d - mid)) lo = mid + 1;
            else hi = mid;
        }
        uint32_t i = lo;
        uint32_t jn = d - lo;
        acc_t acc = S::identity();
        acc_t head = S::identity();
        bool hasend = false;
        uint32_t first_row = 0;
        for (uint32_t k = d; k < d_end; ++k) {
            if (c0.y + jn < s_rowend[i]) {
                S::add(acc, S::combine(s_w[jn], pol.x(s_col[jn], j)));
                ++jn;
            } else {
                if (!hasend) {
                    head = acc;
                    hasend = true;
                    first_row = c0.x + i;
                } else {
                    counted += pol.row_done(c0.x + i, j, S::finalize(acc));
                }
                acc = S::identity();
                ++i;
            }
        }
        s_head[g][j] = head;
        s_tail[g][j] = acc;
        if (j == 0) {
            s_hasend[g] = hasend ? 1u : 0u;
            s_first[g] = first_row;
        }
        __syncthreads();
        if (tid < static_cast<uint32_t>(K)) {
            for (int gg = 0; gg < G; ++gg) {
                if (s_hasend[gg]) {
                    const acc_t whole = S::merge(carry, s_head[gg][tid]);
                    if (!seen_end && first_started_before) {
                        recs[tile].head[tid] = whole;
                    } else {
                        counted += pol.row_done(s_first[gg], tid, S::finalize(whole));
                    }
                    seen_end = true;
                    carry = S::merge(S::identity(), s_tail[gg][tid]);
                } else {
                    carry = S::merge(carry, s_tail[gg][tid]);
                }
            }
        }
        __syncthreads();
    }
    if (tid < static_cast<uint32_t>(K)) {
        recs[tile].carry[tid] = carry;
        if (tid == 0) {
            uint32_t flags = seen_end ? kTileHasEnd : 0u;
            if (seen_end && first_started_before) flags |= kTileDeferred;
            recs[tile].flags = flags;
        }
    }
    if (tile_counts) {
        const uint32_t w = wave_sum(counted);
        if (lane_id() == 0) s_red[tid / kWave] = w;
        __syncthreads();
        if (tid == 0) tile_counts[tile] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
    }
}

// One group of K lanes per tile: finish the deferred first row of each tile (carries of the tiles it
// crosses, left to right, then the tile's head), as spmv_merge_fixup_kernel does per column.
template <typename S, int K, typename P>
__global__ void __launch_bounds__(256) spmm_fixup_kernel(uint32_t num_tiles, const MergeCoord* __restrict__ coords,
                                                        const SpmmTileRec<S, K>* __restrict__ recs, P pol,
                                                        uint32_t* __restrict__ tile_counts) {
    using acc_t = typename S::acc_t;
    const uint32_t t = (blockIdx.x * 256 + threadIdx.x) / K;
    const uint32_t j = threadIdx.x % K;
    const bool live = t < num_tiles && t > 0 && (recs[t].flags & kTileDeferred) != 0u;
    // Whole groups agree on `live` (t is uniform within a group), so row_done is called by full groups.
    if (!live) return;
    uint32_t u = t - 1;
    while (u > 0 && (recs[u].flags & kTileHasEnd) == 0u) --u;
    acc_t acc = S::identity();
    for (uint32_t v = u; v < t; ++v) acc = S::merge(acc, recs[v].carry[j]);
    acc = S::merge(acc, recs[t].head[j]);
    const uint32_t c = pol.row_done(coords[t].x, j, S::finalize(acc));
    if (c && tile_counts) tile_counts[t] += c;  // one group per deferred tile: no other writer
}

// ------------------------------------------------------------------------------------------------
// Batched SSSP / BFS (min-plus). Per vertex: D[v*K + j] the distance of column j, and a mask of the
// columns whose distance changed in the last iteration (mask_prev[v] bit j). A source contributes to
// column j only when that bit is set, so inactive sources cost one 8-byte mask read per nonzero. The
// row's owner applies the result: d = min(D, y); changed bits go to mask_next[v] (written once per row
// per iteration, so the masks need no atomics and no clearing). Rows that changed are counted per tile.
template <int K>
struct MinPlusBatchPolicy {
    uint32_t* D;
    const unsigned long long* mask_prev;
    unsigned long long* mask_next;
    __device__ __forceinline__ uint32_t x(uint32_t u, uint32_t j) const {
        const unsigned long long m = mask_prev[u];
        return ((m >> j) & 1ull) ? D[static_cast<size_t>(u) * K + j] : kInf;
    }
    // Called by the K lanes of a group together (lane j = column j); returns 1 on lane 0 of the group
    // when the row changed in some column (the count of active rows of the next iteration).
    __device__ __forceinline__ uint32_t row_done(uint32_t row, uint32_t j, uint32_t y) const {
        const size_t at = static_cast<size_t>(row) * K + j;
        bool improved = false;
        if (y != kInf) {
            const uint32_t cur = D[at];
            if (y < cur) {
                D[at] = y;
                improved = true;
            }
        }
        const uint64_t bits = group_bits<K>(ballot(improved));
        if (j == 0) mask_next[row] = bits;
        return (j == 0 && bits) ? 1u : 0u;
    }
};

// Batched plus-times: writes y row-major; the PageRank-style apply runs as its own kernel.
template <int K>
struct PlusTimesBatchPolicy {
    const float* X;
    float* Y;
    __device__ __forceinline__ float x(uint32_t u, uint32_t j) const { return X[static_cast<size_t>(u) * K + j]; }
    __device__ __forceinline__ uint32_t row_done(uint32_t row, uint32_t j, float y) const {
        Y[static_cast<size_t>(row) * K + j] = y;
        return 0u;
    }
};

}  // namespace g1
