// Merge-path semiring SpMV for gfx950 (Merrill & Garland, "Merge-based parallel sparse matrix-vector
// multiplication", SC16), the algorithm the reference takes from CUB
// (include/GraphOne/archi/blas/spmv/cuda/merge_based/*), designed here for wave64 and LDS.
//
// The merge path walks the two sorted lists A = row_end[0..n_rows) (row_end[r] = row_offsets[r+1])
// and B = 0..nnz-1 (nonzero indices). Step (i, j) consumes nonzero j when j < row_end[i], otherwise
// it closes row i. The n_rows + nnz steps are cut into tiles of kMergeTile steps (one workgroup of
// kMergeThreads threads, kMergeItems consecutive steps per thread). Tile start coordinates are found
// once per matrix (merge_partition_kernel, the reference's SegmentSearch run once in the dispatcher
// constructor, include/GraphOne/archi/blas/spmv/cuda/cuda.cuh:89-91) and reused by every multiply.
//
// Fold order (what makes plus-times results reproducible): inside a thread a row's products are
// added left to right; a row that crosses thread boundaries is finished by the thread that closes it
// by merging, left to right, the tails of the threads before it back to the last thread that closed
// a row; a row that crosses tiles is finished by the fixup kernel, which merges the carries of the
// tiles it crosses, left to right, followed by the closing tile's head. The order therefore depends
// only on the matrix and the constants below, never on scheduling, so repeated runs and the batched
// SpMM (which uses the same partition) give the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/semiring.hpp"

namespace g1 {

constexpr int kMergeThreads = 256;
constexpr int kMergeItems = 8;
constexpr int kMergeTile = kMergeThreads * kMergeItems;  // 2048 merge steps per workgroup

struct MergeCoord {
    uint32_t x;  // rows closed before this point
    uint32_t y;  // nonzeros consumed before this point
};

// Per-tile record left by the tile kernel for the fixup kernel.
template <typename S>
struct MergeTileRec {
    typename S::acc_t carry;  // fold of the row still open at the tile end (after its last row end)
    typename S::acc_t head;   // fold, inside the tile, of the first row it closes (when deferred)
    uint32_t flags;           // bit 0: the tile closes at least one row; bit 1: head is deferred
};

constexpr uint32_t kTileHasEnd = 1u;
constexpr uint32_t kTileDeferred = 2u;

__host__ __device__ inline uint32_t merge_num_tiles_for(uint32_t n_rows, uint32_t nnz, uint32_t tile_items) {
    const uint64_t items = static_cast<uint64_t>(n_rows) + nnz;
    return static_cast<uint32_t>((items + tile_items - 1) / tile_items);
}
__host__ __device__ inline uint32_t merge_num_tiles(uint32_t n_rows, uint32_t nnz) {
    return merge_num_tiles_for(n_rows, nnz, kMergeTile);
}

// Workspace: coords[num_tiles + 1] then records[num_tiles] (sized for the largest record type).
constexpr size_t kMergeRecBytes = 32;
inline size_t merge_workspace_bytes(uint32_t n_rows, uint32_t nnz) {
    const size_t t = merge_num_tiles(n_rows, nnz);
    const size_t coords = ((t + 1) * sizeof(MergeCoord) + 255) & ~size_t(255);
    return coords + t * kMergeRecBytes + 256;
}
inline MergeCoord* merge_coords(void* ws) { return static_cast<MergeCoord*>(ws); }
template <typename S>
inline MergeTileRec<S>* merge_records(void* ws, uint32_t n_rows, uint32_t nnz) {
    static_assert(sizeof(MergeTileRec<S>) <= kMergeRecBytes, "tile record too large");
    const size_t t = merge_num_tiles(n_rows, nnz);
    const size_t coords = ((t + 1) * sizeof(MergeCoord) + 255) & ~size_t(255);
    return reinterpret_cast<MergeTileRec<S>*>(static_cast<char*>(ws) + coords);
}

// Merge-path search on diagonal `diag`: the smallest i with row_end[i] > diag - i - 1
// (reference: merge_based/thread_search.cuh:11-52).
__device__ __forceinline__ MergeCoord merge_path_search(uint64_t diag, const uint32_t* __restrict__ row_end,
                                                        uint32_t n_rows, uint32_t nnz) {
    uint32_t lo = diag > nnz ? static_cast<uint32_t>(diag - nnz) : 0u;
    uint32_t hi = diag < n_rows ? static_cast<uint32_t>(diag) : n_rows;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (static_cast<uint64_t>(row_end[mid]) + mid + 1 <= diag) lo = mid + 1; else hi = mid;
    }
    return MergeCoord{lo, static_cast<uint32_t>(diag - lo)};
}

static __global__ void __launch_bounds__(256) merge_partition_kernel(const uint32_t* __restrict__ row_offsets,
                                                              uint32_t n_rows, uint32_t nnz,
                                                              uint32_t num_tiles,
                                                              MergeCoord* __restrict__ coords,
                                                              uint32_t tile_items = kMergeTile) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > num_tiles) return;
    const uint64_t total = static_cast<uint64_t>(n_rows) + nnz;
    uint64_t diag = static_cast<uint64_t>(t) * tile_items;
    if (diag > total) diag = total;
    coords[t] = merge_path_search(diag, row_offsets + 1, n_rows, nnz);
}

// A tile's loads, staged in registers: its coordinates, row ends, column indices, values and gathered
// x values (rows + nzs <= kMergeTile, so each thread owns at most kMergeItems row ends and kMergeItems
// nonzeros).
template <typename S>
struct MergeStage {
    MergeCoord c0, c1;
    uint32_t re[kMergeItems];
    uint32_t ci[kMergeItems];
    typename S::value_t wv[kMergeItems];
    typename S::x_t xv[kMergeItems];
};

// Issues the tile's coordinate, row-end, column and value loads.
template <typename S, bool kHasValues>
__device__ __forceinline__ void merge_load_index(uint32_t tile, const uint32_t* __restrict__ row_offsets,
                                                 const uint32_t* __restrict__ col,
                                                 const typename S::value_t* __restrict__ vals,
                                                 const MergeCoord* __restrict__ coords, MergeStage<S>& st) {
    const uint32_t tid = threadIdx.x;
    st.c0 = coords[tile];
    st.c1 = coords[tile + 1];
    const uint32_t rows = st.c1.x - st.c0.x, nzs = st.c1.y - st.c0.y;
    // Every slot loads (a slot past the tile's rows or nonzeros reads the tile's last one and is not
    // used), so no load waits behind another's branch. Column and value loads stream (read once), so the
    // gathered x keeps more of the L2.
    const uint32_t last_nz = nzs > 0 ? st.c0.y + nzs - 1 : 0u;
#pragma unroll
    for (int r = 0; r < kMergeItems; ++r) {
        const uint32_t k = r * kMergeThreads + tid;
        st.re[r] = row_offsets[k < rows ? st.c0.x + 1 + k : st.c1.x];
        st.ci[r] = 0u;
        st.wv[r] = typename S::value_t(1);
        if (nzs > 0) {  // uniform across the workgroup (col and vals may be null without nonzeros)
            const uint32_t e = k < nzs ? st.c0.y + k : last_nz;
            st.ci[r] = __builtin_nontemporal_load(col + e);
            if (kHasValues) st.wv[r] = __builtin_nontemporal_load(vals + e);
        }
    }
}

// Issues the gathers of the tile's x values (its column loads must have arrived). Every gather is issued
// (a slot past the tile's nonzeros reads x[0], unused), so none waits behind another's branch; a tile
// without nonzeros reads nothing (x may be null when the matrix has none).
template <typename S>
__device__ __forceinline__ void merge_load_x(const typename S::x_t* __restrict__ x, MergeStage<S>& st) {
    const uint32_t tid = threadIdx.x;
    const uint32_t nzs = st.c1.y - st.c0.y;
    if (nzs == 0) return;  // uniform across the workgroup
#pragma unroll
    for (int r = 0; r < kMergeItems; ++r) {
        const uint32_t k = r * kMergeThreads + tid;
        st.xv[r] = x[k < nzs ? st.ci[r] : 0u];
    }
}

// Writes the staged row ends and products to LDS.
template <typename S>
__device__ __forceinline__ void merge_to_lds(const MergeStage<S>& st, uint32_t* s_rowend, typename S::y_t* s_prod) {
    const uint32_t tid = threadIdx.x;
    const uint32_t rows = st.c1.x - st.c0.x, nzs = st.c1.y - st.c0.y;
#pragma unroll
    for (int r = 0; r < kMergeItems; ++r) {
        const uint32_t k = r * kMergeThreads + tid;
        if (k < rows) s_rowend[k] = st.re[r];
        if (k < nzs) s_prod[k] = S::combine(st.wv[r], st.xv[r]);
    }
    if (tid == 0) s_rowend[rows] = kInf;  // no further row end inside this tile
}

// The last thread below t that closed a row (-1 if none), from the waves' ballots of the closing threads
// (masks[2w], masks[2w + 1]: the 64 flags of wave w): at most kMergeThreads / 64 words read.
__device__ __forceinline__ int merge_last_end_before(const uint32_t* masks, uint32_t t) {
    for (int w = static_cast<int>((t - 1) / kWave); t > 0 && w >= 0; --w) {
        uint64_t m = (static_cast<uint64_t>(masks[2 * w + 1]) << 32) | masks[2 * w];
        const uint32_t below = t - static_cast<uint32_t>(w) * kWave;  // flags of lanes < below count
        if (below < static_cast<uint32_t>(kWave)) m &= (1ull << below) - 1ull;
        if (m) return w * kWave + 63 - __clzll(m);
    }
    return -1;
}

// The left-to-right fold of the tails of threads [v0, v1): the loads issued eight at a time ahead of their
// merges, so a fold across a long row does not wait on each LDS read in turn (same order, same result).
template <typename S>
__device__ __forceinline__ typename S::acc_t merge_fold_tails(const typename S::acc_t* s_tail, uint32_t v0,
                                                              uint32_t v1) {
    typename S::acc_t chain = S::identity();
    uint32_t v = v0;
    for (; v + 8 <= v1; v += 8) {
        typename S::acc_t t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = s_tail[v + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) chain = S::merge(chain, t[k]);
    }
    for (; v < v1; ++v) chain = S::merge(chain, s_tail[v]);
    return chain;
}

// The walk of one staged tile (LDS filled, after a barrier): each thread's kMergeItems steps, the rows
// it closes, the tails chain, and the tile record. Contains one barrier; every thread must call it.
template <typename S>
__device__ __forceinline__ void merge_walk(uint32_t tile, MergeCoord c0, MergeCoord c1,
                                           const uint32_t* __restrict__ row_offsets, typename S::y_t* __restrict__ y,
                                           MergeTileRec<S>* __restrict__ recs, const uint32_t* s_rowend,
                                           const typename S::y_t* s_prod, typename S::acc_t* s_tail,
                                           uint32_t* s_hasend) {
    using acc_t = typename S::acc_t;
    const uint32_t tid = threadIdx.x;
    const uint32_t rows = c1.x - c0.x;
    const uint32_t nzs = c1.y - c0.y;
    const uint32_t total = rows + nzs;

    // This thread's first step inside the tile, then its kMergeItems steps.
    uint32_t d = tid * kMergeItems;
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
    uint32_t j = d - lo;

    acc_t acc = S::identity();
    acc_t head = S::identity();
    bool hasend = false;
    uint32_t first_row = 0;
    for (uint32_t k = d; k < d_end; ++k) {
        if (c0.y + j < s_rowend[i]) {
            S::add(acc, s_prod[j]);
            ++j;
        } else {
            if (!hasend) {
                head = acc;
                hasend = true;
                first_row = c0.x + i;
            } else {
                y[c0.x + i] = S::finalize(acc);
            }
            acc = S::identity();
            ++i;
        }
    }
    s_tail[tid] = acc;
    // which threads closed a row: one 64-bit ballot per wave, as two words (s_hasend[2w], s_hasend[2w + 1])
    const uint64_t ends = ballot(hasend);
    if (lane_id() == 0) {
        s_hasend[2 * (tid / kWave)] = static_cast<uint32_t>(ends);
        s_hasend[2 * (tid / kWave) + 1] = static_cast<uint32_t>(ends >> 32);
    }
    __syncthreads();

    if (hasend) {
        const int u = merge_last_end_before(s_hasend, tid);
        const acc_t chain = merge_fold_tails<S>(s_tail, u >= 0 ? static_cast<uint32_t>(u) : 0u, tid);
        const acc_t whole = S::merge(chain, head);
        // The tile's first closed row began in an earlier tile: leave it for the fixup kernel.
        if (u < 0 && row_offsets[first_row] < c0.y) {
            recs[tile].head = whole;
        } else {
            y[first_row] = S::finalize(whole);
        }
    }
    if (tid == kMergeThreads - 1) {
        const int u = merge_last_end_before(s_hasend, kMergeThreads);
        const acc_t carry = merge_fold_tails<S>(s_tail, u >= 0 ? static_cast<uint32_t>(u) : 0u, kMergeThreads);
        recs[tile].carry = carry;
        uint32_t flags = (u >= 0) ? kTileHasEnd : 0u;
        // Deferred exactly when the tile's first closed row (row c0.x) began in an earlier tile; the
        // first closing thread made the same test above and stored its head instead of writing y.
        if (u >= 0 && row_offsets[c0.x] < c0.y) flags |= kTileDeferred;
        recs[tile].flags = flags;
    }
}

// One workgroup per tile. LDS: row ends and products of the tile, then one tail per thread.
template <typename S, bool kHasValues>
__global__ void __launch_bounds__(kMergeThreads)
spmv_merge_tile_kernel(uint32_t n_rows, uint32_t nnz, const uint32_t* __restrict__ row_offsets,
                       const uint32_t* __restrict__ col, const typename S::value_t* __restrict__ vals,
                       const typename S::x_t* __restrict__ x, typename S::y_t* __restrict__ y,
                       const MergeCoord* __restrict__ coords, MergeTileRec<S>* __restrict__ recs) {
    using acc_t = typename S::acc_t;
    using y_t = typename S::y_t;
    __shared__ uint32_t s_rowend[kMergeTile + 1];
    __shared__ y_t s_prod[kMergeTile];
    __shared__ acc_t s_tail[kMergeThreads];
    __shared__ uint32_t s_hasend[kMergeThreads];
    const uint32_t tile = blockIdx.x;
    MergeStage<S> st;
    merge_load_index<S, kHasValues>(tile, row_offsets, col, vals, coords, st);
    merge_load_x<S>(x, st);
    merge_to_lds<S>(st, s_rowend, s_prod);
    __syncthreads();
    merge_walk<S>(tile, st.c0, st.c1, row_offsets, y, recs, s_rowend, s_prod, s_tail, s_hasend);
}

// One thread per tile: finish the first closed row of every deferred tile by folding, left to right,
// the carries of the earlier tiles the row crosses and then the tile's head (the reference folds
// these carries with CAS atomics in arbitrary order, agent_segment_fixup.cuh:50-80,257-302).
template <typename S>
__global__ void __launch_bounds__(256) spmv_merge_fixup_kernel(uint32_t num_tiles,
                                                              const MergeCoord* __restrict__ coords,
                                                              const MergeTileRec<S>* __restrict__ recs,
                                                              typename S::y_t* __restrict__ y) {
    using acc_t = typename S::acc_t;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tiles || t == 0) return;
    if ((recs[t].flags & kTileDeferred) == 0u) return;
    const uint32_t row = coords[t].x;
    // the nearest tile u < t that closes a row (or tile 0), eight flags loaded per step: a hub row crosses
    // dozens of tiles, and one dependent load per tile would be as many round trips
    uint32_t u = t - 1;
    for (;;) {
        uint32_t f[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) f[k] = k < u ? recs[u - k].flags : kTileHasEnd;  // tile 0 ends the search
        uint32_t hit = 8;
#pragma unroll
        for (int k = 7; k >= 0; --k)
            if (f[k] & kTileHasEnd) hit = static_cast<uint32_t>(k);
        if (hit < 8) {
            u -= hit;
            break;
        }
        u -= 8;  // u > 7 here: otherwise k = u (tile 0) would have hit
    }
    acc_t acc = S::identity();
    uint32_t v = u;
    for (; v + 8 <= t; v += 8) {
        acc_t cv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) cv[k] = recs[v + k].carry;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = S::merge(acc, cv[k]);
    }
    for (; v < t; ++v) acc = S::merge(acc, recs[v].carry);
    acc = S::merge(acc, recs[t].head);
    y[row] = S::finalize(acc);
}

// One lane's share of a CSR-vector row: entries e, e + STRIDE, ... below end, folded into acc in that
// order. Four column indices are loaded before their four gathers, so a long share keeps four
// independent gathers in flight instead of one (the fold order, and so the result, is unchanged).
template <typename S, bool kHasValues, uint32_t STRIDE>
__device__ __forceinline__ void csrvec_fold(const uint32_t* __restrict__ col,
                                            const typename S::value_t* __restrict__ vals,
                                            const typename S::x_t* __restrict__ x, uint32_t e, uint32_t end,
                                            typename S::acc_t& acc) {
    using value_t = typename S::value_t;
    for (; e < end && end - e > 3 * STRIDE; e += 4 * STRIDE) {
        uint32_t c[4];
        value_t w[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            c[r] = col[e + r * STRIDE];
            w[r] = kHasValues ? vals[e + r * STRIDE] : value_t(1);
        }
        typename S::x_t xv[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) xv[r] = x[c[r]];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) S::add(acc, S::combine(w[r], xv[r]));
    }
    for (; e < end; e += STRIDE) {
        const value_t w = kHasValues ? vals[e] : value_t(1);
        S::add(acc, S::combine(w, x[col[e]]));
    }
}

// CSR-vector: L lanes per row (L a power of two dividing 64), lanes stride over the row, then a
// shuffle tree folds lane results with lower lanes on the left. Rows longer than long_min (a list,
// long_rows[0 .. n_long), built once per matrix) are taken by the first n_long workgroups of the same
// launch, one row each: 256 threads striding over it, each folding its entries left to right, then a
// fixed-order tree across the workgroup (lower threads on the left). A hub row of a skewed graph then no
// longer holds one group of L lanes for its whole length, and it runs beside the short rows.
template <typename S, bool kHasValues, int L>
__global__ void __launch_bounds__(256)
spmv_csrvec_kernel(uint32_t n_rows, const uint32_t* __restrict__ row_offsets, const uint32_t* __restrict__ col,
                   const typename S::value_t* __restrict__ vals, const typename S::x_t* __restrict__ x,
                   typename S::y_t* __restrict__ y, uint32_t long_min, const uint32_t* __restrict__ long_rows,
                   uint32_t n_long) {
    using acc_t = typename S::acc_t;
    if (blockIdx.x < n_long) {
        __shared__ acc_t s_acc[256];
        const uint32_t row = long_rows[blockIdx.x];
        const uint32_t beg = row_offsets[row];
        const uint32_t end = row_offsets[row + 1];
        acc_t acc = S::identity();
        csrvec_fold<S, kHasValues, 256>(col, vals, x, beg + threadIdx.x, end, acc);
        s_acc[threadIdx.x] = acc;
        __syncthreads();
#pragma unroll
        for (uint32_t d = 128; d >= 1; d >>= 1) {
            if (threadIdx.x < d) s_acc[threadIdx.x] = S::merge(s_acc[threadIdx.x], s_acc[threadIdx.x + d]);
            __syncthreads();
        }
        if (threadIdx.x == 0) y[row] = S::finalize(s_acc[0]);
        return;
    }
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x - n_long) * 256 + threadIdx.x;
    const uint64_t row = gid / L;
    const uint32_t lane = static_cast<uint32_t>(gid % L);
    acc_t acc = S::identity();
    bool mine = false;
    if (row < n_rows) {
        const uint32_t beg = row_offsets[row];
        const uint32_t end = row_offsets[row + 1];
        mine = end - beg <= long_min;
        if (mine) csrvec_fold<S, kHasValues, L>(col, vals, x, beg + lane, end, acc);
    }
#pragma unroll
    for (int off = L / 2; off >= 1; off >>= 1) {
        const acc_t other = S::shfl_down(acc, off, L);
        if (lane + off < static_cast<uint32_t>(L)) acc = S::merge(acc, other);
    }
    if (mine && lane == 0) y[row] = S::finalize(acc);
}

}  // namespace g1
