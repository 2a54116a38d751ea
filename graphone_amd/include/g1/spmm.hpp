// Batched multi-source semiring SpMM (k sources as k dense columns, row-major n x k) for gfx950.
//
// Same merge-path tiles as the SpMV of g1/spmv_merge.hpp (same coordinates, kMergeThreads "virtual
// threads" of kMergeItems merge steps per tile); here each virtual thread is executed by a group of K
// lanes, lane j owning column j, so the K lanes walk the same steps in lockstep and load the K values of
// a source row with one coalesced access. A workgroup runs its tile's virtual threads in K rounds of
// G = 256/K. Every value a round needs from global memory (the x values of its nonzeros and, for the
// min-plus apply, the old values of the rows it closes) is loaded one round ahead into registers, with
// a fixed count of loads per round, so that the loads of round r+1 are in flight while round r walks.
//
// Fold order: a row's value is the fold, left to right, of the virtual threads' partial folds, exactly
// as in spmv_merge_tile_kernel: a closing virtual thread merges the tails since the last closing
// virtual thread (across rounds through the open carry), then its head; the tile's open row leaves as
// the tile carry. A batched plus-times column therefore equals the single-source SpMV of that column,
// bit for bit.
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
    if constexpr (K == 64) {
        return wave_ballot;
    } else {
        const uint32_t shift = (lane_id() / K) * K;
        return (wave_ballot >> shift) & ((1ull << K) - 1ull);
    }
}

template <typename S, int K, typename P>
__global__ void __launch_bounds__(256)
spmm_tile_kernel(uint32_t n_rows, uint32_t nnz, const uint32_t* __restrict__ row_offsets,
                 const uint32_t* __restrict__ col, const typename S::value_t* __restrict__ vals,
                 const MergeCoord* __restrict__ coords, SpmmTileRec<S, K>* __restrict__ recs, P pol,
                 unsigned long long* __restrict__ changed) {
    using acc_t = typename S::acc_t;
    using value_t = typename S::value_t;
    using x_t = typename S::x_t;
    using aux_t = typename P::aux_t;
    constexpr int G = 256 / K;                  // virtual threads per round
    constexpr int kRounds = kMergeThreads / G;  // = K
    constexpr int kWavesG = kWave / K;          // groups per wave
    __shared__ uint32_t s_rowend[kMergeTile + 1];
    __shared__ uint32_t s_col[kMergeTile];
    __shared__ value_t s_w[kMergeTile];
    __shared__ aux_t s_aux[P::kHasAux ? kMergeTile : 1];
    __shared__ uint32_t s_vi[kMergeThreads + 1];  // first row (tile-relative) of each virtual thread
    __shared__ uint32_t s_vj[kMergeThreads + 1];  // first nonzero (tile-relative) of each virtual thread
    __shared__ acc_t s_tail[2][G][K];
    __shared__ acc_t s_carry[2][K];
    __shared__ uint32_t s_hm[2][256 / kWave];  // closing groups of the round, one word per wave
    __shared__ unsigned long long s_or;

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
        if constexpr (P::kHasAux) {
            aux_t ax[kMergeItems];
#pragma unroll
            for (int r = 0; r < kMergeItems; ++r) {
                const uint32_t k = r * 256 + tid;
                if (k < nzs) ax[r] = pol.aux(ci[r]);
            }
#pragma unroll
            for (int r = 0; r < kMergeItems; ++r) {
                const uint32_t k = r * 256 + tid;
                if (k < nzs) s_aux[k] = ax[r];
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
    {
        // start coordinate of virtual thread tid (the SpMV thread's search)
        uint32_t d = tid * kMergeItems;
        if (d > total) d = total;
        uint32_t lo = d > nzs ? d - nzs : 0u;
        uint32_t hi = d < rows ? d : rows;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (static_cast<uint64_t>(s_rowend[mid]) + 1 <= static_cast<uint64_t>(c0.y) + (d - mid)) lo = mid + 1;
            else hi = mid;
        }
        s_vi[tid] = lo;
        s_vj[tid] = d - lo;
        if (tid == 0) {
            s_vi[kMergeThreads] = rows;
            s_vj[kMergeThreads] = nzs;
        }
        if (tid < static_cast<uint32_t>(K)) s_carry[0][tid] = S::identity();
    }
    __syncthreads();

    const uint32_t j = tid % K;  // column of this lane
    const uint32_t g = tid / K;  // virtual thread slot within a round
    const uint32_t wave = tid / kWave;
    // The tile's first closed row is deferred to the fixup when it began in an earlier tile.
    const bool first_started_before = (rows > 0) && (row_offsets[c0.x] < c0.y);
    bool seen = false;     // some row closed in an earlier round
    uint64_t counted = 0;  // OR of the changed-column masks this lane's policy calls reported

    // Loads of round r: x values of its nonzeros, old values of the rows it closes (fixed count).
    auto fetch = [&](int r, x_t (&xs)[kMergeItems], x_t (&ds)[kMergeItems]) {
        const uint32_t vt = r * G + g;
        const uint32_t j0 = s_vj[vt], j1 = s_vj[vt + 1];
        const uint32_t i0 = s_vi[vt], i1 = s_vi[vt + 1];
#pragma unroll
        for (int t = 0; t < kMergeItems; ++t) {
            const bool has = j0 + t < j1;
            const uint32_t q = has ? j0 + t : 0u;
            const uint32_t u = has ? s_col[q] : 0u;
            xs[t] = pol.x(u, has ? s_aux[P::kHasAux ? q : 0u] : aux_t(0), j, has);
            if constexpr (P::kRowOld) {
                const bool hr = i0 + t < i1;
                ds[t] = pol.row_old(hr ? c0.x + i0 + t : 0u, j);
            }
        }
    };

    // One round: walk, then the closing virtual threads. xc/dc hold the round's loads.
    auto round = [&](int r, x_t (&xc)[kMergeItems], x_t (&dc)[kMergeItems]) {

        // ---- walk of virtual thread vt (the SpMV thread's loop, over this lane's column)
        const uint32_t vt = r * G + g;
        uint32_t d = vt * kMergeItems;
        if (d > total) d = total;
        uint32_t d_end = d + kMergeItems;
        if (d_end > total) d_end = total;
        uint32_t i = s_vi[vt];
        uint32_t jn = s_vj[vt];
        x_t p[kMergeItems];
#pragma unroll
        for (int t = 0; t < kMergeItems; ++t) {
            uint32_t q = jn + t;
            if (q >= static_cast<uint32_t>(kMergeTile)) q = kMergeTile - 1;
            p[t] = S::combine(s_w[q], xc[t]);
        }
        acc_t acc = S::identity();
        acc_t head = S::identity();
        bool hasend = false;
        uint32_t first_row = 0;
        x_t head_old = x_t(0);
        uint32_t re = s_rowend[i];
#pragma unroll
        for (int k = 0; k < kMergeItems; ++k) {
            if (d + k < d_end) {
                if (c0.y + jn < re) {
                    S::add(acc, p[0]);
#pragma unroll
                    for (int t = 0; t + 1 < kMergeItems; ++t) p[t] = p[t + 1];
                    ++jn;
                } else {
                    if (!hasend) {
                        head = acc;
                        hasend = true;
                        first_row = c0.x + i;
                        head_old = dc[0];
                    } else {
                        counted |= pol.row_done(c0.x + i, j, S::finalize(acc), dc[0]);
                    }
#pragma unroll
                    for (int t = 0; t + 1 < kMergeItems; ++t) dc[t] = dc[t + 1];
                    acc = S::identity();
                    ++i;
                    re = s_rowend[i];
                }
            }
        }
        const int par = r & 1;
        s_tail[par][g][j] = acc;
        {
            const uint64_t b = ballot(j == 0 && hasend);
            if (lane_id() == 0) {
                uint32_t hw = 0;
#pragma unroll
                for (int q = 0; q < kWavesG; ++q) hw |= static_cast<uint32_t>((b >> (q * K)) & 1ull) << q;
                s_hm[par][wave] = hw << (wave * kWavesG);
            }
        }
        __syncthreads();

        // ---- closing virtual threads: fold the tails since the last closer, then the head
        uint32_t hm = 0;
#pragma unroll
        for (int q = 0; q < 256 / kWave; ++q) hm |= s_hm[par][q];
        const acc_t open = s_carry[par][j];
        if (hasend) {
            const uint32_t below = hm & ((g == 31) ? 0x7fffffffu : ((1u << g) - 1u));
            acc_t chain;
            uint32_t v0;
            if (below) {
                v0 = 31u - __clz(below);
                chain = S::identity();
            } else {
                v0 = 0;
                chain = open;
            }
            for (uint32_t v = v0; v < g; ++v) chain = S::merge(chain, s_tail[par][v][j]);
            const acc_t whole = S::merge(chain, head);
            if (!below && !seen && first_started_before) {
                recs[tile].head[j] = whole;
            } else {
                counted |= pol.row_done(first_row, j, S::finalize(whole), head_old);
            }
        }
        if (g == static_cast<uint32_t>(G - 1)) {
            acc_t c;
            uint32_t v0;
            if (hm) {
                v0 = 31u - __clz(hm);
                c = S::identity();
            } else {
                v0 = 0;
                c = open;
            }
            for (uint32_t v = v0; v < static_cast<uint32_t>(G); ++v) c = S::merge(c, s_tail[par][v][j]);
            s_carry[par ^ 1][j] = c;
            if (r + 1 == kRounds) {
                recs[tile].carry[j] = c;
                if (j == 0) {
                    const bool any = seen || hm != 0;
                    uint32_t flags = any ? kTileHasEnd : 0u;
                    if (any && first_started_before) flags |= kTileDeferred;
                    recs[tile].flags = flags;
                }
            }
        }
        seen = seen || hm != 0;
    };

    // Two register buffers in turn (no copies, so the loads of the next round stay in flight while a
    // round walks); every fetch issues the same number of loads (the last ones repeat the final round).
    x_t xa[kMergeItems], da[kMergeItems], xb[kMergeItems], db[kMergeItems];
    fetch(0, xa, da);
    for (int r = 0; r < kRounds; r += 2) {
        fetch(r + 1, xb, db);
        round(r, xa, da);
        fetch(r + 2 < kRounds ? r + 2 : kRounds - 1, xa, da);
        round(r + 1, xb, db);
    }
    // Columns with a changed row: OR them into *changed, with an atomic only when this tile adds bits
    // that are not set yet (a device-wide word hit by every tile would serialize them).
    if (changed) {
        if (tid == 0) s_or = 0ull;
        __syncthreads();
        if (counted) atomicOr(&s_or, static_cast<unsigned long long>(counted));
        __syncthreads();
        if (tid == 0 && s_or) {
            const unsigned long long cur = *reinterpret_cast<volatile unsigned long long*>(changed);
            if (s_or & ~cur) atomicOr(changed, s_or);
        }
    }
}

// Min-plus variant of the tile kernel for order-free semirings (min is exact, so any fold order gives
// the same bits): group g of K lanes walks the contiguous run of K virtual threads g*K .. g*K+K-1 (8K
// merge items) alone, with its running minimum in registers and no workgroup barrier inside the walk;
// the loads of the next virtual thread are in flight while one walks (two register buffers). The
// group's first closed row and its open tail are combined across groups once, after one barrier,
// in group order. Tile records and the fixup are those of spmm_tile_kernel.
template <int K, typename P>
__global__ void __launch_bounds__(256)
spmm_minplus_kernel(uint32_t n_rows, uint32_t nnz, const uint32_t* __restrict__ row_offsets,
                    const uint32_t* __restrict__ col, const uint32_t* __restrict__ vals,
                    const MergeCoord* __restrict__ coords, SpmmTileRec<MinPlusU32, K>* __restrict__ recs, P pol,
                    unsigned long long* __restrict__ changed) {
    using S = MinPlusU32;
    using acc_t = S::acc_t;
    using aux_t = typename P::aux_t;
    constexpr int G = 256 / K;  // groups
    __shared__ uint32_t s_rowend[kMergeTile + 1];
    __shared__ uint32_t s_col[kMergeTile];
    __shared__ uint32_t s_w[kMergeTile];
    __shared__ aux_t s_aux[kMergeTile];
    __shared__ uint32_t s_vi[kMergeThreads + 1];
    __shared__ uint32_t s_vj[kMergeThreads + 1];
    __shared__ uint32_t s_ghead[G][K];
    __shared__ uint32_t s_gtail[G][K];
    __shared__ uint32_t s_gold[G][K];
    __shared__ uint32_t s_grow[G];
    __shared__ uint32_t s_gend[G];
    __shared__ unsigned long long s_or;

    const uint32_t tile = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const MergeCoord c0 = coords[tile];
    const MergeCoord c1 = coords[tile + 1];
    const uint32_t rows = c1.x - c0.x;
    const uint32_t nzs = c1.y - c0.y;
    const uint32_t total = rows + nzs;
    {
        uint32_t re[kMergeItems], ci[kMergeItems], wv[kMergeItems];
        aux_t ax[kMergeItems];
#pragma unroll
        for (int r = 0; r < kMergeItems; ++r) {
            const uint32_t k = r * 256 + tid;
            if (k < rows) re[r] = row_offsets[c0.x + 1 + k];
            if (k < nzs) {
                ci[r] = col[c0.y + k];
                wv[r] = vals ? vals[c0.y + k] : 1u;
            }
        }
#pragma unroll
        for (int r = 0; r < kMergeItems; ++r) {
            const uint32_t k = r * 256 + tid;
            if (k < nzs) ax[r] = pol.aux(ci[r]);
        }
#pragma unroll
        for (int r = 0; r < kMergeItems; ++r) {
            const uint32_t k = r * 256 + tid;
            if (k < rows) s_rowend[k] = re[r];
            if (k < nzs) {
                s_col[k] = ci[r];
                s_w[k] = wv[r];
                s_aux[k] = ax[r];
            }
        }
        if (tid == 0) {
            s_rowend[rows] = kInf;
            s_or = 0ull;
        }
    }
    __syncthreads();
    {
        uint32_t d = tid * kMergeItems;
        if (d > total) d = total;
        uint32_t lo = d > nzs ? d - nzs : 0u;
        uint32_t hi = d < rows ? d : rows;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (static_cast<uint64_t>(s_rowend[mid]) + 1 <= static_cast<uint64_t>(c0.y) + (d - mid)) lo = mid + 1;
            else hi = mid;
        }
        s_vi[tid] = lo;
        s_vj[tid] = d - lo;
        if (tid == 0) {
            s_vi[kMergeThreads] = rows;
            s_vj[kMergeThreads] = nzs;
        }
    }
    __syncthreads();

    const uint32_t j = tid % K;
    const uint32_t g = tid / K;
    uint64_t counted = 0;

    auto fetch = [&](uint32_t vt, uint32_t (&xs)[kMergeItems], uint32_t (&ds)[kMergeItems]) {
        const uint32_t j0 = s_vj[vt], j1 = s_vj[vt + 1];
        const uint32_t i0 = s_vi[vt], i1 = s_vi[vt + 1];
#pragma unroll
        for (int t = 0; t < kMergeItems; ++t) {
            const bool has = j0 + t < j1;
            const uint32_t q = has ? j0 + t : 0u;
            const uint32_t u = has ? s_col[q] : 0u;
            xs[t] = pol.x(u, has ? s_aux[q] : aux_t(0), j, has);
            const bool hr = i0 + t < i1;
            ds[t] = pol.row_old(hr ? c0.x + i0 + t : 0u, j);
        }
    };

    uint32_t acc = kInf;       // running minimum of the open row
    bool gend = false;         // the group closed a row
    uint32_t grow = 0, ghead = kInf, gold = 0;
    uint32_t i = s_vi[g * K];  // current row (tile-relative)
    uint32_t jn = s_vj[g * K];
    uint32_t re = s_rowend[i];
    auto walk = [&](uint32_t vt, uint32_t (&xc)[kMergeItems], uint32_t (&dc)[kMergeItems]) {
        uint32_t d = vt * kMergeItems;
        if (d > total) d = total;
        uint32_t d_end = d + kMergeItems;
        if (d_end > total) d_end = total;
        uint32_t p[kMergeItems];
#pragma unroll
        for (int t = 0; t < kMergeItems; ++t) {
            uint32_t q = jn + t;
            if (q >= static_cast<uint32_t>(kMergeTile)) q = kMergeTile - 1;
            p[t] = S::combine(s_w[q], xc[t]);
        }
#pragma unroll
        for (int k = 0; k < kMergeItems; ++k) {
            if (d + k < d_end) {
                if (c0.y + jn < re) {
                    acc = p[0] < acc ? p[0] : acc;
#pragma unroll
                    for (int t = 0; t + 1 < kMergeItems; ++t) p[t] = p[t + 1];
                    ++jn;
                } else {
                    if (!gend) {
                        gend = true;
                        grow = c0.x + i;
                        ghead = acc;
                        gold = dc[0];
                    } else {
                        counted |= pol.row_done(c0.x + i, j, acc, dc[0]);
                    }
#pragma unroll
                    for (int t = 0; t + 1 < kMergeItems; ++t) dc[t] = dc[t + 1];
                    acc = kInf;
                    ++i;
                    re = s_rowend[i];
                }
            }
        }
    };

    uint32_t xa[kMergeItems], da[kMergeItems], xb[kMergeItems], db[kMergeItems];
    const uint32_t v0 = g * K;
    fetch(v0, xa, da);
    for (uint32_t r = 0; r < static_cast<uint32_t>(K); r += 2) {
        fetch(v0 + r + 1, xb, db);
        walk(v0 + r, xa, da);
        fetch(v0 + (r + 2 < static_cast<uint32_t>(K) ? r + 2 : K - 1), xa, da);
        walk(v0 + r + 1, xb, db);
    }
    s_ghead[g][j] = ghead;
    s_gtail[g][j] = acc;
    s_gold[g][j] = gold;
    if (j == 0) {
        s_grow[g] = grow;
        s_gend[g] = gend ? 1u : 0u;
    }
    __syncthreads();
    // Groups in order (lanes of group 0): a group's first closed row gets the open minimum carried
    // from the groups before it; the tile's first closed row is deferred when it began before the tile.
    const bool first_started_before = (rows > 0) && (row_offsets[c0.x] < c0.y);
    if (g == 0) {
        uint32_t carry = kInf;
        bool seen = false;
        for (int q = 0; q < G; ++q) {
            if (s_gend[q]) {
                const uint32_t h = s_ghead[q][j];
                const uint32_t whole = h < carry ? h : carry;
                if (!seen && first_started_before) {
                    recs[tile].head[j] = acc_t{whole};
                } else {
                    counted |= pol.row_done(s_grow[q], j, whole, s_gold[q][j]);
                }
                seen = true;
                carry = s_gtail[q][j];
            } else {
                const uint32_t t = s_gtail[q][j];
                carry = t < carry ? t : carry;
            }
        }
        recs[tile].carry[j] = acc_t{carry};
        if (j == 0) {
            uint32_t flags = seen ? kTileHasEnd : 0u;
            if (seen && first_started_before) flags |= kTileDeferred;
            recs[tile].flags = flags;
        }
    }
    if (changed) {
        if (counted) atomicOr(&s_or, static_cast<unsigned long long>(counted));
        __syncthreads();
        if (tid == 0 && s_or) {
            const unsigned long long cur = *reinterpret_cast<volatile unsigned long long*>(changed);
            if (s_or & ~cur) atomicOr(changed, s_or);
        }
    }
}

// One group of K lanes per tile: finish the deferred first row of each tile (carries of the tiles it
// crosses, left to right, then the tile's head), as spmv_merge_fixup_kernel does per column.
template <typename S, int K, typename P>
__global__ void __launch_bounds__(256) spmm_fixup_kernel(uint32_t num_tiles, const MergeCoord* __restrict__ coords,
                                                        const SpmmTileRec<S, K>* __restrict__ recs, P pol,
                                                        unsigned long long* __restrict__ changed) {
    using acc_t = typename S::acc_t;
    const uint32_t t = (blockIdx.x * 256 + threadIdx.x) / K;
    const uint32_t j = threadIdx.x % K;
    const bool live = t < num_tiles && t > 0 && (recs[t].flags & kTileDeferred) != 0u;
    // Whole groups agree on `live` (t is uniform within a group), so row_done is called by full groups.
    if (!live) return;
    // The nearest tile u < t that closes a row (or tile 0): the group's K lanes test K tiles per step.
    uint32_t u = t - 1;
    for (;;) {
        const bool has_end = j <= u && (recs[u - j].flags & kTileHasEnd) != 0u;
        const uint64_t hits = group_bits<K>(ballot(has_end));
        if (hits) {
            u -= static_cast<uint32_t>(__builtin_ctzll(hits));
            break;
        }
        if (u < static_cast<uint32_t>(K)) {
            u = 0;
            break;
        }
        u -= K;
    }
    // Carries left to right (the fold order of spmv_merge_fixup_kernel), loaded eight at a time.
    acc_t acc = S::identity();
    uint32_t v = u;
    for (; v + 8 <= t; v += 8) {
        acc_t cv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) cv[i] = recs[v + i].carry[j];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc = S::merge(acc, cv[i]);
    }
    for (; v < t; ++v) acc = S::merge(acc, recs[v].carry[j]);
    acc = S::merge(acc, recs[t].head[j]);
    const uint32_t row = coords[t].x;
    typename S::x_t old = typename S::x_t(0);
    if constexpr (P::kRowOld) old = pol.row_old(row, j);
    const uint64_t c = pol.row_done(row, j, S::finalize(acc), old);
    // Thousands of tiles may report changed columns; the shared word is only written with new bits.
    if (c && changed) {
        const unsigned long long cur = *reinterpret_cast<volatile unsigned long long*>(changed);
        if (c & ~cur) atomicOr(changed, static_cast<unsigned long long>(c));
    }
}

// ------------------------------------------------------------------------------------------------
// Min-plus SpMM at k = 64 with one wavefront per merge-path tile: lane j owns column j.
//
// The tile's nonzeros are taken in chunks of 64 aligned to multiples of 64 (lane t <-> nonzero
// 64c + t): one coalesced load brings the chunk's column indices and weights, one gather the change
// masks of their sources. A source whose mask is empty costs nothing more; for an active one the wave
// reads its index and mask out of the lane that holds them (readlane) and loads the source's 64
// distances with one 256-byte access, the lanes of the unchanged columns masked off. The loads of chunk
// c + 1 (masks) and c + 2 (indices) are issued before those of chunk c, so they are in flight while
// chunk c is folded. Row boundaries inside a chunk come from a per-graph bitmap of row starts (bit q:
// a non-empty row begins at nonzero q): the fold keeps one running minimum and, at each boundary, parks
// the finished row's 64 values in LDS; the rows are then applied in order (d = min(d, y), the columns
// that improved go to mask_next). A chunk with more boundaries than the LDS slots is folded in parts.
//
// The column and weight arrays must be readable up to the next multiple of 64 past the last nonzero
// (the graph keeps kCsrPad = 64 padding entries).
//
// Min is order-free, so the result does not depend on how rows are cut into tiles or chunks; rows that
// cross tiles leave head / carry records in the SpmmTileRec layout and are finished by spmm_fixup_kernel
// with MinPlusBatchPolicy<64>, as for the other min-plus kernels. Distances are updated in place (a
// source row read during the iteration may already hold this iteration's smaller value): each column
// still decreases monotonically to the same fixed point, and an iteration with no change ends the run.
constexpr uint32_t kWaveStage = 16;  // finished rows parked per fold part
constexpr uint32_t kCloseBatch = 4;  // rows applied together (their old distances loaded together)

// Bit q of starts[q >> 6]: a non-empty row of the CSR begins at nonzero q. starts has (nnz >> 6) + 2
// words and must be zero on entry.
static __global__ void __launch_bounds__(256) row_starts_kernel(uint32_t n_rows, const uint32_t* __restrict__ row_offsets,
                                                         unsigned long long* __restrict__ starts) {
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < n_rows; r += gridDim.x * 256) {
        const uint32_t b = row_offsets[r];
        if (row_offsets[r + 1] > b) atomicOr(&starts[b >> 6], 1ull << (b & 63u));
    }
}

// Distances are stored as DT: uint32_t, or int16_t when every finite distance of the run is below
// kCompactLimit (then a source row is one 128-byte line instead of two). 16-bit storage is exact or
// void: -1 is the unreached distance, so a sign-extending load yields G1_INF for it and the distance
// itself otherwise; a relaxation that would store a finite value >= kCompactLimit sets *overflow instead
// (the caller then reruns the batch with 32-bit storage). With 16-bit storage the weight is added with
// unsigned saturation, which equals the semiring's combine because x + w cannot wrap (the caller only
// picks 16-bit storage when every weight is <= 0xFFFF0000).
constexpr uint32_t kCompactLimit = 0x7FFFu;

template <typename DT>
__device__ __forceinline__ uint32_t widen_dist(DT raw) {
    if constexpr (sizeof(DT) == 2) return static_cast<uint32_t>(static_cast<int32_t>(raw));
    return static_cast<uint32_t>(raw);
}

// act[v >> 5] bit (v & 31): mask[v] != 0 (the vertex changed in the last iteration). One wave per 64
// vertices; act has 2 * ceil(n / 64) words.
static __global__ void __launch_bounds__(256) act_bits_kernel(uint32_t n, const unsigned long long* __restrict__ mask,
                                                       uint32_t* __restrict__ act) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t groups = (n + 63u) / 64u;
    for (uint32_t gi = (blockIdx.x * 256u + threadIdx.x) / 64u; gi < groups; gi += gridDim.x * 4u) {
        const uint32_t v = gi * 64u + lane;
        const uint64_t b = ballot(v < n && mask[v] != 0ull);
        if (lane < 2) act[2 * gi + lane] = static_cast<uint32_t>(b >> (32 * lane));
    }
}

template <bool kWeighted, typename DT>
__global__ void __launch_bounds__(256, 5)  // five waves per SIMD: at most 96 VGPRs
spmm_minplus_wave_kernel(uint32_t n_rows, uint32_t num_tiles, const uint32_t* __restrict__ row_offsets,
                         const uint32_t* __restrict__ col, const uint32_t* __restrict__ vals,
                         const unsigned long long* __restrict__ starts, const MergeCoord* __restrict__ coords,
                         SpmmTileRec<MinPlusU32, 64>* __restrict__ recs, DT* __restrict__ D,
                         const uint32_t* __restrict__ act_prev, unsigned long long* __restrict__ mask_next,
                         unsigned long long* __restrict__ changed, uint32_t* __restrict__ overflow) {
    __shared__ uint32_t s_stage[4][kWaveStage][64];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t tile = blockIdx.x * 4u + wave;
    if (tile >= num_tiles) return;
    const uint32_t r0 = coords[tile].x, r_last = coords[tile + 1].x;  // rows [r0, r_last) end in the tile
    const uint32_t q0 = coords[tile].y, q1 = coords[tile + 1].y;      // nonzeros [q0, q1)
    const bool deferred = r0 < r_last && row_offsets[r0] < q0;         // the first row began before the tile

    // Row ends of 64 consecutive rows, one per lane, read out with readlane.
    uint32_t rbase = r0;
    uint32_t rwin = row_offsets[rbase + 1 + lane <= n_rows ? rbase + 1 + lane : n_rows];
    auto row_end = [&](uint32_t r) -> uint32_t {
        if (r - rbase >= 64u) {
            rbase = r;
            rwin = row_offsets[rbase + 1 + lane <= n_rows ? rbase + 1 + lane : n_rows];
        }
        return __builtin_amdgcn_readlane(rwin, r - rbase);
    };

    uint32_t r = r0;
    uint32_t rs = r0 < r_last ? row_offsets[r0] : 0u;  // start of row r
    uint32_t re = r0 < r_last ? row_end(r0) : kInf;    // end of row r (kInf: the tile's open tail)
    uint32_t acc = kInf;                               // running minimum of the open row
    uint64_t counted = 0;

    // Closes every row that ends at or before nonzero position lim, kCloseBatch rows at a time: the old
    // distances of a batch are loaded together, then the rows are applied in order (d = min(d, y), the
    // columns that improved go to mask_next; the deferred first row goes to the tile record). A non-empty
    // row that ends before lim was parked in the next LDS slot; the one that ends at lim is the open
    // minimum; an empty row (or every row when all_inf) is unreached.
    auto close_upto = [&](uint32_t lim, bool all_inf) {
        if (all_inf) {
            // Every row closing here is unreached: nothing to relax and nothing to record (mask_next is
            // cleared before the launch), so whole runs of rows are stepped over through the row-end
            // window; only a deferred first row leaves its (unreached) head.
            if (r == r0 && deferred && re <= lim) {
                recs[tile].head[lane] = MinPlusU32::acc_t{kInf};
                rs = re;
                ++r;
                re = r < r_last ? row_end(r) : kInf;
            }
            while (re <= lim) {
                const uint32_t off = r - rbase;  // < 64: row_end(r) refilled the window
                const uint32_t run = static_cast<uint32_t>(__popcll(ballot(lane >= off && rwin <= lim)));
                uint32_t last = r + (run ? run : 1u) - 1u;
                if (last >= r_last) last = r_last - 1u;
                rs = row_end(last);
                r = last + 1u;
                re = r < r_last ? row_end(r) : kInf;
            }
            return;
        }
        uint32_t k = 0;  // next LDS slot
        while (re <= lim) {
            uint32_t rid[kCloseBatch], src[kCloseBatch], slot[kCloseBatch];  // src: 0 none, 1 unreached, 2 slot, 3 open
            uint32_t open_val = kInf;
#pragma unroll
            for (uint32_t b = 0; b < kCloseBatch; ++b) {
                src[b] = 0;
                rid[b] = 0;
                slot[b] = 0;
                if (re <= lim) {
                    rid[b] = r;
                    if (re == rs || all_inf) {
                        src[b] = 1;
                    } else if (re < lim) {
                        src[b] = 2;
                        slot[b] = k++;
                    } else {
                        src[b] = 3;
                        open_val = acc;
                        acc = kInf;
                    }
                    rs = re;
                    ++r;
                    re = r < r_last ? row_end(r) : kInf;
                }
            }
            uint32_t old[kCloseBatch];
#pragma unroll
            for (uint32_t b = 0; b < kCloseBatch; ++b) {
                old[b] = kInf;
                if (src[b] >= 2 && !(rid[b] == r0 && deferred))
                    old[b] = widen_dist<DT>(D[static_cast<size_t>(rid[b]) * 64 + lane]);
            }
#pragma unroll
            for (uint32_t b = 0; b < kCloseBatch; ++b) {
                if (src[b] == 0) continue;
                const uint32_t val = src[b] == 2 ? s_stage[wave][slot[b]][lane] : (src[b] == 3 ? open_val : kInf);
                if (rid[b] == r0 && deferred) {
                    recs[tile].head[lane] = MinPlusU32::acc_t{val};
                    continue;
                }
                const bool improved = val < old[b];
                if (improved) {
                    if (sizeof(DT) == 2 && val >= kCompactLimit) *overflow = 1u;
                    else D[static_cast<size_t>(rid[b]) * 64 + lane] = static_cast<DT>(val);
                }
                const uint64_t bits = ballot(improved);
                if (lane == 0 && bits) mask_next[rid[b]] = bits;  // cleared before the launch
                counted |= bits;
            }
        }
    };

    close_upto(q0, false);
    if (q0 < q1) {
        const uint32_t c_first = q0 >> 6, c_end = (q1 + 63u) >> 6;
        auto in_tile = [&](uint32_t c) {
            const uint32_t q = c * 64u + lane;
            return q >= q0 && q < q1;
        };
        // Source activity comes from a bitmap of the vertices that changed last iteration (one bit per
        // vertex, small enough to stay in L2): the word of each lane's source, tested at use.
        uint32_t col0 = col[c_first * 64u + lane];
        uint32_t col1 = 0;
        if (c_first + 1 < c_end) col1 = col[(c_first + 1) * 64u + lane];
        uint32_t a0 = in_tile(c_first) ? act_prev[col0 >> 5] : 0u;
        for (uint32_t c = c_first; c < c_end; ++c) {
            const uint64_t act = ballot(((a0 >> (col0 & 31u)) & 1u) != 0u);
            // Prefetch, in this order (vector loads complete in order): activity of chunk c + 1, indices
            // of chunk c + 2; both finish before this chunk's distance loads.
            uint32_t a1 = 0;
            if (c + 1 < c_end && in_tile(c + 1)) a1 = act_prev[col1 >> 5];
            uint32_t col2 = 0;
            if (c + 2 < c_end) col2 = col[(c + 2) * 64u + lane];
            const unsigned long long bnd = starts[c];
            const uint32_t lo = (q0 > c * 64u ? q0 - c * 64u : 0u);
            const uint32_t hi = (q1 - c * 64u < 64u ? q1 - c * 64u : 64u);
            if (act == 0ull && ballot(acc != kInf) == 0ull) {
                close_upto(c * 64u + hi, true);  // nothing to fold: every row closing here is unreached
            } else {
                // The 64 sources' distance rows: an active source's row, or for an inactive one the
                // all-unreached row n_rows (kept in cache), so every entry folds without a test. Every
                // column of an active source is relaxed: a column whose distance did not change last
                // iteration was relaxed through this edge when it last changed, so it cannot improve
                // the row now, and the result (a fixed point of min) is the same.
                const uint32_t w0 = kWeighted ? vals[c * 64u + lane] : 1u;  // weights: active chunks only
                uint32_t x[64];
                const uint32_t lane_off = lane * static_cast<uint32_t>(sizeof(DT));
#pragma unroll
                for (uint32_t t = 0; t < 64; ++t) {
                    uint32_t cu = __builtin_amdgcn_readlane(col0, t);
                    asm volatile("" : "+s"(cu));  // keep the readlane out of a branch: select, not jump
                    const uint32_t u = ((act >> t) & 1ull) ? cu : n_rows;
                    if constexpr (sizeof(DT) == 2) {
                        // 32-bit byte offsets: the caller keeps (n_rows + 1) * 128 below 2^32
                        x[t] = widen_dist<DT>(*reinterpret_cast<const DT*>(reinterpret_cast<const char*>(D) +
                                                                          (u << 7) + lane_off));
                    } else {
                        x[t] = D[static_cast<size_t>(u) * 64 + lane];
                    }
                }
                uint32_t lo_p = lo;
                while (lo_p < hi) {
                    // A fresh copy of the boundary bits for this part's fold, so the bits are tested
                    // where they are used instead of 64 saved predicates being kept alive.
                    uint64_t bnd_f = bnd;
                    asm volatile("" : "+s"(bnd_f));
                    // Boundaries strictly inside the part; at most kWaveStage of them.
                    uint64_t inner = (lo_p + 1 < 64u) ? (bnd >> (lo_p + 1)) << (lo_p + 1) : 0ull;
                    if (hi < 64u) inner &= (1ull << hi) - 1ull;
                    uint32_t hi_p = hi;
                    if (static_cast<uint32_t>(__popcll(inner)) > kWaveStage) {
#pragma unroll
                        for (uint32_t k = 0; k < kWaveStage; ++k) inner &= inner - 1ull;
                        hi_p = static_cast<uint32_t>(__builtin_ctzll(inner));
                    }
                    uint32_t ns = 0;
                    auto fold = [&](uint32_t t, bool first) {
                        if (!first && ((bnd_f >> t) & 1ull)) {
                            s_stage[wave][ns][lane] = acc;
                            ++ns;
                            acc = kInf;
                        }
                        const uint32_t wt = kWeighted ? __builtin_amdgcn_readlane(w0, t) : 1u;
                        const uint32_t p = sizeof(DT) == 2 ? __builtin_elementwise_add_sat(x[t], wt)
                                                           : MinPlusU32::combine(wt, x[t]);
                        acc = p < acc ? p : acc;
                    };
                    if (lo_p == 0 && hi_p == 64) {  // the whole chunk in one part: no range tests
#pragma unroll
                        for (uint32_t t = 0; t < 64; ++t) fold(t, t == 0);
                    } else {
#pragma unroll
                        for (uint32_t t = 0; t < 64; ++t)
                            if (t >= lo_p && t < hi_p) fold(t, t == lo_p);
                    }
                    close_upto(c * 64u + hi_p, false);
                    lo_p = hi_p;
                }
            }
            col0 = col1;
            a0 = a1;
            col1 = col2;
        }
    }
    recs[tile].carry[lane] = MinPlusU32::acc_t{acc};
    if (lane == 0) recs[tile].flags = (r_last > r0 ? kTileHasEnd : 0u) | (deferred ? kTileDeferred : 0u);
    if (counted && lane == 0) {
        const unsigned long long cur = *reinterpret_cast<volatile unsigned long long*>(changed);
        if (counted & ~cur) atomicOr(changed, static_cast<unsigned long long>(counted));
    }
}

// ------------------------------------------------------------------------------------------------
// Batched SSSP / BFS (min-plus). Per vertex: D[v*K + j] the distance of column j, and a mask of the
// columns whose distance changed in the last iteration (mask_prev[v] bit j). A source contributes to
// column j only when that bit is set, so inactive sources cost one 8-byte mask read per nonzero. The
// row's owner applies the result: d = min(D, y); changed bits go to mask_next[v] (written once per row
// per iteration, so the masks need no atomics and no clearing). The columns that changed are ORed into a
// per-iteration word, so each column's convergence is known.
template <int K>
struct MinPlusBatchPolicy {
    using aux_t = unsigned long long;
    static constexpr bool kHasAux = true;   // the source's change mask, staged per nonzero
    static constexpr bool kRowOld = true;   // the apply needs the row's old distance
    uint32_t* D;
    const unsigned long long* mask_prev;
    unsigned long long* mask_next;
    __device__ __forceinline__ aux_t aux(uint32_t u) const { return mask_prev[u]; }
    // Always one load (a resident dummy row for inactive entries) so every round issues the same count.
    __device__ __forceinline__ uint32_t x(uint32_t u, aux_t m, uint32_t j, bool has) const {
        const bool on = has && ((m >> j) & 1ull);
        const uint32_t v = D[on ? static_cast<size_t>(u) * K + j : j];
        return on ? v : kInf;
    }
    __device__ __forceinline__ uint32_t row_old(uint32_t row, uint32_t j) const {
        return D[static_cast<size_t>(row) * K + j];
    }
    // Called by the K lanes of a group together (lane j = column j); returns 1 on lane 0 of the group
    // with the columns of the row that changed (0 on the other lanes).
    __device__ __forceinline__ uint64_t row_done(uint32_t row, uint32_t j, uint32_t y, uint32_t old) const {
        const bool improved = y < old;
        if (improved) D[static_cast<size_t>(row) * K + j] = y;
        const uint64_t bits = group_bits<K>(ballot(improved));
        if (j == 0) mask_next[row] = bits;
        return (j == 0) ? bits : 0ull;
    }
};

// The fixup side of spmm_minplus_wave_kernel with 16-bit distance storage (row_old / row_done only).
struct MinPlusBatchPolicy16 {
    using aux_t = unsigned long long;
    static constexpr bool kHasAux = true;
    static constexpr bool kRowOld = true;
    int16_t* D;
    const unsigned long long* mask_prev;
    unsigned long long* mask_next;
    uint32_t* overflow;
    __device__ __forceinline__ uint32_t row_old(uint32_t row, uint32_t j) const {
        return widen_dist<int16_t>(D[static_cast<size_t>(row) * 64 + j]);
    }
    __device__ __forceinline__ uint64_t row_done(uint32_t row, uint32_t j, uint32_t y, uint32_t old) const {
        const bool improved = y < old;
        if (improved) {
            if (y >= kCompactLimit) *overflow = 1u;
            else D[static_cast<size_t>(row) * 64 + j] = static_cast<int16_t>(y);
        }
        const uint64_t bits = ballot(improved);
        if (j == 0) mask_next[row] = bits;
        return (j == 0) ? bits : 0ull;
    }
};

// Batched plus-times: writes y row-major; the PageRank-style apply runs as its own kernel.
template <int K>
struct PlusTimesBatchPolicy {
    using aux_t = uint32_t;
    static constexpr bool kHasAux = false;
    static constexpr bool kRowOld = false;
    const float* X;
    float* Y;
    __device__ __forceinline__ aux_t aux(uint32_t) const { return 0u; }
    __device__ __forceinline__ float x(uint32_t u, aux_t, uint32_t j, bool) const {
        return X[static_cast<size_t>(u) * K + j];
    }
    __device__ __forceinline__ float row_old(uint32_t, uint32_t) const { return 0.0f; }
    __device__ __forceinline__ uint64_t row_done(uint32_t row, uint32_t j, float y, float) const {
        Y[static_cast<size_t>(row) * K + j] = y;
        return 0ull;
    }
};

// ------------------------------------------------------------------------------------------------
// Batched plus-times (Kahan) at k = 64 with one wavefront per merge-path tile, lane j = column j, the
// walk on the scalar unit. It reproduces the fold of spmv_merge_tile_kernel exactly, per column: the
// tile's 2048 merge steps are cut into the same 256 runs of 8 ("virtual threads"); inside a run a row's
// products are added left to right; a run that closes a row finishes its first closed row by merging,
// left to right, the tails of the runs since the last run that closed a row and then its own head; the
// tile's open row leaves as the carry, a first row that began before the tile as the deferred head
// (finished by spmm_fixup_kernel). Batched personalized PageRank columns therefore stay bit-identical
// to the single-source runs.
//
// The x rows of 32 nonzeros are loaded at a time (one 256-byte access per source, issued a chunk ahead)
// and parked in LDS, so the walk indexes them at run time; weights and column indices come with one
// coalesced load per chunk and are read out of their lanes.
constexpr uint32_t kPprChunk = 32;

template <bool kWeighted>
__global__ void __launch_bounds__(256)
spmm_plustimes_wave_kernel(uint32_t n_rows, uint32_t num_tiles, const uint32_t* __restrict__ row_offsets,
                           const uint32_t* __restrict__ col, const float* __restrict__ vals,
                           const MergeCoord* __restrict__ coords,
                           SpmmTileRec<PlusTimesKahanF32, 64>* __restrict__ recs, const float* __restrict__ X,
                           float* __restrict__ Y) {
    using S = PlusTimesKahanF32;
    using acc_t = S::acc_t;
    __shared__ float s_x[4][kPprChunk][64];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t tile = blockIdx.x * 4u + wave;
    if (tile >= num_tiles) return;
    const uint32_t r0 = coords[tile].x, r_last = coords[tile + 1].x;
    const uint32_t q0 = coords[tile].y, q1 = coords[tile + 1].y;

    uint32_t rbase = r0;
    uint32_t rwin = row_offsets[rbase + 1 + lane <= n_rows ? rbase + 1 + lane : n_rows];
    auto row_end = [&](uint32_t r) -> uint32_t {
        if (r - rbase >= 64u) {
            rbase = r;
            rwin = row_offsets[rbase + 1 + lane <= n_rows ? rbase + 1 + lane : n_rows];
        }
        return __builtin_amdgcn_readlane(rwin, r - rbase);
    };

    uint32_t r = r0;                                    // next row to close
    uint32_t re = r0 < r_last ? row_end(r0) : kInf;     // its end (kInf: no further row end in the tile)
    uint32_t d = 0;                                     // merge steps done in the tile
    acc_t acc = S::identity(), head = S::identity(), chain = S::identity();
    bool hasend = false, any_end = false;
    uint32_t first_row = 0;

    // End of a run of 8 steps (the same folds as merge_walk's, per column).
    auto run_end = [&]() {
        if (hasend) {
            const acc_t whole = S::merge(chain, head);
            if (!any_end && row_offsets[first_row] < q0) {
                recs[tile].head[lane] = whole;
            } else {
                Y[static_cast<size_t>(first_row) * 64 + lane] = S::finalize(whole);
            }
            chain = S::merge(S::identity(), acc);
            any_end = true;
        } else {
            chain = S::merge(chain, acc);
        }
        acc = S::identity();
        head = S::identity();
        hasend = false;
    };
    auto step = [&]() {
        ++d;
        if ((d & (kMergeItems - 1)) == 0) run_end();
    };
    auto close_row = [&]() {
        if (!hasend) {
            head = acc;
            hasend = true;
            first_row = r;
        } else {
            Y[static_cast<size_t>(r) * 64 + lane] = S::finalize(acc);
        }
        acc = S::identity();
        ++r;
        re = r < r_last ? row_end(r) : kInf;
        step();
    };

    if (q0 < q1) {
        const uint32_t c_first = q0 / kPprChunk, c_end = (q1 + kPprChunk - 1) / kPprChunk;
        auto load_cw = [&](uint32_t c, uint32_t& cv, uint32_t& wv) {
            cv = 0;
            wv = __float_as_uint(1.0f);
            const uint32_t q = c * kPprChunk + lane;
            if (lane < kPprChunk && q >= q0 && q < q1) {
                cv = col[q];
                if (kWeighted) wv = __float_as_uint(vals[q]);
            }
        };
        float xr[kPprChunk];
        auto load_x = [&](uint32_t cv) {
#pragma unroll
            for (uint32_t t = 0; t < kPprChunk; ++t) {
                const uint32_t u = __builtin_amdgcn_readlane(cv, t);  // inactive lanes hold source 0
                xr[t] = X[static_cast<size_t>(u) * 64 + lane];
            }
        };
        uint32_t colA, wA, colB = 0, wB = 0;
        load_cw(c_first, colA, wA);
        load_x(colA);
        if (c_first + 1 < c_end) load_cw(c_first + 1, colB, wB);
        for (uint32_t c = c_first; c < c_end; ++c) {
#pragma unroll
            for (uint32_t t = 0; t < kPprChunk; ++t) s_x[wave][t][lane] = xr[t];
            if (c + 1 < c_end) load_x(colB);
            uint32_t colC = 0, wC = 0;
            if (c + 2 < c_end) load_cw(c + 2, colC, wC);
            const uint32_t qa = (q0 > c * kPprChunk) ? q0 : c * kPprChunk;
            const uint32_t qb = (q1 < (c + 1) * kPprChunk) ? q1 : (c + 1) * kPprChunk;
            for (uint32_t q = qa; q < qb; ++q) {
                while (re <= q) close_row();
                const uint32_t t = q - c * kPprChunk;
                const float w = kWeighted ? __uint_as_float(__builtin_amdgcn_readlane(wA, t)) : 1.0f;
                S::add(acc, S::combine(w, s_x[wave][t][lane]));
                step();
            }
            wA = wB;
            colB = colC;
            wB = wC;
        }
    }
    while (r < r_last) close_row();
    // The runs after the tile's last step (the last tile only) are empty: their tails still join the chain.
    if ((d & (kMergeItems - 1)) != 0) run_end();
    for (uint32_t v = (d + kMergeItems - 1) / kMergeItems; v < static_cast<uint32_t>(kMergeThreads); ++v) run_end();
    recs[tile].carry[lane] = chain;
    if (lane == 0) {
        uint32_t flags = any_end ? kTileHasEnd : 0u;
        if (any_end && row_offsets[r0] < q0) flags |= kTileDeferred;
        recs[tile].flags = flags;
    }
}

}  // namespace g1
