// Push-style advance + filter for gfx950, the engine of engine/advance_gc.h:24-283 redesigned.
//
// Frontier representation. A frontier is a queue of vertex ids q[0..nf) plus, for every entry, the
// exclusive prefix sum of the out-degrees qoff[i] = sum_{k<i} deg(q[k]) (so the queue is also the
// row-end list of a merge path over its E = sum deg edges), plus one packed 64-bit counter
// cnt = (E << 32) | nf. Entries are appended with ONE 64-bit atomicAdd per workgroup round that
// reserves queue slots and edge offsets together, so qoff stays sorted in slot order without any scan
// kernel. Vertices with no out-edge are counted but not queued: they have nothing to advance.
//
// Advance. The E edges of the input frontier are cut into tiles of kPushTile consecutive edge slots;
// a workgroup finds the first queue entry of its tile with one wave-wide 64-way search over qoff,
// stages the tile's sources (id, first edge, per-source value) in LDS, then maps each edge slot to
// its source with a binary search in LDS. Every tile holds the same number of edges whatever the
// degree distribution (a 97k-edge hub is spread over ~95 workgroups), unlike the reference's
// per-block balancing (advance_gc.h:119-157).
//
// Filter fused with compaction. The user filter (advance_gc.h:212-244) becomes the policy's relax():
// it returns true exactly once per vertex and iteration (atomicMin + iteration stamp for SSSP,
// atomicCAS on the depth for BFS), so the output needs no INVALID slots and no second pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"

namespace g1 {

constexpr int kPushThreads = 256;
constexpr int kPushEpt = 4;                             // edge slots per thread per tile
constexpr int kPushTile = kPushThreads * kPushEpt;      // 1024 edge slots per tile

struct PushQueues {
    const uint32_t* q_in;
    const uint32_t* qoff_in;
    const unsigned long long* cnt_in;   // (E << 32) | nf of the input frontier
    uint32_t* q_out;
    uint32_t* qoff_out;
    unsigned long long* cnt_out;        // must be zero on entry
    unsigned long long* cnt_clear;      // set to zero by the kernel (the counter of the next round)
    unsigned long long* activated;      // += vertices activated (including those with no out-edge)
};

// Appends the calling thread's candidates (cand[r] != kInf, degree cdeg[r]) to the output queue.
// Block-uniform control flow; every thread of the block must call it.
template <int N>
__device__ __forceinline__ void push_enqueue(const uint32_t (&cand)[N], const uint32_t (&cdeg)[N],
                                             uint32_t my_cnt, uint32_t my_deg, const PushQueues& pq,
                                             unsigned long long* s_scan, unsigned long long* s_base) {
    const unsigned long long mine = (static_cast<unsigned long long>(my_deg) << 32) | my_cnt;
    unsigned long long tot;
    const unsigned long long excl = block_exclusive_sum<kPushThreads>(mine, tot, s_scan);
    if (tot != 0ull) {
        if (threadIdx.x == 0) *s_base = atomicAdd(pq.cnt_out, tot);
        __syncthreads();
        const unsigned long long base = *s_base + excl;
        uint32_t slot = static_cast<uint32_t>(base);
        uint32_t eoff = static_cast<uint32_t>(base >> 32);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            if (cand[r] != kInf) {
                pq.q_out[slot] = cand[r];
                pq.qoff_out[slot] = eoff;
                ++slot;
                eoff += cdeg[r];
            }
        }
        __syncthreads();
    }
}

template <typename P>
static __global__ void __launch_bounds__(kPushThreads)
push_advance_kernel(const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col,
                    const uint32_t* __restrict__ wts, PushQueues pq, P pol) {
    __shared__ uint32_t s_qoff[kPushTile];
    __shared__ uint32_t s_rs[kPushTile];
    __shared__ uint32_t s_u[kPushTile];
    __shared__ typename P::src_t s_su[kPushTile];
    __shared__ uint32_t s_red[kPushThreads / kWave];
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_i0;

    const unsigned long long packed = *pq.cnt_in;
    const uint32_t nf = static_cast<uint32_t>(packed);
    const uint32_t E = static_cast<uint32_t>(packed >> 32);
    if (blockIdx.x == 0 && threadIdx.x == 0) *pq.cnt_clear = 0ull;
    const uint32_t tiles = (E + kPushTile - 1) / kPushTile;
    uint32_t activated = 0;

    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t e0 = tile * kPushTile;
        const uint32_t e1 = (E - e0 < static_cast<uint32_t>(kPushTile)) ? E : e0 + kPushTile;
        if (threadIdx.x < kWave) {
            const uint32_t i0 = wave_search_last_le(pq.qoff_in, nf, e0);
            if (threadIdx.x == 0) s_i0 = i0;
        }
        __syncthreads();
        const uint32_t i0 = s_i0;
        // Stage the tile's sources: entries i0, i0+1, ... whose first edge slot is below e1.
        uint32_t cnt = 0;
        for (uint32_t base = 0; base < static_cast<uint32_t>(kPushTile); base += kPushThreads) {
            const uint32_t k = base + threadIdx.x;
            const uint32_t i = i0 + k;
            bool valid = false;
            if (i < nf) {
                const uint32_t qo = pq.qoff_in[i];
                valid = (k == 0) || (qo < e1);
                if (valid) {
                    const uint32_t u = pq.q_in[i];
                    s_qoff[k] = qo;
                    s_u[k] = u;
                    s_rs[k] = row_off[u];
                    s_su[k] = pol.source_value(u);
                }
            }
            const uint32_t c = block_count<kPushThreads>(valid, s_red);
            cnt += c;
            if (c < static_cast<uint32_t>(kPushThreads)) break;
        }
        // (block_count ends with a barrier, so the staged sources are visible here)
        uint32_t cand[kPushEpt];
        uint32_t cdeg[kPushEpt];
        uint32_t my_cnt = 0, my_deg = 0;
#pragma unroll
        for (int r = 0; r < kPushEpt; ++r) {
            const uint32_t e = e0 + r * kPushThreads + threadIdx.x;
            cand[r] = kInf;
            cdeg[r] = 0;
            if (e < e1) {
                const uint32_t j = search_last_le(s_qoff, cnt, e);
                const uint32_t eu = s_rs[j] + (e - s_qoff[j]);
                const uint32_t v = col[eu];
                const uint32_t w = P::kWeighted ? wts[eu] : 1u;
                if (pol.relax(s_u[j], s_su[j], v, w, eu)) {
                    ++activated;
                    const uint32_t dg = row_off[v + 1] - row_off[v];
                    if (dg > 0) {
                        cand[r] = v;
                        cdeg[r] = dg;
                        ++my_cnt;
                        my_deg += dg;
                    }
                }
            }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, s_scan, &s_base);
        __syncthreads();
    }
    const uint32_t wsum = wave_sum(activated);
    if (lane_id() == 0 && wsum) atomicAdd(pq.activated, static_cast<unsigned long long>(wsum));
}

// SSSP relaxation (apps/sssp_adv.cu:49-71): nd = d[u] + w; atomicMin(&d[v], nd) and keep v when
// nd < old; the stamp makes v enter the next frontier once per iteration (the visited filter).
struct SsspPush {
    using src_t = uint32_t;
    static constexpr bool kWeighted = true;
    uint32_t* dist;
    uint32_t* stamp;
    uint32_t iter;
    __device__ __forceinline__ uint32_t source_value(uint32_t u) const { return dist[u]; }
    __device__ __forceinline__ bool relax(uint32_t, uint32_t du, uint32_t v, uint32_t w, uint32_t) const {
        const uint32_t nd = du + w;
        if (nd >= dist[v]) return false;
        const uint32_t old = atomicMin(&dist[v], nd);
        if (nd >= old) return false;
        return atomicMax(&stamp[v], iter) < iter;
    }
};

// BFS top-down step: the first visitor of v sets its depth; optionally marks v in a bitmap that a
// following bottom-up step reads.
struct BfsPush {
    using src_t = uint32_t;
    static constexpr bool kWeighted = false;
    uint32_t* depth;
    uint32_t* fb_next;  // may be null
    uint32_t next_level;
    __device__ __forceinline__ uint32_t source_value(uint32_t) const { return 0u; }
    __device__ __forceinline__ bool relax(uint32_t, uint32_t, uint32_t v, uint32_t, uint32_t) const {
        if (depth[v] != kInf) return false;
        if (atomicCAS(&depth[v], kInf, next_level) != kInf) return false;
        if (fb_next) atomicOr(&fb_next[v >> 5], 1u << (v & 31u));
        return true;
    }
};

// BFS bottom-up step over the in-edge CSR: every unvisited vertex looks for an in-neighbour in the
// current frontier bitmap and stops at the first one. The next frontier is written both as a bitmap
// (every word fully written by one wave, so it needs no clearing) and as a queue for a following
// top-down step. Blocks cover contiguous runs of 256 vertices.
static __global__ void __launch_bounds__(kPushThreads)
bfs_pull_kernel(uint32_t n, const uint32_t* __restrict__ t_row_off, const uint32_t* __restrict__ t_col,
                const uint32_t* __restrict__ row_off, uint32_t* __restrict__ depth,
                const uint32_t* __restrict__ fb_cur, uint32_t* __restrict__ fb_next, uint32_t next_level,
                PushQueues pq) {
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    if (blockIdx.x == 0 && threadIdx.x == 0) *pq.cnt_clear = 0ull;
    uint32_t activated = 0;
    const uint32_t words = (n + 31) / 32;
    for (uint32_t base = blockIdx.x * kPushThreads; base < n; base += gridDim.x * kPushThreads) {
        const uint32_t v = base + threadIdx.x;
        bool found = false;
        if (v < n && depth[v] == kInf) {
            const uint32_t b = t_row_off[v], e = t_row_off[v + 1];
            for (uint32_t k = b; k < e; ++k) {
                const uint32_t u = t_col[k];
                if ((fb_cur[u >> 5] >> (u & 31u)) & 1u) { found = true; break; }
            }
        }
        const uint64_t m = ballot(found);
        const uint32_t lane = lane_id();
        const uint32_t word = v >> 5;  // lanes 0..31 and 32..63 of a wave each own one word
        if ((lane & 31u) == 0 && word < words)
            fb_next[word] = (lane == 0) ? static_cast<uint32_t>(m) : static_cast<uint32_t>(m >> 32);
        uint32_t cand[1] = {kInf};
        uint32_t cdeg[1] = {0};
        uint32_t my_cnt = 0, my_deg = 0;
        if (found) {
            depth[v] = next_level;
            ++activated;
            const uint32_t dg = row_off[v + 1] - row_off[v];
            if (dg > 0) { cand[0] = v; cdeg[0] = dg; my_cnt = 1; my_deg = dg; }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, s_scan, &s_base);
    }
    const uint32_t wsum = wave_sum(activated);
    if (lane_id() == 0 && wsum) atomicAdd(pq.activated, static_cast<unsigned long long>(wsum));
}

// Sets up a one-vertex frontier {src} for a push engine: q[0] = src, qoff[0] = 0,
// cnt = (deg(src) << 32) | 1 (or 0 when src has no out-edge), clears the other two counters.
static __global__ void push_seed_kernel(const uint32_t* __restrict__ row_off, uint32_t src, uint32_t* q, uint32_t* qoff,
                                 unsigned long long* cnt3) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t dg = row_off[src + 1] - row_off[src];
        q[0] = src;
        qoff[0] = 0;
        cnt3[0] = dg ? ((static_cast<unsigned long long>(dg) << 32) | 1ull) : 0ull;
        cnt3[1] = 0ull;
        cnt3[2] = 0ull;
    }
}

}  // namespace g1
