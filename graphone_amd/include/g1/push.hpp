// Push-style advance + filter for gfx950, the engine of engine/advance_gc.h:24-283 redesigned.
//
// Frontier representation (a segmented queue). A frontier is kSegments independent queue segments.
// Segment s holds vertex ids q[s*cap + 0 .. n_s) and, for each entry, the exclusive prefix sum of the
// out-degrees inside the segment (qoff); its packed counter cnt[s] = (E_s << 32) | n_s. The frontier is
// the concatenation of the segments in order, so entry i of the frontier has the global edge offset
// ebase[s] + qoff[...]: the queue is the row-end list of a merge path over its E = sum E_s edges.
// Producers append to segment (blockIdx.x % kSegments) with ONE 64-bit atomicAdd per workgroup round that
// reserves slots and edge offsets together, so no scan kernel is ever needed; spreading the counters
// over kSegments addresses keeps the device-scope atomics (which serialize per address) off the
// critical path. Vertices with no out-edge are counted but not queued: they have nothing to advance.
//
// Advance. The E edges of the input frontier are cut into tiles of kPushTile consecutive edge slots;
// each workgroup takes a contiguous run of tiles. A tile finds its first queue entry with a search over
// the segment bases (LDS) and a wave-wide 64-way search inside the segment, stages the tile's sources in
// LDS, then maps each edge slot to its source with a binary search in LDS. Every tile holds the same
// number of edges whatever the degree distribution (a 97k-edge hub is spread over 48 to 190 tiles), unlike the
// reference's per-block balancing (advance_gc.h:119-157).
//
// Filter fused with compaction. The user filter (advance_gc.h:212-244) becomes the policy's relax(): it
// returns true exactly once per vertex and iteration (atomicMin + iteration stamp for SSSP, atomicCAS on
// the depth for BFS), so the output needs no INVALID slots and no second pass.
//
// Statistics go to per-workgroup slots (stats[blockIdx.x] += ...), never to a shared counter.
//
// Every queued entry of an input frontier must own at least one edge (the tile staging relies on it);
// a policy with kQueueLeaves = true also queues winners without out-edges, and its caller removes them
// before the next advance.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/spmv_merge.hpp"

namespace g1 {

constexpr int kPushThreads = 256;
constexpr int kPushEpt = 8;                          // edge slots per thread per tile, at most
constexpr int kPushTile = kPushThreads * kPushEpt;   // 2048 edge slots per tile, at most
constexpr uint32_t kPushTileMin = 512;               // the smallest tile a launch cuts its frontier into
constexpr int kSegments = 64;                        // queue segments (one wave scans their counters)
constexpr uint32_t kMaxPushBlocks = 2048;            // producers per launch (bounds segment fill)

// Entries one segment can receive from one launch of at most kMaxPushBlocks workgroups that each take a
// contiguous share of `items` (edges for an advance, vertices for a bottom-up step): a segment is fed by
// at most ceil(G/kSegments) workgroups, each emitting at most one entry per item of its share
// (ceil(items/G) rounded up to whole tiles), and no segment holds a vertex twice.
__host__ __device__ inline uint64_t push_segment_capacity(uint64_t n, uint64_t items) {
    const uint64_t per = (items + kSegments - 1) / kSegments + (items + kMaxPushBlocks - 1) / kMaxPushBlocks +
                         40ull * kPushTile;
    return per < n ? per : n;
}

struct PushQueues {
    const uint32_t* q_in;
    const uint32_t* qoff_in;
    const unsigned long long* cnt_in;  // kSegments packed counters of the input frontier
    uint32_t* q_out;
    uint32_t* qoff_out;
    unsigned long long* cnt_out;       // kSegments counters, zero on entry
    unsigned long long* cnt_clear;     // kSegments counters zeroed by block 0 (the next round's output)
    unsigned long long* stats;         // per-workgroup slots: [b] activated, [kMaxPushBlocks + b] scanned
    uint32_t* overflow;                // set to 1 if a segment would overflow (never expected)
    uint32_t cap;                      // entries per segment
};

// Segment bases of the input frontier in LDS: vb[s], eb[s] exclusive prefixes of the segment counts
// and edge totals, vb[kSegments] = nf, eb[kSegments] = E. Called by wave 0 only.
__device__ __forceinline__ void load_segment_bases(const unsigned long long* __restrict__ cnt, uint32_t* vb,
                                                   uint32_t* eb) {
    const uint32_t lane = lane_id();
    const unsigned long long c = cnt[lane];
    const uint32_t nv = static_cast<uint32_t>(c), ne = static_cast<uint32_t>(c >> 32);
    const uint32_t iv = wave_inclusive_sum(nv), ie = wave_inclusive_sum(ne);
    vb[lane] = iv - nv;
    eb[lane] = ie - ne;
    if (lane == kWave - 1) {
        vb[kSegments] = iv;
        eb[kSegments] = ie;
    }
}

// Appends the calling thread's candidates (cand[r] != kInf, degree cdeg[r]) to segment `seg` of the
// output queue. Block-uniform control flow; every thread of the block must call it.
template <int N>
__device__ __forceinline__ void push_enqueue(const uint32_t (&cand)[N], const uint32_t (&cdeg)[N],
                                             uint32_t my_cnt, uint32_t my_deg, const PushQueues& pq, uint32_t seg,
                                             unsigned long long* s_scan, unsigned long long* s_base) {
    const unsigned long long mine = (static_cast<unsigned long long>(my_deg) << 32) | my_cnt;
    unsigned long long tot;
    const unsigned long long excl = block_exclusive_sum<kPushThreads>(mine, tot, s_scan);
    if (tot != 0ull) {
        if (threadIdx.x == 0) {
            unsigned long long b = atomicAdd(pq.cnt_out + seg, tot);
            if (static_cast<uint64_t>(static_cast<uint32_t>(b)) + static_cast<uint32_t>(tot) > pq.cap) {
                atomicExch(pq.overflow, 1u);
                b = ~0ull;
            }
            *s_base = b;
        }
        __syncthreads();
        const unsigned long long sb = *s_base;
        if (sb != ~0ull) {
            const unsigned long long base = sb + excl;
            uint32_t slot = static_cast<uint32_t>(base);
            uint32_t eoff = static_cast<uint32_t>(base >> 32);
            uint32_t* qo = pq.q_out + static_cast<size_t>(seg) * pq.cap;
            uint32_t* oo = pq.qoff_out + static_cast<size_t>(seg) * pq.cap;
#pragma unroll
            for (int r = 0; r < N; ++r) {
                if (cand[r] != kInf) {
                    qo[slot] = cand[r];
                    oo[slot] = eoff;
                    ++slot;
                    eoff += cdeg[r];
                }
            }
        }
        __syncthreads();
    }
}

// Adds a per-thread count to this workgroup's stats slot (no shared counter).
__device__ __forceinline__ void block_stat_add(unsigned long long* slot, uint32_t v, uint32_t* s_red) {
    const uint32_t w = wave_sum(v);
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kPushThreads / kWave; ++i) t += s_red[i];
        if (t) *slot += t;
    }
}

// The part of each row a push launch walks: where it starts in the column / weight arrays and how many
// edges it has (a queue entry's degree is that count). FullRows: the whole row. Over a copy of the CSR
// whose rows hold their light edges (weight below a bucket width) first and their heavy edges after,
// lc[u] light edges in row u, LightRows walks the light part and HeavyRows the heavy part.
struct FullRows {
    const uint32_t* row_off;
    __host__ __device__ FullRows(const uint32_t* ro) : row_off(ro) {}
    __device__ __forceinline__ uint32_t begin(uint32_t u) const { return row_off[u]; }
    __device__ __forceinline__ uint32_t degree(uint32_t u) const { return row_off[u + 1] - row_off[u]; }
};
struct LightRows {
    const uint32_t* row_off;
    const uint32_t* lc;
    __device__ __forceinline__ uint32_t begin(uint32_t u) const { return row_off[u]; }
    __device__ __forceinline__ uint32_t degree(uint32_t u) const { return lc[u]; }
};
struct HeavyRows {
    const uint32_t* row_off;
    const uint32_t* lc;
    __device__ __forceinline__ uint32_t begin(uint32_t u) const { return row_off[u] + lc[u]; }
    __device__ __forceinline__ uint32_t degree(uint32_t u) const { return row_off[u + 1] - row_off[u] - lc[u]; }
};

// The light/heavy copy of a weighted CSR (delta-stepping): row u of (lcol, lw) holds the row's edges of
// weight below delta first, then the others, each part in the row's own order; lc[u] = its light count.
// One wave per row: a counting pass, then each 64-edge chunk's light and heavy lanes are placed by their
// ranks among the chunk's lanes.
static __global__ void __launch_bounds__(256)
split_rows_kernel(uint32_t n, const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col,
                  const uint32_t* __restrict__ w, uint32_t delta, uint32_t* __restrict__ lcol,
                  uint32_t* __restrict__ lw, uint32_t* __restrict__ lc) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < n; r += waves) {  // wave-uniform
        const uint32_t b = row_off[r], e = row_off[r + 1];
        uint32_t light = 0;
        for (uint32_t k = b; k < e; k += kWave) {
            const bool l = k + lane < e && w[k + lane] < delta;
            light += static_cast<uint32_t>(__popcll(ballot(l)));
        }
        uint32_t lp = b, hp = b + light;
        for (uint32_t k = b; k < e; k += kWave) {
            const uint32_t i = k + lane;
            const bool in = i < e;
            uint32_t c = 0, wt = 0;
            if (in) {
                c = col[i];
                wt = w[i];
            }
            const bool l = in && wt < delta;
            const uint64_t lm = ballot(l), hm = ballot(in && !l);
            if (l) {
                const uint32_t at = lp + mask_prefix(lm);
                lcol[at] = c;
                lw[at] = wt;
            } else if (in) {
                const uint32_t at = hp + mask_prefix(hm);
                lcol[at] = c;
                lw[at] = wt;
            }
            lp += static_cast<uint32_t>(__popcll(lm));
            hp += static_cast<uint32_t>(__popcll(hm));
        }
        if (lane == 0) lc[r] = light;
    }
}

// Processes this workgroup's share of the input frontier's edge tiles (see the file comment).
template <typename P, typename RV>
__device__ __forceinline__ void advance_tiles(const RV& rv, const uint32_t* __restrict__ col,
                                              const uint32_t* __restrict__ wts, const PushQueues& pq, const P& pol,
                                              const uint32_t* s_vb, const uint32_t* s_eb) {
    __shared__ uint32_t s_qoff[kPushTile];
    __shared__ uint32_t s_rs[kPushTile];
    __shared__ typename P::src_t s_su[kPushTile];
    __shared__ uint32_t s_red[kPushThreads / kWave];
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_i0;

    const uint32_t nf = s_vb[kSegments];
    const uint32_t E = s_eb[kSegments];
    // The tile size: the largest power of two from kPushTileMin to kPushTile that still gives every
    // workgroup a tile (uniform across the launch). A frontier of a few hundred thousand edges in
    // 2048-slot tiles left most workgroups idle; a large one keeps the big tiles and their lower
    // per-edge staging cost. RMAT22 step: fixed 2048-slot tiles 1.09 ms; fixed 512: 1.02 ms, but the
    // advance loop's large frontier 0.175 -> 0.20 ms and push-only BFS 0.78 -> 0.90 ms; chosen per launch
    // 0.98 ms (a floor of 256 instead of 512: 0.99-1.01 ms).
    uint32_t T = kPushTile;
    while (T > kPushTileMin && (E + T - 1) / T < gridDim.x) T >>= 1;
    const uint32_t tiles = (E + T - 1) / T;
    const uint32_t per = (tiles + gridDim.x - 1) / gridDim.x;
    const uint32_t t_beg = blockIdx.x * per;
    const uint32_t t_end = (t_beg + per < tiles) ? t_beg + per : tiles;
    const uint32_t seg_out = blockIdx.x % kSegments;
    uint32_t activated = 0;

    // The first queue entry of the next tile, when this tile's staging fixes it (kInf: search for it).
    uint32_t next_i0 = kInf;
    for (uint32_t tile = t_beg; tile < t_end; ++tile) {
        const uint32_t e0 = tile * T;
        const uint32_t e1 = (E - e0 < T) ? E : e0 + T;
        uint32_t i0 = next_i0;  // block-uniform
        if (i0 == kInf) {
            if (threadIdx.x < kWave) {
                const uint32_t s = search_last_le(s_eb, kSegments, e0);
                const uint32_t* qo = pq.qoff_in + static_cast<size_t>(s) * pq.cap;
                const uint32_t nseg = s_vb[s + 1] - s_vb[s];
                const uint32_t loc = wave_search_last_le(qo, nseg, e0 - s_eb[s]);
                if (threadIdx.x == 0) s_i0 = s_vb[s] + loc;
            }
            __syncthreads();
            i0 = s_i0;
        }
        // Stage the tile's sources: entries i0, i0+1, ... whose first edge slot is below e1. The entry whose
        // first slot is e1 exactly, if one is read, starts the next tile; otherwise the next tile starts with
        // this tile's last entry (its edges run on past e1). The entry id is loaded with its offset.
        uint32_t cnt = 0;
        bool exact = false;
        for (uint32_t base = 0; base < T; base += kPushThreads) {
            const uint32_t k = base + threadIdx.x;
            const uint32_t i = i0 + k;
            // the entry's offset and vertex together, then its row start and source value together (a
            // lane past the frontier or past the tile reads vertex 0's), so staging waits twice, not
            // three times
            uint32_t qo = 0, u = 0;
            if (i < nf) {
                const uint32_t s = search_last_le(s_vb, kSegments, i);
                const size_t at = static_cast<size_t>(s) * pq.cap + (i - s_vb[s]);
                qo = s_eb[s] + pq.qoff_in[at];
                u = pq.q_in[at];
            }
            const bool valid = i < nf && ((k == 0) || (qo < e1));
            const bool at_e1 = i < nf && (k != 0) && (qo == e1);
            const uint32_t uc = valid ? u : 0u;
            const uint32_t rs = rv.begin(uc);
            const typename P::src_t su_k = pol.source_value(uc);
            if (valid) {
                s_qoff[k] = qo;
                s_rs[k] = rs;
                s_su[k] = su_k;
            }
            const uint32_t c = block_count_pair<kPushThreads>(valid, at_e1, s_red);
            cnt += c & 0xffffu;
            exact = exact || (c >> 16) != 0u;
            if ((c & 0xffffu) < static_cast<uint32_t>(kPushThreads)) break;
        }
        next_i0 = (cnt < T) ? i0 + cnt - (exact ? 0u : 1u) : kInf;
        // (block_count_pair ends with a barrier, so the staged sources are visible here)
        // Phase 1: map every edge slot of this thread to its source and load (v, w); phase 2: let the
        // policy relax all of them (it issues its gathers and atomics before consuming any result);
        // phase 3: degrees of the winners. Each phase keeps kPushEpt independent accesses in flight.
        uint32_t vv[kPushEpt], ww[kPushEpt], su[kPushEpt], uu[kPushEpt], ee[kPushEpt];
#pragma unroll
        for (int r = 0; r < kPushEpt; ++r) {
            const uint32_t e = e0 + r * kPushThreads + threadIdx.x;
            vv[r] = kInf;
            if (e < e1) {
                const uint32_t j = search_last_le(s_qoff, cnt, e);
                const uint32_t eu = s_rs[j] + (e - s_qoff[j]);
                vv[r] = col[eu];
                ww[r] = P::kWeighted ? wts[eu] : 1u;
                su[r] = s_su[j];
                uu[r] = kInf;  // policies that need the source id return it as their source value
                ee[r] = eu;
            }
        }
        bool won[kPushEpt];
        pol.template relax_many<kPushEpt>(uu, su, vv, ww, ee, won);
        if constexpr (!P::kEnqueue) {
            __syncthreads();  // the tile's staged sources are overwritten by the next tile
            continue;
        }
        uint32_t cand[kPushEpt];
        uint32_t cdeg[kPushEpt];
        uint32_t wd[kPushEpt];
        // every slot loads a degree (a slot that did not win reads vertex 0's), so no load waits behind
        // another's branch
#pragma unroll
        for (int r = 0; r < kPushEpt; ++r) wd[r] = rv.degree(won[r] ? vv[r] : 0u);
        uint32_t my_cnt = 0, my_deg = 0;
#pragma unroll
        for (int r = 0; r < kPushEpt; ++r) {
            cand[r] = kInf;
            cdeg[r] = 0;
            if (won[r]) {
                ++activated;
                const uint32_t dg = wd[r];
                if (dg > 0) {
                    cand[r] = vv[r];
                    cdeg[r] = dg;
                    ++my_cnt;
                    my_deg += dg;
                }
            }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, seg_out, s_scan, &s_base);
        __syncthreads();
    }
    block_stat_add(pq.stats + blockIdx.x, activated, s_red);
}

// Common prologue of every push-engine iteration: segment bases of the input frontier into LDS, the
// counters of the round after next cleared (block 0), and (block 0) the iteration's history record
// hist[2*it] = (E << 32) | nf. Returns false when the input frontier is empty (the launch is a no-op,
// which lets the host enqueue iterations ahead without reading the counters back).
__device__ __forceinline__ bool push_prologue(const PushQueues& pq, uint32_t* s_vb, uint32_t* s_eb) {
    if (threadIdx.x < kWave) load_segment_bases(pq.cnt_in, s_vb, s_eb);
    if (blockIdx.x == 0 && threadIdx.x < kSegments) pq.cnt_clear[threadIdx.x] = 0ull;
    __syncthreads();
    return s_vb[kSegments] != 0;
}

template <typename P, typename RV = FullRows>
static __global__ void __launch_bounds__(kPushThreads)
push_advance_kernel(RV rv, const uint32_t* __restrict__ col, const uint32_t* __restrict__ wts, PushQueues pq, P pol,
                    uint32_t it, unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_vb[kSegments + 1];
    __shared__ uint32_t s_eb[kSegments + 1];
    const bool active = push_prologue(pq, s_vb, s_eb);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        hist[2 * it] = (static_cast<unsigned long long>(s_eb[kSegments]) << 32) | s_vb[kSegments];
        hist[2 * it + 1] = 0ull;
    }
    if (!active) return;
    advance_tiles(rv, col, wts, pq, pol, s_vb, s_eb);
}

// Distance upper bounds for the SSSP target checks: one byte per vertex (b = ceil(d / step), 255 = no
// bound known), 1/8 of the 8-byte distance words, so the array stays in each XCD's L2 where the words
// cannot. A relaxation with nd >= b * step cannot improve its target (b * step >= d) and skips the random
// read of the word; the others read the word as before. A byte is written, with a plain store, after every
// improving atomic: racing stores may leave a larger bound than the final distance allows, never a smaller
// one (each stored bound covers a distance the word has held or holds), so a stale byte only filters less.
struct DistBound {
    uint8_t* b;     // null: no filtering
    uint32_t step;  // distance units per bound step
    __device__ __forceinline__ bool may_improve(uint32_t byte, uint32_t nd) const {
        return byte == 255u || nd < byte * step;
    }
    __device__ __forceinline__ void lower(uint32_t v, uint32_t nd) const {
        if (!b) return;
        const uint32_t q = (nd + step - 1) / step;
        b[v] = static_cast<uint8_t>(q < 255u ? q : 255u);
    }
};

// SSSP relaxation (apps/sssp_adv.cu:49-71): nd = d[u] + w; atomicMin(&d[v], nd) and keep v when
// nd < old; the stamp makes v enter the next frontier once per iteration (the visited filter).
// Distance and stamp share one 64-bit word (dist << 32 | stamp), so ONE atomicMin both lowers the
// distance and claims the enqueue: a strict improvement installs (nd, iter); the first improvement of
// this iteration is the one that saw an older stamp. An equal distance never replaces the word (the
// older stamp is smaller), so a non-improving write cannot claim the vertex.
//
// limit (delta-stepping's light phases): an improvement to nd >= limit lowers the word with stamp 0 instead
// of the iteration's, so it never claims the vertex; a later improvement below limit in the same iteration
// still does.
//
// kAllAtomics: a thread issues all its atomics before testing any (for launches where most relaxations
// improve their target, as in delta-stepping's light phases, where waiting on each in turn serializes
// up to kPushEpt round trips per tile).
template <bool kAllAtomics>
struct SsspPushT {
    using src_t = uint32_t;
    static constexpr bool kWeighted = true;
    static constexpr bool kQueueLeaves = false;
    static constexpr bool kEnqueue = true;
    unsigned long long* dstamp;
    uint32_t iter;
    DistBound db;
    uint32_t limit = kInf;
    __device__ __forceinline__ uint32_t source_value(uint32_t u) const {
        return static_cast<uint32_t>(dstamp[u] >> 32);
    }
    // v[r] == kInf marks an empty slot. The targets' distances are read for every slot first (all of a
    // thread's random reads in flight at once); then the slots that would improve their target take
    // their atomicMin one after another, each tested as it returns (fewer atomics in flight per wave
    // measured faster than issuing them all before testing any). A distance read early may be stale by
    // its atomic: the atomic then returns the smaller value and the slot does not win.
    template <int N>
    __device__ __forceinline__ void relax_many(const uint32_t (&)[N], const uint32_t (&du)[N], const uint32_t (&v)[N],
                                               const uint32_t (&w)[N], const uint32_t (&)[N], bool (&won)[N]) const {
        bool need[N];
#pragma unroll
        for (int r = 0; r < N; ++r) need[r] = v[r] != kInf;
        if (db.b) {
            uint32_t byte[N];
#pragma unroll
            for (int r = 0; r < N; ++r) byte[r] = need[r] ? db.b[v[r]] : 0u;
#pragma unroll
            for (int r = 0; r < N; ++r) need[r] = need[r] && db.may_improve(byte[r], du[r] + w[r]);
        }
        uint32_t cur[N];
#pragma unroll
        for (int r = 0; r < N; ++r) cur[r] = need[r] ? static_cast<uint32_t>(dstamp[v[r]] >> 32) : 0u;
        if constexpr (kAllAtomics) {
            unsigned long long old[N];
#pragma unroll
            for (int r = 0; r < N; ++r) {
                const uint32_t nd = du[r] + w[r];
                need[r] = need[r] && nd < cur[r];
                old[r] = 0ull;
                if (need[r])
                    old[r] = atomicMin(&dstamp[v[r]],
                                       (static_cast<unsigned long long>(nd) << 32) | (nd < limit ? iter : 0u));
            }
#pragma unroll
            for (int r = 0; r < N; ++r) {
                const uint32_t nd = du[r] + w[r];
                won[r] = false;
                if (need[r] && nd < static_cast<uint32_t>(old[r] >> 32)) {
                    db.lower(v[r], nd);
                    won[r] = nd < limit && static_cast<uint32_t>(old[r]) < iter;
                }
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < N; ++r) {
            won[r] = false;
            if (!need[r]) continue;
            const uint32_t nd = du[r] + w[r];
            if (nd >= cur[r]) continue;
            const bool in = nd < limit;
            const unsigned long long old =
                atomicMin(&dstamp[v[r]], (static_cast<unsigned long long>(nd) << 32) | (in ? iter : 0u));
            if (nd < static_cast<uint32_t>(old >> 32)) {
                db.lower(v[r], nd);
                won[r] = in && static_cast<uint32_t>(old) < iter;
            }
        }
    }
    // One slot (the small-frontier kernel): the same relaxation.
    __device__ __forceinline__ bool relax1(uint32_t du, uint32_t v, uint32_t w) const {
        const uint32_t nd = du + w;
        if (nd >= static_cast<uint32_t>(dstamp[v] >> 32)) return false;
        const bool in = nd < limit;
        const unsigned long long old =
            atomicMin(&dstamp[v], (static_cast<unsigned long long>(nd) << 32) | (in ? iter : 0u));
        return in && nd < static_cast<uint32_t>(old >> 32) && static_cast<uint32_t>(old) < iter;
    }
    __device__ __forceinline__ SsspPushT at_iteration(uint32_t it) const {
        return SsspPushT{dstamp, it + 1, db, limit};
    }
    // The source's distance read past the CU's cache: a workgroup that stays resident across iterations
    // must see improvements made by its own atomics in earlier iterations.
    __device__ __forceinline__ uint32_t fresh_source_value(uint32_t u) const {
        return static_cast<uint32_t>(__hip_atomic_load(&dstamp[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
    }
    __device__ __forceinline__ void set_next_bitmap(uint32_t*) {}
};
using SsspPush = SsspPushT<false>;

// Vertex scans that enqueue (sssp_collect_kernel, sssp_pull_apply_kernel): vertices per thread per round.
constexpr uint32_t kScanPerThread = 8;
constexpr uint32_t kScanRound = kScanPerThread * kPushThreads;

// SSSP relaxation of a dense iteration without the winner test: each improving slot lowers its target's
// word with an atomicMin whose result is not waited for, and the iteration's winners -- exactly the
// vertices whose word carries this iteration's stamp afterwards, since a strict improvement installs
// (nd, iter) and nothing else does -- are enqueued by sssp_collect_kernel in one streaming pass. A dense
// iteration spends most of its time waiting on returning atomics one after another; this one does not.
//
// With limit (see SsspPush) an improvement to nd >= limit carries stamp 0, so the collect pass does not
// pick it up.
struct SsspPushNoReturn {
    using src_t = uint32_t;
    static constexpr bool kWeighted = true;
    static constexpr bool kQueueLeaves = false;
    static constexpr bool kEnqueue = false;
    unsigned long long* dstamp;
    uint32_t iter;
    DistBound db;
    uint32_t limit = kInf;
    __device__ __forceinline__ uint32_t source_value(uint32_t u) const {
        return static_cast<uint32_t>(dstamp[u] >> 32);
    }
    // The bound bytes first (L2-resident), then the words of the slots they do not rule out, then the
    // atomics of the improving slots, each followed by its bound byte.
    template <int N>
    __device__ __forceinline__ void relax_many(const uint32_t (&)[N], const uint32_t (&du)[N], const uint32_t (&v)[N],
                                               const uint32_t (&w)[N], const uint32_t (&)[N], bool (&won)[N]) const {
        bool need[N];
#pragma unroll
        for (int r = 0; r < N; ++r) need[r] = v[r] != kInf;
        if (db.b) {
            uint32_t byte[N];
#pragma unroll
            for (int r = 0; r < N; ++r) byte[r] = need[r] ? db.b[v[r]] : 0u;
#pragma unroll
            for (int r = 0; r < N; ++r) need[r] = need[r] && db.may_improve(byte[r], du[r] + w[r]);
        }
        uint32_t cur[N];
#pragma unroll
        for (int r = 0; r < N; ++r) cur[r] = need[r] ? static_cast<uint32_t>(dstamp[v[r]] >> 32) : 0u;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            won[r] = false;
            const uint32_t nd = du[r] + w[r];
            if (need[r] && nd < cur[r]) {
                (void)atomicMin(&dstamp[v[r]], (static_cast<unsigned long long>(nd) << 32) | (nd < limit ? iter : 0u));
                db.lower(v[r], nd);
            }
        }
    }
};

// The winners of a no-return iteration: every vertex whose word carries `stamp` (the iteration's stamp)
// and has out-edges joins the next frontier (segment blockIdx % kSegments of the output queue).
template <typename RV = FullRows>
static __global__ void __launch_bounds__(kPushThreads)
sssp_collect_kernel(uint32_t n, RV rv, const unsigned long long* __restrict__ dstamp, uint32_t stamp, PushQueues pq) {
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_red[kPushThreads / kWave];
    const uint32_t seg_out = blockIdx.x % kSegments;
    uint32_t activated = 0;
    // kScanPerThread vertices per thread and one queue reservation per workgroup round
    for (uint32_t base = blockIdx.x * kScanRound; base < n; base += gridDim.x * kScanRound) {  // uniform
        uint32_t cand[kScanPerThread], cdeg[kScanPerThread];
        uint32_t my_cnt = 0, my_deg = 0;
        bool hit[kScanPerThread];
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            hit[r] = v < n && static_cast<uint32_t>(dstamp[v]) == stamp;
        }
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            cand[r] = kInf;
            cdeg[r] = 0;
            if (hit[r]) {
                ++activated;
                const uint32_t dg = rv.degree(v);
                if (dg > 0) {
                    cand[r] = v;
                    cdeg[r] = dg;
                    ++my_cnt;
                    my_deg += dg;
                }
            }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, seg_out, s_scan, &s_base);
        __syncthreads();
    }
    block_stat_add(pq.stats + blockIdx.x, activated, s_red);
}

// A frontier built from the distances (delta-stepping's bucket scans), in one streaming pass over the
// distance words: every vertex whose distance lies in [lo, hi) takes `stamp` (the stamp of the iteration
// whose input this is, as if that iteration's predecessor had improved it; a pull step selects its input
// by it) and, when its rv part has edges, joins the output queue (segment blockIdx % kSegments). Each
// workgroup also writes mins[blockIdx.x], the smallest distance at or above lo that it saw (kInf if none),
// so the host learns the next non-empty bucket from the same pass.
template <typename RV>
static __global__ void __launch_bounds__(kPushThreads)
dist_range_frontier_kernel(uint32_t n, RV rv, unsigned long long* __restrict__ dstamp, uint32_t lo, uint32_t hi,
                           uint32_t stamp, PushQueues pq, uint32_t* __restrict__ mins) {
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_min[kPushThreads / kWave];
    const uint32_t seg_out = blockIdx.x % kSegments;
    uint32_t mn = kInf;
    for (uint32_t base = blockIdx.x * kScanRound; base < n; base += gridDim.x * kScanRound) {  // uniform
        uint32_t cand[kScanPerThread], cdeg[kScanPerThread];
        uint32_t my_cnt = 0, my_deg = 0;
        // the distance words, then the degrees of the vertices in range, each stage's loads issued
        // together (clamped indices, no load behind a branch)
        uint32_t d[kScanPerThread], dgs[kScanPerThread];
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            const uint32_t dv = static_cast<uint32_t>(dstamp[v < n ? v : n - 1] >> 32);
            d[r] = v < n ? dv : kInf;
        }
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            dgs[r] = rv.degree(d[r] >= lo && d[r] < hi ? v : 0u);
        }
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            cand[r] = kInf;
            cdeg[r] = 0;
            if (d[r] >= lo && d[r] != kInf && d[r] < mn) mn = d[r];
            if (d[r] >= lo && d[r] < hi) {
                dstamp[v] = (static_cast<unsigned long long>(d[r]) << 32) | stamp;
                if (dgs[r] > 0) {
                    cand[r] = v;
                    cdeg[r] = dgs[r];
                    ++my_cnt;
                    my_deg += dgs[r];
                }
            }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, seg_out, s_scan, &s_base);
        __syncthreads();
    }
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_down(mn, off, kWave);
        mn = o < mn ? o : mn;
    }
    if (lane_id() == 0) s_min[threadIdx.x / kWave] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kPushThreads / kWave; ++w) mn = s_min[w] < mn ? s_min[w] : mn;
        mins[blockIdx.x] = mn;
    }
}


// BFS top-down step: the first visitor of v sets its depth; optionally marks v in a bitmap that a
// following bottom-up step reads.
struct BfsPush {
    using src_t = uint32_t;
    static constexpr bool kWeighted = false;
    static constexpr bool kQueueLeaves = false;
    static constexpr bool kEnqueue = true;
    uint32_t* depth;
    uint32_t* fb_next;  // may be null
    uint32_t next_level;
    __device__ __forceinline__ uint32_t source_value(uint32_t) const { return 0u; }
    template <int N>
    __device__ __forceinline__ void relax_many(const uint32_t (&)[N], const uint32_t (&)[N], const uint32_t (&v)[N],
                                               const uint32_t (&)[N], const uint32_t (&)[N], bool (&won)[N]) const {
        // The depth loads go out together (an empty slot reads entry 0: no load behind a branch), then
        // the claims; the bitmap bits only once every claim has returned, so no wait sits between them.
        uint32_t cur[N];
#pragma unroll
        for (int r = 0; r < N; ++r) cur[r] = depth[v[r] != kInf ? v[r] : 0u];
        uint32_t old[N];
#pragma unroll
        for (int r = 0; r < N; ++r) {
            old[r] = 0u;
            if (v[r] != kInf && cur[r] == kInf) old[r] = atomicCAS(&depth[v[r]], kInf, next_level);
        }
#pragma unroll
        for (int r = 0; r < N; ++r) won[r] = v[r] != kInf && cur[r] == kInf && old[r] == kInf;
        if (fb_next) {
#pragma unroll
            for (int r = 0; r < N; ++r)
                if (won[r]) atomicOr(&fb_next[v[r] >> 5], 1u << (v[r] & 31u));
        }
    }
    __device__ __forceinline__ bool relax1(uint32_t, uint32_t v, uint32_t) const {
        if (depth[v] != kInf || atomicCAS(&depth[v], kInf, next_level) != kInf) return false;
        if (fb_next) atomicOr(&fb_next[v >> 5], 1u << (v & 31u));
        return true;
    }
    __device__ __forceinline__ BfsPush at_iteration(uint32_t it) const { return BfsPush{depth, nullptr, it + 1}; }
    __device__ __forceinline__ void set_next_bitmap(uint32_t* b) { fb_next = b; }
    __device__ __forceinline__ uint32_t fresh_source_value(uint32_t) const { return 0u; }
};

// SSSP pull step, for iterations whose frontier has too many edges for the push kernel to pay: the
// frontier's distances are scattered into a vector x (kInf elsewhere), one min-plus SpMV over the in-edge
// CSR gives y[v] = min over in-edges (u, v) of x[u] + w, and the apply relaxes and enqueues every vertex
// that improved. One iteration of the same relaxation the push kernel does (the frontier's out-edges),
// in one pass over the matrix instead of random atomics per edge.
//
// x: the input frontier of iteration `it` is the set of vertices whose stamp is `it` (improved during
// iteration it - 1; every push and pull step stamps its winners it + 1), so x[v] = dist(v) for those and
// kInf for the rest, in one streaming pass (vertices without out-edges are never read through x). Block 0
// also does the push prologue's bookkeeping: hist[2*it] = (E << 32) | nf of the input frontier, and the
// counters of the round after next cleared.
static __global__ void __launch_bounds__(kPushThreads)
sssp_pull_x_kernel(uint32_t n, const unsigned long long* __restrict__ dstamp, uint32_t it, uint32_t* __restrict__ x,
                   PushQueues pq, unsigned long long* __restrict__ hist) {
    if (blockIdx.x == 0) {
        __shared__ uint32_t s_vb[kSegments + 1];
        __shared__ uint32_t s_eb[kSegments + 1];
        push_prologue(pq, s_vb, s_eb);
        if (threadIdx.x == 0) {
            hist[2 * it] = (static_cast<unsigned long long>(s_eb[kSegments]) << 32) | s_vb[kSegments];
            hist[2 * it + 1] = 0ull;
        }
    }
    for (uint32_t v = blockIdx.x * kPushThreads + threadIdx.x; v < n; v += gridDim.x * kPushThreads) {
        const unsigned long long ds = dstamp[v];
        x[v] = static_cast<uint32_t>(ds) == it ? static_cast<uint32_t>(ds >> 32) : kInf;
    }
}

// The byte x of the delta engine's heavy pull step (MinPlusU8X): xb[v] = d(v) - lo for the bucket [lo, hi)
// (hi - lo <= 254), 254 below it, 255 at or past hi (or no distance yet); the history record as
// sssp_pull_x_kernel writes it. The byte also tells which rows the pull visits (sssp_pull_rows_kernel).
static __global__ void __launch_bounds__(kPushThreads)
sssp_pull_xb_kernel(uint32_t n, const unsigned long long* __restrict__ dstamp, uint32_t lo, uint32_t hi, uint32_t it,
                    uint8_t* __restrict__ xb, PushQueues pq, unsigned long long* __restrict__ hist) {
    if (blockIdx.x == 0) {
        __shared__ uint32_t s_vb[kSegments + 1];
        __shared__ uint32_t s_eb[kSegments + 1];
        push_prologue(pq, s_vb, s_eb);
        if (threadIdx.x == 0) {
            hist[2 * it] = (static_cast<unsigned long long>(s_eb[kSegments]) << 32) | s_vb[kSegments];
            hist[2 * it + 1] = 0ull;
        }
    }
    // four vertices per thread: one 32-bit store of four bytes
    const uint32_t quads = (n + 3) / 4;
    for (uint32_t q = blockIdx.x * kPushThreads + threadIdx.x; q < quads; q += gridDim.x * kPushThreads) {
        // the quad's four distance words in two 16-byte loads (the last, partial quad word by word)
        uint32_t d[4];
        if (q * 4 + 3 < n) {
            const ulonglong2* p = reinterpret_cast<const ulonglong2*>(dstamp + static_cast<size_t>(q) * 4);
            const ulonglong2 a = p[0], b = p[1];
            d[0] = static_cast<uint32_t>(a.x >> 32);
            d[1] = static_cast<uint32_t>(a.y >> 32);
            d[2] = static_cast<uint32_t>(b.x >> 32);
            d[3] = static_cast<uint32_t>(b.y >> 32);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t v = q * 4 + k;
                d[k] = v < n ? static_cast<uint32_t>(dstamp[v] >> 32) : kInf;
            }
        }
        uint32_t packed = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t b = d[k] < hi ? (d[k] >= lo ? d[k] - lo : 254u) : 255u;
            packed |= b << (8 * k);
        }
        reinterpret_cast<uint32_t*>(xb)[q] = packed;
    }
}

// The delta engine's heavy pull over the rows it can improve. After a bucket [lo, hi)'s light iterations
// every light edge out of the bucket is relaxed, and a heavy edge (w >= hi - lo) out of it offers at least
// hi, so only a vertex whose distance is at or past hi (byte 255) can improve: y[v] = min over v's
// in-edges (u, v) of w + xb[u] (MinPlusU8X) is computed for those rows only; every other row gets
// y = infinity. On RMAT22 from the hub the first bucket's 0.5 M vertices are the targets of 87 % of the
// in-edges; 88 % of the rows are visited, 41 % of those have in-edges (5.4 on average).
//
// Each wave takes kPullRowsPerWave consecutive rows: it reads their bytes and, for the visited rows, their
// in-edge ranges, and lists the visited rows with in-edges (at most huge_min each) in LDS, ballot-
// compacted, with the prefix sums of their lengths: the wave's in-edges as one range. Every other row
// (but a workgroup's) gets y = infinity at once. The wave then walks that range E edges per lane per round
// (lane l takes E consecutive edges: one search of the list, then a step to the next row where one
// ends), issues every column and weight load of the round, then every x gather, folds each run of one
// row in registers and lowers the row's LDS minimum once per run. Rows longer than huge_min are the
// first n_huge entries of the long-row list (longest first), one workgroup each at the start of the
// launch. The minimum does not depend on the fold order.
// Rows per lane: 1 (64 rows per wave). The RMAT22 heavy pull took 170 / 172 / 179 / 219 us at 1 / 2 / 4 /
// 8 rows per lane.
constexpr uint32_t kPullRowsPerLane = 1;
constexpr uint32_t kPullRowsPerWave = kPullRowsPerLane * kWave;
constexpr uint32_t kPullRowsWaves = kPushThreads / kWave;
constexpr uint32_t kPullRowsPerBlock = kPullRowsPerWave * kPullRowsWaves;
constexpr uint32_t kPullRowsWaveMax = 4096;  // rows up to this length stay with a wave

template <int E>
__global__ void __launch_bounds__(kPushThreads)
sssp_pull_rows_kernel(uint32_t n, const uint32_t* __restrict__ t_row_off, const uint32_t* __restrict__ t_col,
                      const uint32_t* __restrict__ t_w, const uint8_t* __restrict__ xb, uint32_t* __restrict__ y,
                      uint32_t huge_min, const uint32_t* __restrict__ huge_rows, uint32_t n_huge) {
    using S = MinPlusU8X;
    using acc_t = S::acc_t;
    if (blockIdx.x < n_huge) {
        __shared__ acc_t s_acc[kPushThreads];
        const uint32_t row = huge_rows[blockIdx.x];
        if (xb[row] != 255u) return;  // not visited (uniform across the workgroup)
        acc_t acc = S::identity();
        csrvec_fold<S, true, kPushThreads>(t_col, t_w, xb, t_row_off[row] + threadIdx.x, t_row_off[row + 1], acc);
        s_acc[threadIdx.x] = acc;
        __syncthreads();
#pragma unroll
        for (uint32_t d = kPushThreads / 2; d >= 1; d >>= 1) {
            if (threadIdx.x < d) s_acc[threadIdx.x] = S::merge(s_acc[threadIdx.x], s_acc[threadIdx.x + d]);
            __syncthreads();
        }
        if (threadIdx.x == 0) y[row] = S::finalize(s_acc[0]);
        return;
    }
    __shared__ uint32_t s_row[kPullRowsWaves][kPullRowsPerWave];
    __shared__ uint32_t s_beg[kPullRowsWaves][kPullRowsPerWave];
    __shared__ uint32_t s_off[kPullRowsWaves][kPullRowsPerWave + 1];
    __shared__ uint32_t s_y[kPullRowsWaves][kPullRowsPerWave];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t base = (blockIdx.x - n_huge) * kPullRowsPerBlock + wave * kPullRowsPerWave;
    // every row's byte and in-edge range, loaded without branches (a lane past n reads row n - 1 and
    // visits nothing), so the loads are in flight together
    uint32_t byte[kPullRowsPerLane], beg[kPullRowsPerLane], end[kPullRowsPerLane];
#pragma unroll
    for (uint32_t r = 0; r < kPullRowsPerLane; ++r) {
        const uint32_t v = base + r * kWave + lane;
        const uint32_t vc = v < n ? v : n - 1;
        byte[r] = xb[vc];
        beg[r] = t_row_off[vc];
        end[r] = t_row_off[vc + 1];
    }
    uint32_t n_rows = 0, n_edges = 0;  // wave-uniform
#pragma unroll
    for (uint32_t r = 0; r < kPullRowsPerLane; ++r) {
        const uint32_t v = base + r * kWave + lane;
        const bool visit = v < n && byte[r] == 255u;
        const uint32_t dg = end[r] - beg[r];
        const bool listed = visit && dg > 0 && dg <= huge_min;
        if (v < n && !listed && !(visit && dg > huge_min)) y[v] = kInf;
        const uint64_t m = ballot(listed);
        const uint32_t len = listed ? dg : 0u;
        const uint32_t incl = wave_inclusive_sum(len);
        if (listed) {
            const uint32_t k = n_rows + mask_prefix(m);
            s_row[wave][k] = v;
            s_beg[wave][k] = beg[r];
            s_off[wave][k] = n_edges + incl - len;
            s_y[wave][k] = kInf;
        }
        n_rows += static_cast<uint32_t>(__popcll(m));
        n_edges += __shfl(incl, kWave - 1);
    }
    if (lane == 0) s_off[wave][n_rows] = n_edges;
    __syncthreads();
    for (uint32_t f0 = 0; f0 < n_edges; f0 += kWave * E) {  // wave-uniform
        const uint32_t f1 = f0 + lane * E;
        if (f1 < n_edges) {
            // the listed row holding edge f1: the last k with s_off[k] <= f1
            uint32_t lo = 0, hi = n_rows;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_off[wave][mid] <= f1) lo = mid;
                else hi = mid;
            }
            // Edge positions first (LDS only); then every load of the round without a branch between
            // them (a position past the range reads its row's first edge and is not folded), so they
            // are all in flight together.
            uint32_t k = lo;
            uint32_t kk[E], e[E], c[E], w[E];
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const uint32_t f = f1 + i;
                const bool valid = f < n_edges;
                if (valid)
                    while (s_off[wave][k + 1] <= f) ++k;
                kk[i] = valid ? k : kInf;
                e[i] = s_beg[wave][k] + (valid ? f - s_off[wave][k] : 0u);
            }
#pragma unroll
            for (int i = 0; i < E; ++i) {
                c[i] = t_col[e[i]];
                w[i] = t_w[e[i]];
            }
            uint8_t xv[E];
#pragma unroll
            for (int i = 0; i < E; ++i) xv[i] = xb[c[i]];
            uint32_t run_k = kk[0], run = kInf;
#pragma unroll
            for (int i = 0; i < E; ++i) {
                if (kk[i] == kInf) break;
                const uint32_t cand = S::combine(w[i], xv[i]);
                if (kk[i] != run_k) {
                    if (run != kInf) atomicMin(&s_y[wave][run_k], run);
                    run_k = kk[i];
                    run = cand;
                } else {
                    run = cand < run ? cand : run;
                }
            }
            if (run != kInf) atomicMin(&s_y[wave][run_k], run);
        }
    }
    __syncthreads();
    for (uint32_t k = lane; k < n_rows; k += kWave) y[s_row[wave][k]] = s_y[wave][k];
}

// Profiling only: the rows the next heavy pull visits (distance at or past hi) and their in-edges, per
// workgroup into part[2 * blockIdx.x], part[2 * blockIdx.x + 1].
static __global__ void __launch_bounds__(kPushThreads)
sssp_pull_rows_count_kernel(uint32_t n, const unsigned long long* __restrict__ dstamp, uint32_t hi,
                            const uint32_t* __restrict__ t_row_off, uint32_t* __restrict__ part) {
    __shared__ uint32_t s_r[kPushThreads / kWave], s_e[kPushThreads / kWave];
    uint32_t rows = 0, edges = 0;
    for (uint32_t v = blockIdx.x * kPushThreads + threadIdx.x; v < n; v += gridDim.x * kPushThreads) {
        if (static_cast<uint32_t>(dstamp[v] >> 32) >= hi) {
            ++rows;
            edges += t_row_off[v + 1] - t_row_off[v];
        }
    }
    for (int off = kWave / 2; off >= 1; off >>= 1) {
        rows += __shfl_down(rows, off);
        edges += __shfl_down(edges, off);
    }
    if (threadIdx.x % kWave == 0) {
        s_r[threadIdx.x / kWave] = rows;
        s_e[threadIdx.x / kWave] = edges;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = 0, e = 0;
        for (uint32_t w = 0; w < kPushThreads / kWave; ++w) {
            r += s_r[w];
            e += s_e[w];
        }
        part[2 * blockIdx.x] = r;
        part[2 * blockIdx.x + 1] = e;
    }
}

// Apply: every vertex whose y (plus y_add, saturating) beats its distance takes (y + y_add, stamp) and, with
// out-edges, joins the next frontier (segment blockIdx % kSegments of the output queue). Each vertex has
// one owner, so no atomics. With frontier_lo below kInf (the delta engine's last heavy phase) the next
// frontier is instead every vertex whose distance, after the apply, is finite and at least frontier_lo,
// each restamped with `stamp`: the frontier its distance scan would build.
static __global__ void __launch_bounds__(kPushThreads)
sssp_pull_apply_kernel(uint32_t n, const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ y,
                       unsigned long long* __restrict__ dstamp, uint32_t stamp, PushQueues pq, DistBound db,
                       uint32_t y_add = 0, uint32_t frontier_lo = kInf) {
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_red[kPushThreads / kWave];
    const uint32_t seg_out = blockIdx.x % kSegments;
    uint32_t activated = 0;
    for (uint32_t base = blockIdx.x * kScanRound; base < n; base += gridDim.x * kScanRound) {  // uniform
        // Loads go out a stage at a time with clamped indices (a lane past n reads vertex n - 1 and does
        // nothing with it), so no load waits behind another's branch: y and the distance words, then the
        // degrees of the vertices to enqueue.
        uint32_t cand[kScanPerThread], cdeg[kScanPerThread];
        uint32_t my_cnt = 0, my_deg = 0;
        uint32_t yv[kScanPerThread], d_new[kScanPerThread];
        unsigned long long ds[kScanPerThread];
        bool improved[kScanPerThread], enq[kScanPerThread];
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            const uint32_t vc = v < n ? v : n - 1;
            yv[r] = y[vc];
            ds[r] = dstamp[vc];
        }
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            yv[r] = (yv[r] >= kInf - y_add) ? kInf : yv[r] + y_add;
            const uint32_t d_old = static_cast<uint32_t>(ds[r] >> 32);
            improved[r] = v < n && yv[r] < d_old;
            d_new[r] = improved[r] ? yv[r] : d_old;
            enq[r] = frontier_lo == kInf ? improved[r] : (v < n && d_new[r] >= frontier_lo && d_new[r] != kInf);
        }
        uint32_t dgs[kScanPerThread];
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            const uint32_t vc = enq[r] ? v : 0u;
            dgs[r] = row_off[vc + 1] - row_off[vc];
        }
#pragma unroll
        for (uint32_t r = 0; r < kScanPerThread; ++r) {
            const uint32_t v = base + r * kPushThreads + threadIdx.x;
            cand[r] = kInf;
            cdeg[r] = 0;
            if (improved[r]) {
                db.lower(v, yv[r]);
                ++activated;
            }
            if (improved[r] || enq[r]) dstamp[v] = (static_cast<unsigned long long>(d_new[r]) << 32) | stamp;
            if (enq[r] && dgs[r] > 0) {
                cand[r] = v;
                cdeg[r] = dgs[r];
                ++my_cnt;
                my_deg += dgs[r];
            }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, seg_out, s_scan, &s_base);
        __syncthreads();
    }
    block_stat_add(pq.stats + blockIdx.x, activated, s_red);
}

// BFS bottom-up step over the in-edge CSR: every unvisited vertex looks for an in-neighbour in the
// current frontier bitmap and stops at the first one. The next frontier is written both as a bitmap
// (every word fully written by one wave, so it needs no clearing) and as a queue for a following
// top-down step. Each workgroup takes a contiguous run of 256-vertex groups.
constexpr uint32_t kPullGroups = 8;
__device__ __forceinline__ void pull_groups(uint32_t n, const uint32_t* __restrict__ t_row_off,
                                            const uint32_t* __restrict__ t_col, const uint32_t* __restrict__ row_off,
                                            uint32_t* __restrict__ depth, const uint32_t* __restrict__ fb_cur,
                                            uint32_t* __restrict__ fb_next, uint32_t next_level, const PushQueues& pq) {
    __shared__ unsigned long long s_scan[kPushThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_red[kPushThreads / kWave];
    uint32_t activated = 0;
    uint32_t scanned = 0;
    const uint32_t words = (n + 31) / 32;
    const uint32_t groups = (n + kPushThreads - 1) / kPushThreads;
    const uint32_t per = (groups + gridDim.x - 1) / gridDim.x;
    const uint32_t g_beg = blockIdx.x * per;
    const uint32_t g_end = (g_beg + per < groups) ? g_beg + per : groups;
    const uint32_t seg_out = blockIdx.x % kSegments;
    // kPullGroups groups per round and one queue reservation per round: waves run through their vertices
    // without waiting for each other group by group
    for (uint32_t g0 = g_beg; g0 < g_end; g0 += kPullGroups) {  // uniform per workgroup
        // Every load of a stage is issued before any is waited on, with indices clamped rather than loads
        // behind branches (a lane with nothing to load reads entry 0): the groups' depths, then the
        // unvisited vertices' in-edge ranges, then the first in-edge of every group and its bitmap
        // word (one per vertex keeps 71 VGPRs, six waves per SIMD; two took 85 and five, four 115 and
        // four). Only vertices with no hit there and more in-edges go on alone.
        uint32_t vv[kPullGroups], rb[kPullGroups], re[kPullGroups], dep[kPullGroups];
        bool found[kPullGroups];
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
            vv[r] = (g0 + r) * kPushThreads + threadIdx.x;
            const bool in = g0 + r < g_end && vv[r] < n;
            dep[r] = depth[in ? vv[r] : 0u];
            found[r] = in;
        }
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
            found[r] = found[r] && dep[r] == kInf;  // unvisited (scan below)
            const uint32_t vc = found[r] ? vv[r] : 0u;
            rb[r] = t_row_off[vc];
            re[r] = t_row_off[vc + 1];
            if (!found[r]) re[r] = rb[r];
        }
        constexpr int kFirst = 1;
        uint32_t u0[kPullGroups][kFirst], w0[kPullGroups][kFirst];
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
#pragma unroll
            for (int q = 0; q < kFirst; ++q) {
                const uint32_t k = rb[r] + q;
                const uint32_t c = t_col[k < re[r] ? k : 0u];
                u0[r][q] = k < re[r] ? c : kInf;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
#pragma unroll
            for (int q = 0; q < kFirst; ++q) w0[r][q] = fb_cur[(u0[r][q] != kInf ? u0[r][q] : 0u) >> 5];
        }
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
            bool hit = false;
#pragma unroll
            for (int q = 0; q < kFirst; ++q) {
                if (!hit && u0[r][q] != kInf) {
                    ++scanned;
                    if ((w0[r][q] >> (u0[r][q] & 31u)) & 1u) hit = true;
                }
            }
            // the rest of the in-edges, four per step, their bitmap words loaded together
            for (uint32_t k = rb[r] + kFirst; k < re[r] && !hit; k += 4) {
                uint32_t u[4], wd[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) u[q] = (k + q < re[r]) ? t_col[k + q] : kInf;
#pragma unroll
                for (int q = 0; q < 4; ++q) wd[q] = (u[q] != kInf) ? fb_cur[u[q] >> 5] : 0u;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (!hit && u[q] != kInf) {
                        ++scanned;
                        if ((wd[q] >> (u[q] & 31u)) & 1u) hit = true;
                    }
                }
            }
            found[r] = hit;
        }
        uint32_t dgs[kPullGroups];
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
            const uint32_t vc = found[r] ? vv[r] : 0u;
            dgs[r] = row_off[vc + 1] - row_off[vc];
        }
        uint32_t cand[kPullGroups], cdeg[kPullGroups];
        uint32_t my_cnt = 0, my_deg = 0;
        const uint32_t lane = lane_id();
#pragma unroll
        for (uint32_t r = 0; r < kPullGroups; ++r) {
            cand[r] = kInf;
            cdeg[r] = 0;
            if (g0 + r >= g_end) continue;  // uniform
            const uint64_t mk = ballot(found[r]);
            const uint32_t word = vv[r] >> 5;  // lanes 0..31 and 32..63 of a wave each own one word
            if ((lane & 31u) == 0 && word < words)
                fb_next[word] = (lane == 0) ? static_cast<uint32_t>(mk) : static_cast<uint32_t>(mk >> 32);
            if (found[r]) {
                depth[vv[r]] = next_level;
                ++activated;
                const uint32_t dg = dgs[r];
                if (dg > 0) {
                    cand[r] = vv[r];
                    cdeg[r] = dg;
                    ++my_cnt;
                    my_deg += dg;
                }
            }
        }
        push_enqueue(cand, cdeg, my_cnt, my_deg, pq, seg_out, s_scan, &s_base);
    }
    block_stat_add(pq.stats + blockIdx.x, activated, s_red);
    __syncthreads();
    block_stat_add(pq.stats + kMaxPushBlocks + blockIdx.x, scanned, s_red);
}

// One BFS level. With kDirOpt the level runs top-down (advance) or bottom-up (pull) by Beamer's rule,
// decided identically by every workgroup from the input counters and the previous level's history:
// pull when the frontier's out-edges exceed 1/14 of the edges not yet explored, push again when the
// frontier falls below n/24 vertices. History: hist[2*it] = (E << 32) | nf, hist[2*it + 1] = edges
// explored through this level | (pull << 63). Frontier bitmaps rotate over three buffers fb0..fb2
// (`words` words each): read fb[it%3], write fb[(it+1)%3], clear fb[(it+2)%3].
template <bool kDirOpt>
static __global__ void __launch_bounds__(kPushThreads)
bfs_step_kernel(uint32_t n, uint32_t m, const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col,
                const uint32_t* __restrict__ t_row_off, const uint32_t* __restrict__ t_col,
                uint32_t* __restrict__ depth, uint32_t* __restrict__ fb, uint32_t words, PushQueues pq, uint32_t it,
                unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_vb[kSegments + 1];
    __shared__ uint32_t s_eb[kSegments + 1];
    const bool active = push_prologue(pq, s_vb, s_eb);
    const uint32_t nf = s_vb[kSegments];
    const uint32_t E = s_eb[kSegments];
    bool pull = false;
    unsigned long long explored = 0;
    if (kDirOpt) {
        uint32_t* fb_clear = fb + static_cast<size_t>((it + 2) % 3) * words;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x)
            fb_clear[i] = 0u;
        bool pull_prev = false;
        if (it > 0) {
            const unsigned long long h = hist[2 * (it - 1) + 1];
            explored = h & ~(1ull << 63);
            pull_prev = (h >> 63) != 0;
        }
        const unsigned long long unexplored = (m > explored + E) ? m - explored - E : 0ull;
        pull = pull_prev ? (static_cast<unsigned long long>(nf) * 24 >= n)
                         : (static_cast<unsigned long long>(E) * 14 > unexplored);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        hist[2 * it] = (static_cast<unsigned long long>(E) << 32) | nf;
        hist[2 * it + 1] = (explored + E) | (pull ? (1ull << 63) : 0ull);
    }
    if (!active) return;
    uint32_t* fb_cur = fb + static_cast<size_t>(it % 3) * words;
    uint32_t* fb_next = fb + static_cast<size_t>((it + 1) % 3) * words;
    if (kDirOpt && pull) {
        pull_groups(n, t_row_off, t_col, row_off, depth, fb_cur, fb_next, it + 1, pq);
    } else {
        BfsPush pol{depth, kDirOpt ? fb_next : nullptr, it + 1};
        advance_tiles(FullRows(row_off), col, nullptr, pq, pol, s_vb, s_eb);
    }
}

// Sets up a one-vertex frontier {src} for a push engine: segment 0 holds src (when it has out-edges),
// every counter of the three rotating counter sets is cleared otherwise.
static __global__ void push_seed_kernel(const uint32_t* __restrict__ row_off, uint32_t src, uint32_t* q,
                                        uint32_t* qoff, unsigned long long* cnt3) {
    const uint32_t t = threadIdx.x;
    if (blockIdx.x != 0) return;
    for (uint32_t i = t; i < 3u * kSegments; i += blockDim.x) cnt3[i] = 0ull;
    __syncthreads();
    if (t == 0) {
        const uint32_t dg = row_off[src + 1] - row_off[src];
        q[0] = src;
        qoff[0] = 0;
        cnt3[0] = dg ? ((static_cast<unsigned long long>(dg) << 32) | 1ull) : 0ull;
    }
}

// Start of a push-engine run in one launch: the per-vertex state (SSSP: distance words (dist << 32 |
// stamp), all unreached but src; BFS: depths, and the direction-optimizing engine's first two frontier
// bitmaps with src set in the first), and, in block 0, the seeded input queue (push_seed_kernel) and the
// run counters ctr[0..8) cleared. With lc (delta-stepping) the seed's degree is its light edge count.
template <bool kSssp>
static __global__ void __launch_bounds__(256)
push_start_kernel(uint32_t n, uint32_t src, unsigned long long* __restrict__ dstamp, uint32_t* __restrict__ depth,
                  uint32_t* __restrict__ fb, uint32_t words, const uint32_t* __restrict__ row_off,
                  const uint32_t* __restrict__ lc, uint32_t* q, uint32_t* qoff, unsigned long long* cnt3,
                  unsigned long long* ctr, uint8_t* __restrict__ dbound) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t v = t0; v < n; v += stride) {
        if (kSssp) dstamp[v] = (v == src) ? 0ull : (static_cast<unsigned long long>(kInf) << 32);
        else depth[v] = (v == src) ? 0u : kInf;
        if (kSssp && dbound) dbound[v] = (v == src) ? 0u : 255u;
    }
    if (!kSssp && fb)
        for (uint32_t i = t0; i < 2 * words; i += stride) fb[i] = (i == (src >> 5)) ? (1u << (src & 31u)) : 0u;
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < 3u * kSegments; i += blockDim.x) cnt3[i] = 0ull;
        if (threadIdx.x < 8) ctr[threadIdx.x] = 0ull;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t dg = lc ? lc[src] : row_off[src + 1] - row_off[src];
            q[0] = src;
            qoff[0] = 0;
            cnt3[0] = dg ? ((static_cast<unsigned long long>(dg) << 32) | 1ull) : 0ull;
        }
    }
}

// End of a push-engine run, one workgroup: folds (and clears) the four stats slot regions -- activated and
// scanned of the run's launches, reached vertices and their out-edges of the result -- and writes the run's
// record to host memory: out[0..8) the counters ctr[0..8) with [3] += activated, [4] = reached,
// [5] = reached out-edges, [7] += scanned; out[8] the overflow flag; hist_out[0 .. 2 * launched) the
// iteration history.
static __global__ void __launch_bounds__(256)
push_finish_kernel(unsigned long long* __restrict__ slots, const unsigned long long* __restrict__ ctr,
                   const uint32_t* __restrict__ overflow, const unsigned long long* __restrict__ hist, uint32_t launched,
                   unsigned long long* out, unsigned long long* hist_out) {
    __shared__ unsigned long long s[4][256 / kWave];
    unsigned long long acc[4] = {0ull, 0ull, 0ull, 0ull};
    for (uint32_t i = threadIdx.x; i < kMaxPushBlocks; i += 256) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc[k] += slots[k * kMaxPushBlocks + i];
            slots[k * kMaxPushBlocks + i] = 0ull;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        acc[k] = wave_sum(acc[k]);
        if (lane_id() == 0) s[k][threadIdx.x / kWave] = acc[k];
    }
    for (uint32_t i = threadIdx.x; i < 2 * launched; i += 256) hist_out[i] = hist[i];
    __syncthreads();
    if (threadIdx.x < 8) {
        unsigned long long v = ctr[threadIdx.x];
        const int k = threadIdx.x == 3 ? 0 : threadIdx.x == 7 ? 1 : threadIdx.x == 4 ? 2 : threadIdx.x == 5 ? 3 : -1;
        if (k >= 0) {
            const unsigned long long t = s[k][0] + s[k][1] + s[k][2] + s[k][3];
            v = (k >= 2) ? t : v + t;
        }
        out[threadIdx.x] = v;
    }
    if (threadIdx.x == 8) out[8] = *overflow;
}

// Adds per-workgroup stats slots [0, count) to out[0] (one block, fixed order) and clears them.
static __global__ void __launch_bounds__(256) sum_slots_kernel(unsigned long long* __restrict__ slots, uint32_t count,
                                                                unsigned long long* __restrict__ out) {
    __shared__ unsigned long long s[256];
    unsigned long long acc = 0;
    for (uint32_t i = threadIdx.x; i < count; i += 256) {
        acc += slots[i];
        slots[i] = 0ull;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < 256; ++i) t += s[i];
        out[0] += t;
    }
}

// ------------------------------------------------------------------------------------------------
// Small frontiers: one persistent workgroup of 1024 threads runs consecutive iterations while the
// frontier's out-edges fit in LDS (kSmallCap), with workgroup barriers between iterations instead of
// kernel launches. A high-diameter graph (a road network or grid: thousands of levels of a few thousand
// vertices) otherwise pays a full launch of the 2048-workgroup kernel and its chain of dependent loads
// per level. The iterations are the advance kernel's, with the same relaxation policies (iteration
// stamps, depths, bitmaps and history records), so the engines can switch between the two kernels at
// any iteration.
//
// Input: the frontier of iteration it0 in the segmented queue (counter set it0 % 3, entries in q[it0&1];
// the caller guarantees nf <= kSmallCap). Per iteration: degrees and their exclusive prefix (LDS), then
// one thread per edge slot: relax, and winners append to the other LDS buffer (winners <= edges <=
// kSmallCap). Stops when the frontier is empty (result[1] = 0) or when the next input has more than
// kSmallCap edges or reaches it_limit: that frontier is written back to the segmented queue as the
// input of iteration result[0] (result[1] = 1), entries without out-edges dropped, all in segment 0.
// BFS direction-optimizing state is kept as the advance kernel keeps it in push mode: the frontier
// bitmap of the next level (policy's fb_next), the bitmap after next cleared, hist[2*it + 1] = edges
// explored through the level.
constexpr int kSmallThreads = 1024;
constexpr uint32_t kSmallCap = 8192;
constexpr uint32_t kSmallPer = kSmallCap / kSmallThreads;

template <typename P, typename RV = FullRows>
static __global__ void __launch_bounds__(kSmallThreads)
push_small_kernel(RV rv, const uint32_t* __restrict__ col,
                  const uint32_t* __restrict__ wts, P pol, uint32_t* q0, uint32_t* q1, uint32_t* qoff0,
                  uint32_t* qoff1, unsigned long long* qcnt, uint32_t cap, uint32_t it0, uint32_t it_limit,
                  uint32_t* fb, uint32_t words, unsigned long long* __restrict__ hist,
                  unsigned long long* __restrict__ stats, uint32_t* __restrict__ result) {
    __shared__ uint32_t s_q[2][kSmallCap];
    __shared__ uint32_t s_pre[kSmallCap + 1];
    __shared__ uint32_t s_rb[kSmallCap];
    __shared__ uint32_t s_vb[kSegments + 1];
    __shared__ uint32_t s_eb[kSegments + 1];
    __shared__ uint32_t s_n[2];
    __shared__ uint32_t s_scan[kSmallThreads / kWave];
    const uint32_t tid = threadIdx.x;

    // input frontier of iteration it0 from the segmented queue
    {
        const unsigned long long* cnt = qcnt + (it0 % 3) * kSegments;
        const uint32_t* qin = (it0 & 1) ? q1 : q0;
        if (tid < kWave) load_segment_bases(cnt, s_vb, s_eb);
        if (tid == 0) s_n[1] = 0;
        __syncthreads();
        const uint32_t nf = s_vb[kSegments] < kSmallCap ? s_vb[kSegments] : kSmallCap;
        for (uint32_t i = tid; i < nf; i += kSmallThreads) {
            const uint32_t sg = search_last_le(s_vb, kSegments, i);
            s_q[0][i] = qin[static_cast<size_t>(sg) * cap + (i - s_vb[sg])];
        }
        if (tid == 0) s_n[0] = nf;
        __syncthreads();
    }
    unsigned long long explored = 0;
    if (it0 > 0) explored = hist[2 * (it0 - 1) + 1] & ~(1ull << 63);
    uint32_t it = it0, cur = 0, activated = 0;
    uint32_t status = 0;
    for (;;) {
        const uint32_t n = s_n[cur];
        if (n == 0) break;
        // degrees and their exclusive prefix; kSmallPer consecutive entries per thread
        uint32_t rb[kSmallPer], dg[kSmallPer], local = 0;
#pragma unroll
        for (uint32_t r = 0; r < kSmallPer; ++r) {
            const uint32_t i = tid * kSmallPer + r;
            rb[r] = 0;
            dg[r] = 0;
            if (i < n) {
                const uint32_t u = s_q[cur][i];
                rb[r] = rv.begin(u);
                dg[r] = rv.degree(u);
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kSmallPer; ++r) local += dg[r];
        uint32_t E;
        uint32_t run = block_exclusive_sum<kSmallThreads>(local, E, s_scan);
#pragma unroll
        for (uint32_t r = 0; r < kSmallPer; ++r) {
            const uint32_t i = tid * kSmallPer + r;
            if (i < n) {
                s_pre[i] = run;
                s_rb[i] = rb[r];
            }
            run += dg[r];
        }
        if (tid == 0) s_pre[n] = E;
        if (E > kSmallCap || it >= it_limit) {
            // hand the frontier of iteration `it` back to the segmented queue (segment 0)
            __syncthreads();
            uint32_t* qo = (it & 1) ? q1 : q0;
            uint32_t* oo = (it & 1) ? qoff1 : qoff0;
            uint32_t keep = 0;
#pragma unroll
            for (uint32_t r = 0; r < kSmallPer; ++r) keep += (tid * kSmallPer + r < n && dg[r] > 0) ? 1u : 0u;
            uint32_t kept;
            uint32_t at = block_exclusive_sum<kSmallThreads>(keep, kept, s_scan);
#pragma unroll
            for (uint32_t r = 0; r < kSmallPer; ++r) {
                const uint32_t i = tid * kSmallPer + r;
                if (i < n && dg[r] > 0) {
                    qo[at] = s_q[cur][i];
                    oo[at] = s_pre[i];
                    ++at;
                }
            }
            unsigned long long* cin = qcnt + (it % 3) * kSegments;
            unsigned long long* cout = qcnt + ((it + 1) % 3) * kSegments;
            if (tid < kSegments) {
                cin[tid] = (tid == 0) ? ((static_cast<unsigned long long>(E) << 32) | kept) : 0ull;
                cout[tid] = 0ull;
            }
            status = 1;
            break;
        }
        if (tid == 0) {
            hist[2 * it] = (static_cast<unsigned long long>(E) << 32) | n;
            hist[2 * it + 1] = explored + E;
        }
        explored += E;
        P p = pol.at_iteration(it);
        if (fb) {
            // the advance kernel's bitmap rotation: write fb[(it+1)%3], clear fb[(it+2)%3]
            p.set_next_bitmap(fb + static_cast<size_t>((it + 1) % 3) * words);
            uint32_t* fc = fb + static_cast<size_t>((it + 2) % 3) * words;
            for (uint32_t i = tid; i < words; i += kSmallThreads) fc[i] = 0u;
        }
        __syncthreads();  // prefixes visible; bitmap cleared before any winner sets a bit in it
        const uint32_t nxt = cur ^ 1;
        for (uint32_t e = tid; e < E; e += kSmallThreads) {
            const uint32_t j = search_last_le(s_pre, n, e);
            const uint32_t u = s_q[cur][j];
            const uint32_t k = s_rb[j] + (e - s_pre[j]);
            const uint32_t v = col[k];
            const uint32_t w = P::kWeighted ? wts[k] : 1u;
            if (p.relax1(p.fresh_source_value(u), v, w)) {
                ++activated;
                const uint32_t pos = atomicAdd(&s_n[nxt], 1u);
                s_q[nxt][pos] = v;
            }
        }
        __syncthreads();
        if (tid == 0) s_n[cur] = 0;
        cur = nxt;
        ++it;
    }
    __syncthreads();
    uint32_t total;
    (void)block_exclusive_sum<kSmallThreads>(activated, total, s_scan);
    if (tid == 0) {
        stats[0] += total;
        result[0] = it;
        result[1] = status;
    }
}

}  // namespace g1
