// Dense SSSP iterations by vertex sweep, without a frontier queue (the advance of engine/advance_gc.h:24-158
// with the relax of apps/sssp_adv.cu:49-60, for frontiers that hold a large share of the graph's edges).
//
// The frontier of iteration `it` is the set of vertices whose (dist << 32 | stamp) word carries stamp `it`
// (improved during iteration it - 1), taken as a bitmap before the launch (sweep_frontier_bits_kernel) so
// that membership stays fixed while the launch restamps the vertices it improves -- the same frontier a
// queue would hold. One launch relaxes every out-edge of those vertices without waiting on its atomics
// (the no-return mode of the push engine: the winners are the vertices stamped it + 1 afterwards, enqueued
// by sssp_collect_kernel), each source's distance read live when its row (or piece) starts, as the queue
// advance reads it.
//
// Work items, one wavefront each, dealt round-robin over the grid's waves:
//  - the pieces of the long rows (more than kSweepLong out-edges), cut into runs of at most kSweepPiece
//    edges once per graph (ensure_sweep_pieces): a wave streams a piece 64 consecutive edges per step,
//    with no search at all (on RMAT22 the rows past 128 edges hold 73 % of the edges);
//  - groups of 64 consecutive vertices: the group's short rows that are in the frontier are relaxed by
//    its wave, their edges dealt out 64 per step (wave_expand).
// Every load of the row structure is a coalesced stream; the random accesses left are the target reads
// and the atomics of the relaxations themselves.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/push.hpp"

namespace g1 {

constexpr uint32_t kSweepLong = 128;    // rows longer than this are cut into pieces
constexpr uint32_t kSweepPiece = 1024;  // edges per piece
constexpr int kSweepUnroll = 4;         // 64-edge steps in flight per wave

// Piece of a long row: edges [e0, e1) of row v.
struct SweepPiece {
    uint32_t v;
    uint32_t e0;
    uint32_t e1;
    uint32_t pad;
};

// pieces per row (0 for short rows)
__host__ __device__ inline uint32_t sweep_row_pieces(uint32_t deg) {
    return deg > kSweepLong ? (deg + kSweepPiece - 1) / kSweepPiece : 0u;
}

static __global__ void __launch_bounds__(256)
sweep_piece_count_kernel(uint32_t n, const uint32_t* __restrict__ row_off, uint32_t* __restrict__ cnt) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) cnt[v] = sweep_row_pieces(row_off[v + 1] - row_off[v]);
}

// pos: exclusive prefix of the per-row piece counts
static __global__ void __launch_bounds__(256)
sweep_piece_write_kernel(uint32_t n, const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ pos,
                         SweepPiece* __restrict__ pieces) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint32_t rs = row_off[v], re = row_off[v + 1];
    const uint32_t k = sweep_row_pieces(re - rs);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t e0 = rs + i * kSweepPiece;
        pieces[pos[v] + i] = SweepPiece{v, e0, (re - e0 > kSweepPiece) ? e0 + kSweepPiece : re, 0u};
    }
}

// bits[w] bit j = vertex 64w + j carries stamp `it` (one wave per 64-vertex word).
static __global__ void __launch_bounds__(256)
sweep_frontier_bits_kernel(uint32_t n, const unsigned long long* __restrict__ dstamp, uint32_t it,
                           unsigned long long* __restrict__ bits) {
    const uint32_t words = (n + kWave - 1) / kWave;
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; w < words; w += waves) {
        const uint32_t v = w * kWave + lane_id();
        const uint64_t b = ballot(v < n && static_cast<uint32_t>(dstamp[v]) == it);
        if (lane_id() == 0) bits[w] = b;
    }
}

// XCD ownership of the targets (kOwned sweeps): vertex v belongs to XCD (v >> kOwnShift) % 8, so each XCD's
// share of the distance words is 1/8 of them in 32 KB runs, about the size of its 4 MB L2 on RMAT22.
constexpr uint32_t kOwnShift = 12;
constexpr uint32_t kXcds = 8;

// Relaxes up to kSweepUnroll x 64 edges held one per lane per step (ok[u]: slot u holds an edge e[u]
// whose source distance is du[u]); the atomics are not waited for. kOwned: only the edges whose target
// this workgroup's XCD owns (the column is read for every edge; weight, target read and atomic only for
// the owned ones).
template <bool kOwned>
__device__ __forceinline__ void sweep_relax(const uint32_t* __restrict__ col, const uint32_t* __restrict__ wts,
                                            unsigned long long* dstamp, uint32_t iter, uint32_t xcd,
                                            const uint32_t (&e)[kSweepUnroll], const uint32_t (&du)[kSweepUnroll],
                                            const bool (&ok)[kSweepUnroll]) {
    uint32_t v[kSweepUnroll], w[kSweepUnroll], cur[kSweepUnroll];
    bool on[kSweepUnroll];
#pragma unroll
    for (int u = 0; u < kSweepUnroll; ++u) v[u] = ok[u] ? col[e[u]] : 0u;
#pragma unroll
    for (int u = 0; u < kSweepUnroll; ++u) {
        on[u] = ok[u] && (!kOwned || ((v[u] >> kOwnShift) % kXcds) == xcd);
        if (on[u]) w[u] = wts[e[u]];
    }
    const uint32_t* dhi = reinterpret_cast<const uint32_t*>(dstamp) + 1;  // the distance halves
#pragma unroll
    for (int u = 0; u < kSweepUnroll; ++u) cur[u] = on[u] ? dhi[2ull * v[u]] : 0u;
#pragma unroll
    for (int u = 0; u < kSweepUnroll; ++u) {
        const uint32_t nd = du[u] + w[u];
        if (on[u] && nd < cur[u]) (void)atomicMin(&dstamp[v[u]], (static_cast<unsigned long long>(nd) << 32) | iter);
    }
}

// Deals the out-edges of the rows held one per lane (row start rs, length deg, may be 0; value tag) to
// the lanes of the wave, 64 consecutive edge slots per step and kSweepUnroll steps at a time: slot s
// belongs to the last row whose first slot is <= s. Rows with edges are first moved to the low lanes
// (so their first slots strictly increase); per step, the rows starting inside the step mark their
// offset in a 64-entry LDS array of the wave, and a lane's row is the number of marks at or below it
// plus the rows started before the step. `s_mark` is the wave's own 64 words. Wave-uniform control flow.
template <typename Relax>
__device__ __forceinline__ void wave_expand(uint32_t rs, uint32_t deg, uint32_t tag, uint32_t* s_mark,
                                            Relax&& relax) {
    const uint32_t lane = lane_id();
    const uint64_t has = ballot(deg > 0);
    if (has == 0ull) return;
    const uint32_t P = static_cast<uint32_t>(__popcll(has));
    const uint32_t dst = deg > 0 ? mask_prefix(has) : P + mask_prefix(~has);
    rs = static_cast<uint32_t>(__builtin_amdgcn_ds_permute(static_cast<int>(dst * 4), static_cast<int>(rs)));
    deg = static_cast<uint32_t>(__builtin_amdgcn_ds_permute(static_cast<int>(dst * 4), static_cast<int>(deg)));
    tag = static_cast<uint32_t>(__builtin_amdgcn_ds_permute(static_cast<int>(dst * 4), static_cast<int>(tag)));
    const uint32_t incl = wave_inclusive_sum(deg);
    const uint32_t pre = incl - deg;
    const uint32_t total = static_cast<uint32_t>(__shfl(static_cast<int>(incl), kWave - 1));
    const uint32_t delta = rs - pre;  // edge index of slot s of this row = delta + s
    const uint64_t le_mask = (lane == kWave - 1) ? ~0ull : ((2ull << lane) - 1ull);
    for (uint32_t base = 0; base < total; base += kWave * kSweepUnroll) {  // wave-uniform
        uint32_t e[kSweepUnroll], du[kSweepUnroll];
        bool ok[kSweepUnroll];
#pragma unroll
        for (int u = 0; u < kSweepUnroll; ++u) {
            const uint32_t b = base + u * kWave;
            ok[u] = false;
            e[u] = 0;
            du[u] = 0;
            if (b < total) {  // wave-uniform
                s_mark[lane] = 0u;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (lane < P && pre >= b && pre - b < static_cast<uint32_t>(kWave)) s_mark[pre - b] = 1u;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const uint64_t marks = ballot(s_mark[lane] != 0u);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const uint32_t before = static_cast<uint32_t>(__popcll(ballot(lane < P && pre < b)));
                const uint32_t owner = before + static_cast<uint32_t>(__popcll(marks & le_mask)) - 1u;
                const uint32_t s = b + lane;
                const uint32_t d = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(static_cast<int>(owner * 4),
                                                                                       static_cast<int>(delta)));
                const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(static_cast<int>(owner * 4),
                                                                                       static_cast<int>(tag)));
                ok[u] = s < total;
                e[u] = d + s;
                du[u] = t;
            }
        }
        relax(e, du, ok);
    }
}

// One dense iteration (see the file comment). kOwned: every XCD (workgroup b runs on XCD b % 8) walks all
// the work items and relaxes only the edges whose targets it owns, so the random reads of the distance
// words stay in its own L2; the columns are streamed by all eight. The grid is then a multiple of 8.
// Block 0 also does the push prologue's bookkeeping: the
// iteration's history record hist[2*it] = (E << 32) | nf of the input frontier (from its queue counters,
// which the previous iteration's collect or advance filled) and the counters of the round after next
// cleared.
template <bool kOwned>
static __global__ void __launch_bounds__(256)
sssp_sweep_kernel(uint32_t n, const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col,
                  const uint32_t* __restrict__ wts, unsigned long long* dstamp, uint32_t it,
                  const unsigned long long* __restrict__ fbits, const SweepPiece* __restrict__ pieces,
                  uint32_t n_pieces, PushQueues pq, unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_mark[256];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    if (blockIdx.x == 0) {
        if (threadIdx.x < kSegments) pq.cnt_clear[threadIdx.x] = 0ull;
        if (wave == 0) {
            const unsigned long long c = pq.cnt_in[lane];
            const unsigned long long nf = wave_sum<unsigned long long>(c & 0xFFFFFFFFull);
            const unsigned long long E = wave_sum<unsigned long long>(c >> 32);
            if (lane == 0) {
                hist[2 * it] = (E << 32) | nf;
                hist[2 * it + 1] = 0ull;
            }
        }
    }
    const uint32_t iter = it + 1;
    const uint32_t groups = (n + kWave - 1) / kWave;
    const uint32_t items = n_pieces + groups;
    const uint32_t wpb = blockDim.x / kWave;
    // the waves that share the items: all of the grid's, or (kOwned) those of this workgroup's XCD
    const uint32_t xcd = kOwned ? blockIdx.x % kXcds : 0u;
    const uint32_t waves = kOwned ? (gridDim.x / kXcds) * wpb : gridDim.x * wpb;
    const uint32_t first = kOwned ? (blockIdx.x / kXcds) * wpb + wave : blockIdx.x * wpb + wave;
    uint32_t* mark = s_mark + wave * kWave;
    auto relax = [&](const uint32_t (&e)[kSweepUnroll], const uint32_t (&du)[kSweepUnroll],
                     const bool (&ok)[kSweepUnroll]) { sweep_relax<kOwned>(col, wts, dstamp, iter, xcd, e, du, ok); };
    for (uint32_t item = first; item < items; item += waves) {  // wave-uniform
        if (item < n_pieces) {
            const SweepPiece pc = pieces[item];
            if (((fbits[pc.v / kWave] >> (pc.v % kWave)) & 1ull) == 0ull) continue;
            const uint32_t du = static_cast<uint32_t>(dstamp[pc.v] >> 32);
            for (uint32_t e0 = pc.e0; e0 < pc.e1; e0 += kWave * kSweepUnroll) {
                uint32_t e[kSweepUnroll], d[kSweepUnroll];
                bool ok[kSweepUnroll];
#pragma unroll
                for (int u = 0; u < kSweepUnroll; ++u) {
                    e[u] = e0 + u * kWave + lane;
                    d[u] = du;
                    ok[u] = e[u] < pc.e1;
                }
                relax(e, d, ok);
            }
        } else {
            const uint32_t grp = item - n_pieces;
            const unsigned long long word = fbits[grp];
            if (word == 0ull) continue;  // no frontier vertex in the group
            const uint32_t v = grp * kWave + lane;
            uint32_t rs = 0, deg = 0, du = 0;
            if ((word >> lane) & 1ull) {
                const uint32_t a = row_off[v], b = row_off[v + 1];
                if (b - a <= kSweepLong) {
                    rs = a;
                    deg = b - a;
                    du = static_cast<uint32_t>(dstamp[v] >> 32);
                }
            }
            wave_expand(rs, deg, du, mark, relax);
        }
    }
}

}  // namespace g1
