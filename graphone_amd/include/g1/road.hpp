// Whole queries on high-diameter, low-degree graphs (road networks, grids): one workgroup of 1024
// threads runs one single-source BFS or SSSP from start to end, with workgroup barriers between
// iterations. Such a graph needs thousands of iterations of a few thousand edges each; as separate
// launches of the 2048-workgroup advance kernel each iteration would cost its launch and its chain of
// dependent loads, and a Bellman-Ford frontier re-relaxes every vertex about a hundred times on a
// weighted grid. Here an iteration costs a handful of dependent memory round trips, and SSSP runs
// near-far (Davidson et al., "Work-efficient parallel GPU methods for single-source shortest paths",
// IPDPS 2014): only vertices below a threshold T are expanded; the others wait in a far list until the
// near frontier is empty, then T grows by delta and the far list is split. One launch runs any number of
// independent queries, one per workgroup (the batched road-graph engine of the query server).
//
// The relaxations are the push engine's (SsspPush / BfsPush: 64-bit (dist, stamp) atomicMin, depth CAS),
// so results are bit-identical to every other engine's. Near and far lists live in global memory (per
// query), the per-iteration degree prefix in LDS (chunks of kRoadChunk entries).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/push.hpp"

namespace g1 {

constexpr int kRoadThreads = 1024;
constexpr uint32_t kRoadChunk = 8192;
constexpr uint32_t kRoadPer = kRoadChunk / kRoadThreads;

// Per-query device state. dstamp (SSSP) or depth (BFS) must be initialized by the caller (source set).
struct RoadQuery {
    uint32_t sssp;               // 1: SSSP (dstamp), 0: BFS (depth)
    unsigned long long* dstamp;  // SSSP: (dist << 32 | stamp), n words
    uint32_t* depth;             // BFS: n words
    uint32_t* q[2];              // near frontier buffers, q_cap (>= n) entries each
    uint32_t* far[2];            // far lists (SSSP), far_cap entries each
    uint32_t src;
    uint32_t status;             // out: 0 done, 2 far list overflow (run the query elsewhere)
    uint32_t iterations;         // out: expanded iterations
    uint32_t splits;             // out: far-list splits
    unsigned long long frontier_sum;  // out: vertices expanded
    unsigned long long edges_sum;     // out: relaxations
};

static __global__ void __launch_bounds__(kRoadThreads)
road_query_kernel(const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col,
                  const uint32_t* __restrict__ wts, RoadQuery* __restrict__ queries, uint32_t delta, uint32_t far_cap,
                  uint32_t q_cap, uint32_t it_limit) {
    __shared__ uint32_t s_pre[kRoadChunk + 1];
    __shared__ uint32_t s_rb[kRoadChunk];
    __shared__ uint32_t s_u[kRoadChunk];
    __shared__ uint32_t s_scan[kRoadThreads / kWave];
    __shared__ uint32_t s_n[2];
    __shared__ uint32_t s_far[2];
    __shared__ uint32_t s_bad;
    RoadQuery& Q = queries[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const bool kSssp = Q.sssp != 0;  // uniform per workgroup
    // SSSP: a relaxation that wins (the vertex's first improvement of the iteration) files the vertex by
    // its new distance: near frontier (q[cur^1]) below T, far list above. A later improvement in the same
    // iteration does not win, except one that crosses T from above (the vertex was filed far, or not at
    // all, by a value >= T): it files the vertex near as well. A near entry is expanded with the distance it
    // has when expanded. BFS: plain alternating frontiers.
    if (tid == 0) {
        Q.q[0][0] = Q.src;
        s_n[0] = 1;
        s_n[1] = 0;
        s_far[0] = 0;
        s_far[1] = 0;
        s_bad = 0;
    }
    __syncthreads();
    uint32_t cur = 0, fcur = 0, it = 0, splits = 0;
    uint32_t T = delta;  // near-far threshold (SSSP)
    unsigned long long fsum = 0, esum = 0;
    for (;;) {
        const uint32_t n = s_n[cur];
        if (n == 0) {
            if (!kSssp || s_far[fcur] == 0 || s_bad) break;
            // Split: raise T. A far entry whose distance is now in [T_old, T) becomes a candidate; one
            // below T_old was already expanded (every improvement made it a candidate, and a candidate
            // below the threshold of its time is expanded, while one above it was filed far again and
            // released by the split that passed it); the rest stay far. Copies of one vertex may be
            // expanded twice: that only repeats relaxations that cannot win.
            const uint32_t nfar = s_far[fcur];
            const uint32_t t_old = T;
            T = (T > 0xFFFFFFFFu - delta) ? 0xFFFFFFFFu : T + delta;
            for (uint32_t i = tid; i < nfar; i += kRoadThreads) {
                const uint32_t v = Q.far[fcur][i];
                const uint32_t d = static_cast<uint32_t>(
                    __hip_atomic_load(&Q.dstamp[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
                if (d < t_old) continue;
                if (d < T) {
                    const uint32_t pos = atomicAdd(&s_n[cur], 1u);
                    if (pos < q_cap) Q.q[cur][pos] = v;
                    else s_bad = 1;
                } else {
                    const uint32_t pos = atomicAdd(&s_far[fcur ^ 1], 1u);
                    if (pos < far_cap) Q.far[fcur ^ 1][pos] = v;
                    else s_bad = 1;
                }
            }
            __syncthreads();
            if (tid == 0) {
                s_far[fcur] = 0;
                if (s_n[cur] > q_cap) s_n[cur] = q_cap;  // (flagged above; the run is abandoned)
            }
            fcur ^= 1;
            ++splits;
            __syncthreads();
            if (s_bad) break;
            continue;
        }
        if (it >= it_limit) {
            if (tid == 0) s_bad = 1;
            break;
        }
        uint32_t* near = Q.q[cur];
        const uint32_t out = cur ^ 1;
        fsum += n;
        // The frontier in chunks of kRoadChunk entries: degree prefix in LDS, then one thread per edge.
        for (uint32_t c0 = 0; c0 < n; c0 += kRoadChunk) {
            const uint32_t cn = (n - c0 < kRoadChunk) ? n - c0 : kRoadChunk;
            uint32_t rb[kRoadPer], dg[kRoadPer], local = 0;
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) {
                const uint32_t i = tid * kRoadPer + r;
                rb[r] = 0;
                dg[r] = 0;
                if (i < cn) {
                    const uint32_t u = near[c0 + i];
                    s_u[i] = u;
                    rb[r] = row_off[u];
                    dg[r] = row_off[u + 1] - rb[r];
                }
            }
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) local += dg[r];
            uint32_t E;
            uint32_t run = block_exclusive_sum<kRoadThreads>(local, E, s_scan);
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) {
                const uint32_t i = tid * kRoadPer + r;
                if (i < cn) {
                    s_pre[i] = run;
                    s_rb[i] = rb[r];
                }
                run += dg[r];
            }
            if (tid == 0) s_pre[cn] = E;
            __syncthreads();
            esum += E;
            for (uint32_t e = tid; e < E; e += kRoadThreads) {
                const uint32_t j = search_last_le(s_pre, cn, e);
                const uint32_t u = s_u[j];
                const uint32_t k = s_rb[j] + (e - s_pre[j]);
                const uint32_t v = col[k];
                if (kSssp) {
                    const uint32_t du = static_cast<uint32_t>(
                        __hip_atomic_load(&Q.dstamp[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
                    const uint32_t w = wts[k];
                    const uint32_t nd = du + w;
                    if (nd < static_cast<uint32_t>(Q.dstamp[v] >> 32)) {
                        const unsigned long long old =
                            atomicMin(&Q.dstamp[v], (static_cast<unsigned long long>(nd) << 32) | (it + 1));
                        const uint32_t od = static_cast<uint32_t>(old >> 32);
                        if (nd < od) {
                            const bool first = static_cast<uint32_t>(old) < it + 1;
                            if (nd < T && (first || od >= T)) {
                                const uint32_t pos = atomicAdd(&s_n[out], 1u);
                                if (pos < q_cap) Q.q[out][pos] = v;
                                else s_bad = 1;
                            } else if (nd >= T && first) {
                                const uint32_t pos = atomicAdd(&s_far[fcur], 1u);
                                if (pos < far_cap) Q.far[fcur][pos] = v;
                                else s_bad = 1;
                            }
                        }
                    }
                } else {
                    if (BfsPush{Q.depth, nullptr, it + 1}.relax1(0u, v, 1u)) {
                        const uint32_t pos = atomicAdd(&s_n[out], 1u);
                        if (pos < q_cap) Q.q[out][pos] = v;
                        else s_bad = 1;
                    }
                }
            }
            __syncthreads();
        }
        if (tid == 0) s_n[cur] = 0;
        cur = out;
        ++it;
        __syncthreads();
        if (s_bad) break;
    }
    __syncthreads();
    if (tid == 0) {
        Q.status = s_bad ? 2u : 0u;
        Q.iterations = it;
        Q.splits = splits;
        Q.frontier_sum = fsum;  // uniform: every thread saw the same n and E
        Q.edges_sum = esum;
    }
}

}  // namespace g1
