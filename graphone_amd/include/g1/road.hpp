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
// so results are bit-identical to every other engine's. An iteration is a chain of three global round
// trips: the frontier's row offsets (and, for SSSP, the sources' distances, read past the CU's cache),
// then the edges' targets and weights, then the atomics (SSSP goes straight to the atomicMin; a read first
// would add a round trip). The first kRoadChunk entries of each near frontier live in LDS, the rest and the
// far lists in global memory (per query); the per-chunk degree prefix is in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/push.hpp"

namespace g1 {

constexpr int kRoadThreads = 1024;
constexpr uint32_t kRoadChunk = 4096;
constexpr uint32_t kRoadPer = kRoadChunk / kRoadThreads;
constexpr uint32_t kRoadEdges = 4;  // edges in flight per thread

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
    __shared__ uint32_t s_du[kRoadChunk];
    __shared__ uint32_t s_u[kRoadChunk];
    __shared__ uint32_t s_q[2][kRoadChunk];  // entries 0 .. kRoadChunk-1 of the two near frontiers
    __shared__ uint32_t s_scan[kRoadThreads / kWave];
    __shared__ uint32_t s_n[2];
    __shared__ uint32_t s_far[2];
    __shared__ uint32_t s_bad;
    RoadQuery& Q = queries[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const bool kSssp = Q.sssp != 0;  // uniform per workgroup
    // Appends v to near frontier b: LDS for the first kRoadChunk entries, the query's buffer after them.
    auto push_near = [&](uint32_t b, uint32_t v) {
        const uint32_t pos = atomicAdd(&s_n[b], 1u);
        if (pos >= q_cap) s_bad = 1;
        else if (pos < kRoadChunk) s_q[b][pos] = v;
        else Q.q[b][pos] = v;
    };
    // SSSP: a relaxation that wins (the vertex's first improvement of the iteration) files the vertex by
    // its new distance: near frontier (out) below T, far list above. A later improvement in the same
    // iteration does not win, except one that crosses T from above (the vertex was filed far, or not at
    // all, by a value >= T): it files the vertex near as well. A near entry is expanded with the distance it
    // has when its chunk is staged (an improvement after that re-files it for the next iteration). BFS:
    // plain alternating frontiers.
    if (tid == 0) {
        s_q[0][0] = Q.src;
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
                    push_near(cur, v);
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
        const uint32_t* near = Q.q[cur];
        const uint32_t out = cur ^ 1;
        fsum += n;
        // The frontier in chunks of kRoadChunk entries: degree prefix in LDS, then kRoadEdges edges per
        // thread at a time.
        for (uint32_t c0 = 0; c0 < n; c0 += kRoadChunk) {
            const uint32_t cn = (n - c0 < kRoadChunk) ? n - c0 : kRoadChunk;
            uint32_t u[kRoadPer], rb[kRoadPer], re[kRoadPer], du[kRoadPer], local = 0;
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) {
                const uint32_t i = tid * kRoadPer + r;
                u[r] = kInf;
                if (i < cn) u[r] = c0 == 0 ? s_q[cur][i] : near[c0 + i];
            }
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) {
                rb[r] = 0;
                re[r] = 0;
                du[r] = 0;
                if (u[r] != kInf) {
                    rb[r] = row_off[u[r]];
                    re[r] = row_off[u[r] + 1];
                    if (kSssp)
                        du[r] = static_cast<uint32_t>(
                            __hip_atomic_load(&Q.dstamp[u[r]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32);
                }
            }
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) local += re[r] - rb[r];
            uint32_t E;
            uint32_t run = block_exclusive_sum<kRoadThreads>(local, E, s_scan);
#pragma unroll
            for (uint32_t r = 0; r < kRoadPer; ++r) {
                const uint32_t i = tid * kRoadPer + r;
                if (i < cn) {
                    s_pre[i] = run;
                    s_rb[i] = rb[r];
                    s_u[i] = u[r];
                    s_du[i] = du[r];
                }
                run += re[r] - rb[r];
            }
            if (tid == 0) s_pre[cn] = E;
            __syncthreads();
            esum += E;
            for (uint32_t e0 = 0; e0 < E; e0 += kRoadEdges * kRoadThreads) {
                uint32_t v[kRoadEdges], nd[kRoadEdges];
#pragma unroll
                for (uint32_t r = 0; r < kRoadEdges; ++r) {
                    const uint32_t e = e0 + r * kRoadThreads + tid;
                    v[r] = kInf;
                    nd[r] = 0;
                    if (e < E) {
                        const uint32_t j = search_last_le(s_pre, cn, e);
                        const uint32_t k = s_rb[j] + (e - s_pre[j]);
                        v[r] = col[k];
                        if (kSssp) nd[r] = s_du[j] + wts[k];
                    }
                }
                if (kSssp) {
                    unsigned long long old[kRoadEdges];
#pragma unroll
                    for (uint32_t r = 0; r < kRoadEdges; ++r)
                        old[r] = v[r] != kInf ? atomicMin(&Q.dstamp[v[r]],
                                                          (static_cast<unsigned long long>(nd[r]) << 32) | (it + 1))
                                              : 0ull;
#pragma unroll
                    for (uint32_t r = 0; r < kRoadEdges; ++r) {
                        const uint32_t od = static_cast<uint32_t>(old[r] >> 32);
                        if (v[r] == kInf || nd[r] >= od) continue;
                        const bool first = static_cast<uint32_t>(old[r]) < it + 1;
                        if (nd[r] < T && (first || od >= T)) {
                            push_near(out, v[r]);
                        } else if (nd[r] >= T && first) {
                            const uint32_t pos = atomicAdd(&s_far[fcur], 1u);
                            if (pos < far_cap) Q.far[fcur][pos] = v[r];
                            else s_bad = 1;
                        }
                    }
                } else {
#pragma unroll
                    for (uint32_t r = 0; r < kRoadEdges; ++r)
                        if (v[r] != kInf && BfsPush{Q.depth, nullptr, it + 1}.relax1(0u, v[r], 1u)) push_near(out, v[r]);
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

// SSSP queries on small road-like graphs (n <= kRoadLdsMaxN): the same near-far iteration with the
// distances kept in LDS as 16-bit values (vertex v in half v & 1 of word v >> 1, 0xFFFF = unreached), so a
// relaxation is an LDS compare-and-swap instead of a device-scope atomic on a distance word, and a
// source's distance an LDS read. "The vertex's first improvement of this iteration" is a bit of an LDS
// bitmap cleared after every iteration (the stamp of the global kernel). Distances are exact while they
// stay below 0xFFFF; a relaxation that would reach it ends the query with status 2, as does a far-list
// overflow, leaving Q.dstamp untouched so the caller can run the query on road_query_kernel instead. On
// success the distances are written to Q.dstamp (dist << 32). Same results, bit for bit, as every engine.
constexpr uint32_t kRoadLdsMaxN = 46080;  // 16-bit distances: 90 KB of LDS
constexpr uint32_t kRoadLdsChunk = 2048;
constexpr uint32_t kRoadLdsPer = kRoadLdsChunk / kRoadThreads;

static __global__ void __launch_bounds__(kRoadThreads)
road_lds_sssp_kernel(uint32_t n, const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col,
                     const uint32_t* __restrict__ wts, RoadQuery* __restrict__ queries, uint32_t delta,
                     uint32_t far_cap, uint32_t q_cap, uint32_t it_limit) {
    __shared__ uint32_t s_d[kRoadLdsMaxN / 2];
    __shared__ uint32_t s_first[kRoadLdsMaxN / 32];
    __shared__ uint32_t s_pre[kRoadLdsChunk + 1];
    __shared__ uint32_t s_rb[kRoadLdsChunk];
    __shared__ uint32_t s_du[kRoadLdsChunk];
    __shared__ uint32_t s_q[2][kRoadLdsChunk];
    __shared__ uint32_t s_scan[kRoadThreads / kWave];
    __shared__ uint32_t s_n[2];
    __shared__ uint32_t s_far[2];
    __shared__ uint32_t s_bad;
    RoadQuery& Q = queries[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const uint32_t words = (n + 1) / 2, fwords = (n + 31) / 32;
    for (uint32_t i = tid; i < words; i += kRoadThreads) s_d[i] = 0xFFFFFFFFu;
    for (uint32_t i = tid; i < fwords; i += kRoadThreads) s_first[i] = 0u;
    if (tid == 0) {
        s_q[0][0] = Q.src;
        s_n[0] = 1;
        s_n[1] = 0;
        s_far[0] = 0;
        s_far[1] = 0;
        s_bad = 0;
    }
    __syncthreads();
    if (tid == 0) s_d[Q.src >> 1] &= ~(0xFFFFu << (16u * (Q.src & 1u)));  // distance 0
    __syncthreads();
    auto dist = [&](uint32_t v) -> uint32_t { return (s_d[v >> 1] >> (16u * (v & 1u))) & 0xFFFFu; };
    auto push_near = [&](uint32_t b, uint32_t v) {
        const uint32_t pos = atomicAdd(&s_n[b], 1u);
        if (pos >= q_cap) s_bad = 1;
        else if (pos < kRoadLdsChunk) s_q[b][pos] = v;
        else Q.q[b][pos] = v;
    };
    auto push_far = [&](uint32_t f, uint32_t v) {
        const uint32_t pos = atomicAdd(&s_far[f], 1u);
        if (pos < far_cap) Q.far[f][pos] = v;
        else s_bad = 1;
    };
    uint32_t cur = 0, fcur = 0, it = 0, splits = 0;
    uint32_t T = delta;
    unsigned long long fsum = 0, esum = 0;
    for (;;) {
        const uint32_t nn = s_n[cur];
        if (nn == 0) {
            if (s_far[fcur] == 0 || s_bad) break;
            const uint32_t nfar = s_far[fcur];
            const uint32_t t_old = T;
            T = (T > 0xFFFFFFFFu - delta) ? 0xFFFFFFFFu : T + delta;
            for (uint32_t i = tid; i < nfar; i += kRoadThreads) {
                const uint32_t v = Q.far[fcur][i];
                const uint32_t d = dist(v);
                if (d < t_old) continue;
                if (d < T) push_near(cur, v);
                else push_far(fcur ^ 1, v);
            }
            __syncthreads();
            if (tid == 0) {
                s_far[fcur] = 0;
                if (s_n[cur] > q_cap) s_n[cur] = q_cap;
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
        const uint32_t* near = Q.q[cur];
        const uint32_t out = cur ^ 1;
        fsum += nn;
        for (uint32_t c0 = 0; c0 < nn; c0 += kRoadLdsChunk) {
            const uint32_t cn = (nn - c0 < kRoadLdsChunk) ? nn - c0 : kRoadLdsChunk;
            uint32_t rb[kRoadLdsPer], re[kRoadLdsPer], du[kRoadLdsPer], local = 0;
#pragma unroll
            for (uint32_t r = 0; r < kRoadLdsPer; ++r) {
                const uint32_t i = tid * kRoadLdsPer + r;
                rb[r] = 0;
                re[r] = 0;
                du[r] = 0;
                if (i < cn) {
                    const uint32_t u = c0 == 0 ? s_q[cur][i] : near[c0 + i];
                    rb[r] = row_off[u];
                    re[r] = row_off[u + 1];
                    du[r] = dist(u);
                }
            }
#pragma unroll
            for (uint32_t r = 0; r < kRoadLdsPer; ++r) local += re[r] - rb[r];
            uint32_t E;
            uint32_t run = block_exclusive_sum<kRoadThreads>(local, E, s_scan);
#pragma unroll
            for (uint32_t r = 0; r < kRoadLdsPer; ++r) {
                const uint32_t i = tid * kRoadLdsPer + r;
                if (i < cn) {
                    s_pre[i] = run;
                    s_rb[i] = rb[r];
                    s_du[i] = du[r];
                }
                run += re[r] - rb[r];
            }
            if (tid == 0) s_pre[cn] = E;
            __syncthreads();
            esum += E;
            for (uint32_t e0 = 0; e0 < E; e0 += kRoadEdges * kRoadThreads) {
                uint32_t v[kRoadEdges], w[kRoadEdges], d0[kRoadEdges];
#pragma unroll
                for (uint32_t r = 0; r < kRoadEdges; ++r) {
                    const uint32_t e = e0 + r * kRoadThreads + tid;
                    v[r] = kInf;
                    if (e < E) {
                        const uint32_t j = search_last_le(s_pre, cn, e);
                        const uint32_t k = s_rb[j] + (e - s_pre[j]);
                        v[r] = col[k];
                        w[r] = wts[k];
                        d0[r] = s_du[j];
                    }
                }
#pragma unroll
                for (uint32_t r = 0; r < kRoadEdges; ++r) {
                    if (v[r] == kInf) continue;
                    if (w[r] >= 0xFFFFu - d0[r]) {  // d0 + w would reach 0xFFFF: not representable here
                        s_bad = 1;
                        continue;
                    }
                    const uint32_t nd = d0[r] + w[r];
                    uint32_t* word = &s_d[v[r] >> 1];
                    const uint32_t sh = 16u * (v[r] & 1u);
                    uint32_t old = *word, od = 0;
                    bool improved = false;
                    for (;;) {
                        od = (old >> sh) & 0xFFFFu;
                        if (nd >= od) break;
                        const uint32_t prev = atomicCAS(word, old, (old & ~(0xFFFFu << sh)) | (nd << sh));
                        if (prev == old) {
                            improved = true;
                            break;
                        }
                        old = prev;
                    }
                    if (!improved) continue;
                    const uint32_t bit = 1u << (v[r] & 31u);
                    const bool first = (atomicOr(&s_first[v[r] >> 5], bit) & bit) == 0u;
                    const uint32_t odf = od == 0xFFFFu ? kInf : od;
                    if (nd < T && (first || odf >= T)) push_near(out, v[r]);
                    else if (nd >= T && first) push_far(fcur, v[r]);
                }
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < fwords; i += kRoadThreads) s_first[i] = 0u;
        if (tid == 0) s_n[cur] = 0;
        cur = out;
        ++it;
        __syncthreads();
        if (s_bad) break;
    }
    __syncthreads();
    if (!s_bad) {
        for (uint32_t v = tid; v < n; v += kRoadThreads) {
            const uint32_t d = dist(v);
            Q.dstamp[v] = static_cast<unsigned long long>(d == 0xFFFFu ? kInf : d) << 32;
        }
    }
    if (tid == 0) {
        Q.status = s_bad ? 2u : 0u;
        Q.iterations = it;
        Q.splits = splits;
        Q.frontier_sum = fsum;
        Q.edges_sum = esum;
    }
}

}  // namespace g1
