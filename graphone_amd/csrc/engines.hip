// Whole-algorithm runs of the C-ABI: the Run<F>(comp, frontier) loop (include/GraphOne/base/run.h:8-39)
// with the functor fixed to BFS, SSSP, PageRank or personalized PageRank, over the push engine
// (engine/advance_gc.h) or the pull SpMV engine (engine/blas_gc.h).
//
// Host loop: one launch per engine phase, then an 8-byte readback of the device frontier counter
// (the convergence test the reference does with frontier.IsConvergent()/a D2H copy per iteration).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <immintrin.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "g1/apply.hpp"
#include "g1/push.hpp"
#include "g1/road.hpp"
#include "g1/sweep.hpp"
#include "g1/spmv_merge.hpp"
#include "graph.hpp"
#include "spmv_launch.hpp"

namespace g1 {
namespace {

constexpr uint32_t kMaxIterations = 1u << 20;

uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t cap) {
    uint64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return static_cast<uint32_t>(b);
}

uint32_t persistent_cap() { return static_cast<uint32_t>(device_cu_count()) * 8; }

g1_status read_ctr(g1_graph* g, hipStream_t s, uint32_t first, uint32_t count) {
    G1_HIP(hipMemcpyAsync(g->h_ctr + first, g->ctr + first, count * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    G1_HIP(hipStreamSynchronize(s));
    return G1_OK;
}

__global__ void set_u32_kernel(uint32_t* p, uint32_t idx, uint32_t value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[idx] = value;
}

__global__ void clear_u64_kernel(unsigned long long* p, uint32_t count) {
    if (threadIdx.x < count && blockIdx.x == 0) p[threadIdx.x] = 0ull;
}

g1_status finish_stats(g1_graph* g, const uint32_t* d, g1_run_stats* st, hipStream_t s) {
    G1_TRY(ensure_push_work(g, s));  // the stats slots live with the push work areas
    clear_u64_kernel<<<1, 64, 0, s>>>(g->ctr + 4, 2);
    G1_LAUNCHED();
    reached_stats_kernel<<<grid_for(g->n, 256, kMaxPushBlocks), 256, 0, s>>>(g->n, d, g->row_off, g->slots,
                                                                            g->slots + kMaxPushBlocks);
    G1_LAUNCHED();
    sum_slots_kernel<<<1, 256, 0, s>>>(g->slots, kMaxPushBlocks, g->ctr + 4);
    G1_LAUNCHED();
    sum_slots_kernel<<<1, 256, 0, s>>>(g->slots + kMaxPushBlocks, kMaxPushBlocks, g->ctr + 5);
    G1_LAUNCHED();
    G1_TRY(read_ctr(g, s, 4, 2));
    st->reached = g->h_ctr[4];
    st->reached_edges = g->h_ctr[5];
    return G1_OK;
}

g1_status elapsed(g1_graph* g, g1_run_stats* st) {
    G1_HIP(hipEventSynchronize(g->ev[1]));
    float ms = 0.0f;
    G1_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
    st->elapsed_ms = ms;
    return G1_OK;
}

// ---------------------------------------------------------------- push engines (BFS / SSSP)

PushQueues make_queues(g1_graph* g, uint32_t k) {
    PushQueues pq;
    pq.q_in = g->q[k & 1];
    pq.qoff_in = g->qoff[k & 1];
    pq.cnt_in = g->qcnt + (k % 3) * kSegments;
    pq.q_out = g->q[(k + 1) & 1];
    pq.qoff_out = g->qoff[(k + 1) & 1];
    pq.cnt_out = g->qcnt + ((k + 1) % 3) * kSegments;
    pq.cnt_clear = g->qcnt + ((k + 2) % 3) * kSegments;
    pq.stats = g->slots;
    pq.overflow = g->overflow;
    pq.cap = g->seg_cap;
    return pq;
}

// The SSSP push engine relaxes a frontier of at least G1_SSSP_NORETURN (default 0.3) times the graph's
// edges without waiting on its atomics; G1_SSSP_NORETURN=0 disables it.
double sssp_noreturn_fraction() {
    static const double f = [] {
        const char* e = getenv("G1_SSSP_NORETURN");
        return e ? atof(e) : 0.3;
    }();
    return f;
}

// Dense SSSP iterations run on the queue-driven no-return advance. G1_SSSP_SWEEP=1 runs them as a vertex
// sweep instead (sssp_sweep_kernel, frontier bitmap, long rows in pieces), =2 as the sweep with every XCD
// relaxing only the targets it owns. Measured on the RMAT22 step (20 steps each): queue 2.43 ms, sweep
// 2.88-2.91 ms, owned sweep 7.74 ms; the results are the same.
int sssp_sweep_mode() {
    static const int m = [] {
        const char* e = getenv("G1_SSSP_SWEEP");
        const int v = e ? atoi(e) : 0;
        return (v >= 0 && v <= 2) ? v : 0;
    }();
    return m;
}
bool sssp_sweep_enabled() { return sssp_sweep_mode() != 0; }

// The SSSP push engines check a one-byte distance bound before reading a target's word (DistBound),
// unless G1_SSSP_DBOUND=0. Step: the largest weight / 64, rounded up (bounds cover distances up to
// 254 steps, about four times the largest weight; a longer distance keeps "no bound").
DistBound sssp_dist_bound(const g1_graph* g) {
    static const bool on = [] {
        const char* e = getenv("G1_SSSP_DBOUND");
        return !(e && atoi(e) == 0);
    }();
    const uint32_t step = std::max<uint32_t>(1u, (g->max_w + 63u) / 64u);
    return DistBound{on ? g->dbound : nullptr, step};
}

// Workgroups of a sweep launch: 1024 (four per CU), or G1_SSSP_SWEEP_BLOCKS (8..8192, rounded down to a
// multiple of 8: the owned sweep splits the grid by XCD).
uint32_t sssp_sweep_blocks() {
    static const uint32_t b = [] {
        const char* e = getenv("G1_SSSP_SWEEP_BLOCKS");
        const long v = e ? atol(e) : 0;
        return (v >= 8 && v <= 8192) ? static_cast<uint32_t>(v / 8 * 8) : 1024u;
    }();
    return b;
}

// Workgroup count of a push launch from the environment variable `name`, or `dflt`. Valid counts are the
// multiples of kSegments from kSegments to kMaxPushBlocks: push_segment_capacity() bounds a segment's
// entries by items / kSegments plus slack, which holds when every segment has the same number of
// producers (G / kSegments) each taking 1/G of the items. Anything else falls back to `dflt`.
uint32_t push_blocks_from_env(const char* name, uint32_t dflt) {
    const char* e = getenv(name);
    const long v = e ? atol(e) : 0;
    if (v >= static_cast<long>(kSegments) && v <= static_cast<long>(kMaxPushBlocks) && v % kSegments == 0)
        return static_cast<uint32_t>(v);
    return dflt;
}

// Workgroups of a no-return advance launch: 1024 (four per CU), or G1_SSSP_NR_BLOCKS. RMAT22 step, 10
// steps each: 2.44 ms at 1024 and at 768, 2.48 at 1536 and 2048, 2.55 at 896, 2.66 at 512 (measured;
// counts that are not a multiple of the 256 CUs did worse).
uint32_t sssp_noreturn_blocks() {
    static const uint32_t b = push_blocks_from_env("G1_SSSP_NR_BLOCKS", 1024u);
    return b;
}

// Workgroups of a returning SSSP advance launch: kMaxPushBlocks, or G1_SSSP_PUSH_BLOCKS. On the RMAT22
// step 1024 and 2048 are within the run-to-run spread (2.43-2.47 ms); 512 is no better.
uint32_t sssp_push_blocks() {
    static const uint32_t b = push_blocks_from_env("G1_SSSP_PUSH_BLOCKS", kMaxPushBlocks);
    return b;
}

// Workgroups of a BFS step launch: 1024 (four per CU), or G1_BFS_BLOCKS. RMAT22 BFS from the hub, 10
// steps each, twice: 0.228 / 0.226 ms at 1024, 0.230 / 0.232 at 2048, 0.255 at 512.
uint32_t bfs_step_blocks() {
    static const uint32_t b = push_blocks_from_env("G1_BFS_BLOCKS", 1024u);
    return b;
}

// The push SSSP engines switch an iteration to a pull step when its frontier has at least
// G1_SSSP_PULL (default 0.8) times the graph's edges; G1_SSSP_PULL=0 disables the switch. Read at every run.
double sssp_pull_fraction() {
    const char* e = getenv("G1_SSSP_PULL");
    return e ? atof(e) : 0.8;
}

// One wave sums the kSegments counters of a frontier and posts nf and E to the graph's coherent pinned
// mailbox, then the sequence number with a system-scope release store, so the host learns a frontier's
// size by watching host memory instead of a device-to-host copy plus a stream synchronization. With
// mins (nmins per-workgroup minima of a distance scan) their minimum is posted too.
__global__ void __launch_bounds__(64) frontier_post_kernel(const unsigned long long* __restrict__ cnt,
                                                           unsigned long long* mbox, unsigned long long seq,
                                                           const uint32_t* __restrict__ mins, uint32_t nmins) {
    const unsigned long long c = cnt[lane_id()];
    const unsigned long long nf = wave_sum<unsigned long long>(c & 0xFFFFFFFFull);
    const unsigned long long E = wave_sum<unsigned long long>(c >> 32);
    uint32_t mn = kInf;
    for (uint32_t i = lane_id(); i < nmins; i += kWave) mn = mins[i] < mn ? mins[i] : mn;
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_down(mn, off, kWave);
        mn = o < mn ? o : mn;
    }
    if (lane_id() == 0) {
        __hip_atomic_store(mbox + 1, nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(mbox + 2, E, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(mbox + 3, static_cast<unsigned long long>(mn), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(mbox, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Returns (E << 32) | nf of the frontier in counter set `set`, once the work enqueued on s before it is
// done. The host spins on the mailbox, yielding its core between polls after the first 2048 (a dense
// iteration runs for hundreds of microseconds, and the query server polls from many threads at once);
// every 1024 polls it asks the stream, and a stream that has drained (or failed) without the post showing
// is an error.
g1_status read_frontier(g1_graph* g, uint32_t set, hipStream_t s, unsigned long long* packed,
                        const uint32_t* mins = nullptr, uint32_t nmins = 0, uint32_t* min_out = nullptr) {
    const unsigned long long seq = ++g->post_seq;
    frontier_post_kernel<<<1, 64, 0, s>>>(g->qcnt + set * kSegments, g->h_mbox, seq, mins, nmins);
    G1_LAUNCHED();
    uint32_t polls = 0;
    while (__atomic_load_n(g->h_mbox, __ATOMIC_ACQUIRE) != seq) {
        if ((++polls & 1023u) == 0) {
            const hipError_t q = hipStreamQuery(s);
            if (q != hipErrorNotReady) {
                if (__atomic_load_n(g->h_mbox, __ATOMIC_ACQUIRE) == seq) break;
                G1_HIP(q);
                G1_HIP(hipStreamSynchronize(s));
                G1_CHECK(__atomic_load_n(g->h_mbox, __ATOMIC_ACQUIRE) == seq, G1_ERR_INTERNAL,
                         "frontier post missing after the stream drained");
                break;
            }
        }
        if (polls < 2048) _mm_pause();
        else std::this_thread::yield();
    }
    const unsigned long long nf = __atomic_load_n(g->h_mbox + 1, __ATOMIC_ACQUIRE);
    const unsigned long long E = __atomic_load_n(g->h_mbox + 2, __ATOMIC_ACQUIRE);
    G1_CHECK(E < 0xFFFFFFFFull && nf <= 0xFFFFFFFFull, G1_ERR_INTERNAL, "frontier edge count overflow");
    *packed = (E << 32) | nf;
    if (min_out) *min_out = static_cast<uint32_t>(__atomic_load_n(g->h_mbox + 3, __ATOMIC_ACQUIRE));
    return G1_OK;
}

// Folds the per-workgroup stats slots into ctr[3] (activated) and ctr[7] (scanned) and reads them.
g1_status fold_slots(g1_graph* g, hipStream_t s, bool read) {
    sum_slots_kernel<<<1, 256, 0, s>>>(g->slots, kMaxPushBlocks, g->ctr + 3);
    G1_LAUNCHED();
    sum_slots_kernel<<<1, 256, 0, s>>>(g->slots + kMaxPushBlocks, kMaxPushBlocks, g->ctr + 7);
    G1_LAUNCHED();
    if (read) G1_TRY(read_ctr(g, s, 0, 8));
    return G1_OK;
}

// Launches to enqueue for a frontier of `edges` out-edges that is already shrinking: enough for it to
// fall to the single-workgroup kernel's size if each iteration keeps a quarter of the edges (a shrinking
// RMAT frontier keeps 10-25 %), at most 32. Launches past the frontier's end are no-ops but not free (a
// few microseconds each, and as long again for the host to enqueue).
uint32_t tail_batch(uint64_t edges) {
    uint32_t k = 1;
    while (edges > kSmallCap && k < 32) {
        edges /= 4;
        ++k;
    }
    return k;
}

// Device-driven iteration loop of the push engines. Every launch reads its input frontier's counters
// on the device and is a no-op once the frontier is empty, so the host enqueues iterations in batches
// and reads back only the output counters of the last launch of a batch: 2, 4, 8, ... launches while the
// frontier grows, tail_batch() once it shrinks. With profiling on, the batch is one launch so each
// launch's counts can be attributed to it.
//
// With pull_min_edges > 0 an iteration whose frontier has at least that many out-edges runs as one pull
// step instead (pull(it): scatter, SpMV, apply). The host then needs each frontier's edge count before
// its iteration, so until the frontier shrinks below a quarter of that size iterations go one at a time;
// single_min_edges (when not 0) sets that size instead, for a launch whose kind depends on its edge count.
// it0 is the first iteration; first_edges, when known, the edge count of its frontier.
template <typename Launch, typename AfterLaunch, typename Small, typename Pull>
g1_status drive_push(g1_graph* g, hipStream_t s, Launch&& launch, AfterLaunch&& after, Small&& small,
                     bool small_first, uint32_t* launched, Pull&& pull, uint64_t pull_min_edges, uint32_t it0 = 0,
                     uint64_t first_edges = ~0ull, uint64_t single_min_edges = 0, uint32_t batch0 = 2,
                     bool lazy_profile = false) {
    if (single_min_edges == 0) single_min_edges = pull_min_edges;
    // With profiling on, launches go one at a time so that `after` can attribute each one's counts, unless
    // the caller attributes them afterwards from the iteration history (lazy_profile): then the launches
    // keep their batches and their event times are those of the unprofiled run.
    const bool one_by_one = g->profile && !lazy_profile;
    uint32_t it = it0, batch = batch0;
    bool try_small = small_first;
    bool known = first_edges != ~0ull;  // next_edges holds the edge count of the frontier of iteration `it`
    bool shrinking = false;             // ... and it is below the count read before it
    uint64_t next_edges = known ? first_edges : 0;
    for (;;) {
        // A frontier whose out-edges fit the single-workgroup kernel runs there (see push_small_kernel)
        // until it grows past it or ends; profiling keeps one launch per iteration instead.
        if (try_small && !one_by_one) {
            uint32_t res[2];
            G1_TRY(small(it, res));
            it = res[0];
            if (res[1] == 0) break;
            batch = batch0;
            shrinking = false;
            // the frontier handed back is past the small kernel's size: read it so that this iteration can
            // already be a pull step or a dense launch
            unsigned long long back = 0;
            G1_TRY(read_frontier(g, it % 3, s, &back));
            next_edges = back >> 32;
            known = true;
        }
        unsigned long long next = 0;
        if (pull_min_edges && known && next_edges >= pull_min_edges) {
            G1_CHECK(it + 1 <= kHistIters, G1_ERR_INTERNAL, "push engine exceeded %u iterations", kHistIters);
            G1_TRY(pull(it));
            ++it;
        } else {
            uint32_t b = shrinking ? tail_batch(next_edges) : batch;
            if (one_by_one) b = 1;
            if (single_min_edges && (!known || !shrinking || next_edges >= single_min_edges / 4)) b = 1;
            G1_CHECK(it + b <= kHistIters, G1_ERR_INTERNAL, "push engine exceeded %u iterations", kHistIters);
            for (uint32_t k = 0; k < b; ++k) {
                G1_TRY(launch(it, k == 0 && known ? next_edges : 0));
                G1_TRY(after(it));
                ++it;
            }
            if (batch < 32) batch *= 2;
        }
        G1_TRY(read_frontier(g, it % 3, s, &next));
        if (static_cast<uint32_t>(next) == 0) break;
        shrinking = known && (next >> 32) < next_edges;
        next_edges = next >> 32;
        known = true;
        try_small = next_edges <= kSmallCap;
    }
    *launched = it;
    return G1_OK;
}

template <typename Launch, typename AfterLaunch, typename Small>
g1_status drive_push(g1_graph* g, hipStream_t s, Launch&& launch, AfterLaunch&& after, Small&& small,
                     bool small_first, uint32_t* launched) {
    auto no_pull = [](uint32_t) -> g1_status { return G1_OK; };
    return drive_push(g, s, launch, after, small, small_first, launched, no_pull, 0);
}

// Runs push_small_kernel from iteration it0 and reads its result: res[0] the next iteration, res[1] 1
// if the frontier was handed back to the segmented queue (0: the run is over).
template <typename P, typename RV = FullRows>
g1_status run_small(g1_graph* g, hipStream_t s, const P& pol, const uint32_t* wts, uint32_t* fb, uint32_t words,
                    uint32_t it0, uint32_t* res, RV rv = FullRows(nullptr), const uint32_t* col = nullptr) {
    uint32_t* dres = reinterpret_cast<uint32_t*>(g->ctr + 10);
    if constexpr (std::is_same<RV, FullRows>::value) rv = FullRows(g->row_off);
    if (!col) col = g->col;
    push_small_kernel<P, RV><<<1, kSmallThreads, 0, s>>>(rv, col, wts, pol, g->q[0], g->q[1], g->qoff[0],
                                                      g->qoff[1], g->qcnt, g->seg_cap, it0, kHistIters - 1, fb,
                                                      words, g->hist, g->slots, dres);
    G1_LAUNCHED();
    G1_HIP(hipMemcpyAsync(g->h_ctr + 10, dres, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    G1_HIP(hipStreamSynchronize(s));
    const uint32_t* h = reinterpret_cast<const uint32_t*>(g->h_ctr + 10);
    res[0] = h[0];
    res[1] = h[1];
    return G1_OK;
}

// The end of a push-engine run (BFS or SSSP) in two launches and one stream synchronization: the reached
// vertices of the result d are counted into the third and fourth stats-slot regions, then
// push_finish_kernel folds every slot region and writes the counters, the overflow flag and the
// iteration history to pinned memory. The run's end event must already be recorded.
g1_status collect_push_run(g1_graph* g, hipStream_t s, uint32_t launched, const uint32_t* d, g1_run_stats* st,
                           bool count_edges, bool skip_empty = false, bool stats_done = false) {
    if (!stats_done) {  // (done: a run on the hub-first copy counted them while mapping its result back)
        reached_stats_kernel<<<grid_for(g->n, 256, kMaxPushBlocks), 256, 0, s>>>(
            g->n, d, g->row_off, g->slots + 2 * kMaxPushBlocks, g->slots + 3 * kMaxPushBlocks);
        G1_LAUNCHED();
    }
    if (!g->h_hist) {
        G1_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->h_hist), 2ull * kHistIters * sizeof(unsigned long long)));
    }
    push_finish_kernel<<<1, 256, 0, s>>>(g->slots, g->ctr, g->overflow, g->hist, launched, g->h_ctr, g->h_hist);
    G1_LAUNCHED();
    G1_HIP(hipStreamSynchronize(s));
    G1_TRY(elapsed(g, st));
    G1_TRY(prof_flush(g));
    uint32_t flag = 0;
    memcpy(&flag, g->h_ctr + 8, sizeof(flag));
    if (flag != 0) {  // cleared, so the next run on this graph starts clean
        G1_TRY(g1_fill_u32(g->overflow, 0u, 1, reinterpret_cast<g1_stream_t>(s)));
        return set_error(G1_ERR_INTERNAL, "frontier queue segment overflow");
    }
    uint32_t iters = 0;
    for (uint32_t i = 0; i < launched; ++i) {
        const unsigned long long* h = g->h_hist + 2ull * i;
        if (static_cast<uint32_t>(h[0]) == 0) {
            if (skip_empty) continue;  // a no-op launch between the delta engine's phases
            break;
        }
        ++iters;
        const bool pull = (h[1] >> 63) != 0;
        if (pull) ++st->pull_iterations; else ++st->push_iterations;
        if (count_edges && !pull) st->edges_sum += h[0] >> 32;
    }
    st->iterations = iters;
    st->reached = g->h_ctr[4];
    st->reached_edges = g->h_ctr[5];
    return G1_OK;
}

g1_status run_bfs_road(g1_graph* g, uint32_t src, uint32_t* depth, g1_run_stats* st, hipStream_t s, bool* ok);
bool road_graph(const g1_graph* g);

// The delta-stepping SSSP and direction-optimizing BFS run on the hub-first copy of a graph (ensure_hub_copy)
// from 2^20 vertices up, or on any graph whose copy was prepared.
constexpr uint32_t kHubMinVertices = 1u << 20;
bool hub_wanted(const g1_graph* g) { return hub_eligible(g) && (g->hub != nullptr || g->n >= kHubMinVertices); }

// A run on the hub-first copy: its result (u32 depths, or the distance words) into out in the ids of the
// graph the caller holds, with the reached-vertex stats collect_push_run would count (it is then told
// they are done).
template <bool kWord>
g1_status map_back(g1_graph* g, const void* src, uint32_t* out, hipStream_t s) {
    const int tm = prof_begin(g, "hub_map_back", s);
    map_back_stats_kernel<kWord><<<grid_for((g->n + 1023) / 1024, 1, 1u << 20), 256, 0, s>>>(
        g->n, g->out_inv, src, out, g->out_row_off, g->slots + 2 * kMaxPushBlocks, g->slots + 3 * kMaxPushBlocks,
        kMaxPushBlocks);
    G1_LAUNCHED();
    // the map, the results gathered, the row offsets read, the results written
    prof_end(g, tm, s, (kWord ? 20.0 : 16.0) * g->n);
    return G1_OK;
}

g1_status run_bfs_push(g1_graph* g, uint32_t src, uint32_t* depth, bool diropt, g1_run_stats* st, hipStream_t s) {
    if (road_graph(g)) {
        bool ok = false;
        G1_TRY(run_bfs_road(g, src, depth, st, s, &ok));
        if (ok) return G1_OK;
        memset(st, 0, sizeof(*st));
    }
    if (diropt && g->has_t && hub_wanted(g)) {
        G1_TRY(ensure_hub_copy(g, s));
        g->hub->profile = g->profile;
        return run_bfs_push(g->hub, g->hub_inv_h[src], depth, true, st, s);
    }
    uint32_t* const out = depth;
    if (g->out_inv) depth = g->out_work;  // a run on the hub-first copy: its depths, mapped into out at the end
    G1_TRY(ensure_push_work(g, s));
    if (diropt) G1_CHECK(g->has_t, G1_ERR_INVALID_ARG, "g1_bfs: direction-optimizing BFS needs the transpose");
    const uint32_t n = g->n;
    const uint32_t words = (n + 31) / 32;
    push_start_kernel<false><<<grid_for(n, 256, kMaxPushBlocks), 256, 0, s>>>(
        n, src, nullptr, depth, diropt ? g->fb3 : nullptr, words, g->row_off, nullptr, g->q[0], g->qoff[0], g->qcnt,
        g->ctr, nullptr);
    G1_LAUNCHED();
    memset(g->h_ctr, 0, 8 * sizeof(unsigned long long));  // the host mirror of the cleared counters
    G1_HIP(hipEventRecord(g->ev[0], s));
    int tok = -1;
    auto launch = [&](uint32_t it, uint64_t) -> g1_status {
        PushQueues pq = make_queues(g, it);
        tok = prof_begin(g, diropt ? "bfs_step_diropt" : "bfs_advance", s);
        if (diropt) {
            bfs_step_kernel<true><<<bfs_step_blocks(), kPushThreads, 0, s>>>(
                n, g->m, g->row_off, g->col, g->t_row_off, g->t_col, depth, g->fb3, words, pq, it, g->hist);
        } else {
            bfs_step_kernel<false><<<bfs_step_blocks(), kPushThreads, 0, s>>>(
                n, g->m, g->row_off, g->col, g->t_row_off, g->t_col, depth, g->fb3, words, pq, it, g->hist);
        }
        G1_LAUNCHED();
        prof_end(g, tok, s, 0.0);
        return G1_OK;
    };
    auto after = [&](uint32_t it) -> g1_status {
        if (!g->profile) return G1_OK;
        // attribute this launch's algorithmic bytes from its history record and stats slots
        const uint64_t act0 = g->h_ctr[3], scan0 = g->h_ctr[7];
        G1_TRY(fold_slots(g, s, true));
        unsigned long long h[2];
        G1_HIP(hipMemcpy(h, g->hist + 2 * it, sizeof(h), hipMemcpyDeviceToHost));
        const double nf = static_cast<double>(static_cast<uint32_t>(h[0]));
        const double E = static_cast<double>(h[0] >> 32);
        const double act = static_cast<double>(g->h_ctr[3] - act0);
        const double scanned = static_cast<double>(g->h_ctr[7] - scan0);
        if (nf == 0.0) prof_set_bytes(g, tok, 0.0);
        else if (h[1] >> 63)  // bottom-up: depth and in-edge offsets of every vertex, scanned in-edges,
                              // the next bitmap, the new depths
            prof_set_bytes(g, tok, 4.0 * n + 4.0 * (n + 1.0) + 4.0 * scanned + n / 8.0 + 4.0 * act);
        else
            prof_set_bytes(g, tok, 16.0 * nf + 8.0 * E + 4.0 * act);
        return G1_OK;
    };
    auto small = [&](uint32_t it0, uint32_t* res) -> g1_status {
        BfsPush pol{depth, nullptr, 0};
        return run_small(g, s, pol, nullptr, diropt ? g->fb3 : nullptr, words, it0, res);
    };
    uint32_t launched = 0;
    G1_TRY(drive_push(g, s, launch, after, small, g->max_out_deg <= kSmallCap, &launched));
    if (g->out_inv) G1_TRY(map_back<false>(g, depth, out, s));
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(collect_push_run(g, s, launched, out, st, true, false, g->out_inv != nullptr));
    st->frontier_sum = 1 + g->h_ctr[3];
    st->edges_sum += g->h_ctr[7];
    return G1_OK;
}

__global__ void dstamp_init_kernel(uint32_t n, uint32_t src, unsigned long long* __restrict__ ds) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        ds[v] = (v == src) ? 0ull : (static_cast<unsigned long long>(kInf) << 32);
}

__global__ void dstamp_extract_kernel(uint32_t n, const unsigned long long* __restrict__ ds,
                                      uint32_t* __restrict__ dist) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        dist[v] = static_cast<uint32_t>(ds[v] >> 32);
}



// ---------------------------------------------------------------- road-like graphs (one workgroup)

// Low-degree graphs (at most kRoadMaxDeg out-edges per vertex: grids, road networks) take the
// single-workgroup engine of g1/road.hpp; it reports an overflow of its far list (never expected) and the
// query then runs on the push engine.
constexpr uint32_t kRoadMaxDeg = 16;

bool road_graph(const g1_graph* g) { return g->n > 0 && g->max_out_deg <= kRoadMaxDeg; }

// near-far bucket width: sixteen heaviest edges (a wide bucket: few splits, few extra relaxations)
// Bucket width of the road engine's near-far split, a multiple of the largest weight: 16 for the LDS kernel
// (a wider bucket saves iterations, which cost ~3 us each there), 8 for the global-memory kernel (whose
// extra relaxations cost more). Measured from vertex 1 on p = 0.6 grids: 213 x 213, weights [1, 64], LDS
// kernel 1.60 ms at 16 (1.67 at 8, 1.63 at 32); 2048 x 2048, weights [1, 255], global kernel 68.3 ms at 8
// (76.3 at 16, 68.1 at 4). G1_ROAD_DELTA sets one multiple for both.
uint32_t road_delta(const g1_graph* g, bool lds) {
    static const int forced = [] {
        const char* e = getenv("G1_ROAD_DELTA");
        return e ? atoi(e) : 0;
    }();
    const uint32_t mult = forced > 0 ? static_cast<uint32_t>(forced) : (lds ? 16u : 8u);
    return std::max<uint32_t>(mult * std::max<uint32_t>(g->max_w, 1u), 16u);
}

constexpr uint32_t kRoadSlots = 64;  // queries per batched launch

uint64_t road_work_bytes(const g1_graph* g) {
    if (g->road_q) return 0;
    const uint64_t cap = 2ull * g->n + 1024;
    return 2 * cap * sizeof(uint32_t) + kRoadSlots * sizeof(RoadQuery) + kRoadSlots * sizeof(uint32_t*);
}

g1_status ensure_road_work(g1_graph* g, hipStream_t s) {
    G1_TRY(ensure_push_work(g, s));
    if (g->road_q) return G1_OK;
    const uint64_t cap = 2ull * g->n + 1024;
    G1_CHECK(cap < 0xFFFFFFFFull, G1_ERR_INVALID_ARG, "graph too large for the road engine");
    g->road_far_cap = static_cast<uint32_t>(cap);
    for (int i = 0; i < 2; ++i) {
        void* p = nullptr;
        G1_TRY(graph_alloc(g, &p, cap * sizeof(uint32_t)));
        g->road_far[i] = static_cast<uint32_t*>(p);
    }
    G1_TRY(graph_alloc(g, &g->road_q, kRoadSlots * sizeof(RoadQuery)));
    void* p = nullptr;
    G1_TRY(graph_alloc(g, &p, kRoadSlots * sizeof(uint32_t*)));
    g->road_outs = static_cast<uint32_t**>(p);
    return G1_OK;
}

// Runs one query on the road engine; *ok = false when it must be run elsewhere.
template <bool kSssp>
g1_status run_road(g1_graph* g, uint32_t src, uint32_t* depth, g1_run_stats* st, hipStream_t s, bool* ok) {
    G1_TRY(ensure_road_work(g, s));
    RoadQuery* hq = reinterpret_cast<RoadQuery*>(g->h_ctr + 128);  // pinned staging
    memset(hq, 0, sizeof(RoadQuery));
    G1_CHECK(kSssp || depth != nullptr, G1_ERR_INTERNAL, "road engine: BFS without a depth array");
    hq->sssp = kSssp ? 1u : 0u;
    hq->dstamp = g->dstamp;
    hq->depth = depth;
    hq->q[0] = g->q[0];
    hq->q[1] = g->q[1];
    hq->far[0] = g->road_far[0];
    hq->far[1] = g->road_far[1];
    hq->src = src;
    const RoadQuery staged = *hq;
    const uint64_t qcap = std::min<uint64_t>(static_cast<uint64_t>(g->seg_cap) * kSegments, 0xFFFFFFFFull);
    // A small graph's SSSP runs with its distances in LDS (road_lds_sssp_kernel); if that reports status 2
    // (a distance past 16 bits, or a far-list overflow) it leaves the distance words untouched and the
    // query runs on the global-memory kernel.
    int tok = -1;
    for (int pass = (kSssp && g->n <= kRoadLdsMaxN) ? 0 : 1; pass < 2; ++pass) {
        *hq = staged;
        G1_HIP(hipMemcpyAsync(g->road_q, hq, sizeof(RoadQuery), hipMemcpyHostToDevice, s));
        tok = prof_begin(g, kSssp ? "road_sssp" : "road_bfs", s);
        RoadQuery* dq = static_cast<RoadQuery*>(g->road_q);
        if (pass == 0) {
            road_lds_sssp_kernel<<<1, kRoadThreads, 0, s>>>(g->n, g->row_off, g->col, g->w, dq, road_delta(g, true),
                                                            g->road_far_cap, static_cast<uint32_t>(qcap), kHistIters);
        } else {
            road_query_kernel<<<1, kRoadThreads, 0, s>>>(g->row_off, g->col, g->w, dq, road_delta(g, false), g->road_far_cap,
                                                         static_cast<uint32_t>(qcap), kHistIters);
        }
        G1_LAUNCHED();
        prof_end(g, tok, s, 0.0);
        G1_HIP(hipMemcpyAsync(hq, g->road_q, sizeof(RoadQuery), hipMemcpyDeviceToHost, s));
        G1_HIP(hipStreamSynchronize(s));
        if (hq->status == 0) break;
    }
    *ok = hq->status == 0;
    if (!*ok) return G1_OK;
    prof_set_bytes(g, tok, 16.0 * static_cast<double>(hq->frontier_sum) + 12.0 * static_cast<double>(hq->edges_sum));
    st->iterations = hq->iterations;
    st->push_iterations = hq->iterations;
    st->frontier_sum = hq->frontier_sum;
    st->edges_sum = hq->edges_sum;
    return G1_OK;
}

// Batched road-graph runs: query i (kinds[i]: 1 SSSP, 0 BFS) from sources[i] into outs[i] (n u32), one
// workgroup per query, up to kRoadSlots per launch. Slot state (SSSP distance words, near buffers, far
// lists) is allocated on first use and kept with the graph. ok[i] = 0 asks the caller to rerun query i
// on the push engine.
__global__ void road_init_kernel(uint32_t n, RoadQuery* __restrict__ qs, uint32_t count) {
    const uint32_t i = blockIdx.y;
    if (i >= count) return;
    const RoadQuery& Q = qs[i];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        if (Q.sssp) Q.dstamp[v] = (v == Q.src) ? 0ull : (static_cast<unsigned long long>(kInf) << 32);
        else Q.depth[v] = (v == Q.src) ? 0u : kInf;
    }
}

__global__ void road_extract_kernel(uint32_t n, const RoadQuery* __restrict__ qs, uint32_t* const* __restrict__ outs,
                                    uint32_t count) {
    const uint32_t i = blockIdx.y;
    if (i >= count || !qs[i].sssp) return;
    const unsigned long long* ds = qs[i].dstamp;
    uint32_t* out = outs[i];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) out[v] = static_cast<uint32_t>(ds[v] >> 32);
}

g1_status road_batch_run(g1_graph* g, const uint32_t* kinds, const uint32_t* sources, uint32_t count, uint32_t** outs,
                         uint8_t* ok, uint32_t* iters, hipStream_t s) {
    G1_TRY(ensure_road_work(g, s));
    const size_t n = g->n;
    const size_t per = n * sizeof(unsigned long long) + 2 * n * sizeof(uint32_t) +
                       2ull * g->road_far_cap * sizeof(uint32_t);
    while (g->road_slots.size() < std::min<uint32_t>(count, kRoadSlots)) {
        void* p = nullptr;
        G1_TRY(graph_alloc(g, &p, per));
        g->road_slots.push_back(p);
    }
    const uint64_t qcap = n;
    for (uint32_t b0 = 0; b0 < count; b0 += kRoadSlots) {
        const uint32_t c = std::min<uint32_t>(kRoadSlots, count - b0);
        std::vector<RoadQuery> hq(c);
        std::vector<uint32_t*> hout(c);
        for (uint32_t i = 0; i < c; ++i) {
            uint8_t* base = static_cast<uint8_t*>(g->road_slots[i]);
            RoadQuery& Q = hq[i];
            memset(&Q, 0, sizeof(Q));
            Q.sssp = kinds[b0 + i] ? 1u : 0u;
            Q.dstamp = reinterpret_cast<unsigned long long*>(base);
            uint32_t* u32 = reinterpret_cast<uint32_t*>(base + n * sizeof(unsigned long long));
            Q.q[0] = u32;
            Q.q[1] = u32 + n;
            Q.far[0] = u32 + 2 * n;
            Q.far[1] = u32 + 2 * n + g->road_far_cap;
            Q.depth = outs[b0 + i];
            Q.src = sources[b0 + i];
            hout[i] = outs[b0 + i];
        }
        RoadQuery* dq = static_cast<RoadQuery*>(g->road_q);
        G1_HIP(hipMemcpyAsync(dq, hq.data(), c * sizeof(RoadQuery), hipMemcpyHostToDevice, s));
        uint32_t** douts = g->road_outs;
        G1_HIP(hipMemcpyAsync(douts, hout.data(), c * sizeof(uint32_t*), hipMemcpyHostToDevice, s));
        road_init_kernel<<<dim3(std::min<uint32_t>(grid_for(g->n, 256, 1024), 1024), c), 256, 0, s>>>(g->n, dq, c);
        G1_LAUNCHED();
        road_query_kernel<<<c, kRoadThreads, 0, s>>>(g->row_off, g->col, g->w, dq, road_delta(g, false), g->road_far_cap,
                                                     static_cast<uint32_t>(qcap), kHistIters);
        G1_LAUNCHED();
        road_extract_kernel<<<dim3(std::min<uint32_t>(grid_for(g->n, 256, 1024), 1024), c), 256, 0, s>>>(g->n, dq, douts, c);
        G1_LAUNCHED();
        G1_HIP(hipMemcpyAsync(hq.data(), dq, c * sizeof(RoadQuery), hipMemcpyDeviceToHost, s));
        G1_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < c; ++i) {
            ok[b0 + i] = hq[i].status == 0 ? 1 : 0;
            if (iters) iters[b0 + i] = hq[i].iterations;
        }
    }
    return G1_OK;
}

g1_status run_sssp_push(g1_graph* g, uint32_t src, uint32_t* dist, g1_run_stats* st, hipStream_t s,
                        bool allow_road = true);

g1_status run_sssp_road(g1_graph* g, uint32_t src, uint32_t* dist, g1_run_stats* st, hipStream_t s) {
    G1_TRY(ensure_road_work(g, s));
    const uint32_t n = g->n;
    dstamp_init_kernel<<<grid_for(n, 256, kMaxPushBlocks), 256, 0, s>>>(n, src, g->dstamp);
    G1_LAUNCHED();
    G1_HIP(hipEventRecord(g->ev[0], s));
    bool ok = false;
    G1_TRY(run_road<true>(g, src, nullptr, st, s, &ok));
    if (!ok) return run_sssp_push(g, src, dist, st, s, false);
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(elapsed(g, st));
    G1_TRY(prof_flush(g));
    dstamp_extract_kernel<<<grid_for(n, 256, kMaxPushBlocks), 256, 0, s>>>(n, g->dstamp, dist);
    G1_LAUNCHED();
    return finish_stats(g, dist, st, s);
}

g1_status run_bfs_road(g1_graph* g, uint32_t src, uint32_t* depth, g1_run_stats* st, hipStream_t s, bool* ok) {
    G1_TRY(ensure_road_work(g, s));
    G1_TRY(g1_fill_u32(depth, kInf, g->n, s));
    set_u32_kernel<<<1, 64, 0, s>>>(depth, src, 0u);
    G1_LAUNCHED();
    G1_HIP(hipEventRecord(g->ev[0], s));
    G1_TRY(run_road<false>(g, src, depth, st, s, ok));
    if (!*ok) return G1_OK;
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(elapsed(g, st));
    G1_TRY(prof_flush(g));
    return finish_stats(g, depth, st, s);
}

// The push iterations of the SSSP advance engine from iteration it0, whose input frontier is in the
// segmented queue (counter set it0 % 3; with a pull step possible at it0, its vertices carry stamp it0):
// returning and no-return advance launches, the single-workgroup kernel for small frontiers, and pull
// steps for dense ones, driven by drive_push. *launched gets the next iteration index; no-return and pull
// iterations are counted into *dense and *pulls.
g1_status sssp_push_iterations(g1_graph* g, hipStream_t s, const DistBound& db, uint32_t it0, bool small_first,
                               uint32_t* launched, uint32_t* dense, uint32_t* pulls, uint64_t first_edges = ~0ull) {
    const uint32_t n = g->n;
    int tok = -1;
    // A frontier known to hold at least G1_SSSP_NORETURN x m out-edges (0.3 by default) relaxes without
    // the winner test (SsspPushNoReturn) and its winners are collected from the stamps afterwards.
    const uint64_t noreturn_min =
        sssp_noreturn_fraction() > 0.0 ? std::max<uint64_t>(static_cast<uint64_t>(sssp_noreturn_fraction() * g->m), 1)
                                       : ~0ull;
    int tok_c = -1;  // the collect launch of a no-return iteration (profiling)
    auto launch = [&](uint32_t it, uint64_t edges) -> g1_status {
        PushQueues pq = make_queues(g, it);
        tok_c = -1;
        if (edges >= noreturn_min) {
            ++*dense;
            if (sssp_sweep_enabled()) {
                // the frontier's bitmap (g->fb[0] holds ceil(n / 64) 64-bit words), then the sweep
                unsigned long long* fbits = reinterpret_cast<unsigned long long*>(g->fb[0]);
                tok = prof_begin(g, "sssp_sweep", s);
                sweep_frontier_bits_kernel<<<grid_for((n + 63) / 64 * 64ull, 256, kMaxPushBlocks), 256, 0, s>>>(
                    n, g->dstamp, it, fbits);
                if (sssp_sweep_mode() == 2)
                    sssp_sweep_kernel<true><<<sssp_sweep_blocks(), 256, 0, s>>>(
                        n, g->row_off, g->col, g->w, g->dstamp, it, fbits,
                        static_cast<const SweepPiece*>(g->sweep_pieces), g->n_sweep_pieces, pq, g->hist);
                else
                    sssp_sweep_kernel<false><<<sssp_sweep_blocks(), 256, 0, s>>>(
                        n, g->row_off, g->col, g->w, g->dstamp, it, fbits,
                        static_cast<const SweepPiece*>(g->sweep_pieces), g->n_sweep_pieces, pq, g->hist);
            } else {
                SsspPushNoReturn pol{g->dstamp, it + 1, db};
                tok = prof_begin(g, "sssp_advance_nr", s);
                push_advance_kernel<SsspPushNoReturn><<<sssp_noreturn_blocks(), kPushThreads, 0, s>>>(
                    FullRows(g->row_off), g->col, g->w, pq, pol, it, g->hist);
            }
            G1_LAUNCHED();
            prof_end(g, tok, s, 0.0);
            tok_c = prof_begin(g, "sssp_collect", s);
            sssp_collect_kernel<<<kMaxPushBlocks, kPushThreads, 0, s>>>(n, FullRows(g->row_off), g->dstamp, it + 1, pq);
            G1_LAUNCHED();
            prof_end(g, tok_c, s, 0.0);
        } else {
            SsspPush pol{g->dstamp, it + 1, db};
            tok = prof_begin(g, "sssp_advance", s);
            push_advance_kernel<SsspPush><<<sssp_push_blocks(), kPushThreads, 0, s>>>(FullRows(g->row_off), g->col, g->w,
                                                                                     pq, pol, it, g->hist);
            G1_LAUNCHED();
            prof_end(g, tok, s, 0.0);
        }
        return G1_OK;
    };
    auto after = [&](uint32_t it) -> g1_status {
        if (!g->profile) return G1_OK;
        const uint64_t act0 = g->h_ctr[3];
        G1_TRY(fold_slots(g, s, true));
        unsigned long long h[1];
        G1_HIP(hipMemcpy(h, g->hist + 2 * it, sizeof(h), hipMemcpyDeviceToHost));
        const double nf = static_cast<double>(static_cast<uint32_t>(h[0]));
        const double E = static_cast<double>(h[0] >> 32);
        const double act = static_cast<double>(g->h_ctr[3] - act0);
        // advance: frontier entries (id, offset, row start, source distance) + column, weight and target word
        // per edge; its winners' entries (4 B each) unless a collect launch enqueued them, which reads every
        // distance word and the winners' degrees and writes their entries
        if (tok_c >= 0) {
            prof_set_bytes(g, tok, nf == 0.0 ? 0.0 : 16.0 * nf + 12.0 * E);
            prof_set_bytes(g, tok_c, 8.0 * n + 16.0 * act);
        } else {
            prof_set_bytes(g, tok, nf == 0.0 ? 0.0 : 16.0 * nf + 12.0 * E + 4.0 * act);
        }
        return G1_OK;
    };
    auto small = [&](uint32_t it0, uint32_t* res) -> g1_status {
        SsspPush pol{g->dstamp, 0, db};
        return run_small(g, s, pol, g->w, nullptr, 0, it0, res);
    };
    // A dense frontier (G1_SSSP_PULL x m edges, 0.8 by default) runs as a pull step (sssp_pull_x_kernel):
    // one streaming pass over the in-edge CSR costs less than that many random relaxations.
    const bool can_pull = g->has_t && g->t_w != nullptr && sssp_pull_fraction() > 0.0;
    if (can_pull) G1_TRY(ensure_pull_work(g, false));
    auto pull = [&](uint32_t it) -> g1_status {
        PushQueues pq = make_queues(g, it);
        tok = prof_begin(g, "sssp_pull", s);
        sssp_pull_x_kernel<<<grid_for(n, kPushThreads, kMaxPushBlocks), kPushThreads, 0, s>>>(n, g->dstamp, it, g->xu,
                                                                                           pq, g->hist);
        G1_LAUNCHED();
        G1_TRY(launch_spmv_merge<MinPlusU32>(n, g->m, g->t_row_off, g->t_col, g->t_w, g->xu, g->yu, g->t_ws, s));
        sssp_pull_apply_kernel<<<kMaxPushBlocks, kPushThreads, 0, s>>>(n, g->row_off, g->yu, g->dstamp, it + 1, pq, db);
        G1_LAUNCHED();
        // x 12n + SpMV 4(n+1) + 8m + 4n (x) + 4n (y) + apply 4n (y) + 16n (dist word) + 8n (degrees)
        prof_end(g, tok, s, 4.0 * (n + 1.0) + 8.0 * g->m + 44.0 * n);
        ++*pulls;
        return G1_OK;
    };
    return drive_push(g, s, launch, after, small, small_first, launched, pull,
                      can_pull ? std::max<uint64_t>(static_cast<uint64_t>(sssp_pull_fraction() * g->m), 1) : 0, it0,
                      first_edges);
}

g1_status run_sssp_push(g1_graph* g, uint32_t src, uint32_t* dist, g1_run_stats* st, hipStream_t s, bool allow_road) {
    if (allow_road && road_graph(g) && g->w) return run_sssp_road(g, src, dist, st, s);
    G1_TRY(ensure_push_work(g, s));
    G1_CHECK(g->w != nullptr, G1_ERR_INVALID_ARG, "g1_sssp: graph has no weights");
    if (sssp_sweep_enabled()) G1_TRY(ensure_sweep_pieces(g, s));
    const uint32_t n = g->n;
    const uint32_t fill_blocks = grid_for(n, 256, kMaxPushBlocks);
    const DistBound db = sssp_dist_bound(g);
    push_start_kernel<true><<<fill_blocks, 256, 0, s>>>(n, src, g->dstamp, nullptr, nullptr, 0, g->row_off, nullptr,
                                                        g->q[0], g->qoff[0], g->qcnt, g->ctr, db.b);
    G1_LAUNCHED();
    memset(g->h_ctr, 0, 8 * sizeof(unsigned long long));  // the host mirror of the cleared counters
    G1_HIP(hipEventRecord(g->ev[0], s));
    uint32_t launched = 0, dense = 0, pulls = 0;
    G1_TRY(sssp_push_iterations(g, s, db, 0, g->max_out_deg <= kSmallCap, &launched, &dense, &pulls));
    dstamp_extract_kernel<<<fill_blocks, 256, 0, s>>>(n, g->dstamp, dist);
    G1_LAUNCHED();
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(collect_push_run(g, s, launched, dist, st, true));
    st->pull_iterations += pulls;
    st->push_iterations -= std::min(st->push_iterations, pulls);
    st->dense_iterations = dense;
    st->frontier_sum = 1 + g->h_ctr[3];
    return G1_OK;
}

// ---------------------------------------------------------------- delta-stepping SSSP

// Bucket width of the delta-stepping engine: the largest weight / 16, rounded up (at least 1), or
// G1_SSSP_DELTA. RMAT22 (weights 1..255) from the hub, 10 runs each: width 4 1.57 ms (104 M relaxations),
// 6 1.34, 8 1.34, 12 1.33, 16 1.32 (68 M), 24 1.39, 32 1.47, 48 1.62, 64 1.74-1.80 ms (85 M); the advance
// engine 2.18 ms (170 M). A simulation of the schedule from five sources gave 74-81 M relaxations at 32
// against 208-327 M for the plain frontier loop.
uint32_t sssp_delta_width(const g1_graph* g) {
    const char* e = getenv("G1_SSSP_DELTA");
    const long v = e ? atol(e) : 0;
    if (v > 0 && v <= 0x7FFFFFFFL) return static_cast<uint32_t>(v);
    return std::max<uint32_t>(1u, (g->max_w + 15u) / 16u);
}

// The delta engine leaves the buckets for plain push iterations once a heavy phase has relaxed at least
// G1_SSSP_DELTA_SWITCH (default 0.01) times the graph's edges, or after kDeltaMaxBuckets buckets. Both
// variables are read at every run.
double sssp_delta_switch() {
    const char* e = getenv("G1_SSSP_DELTA_SWITCH");
    const double v = e ? atof(e) : 0.01;
    return v >= 0.0 ? v : 0.01;
}
constexpr uint32_t kDeltaMaxBuckets = 64;

// A light-phase iteration of at least G1_SSSP_DELTA_NORETURN (default 0.3) times the graph's edges relaxes
// without the winner test and collects its winners afterwards (SsspPushNoReturn + sssp_collect_kernel).
// Light frontiers are smaller than the advance engine's dense ones; lower thresholds measured slower
// (RMAT22, width 16: 1.40 ms at 0.005, 1.37 at 0.02, 1.33 at 0.3, with the host reading every frontier).
// Light iterations run in device-driven batches (the kind of each launch after the first of a batch
// is then decided without its edge count): 1.27 ms against 1.28 one at a time.
double sssp_delta_noreturn() {
    const char* e = getenv("G1_SSSP_DELTA_NORETURN");
    const double v = e ? atof(e) : 0.3;
    return v >= 0.0 ? v : 0.3;
}

// The heavy pull's form: over the rows whose distance is at or past the bucket's end only
// (sssp_pull_rows_kernel, G1_SSSP_DELTA_PULL_EDGES = 1, 4, 8 or 16 edges per lane per round, default 1:
// each round's column and weight loads are then fully coalesced; on RMAT22 from the hub the kernel took
// 102 / 113 / 125 us at 1 / 4 / 8), or with G1_SSSP_DELTA_PULL_ROWS=0 the merge-path SpMV over every row.
bool sssp_delta_pull_rows() {
    const char* e = getenv("G1_SSSP_DELTA_PULL_ROWS");
    return !(e && atoi(e) == 0);
}
uint32_t sssp_delta_pull_edges() {
    const char* e = getenv("G1_SSSP_DELTA_PULL_EDGES");
    const int v = e ? atoi(e) : 0;
    return (v == 1 || v == 4 || v == 8 || v == 16) ? static_cast<uint32_t>(v) : 1u;
}

// The delta engine's heavy pull over the rows at or past the bucket's end (sssp_pull_rows_kernel): x in
// dl_xb, y in yu; rows longer than huge_min are the long-row list's first t_nhuge entries.
g1_status launch_pull_rows(g1_graph* g, uint32_t edges, hipStream_t s) {
    const uint32_t blocks = g->t_nhuge + (g->n + kPullRowsPerBlock - 1) / kPullRowsPerBlock;
#define G1_PULL_ROWS_CASE(E_)                                                                                     \
    case E_:                                                                                                     \
        sssp_pull_rows_kernel<E_><<<blocks, kPushThreads, 0, s>>>(g->n, g->t_row_off, g->t_col, g->t_w, g->dl_xb, \
                                                                 g->yu, g->huge_min, g->t_long, g->t_nhuge);     \
        break
    switch (edges) {
        G1_PULL_ROWS_CASE(4);
        G1_PULL_ROWS_CASE(8);
        G1_PULL_ROWS_CASE(16);
        default: G1_PULL_ROWS_CASE(1);
    }
#undef G1_PULL_ROWS_CASE
    G1_LAUNCHED();
    return G1_OK;
}

// Zeroes queue counter sets a and b (kSegments counters each).
__global__ void __launch_bounds__(128) clear_sets_kernel(unsigned long long* qcnt, uint32_t a, uint32_t b) {
    const uint32_t t = threadIdx.x;
    if (t < kSegments) qcnt[a * kSegments + t] = 0ull;
    else qcnt[b * kSegments + (t - kSegments)] = 0ull;
}

// SSSP by delta-stepping over the push engine's kernels (Meyer and Sanders' light/heavy buckets, each
// phase an advance + filter over a frontier queue). Bucket b holds the distances in [lo, lo + delta).
// Light phase: push iterations over the light parts of the rows (ensure_delta_split), the bucket's limit
// keeping improvements past it out of the frontier (they carry stamp 0); when the frontier empties, every
// distance in the bucket is final. Heavy phase: one relaxation of the heavy parts of the bucket's rows (a
// pull step when they hold enough edges). Then the next non-empty bucket, both frontiers built by a scan
// of the distances (dist_range_frontier_kernel), which also finds the next bucket. Once a heavy phase
// relaxes G1_SSSP_DELTA_SWITCH x m edges (the frontier has reached the graph's bulk), the remaining
// vertices with finite distances form one frontier for the plain push iterations (sssp_push_iterations),
// which finish the run. The distances are those of any other engine: each bucket's vertices relax their
// light edges until nothing in the bucket improves, so they are final (a shortest path's last edge into
// the bucket from inside it is light) before their heavy edges are relaxed once.
g1_status run_sssp_delta(g1_graph* g, uint32_t src, uint32_t* dist, g1_run_stats* st, hipStream_t s) {
    if (road_graph(g) && g->w) return run_sssp_road(g, src, dist, st, s);
    if (g->w && hub_wanted(g)) {
        G1_TRY(ensure_hub_copy(g, s));
        g->hub->profile = g->profile;
        return run_sssp_delta(g->hub, g->hub_inv_h[src], dist, st, s);
    }
    G1_TRY(ensure_push_work(g, s));
    G1_CHECK(g->w != nullptr, G1_ERR_INVALID_ARG, "g1_sssp: graph has no weights");
    const uint32_t delta = sssp_delta_width(g);
    G1_TRY(ensure_delta_split(g, delta, s));
    if (sssp_sweep_enabled()) G1_TRY(ensure_sweep_pieces(g, s));
    const uint32_t n = g->n;
    const uint32_t fill_blocks = grid_for(n, 256, kMaxPushBlocks);
    const DistBound db = sssp_dist_bound(g);
    const LightRows lrows{g->row_off, g->dl_lc};
    const HeavyRows hrows{g->row_off, g->dl_lc};
    const bool can_pull = g->has_t && g->t_w != nullptr && sssp_pull_fraction() > 0.0;
    const bool pull_rows = sssp_delta_pull_rows();
    const uint32_t pull_edges = sssp_delta_pull_edges();
    if (can_pull) G1_TRY(ensure_pull_work(g, false));
    if (can_pull && pull_rows) G1_TRY(ensure_long_rows(g, false, s));
    const uint64_t pull_min =
        can_pull ? std::max<uint64_t>(static_cast<uint64_t>(sssp_pull_fraction() * g->m), 1) : ~0ull;
    const uint64_t noreturn_min =
        sssp_delta_noreturn() > 0.0 ? std::max<uint64_t>(static_cast<uint64_t>(sssp_delta_noreturn() * g->m), 1)
                                    : ~0ull;
    const uint64_t switch_edges = static_cast<uint64_t>(sssp_delta_switch() * g->m);

    push_start_kernel<true><<<fill_blocks, 256, 0, s>>>(n, src, g->dstamp, nullptr, nullptr, 0, g->row_off, g->dl_lc,
                                                        g->q[0], g->qoff[0], g->qcnt, g->ctr, db.b);
    G1_LAUNCHED();
    memset(g->h_ctr, 0, 8 * sizeof(unsigned long long));  // the host mirror of the cleared counters
    G1_HIP(hipEventRecord(g->ev[0], s));

    uint32_t it = 0, dense = 0, pulls = 0, buckets = 0;
    uint32_t limit = 0;
    int tok = -1, tok_c = -1;
    auto light_launch = [&](uint32_t i, uint64_t edges) -> g1_status {
        PushQueues pq = make_queues(g, i);
        tok_c = -1;
        if (edges >= noreturn_min) {
            ++dense;
            SsspPushNoReturn pol{g->dstamp, i + 1, db, limit};
            tok = prof_begin(g, "sssp_delta_light_nr", s);
            push_advance_kernel<SsspPushNoReturn, LightRows><<<sssp_noreturn_blocks(), kPushThreads, 0, s>>>(
                lrows, g->dl_col, g->dl_w, pq, pol, i, g->hist);
            G1_LAUNCHED();
            prof_end(g, tok, s, 0.0);
            tok_c = prof_begin(g, "sssp_delta_collect", s);
            sssp_collect_kernel<LightRows><<<kMaxPushBlocks, kPushThreads, 0, s>>>(n, lrows, g->dstamp, i + 1, pq);
            G1_LAUNCHED();
            prof_end(g, tok_c, s, 0.0);
        } else {
            SsspPushT<true> pol{g->dstamp, i + 1, db, limit};
            tok = prof_begin(g, "sssp_delta_light", s);
            push_advance_kernel<SsspPushT<true>, LightRows><<<sssp_push_blocks(), kPushThreads, 0, s>>>(
                lrows, g->dl_col, g->dl_w, pq, pol, i, g->hist);
            G1_LAUNCHED();
            prof_end(g, tok, s, 0.0);
        }
        return G1_OK;
    };
    // profiling: the launch's algorithmic bytes, counted as for the advance engine's launches
    // Profiling: each light launch's algorithmic bytes (as counted for the advance engine's launches) are
    // set after its phase, from the iteration history: its input (nf, E) and its output queue (the next
    // iteration's nf); no synchronization between the launches.
    struct LightTok {
        int tok, tok_c;
        uint32_t it;
    };
    std::vector<LightTok> light_toks;
    auto light_after = [&](uint32_t i) -> g1_status {
        if (g->profile) light_toks.push_back(LightTok{tok, tok_c, i});
        return G1_OK;
    };
    auto light_bytes = [&](uint32_t it_end) -> g1_status {
        if (light_toks.empty()) return G1_OK;
        const uint32_t i0 = light_toks.front().it;
        std::vector<unsigned long long> h(2ull * (it_end - i0 + 1), 0ull);
        G1_HIP(hipMemcpy(h.data(), g->hist + 2ull * i0, sizeof(unsigned long long) * 2ull * (it_end - i0),
                         hipMemcpyDeviceToHost));
        for (const LightTok& t : light_toks) {
            const unsigned long long hi = h[2ull * (t.it - i0)];
            const double nf = static_cast<double>(static_cast<uint32_t>(hi));
            const double E = static_cast<double>(hi >> 32);
            const double out = t.it + 1 < it_end ? static_cast<double>(static_cast<uint32_t>(h[2ull * (t.it + 1 - i0)]))
                                                 : 0.0;
            if (t.tok_c >= 0) {
                prof_set_bytes(g, t.tok, nf == 0.0 ? 0.0 : 16.0 * nf + 12.0 * E);
                prof_set_bytes(g, t.tok_c, 8.0 * n + 16.0 * out);
            } else {
                prof_set_bytes(g, t.tok, nf == 0.0 ? 0.0 : 16.0 * nf + 12.0 * E + 4.0 * out);
            }
        }
        light_toks.clear();
        return G1_OK;
    };
    auto light_small = [&](uint32_t it0, uint32_t* res) -> g1_status {
        SsspPush pol{g->dstamp, 0, db, limit};
        tok = prof_begin(g, "sssp_delta_light_small", s);
        const g1_status st_small = run_small(g, s, pol, g->dl_w, nullptr, 0, it0, res, lrows, g->dl_col);
        prof_end(g, tok, s, 0.0);
        return st_small;
    };
    auto no_pull = [](uint32_t) -> g1_status { return G1_OK; };
    // A distance scan that writes the input frontier of iteration `it` (the queues of iteration it - 1,
    // taken mod 6 so both rotations agree) with stamp `it`, and reads its size and the smallest distance
    // at or above lo.
    auto scan = [&](auto rv, uint64_t lo, uint64_t hi, unsigned long long* fr, uint32_t* mn) -> g1_status {
        clear_sets_kernel<<<1, 2 * kSegments, 0, s>>>(g->qcnt, it % 3, (it + 1) % 3);
        G1_LAUNCHED();
        tok = prof_begin(g, "sssp_delta_scan", s);
        dist_range_frontier_kernel<decltype(rv)><<<kMaxPushBlocks, kPushThreads, 0, s>>>(
            n, rv, g->dstamp, static_cast<uint32_t>(std::min<uint64_t>(lo, kInf)),
            static_cast<uint32_t>(std::min<uint64_t>(hi, kInf)), it, make_queues(g, it + 5), g->dl_mins);
        G1_LAUNCHED();
        prof_end(g, tok, s, 8.0 * n);
        return read_frontier(g, it % 3, s, fr, g->dl_mins, kMaxPushBlocks, mn);
    };

    uint64_t lo = 0;
    // The seed's bucket starts on the advance kernel, whatever the seed's degree (from a hub, the
    // single-workgroup kernel spends ~45 us on the first iteration's few thousand edges before handing the
    // frontier back), and light iterations go in device-driven batches of kLightBatch launches (a
    // readback between batches leaves the GPU idle ~25 us; a launch past the frontier's end costs ~5 us).
    // RMAT22 from the hub, SSSP: 1.238 ms with the small kernel first and batches of 2, 1.191 without it,
    // 1.176 / 1.169 / 1.164 with batches of 4 / 8 / 12.
    bool small_first = false;
    constexpr uint32_t kLightBatch = 8;
    uint64_t first_edges = ~0ull;    // the light frontier's edge count, when a scan read it
    unsigned long long fr = 0;
    uint32_t mn = kInf;
    for (;;) {
        limit = static_cast<uint32_t>(std::min<uint64_t>(lo + delta, kInf));
        ++buckets;
        uint32_t launched = it;
        if (first_edges != 0)
            G1_TRY(drive_push(g, s, light_launch, light_after, light_small, small_first, &launched, no_pull, 0, it,
                              first_edges, 0, kLightBatch, true));
            G1_TRY(light_bytes(launched));
        it = launched;
        G1_CHECK(it + 2 <= kHistIters, G1_ERR_INTERNAL, "delta engine exceeded %u iterations", kHistIters);
        // heavy phase: the bucket's vertices with heavy edges
        G1_TRY(scan(hrows, lo, limit, &fr, &mn));
        const uint64_t heavy_edges = fr >> 32;
        // after this heavy phase the engine leaves the buckets for the plain loop (see below); a pull step
        // then builds that loop's first frontier in its apply instead of a separate scan
        const bool leaving = limit != kInf && (heavy_edges >= switch_edges || buckets >= kDeltaMaxBuckets);
        bool fused = false;
        if (static_cast<uint32_t>(fr) != 0) {
            PushQueues pq = make_queues(g, it);
            if (heavy_edges >= pull_min && limit - lo <= 254u) {
                // x = the bucket's distances less lo, one byte each (254 below the bucket, 255 past it),
                // y = A^T (min,+) x over the rows at or past the bucket's end, apply with lo added back
                double bytes = 4.0 * (n + 1.0) + 8.0 * g->m + 42.0 * n;
                if (g->profile && pull_rows && n >= 512) {  // 2 words per counting workgroup fit in y
                    // the rows the pull visits and their in-edges, counted outside its timed region
                    const uint32_t cb = grid_for(n, kPushThreads, kMaxPushBlocks);
                    sssp_pull_rows_count_kernel<<<cb, kPushThreads, 0, s>>>(n, g->dstamp, limit, g->t_row_off,
                                                                             g->yu);
                    G1_LAUNCHED();
                    std::vector<uint32_t> part(2ull * cb);
                    G1_HIP(hipMemcpyAsync(part.data(), g->yu, sizeof(uint32_t) * part.size(), hipMemcpyDeviceToHost,
                                          s));
                    G1_HIP(hipStreamSynchronize(s));
                    double rows = 0.0, edges = 0.0;
                    for (uint32_t b = 0; b < cb; ++b) {
                        rows += part[2ull * b];
                        edges += part[2ull * b + 1];
                    }
                    // x 8n + n; pull n (row bytes) + 4n (y), 8 per row visited (in-edge range), 9 per
                    // in-edge (column, weight, x); apply 4n (y) + 16n (dist word) + 8n (degrees)
                    bytes = 42.0 * n + 8.0 * rows + 9.0 * edges;
                }
                tok = prof_begin(g, "sssp_delta_pull", s);
                sssp_pull_xb_kernel<<<grid_for((n + 3) / 4, kPushThreads, kMaxPushBlocks), kPushThreads, 0, s>>>(
                    n, g->dstamp, static_cast<uint32_t>(lo), limit, it, g->dl_xb, pq, g->hist);
                G1_LAUNCHED();
                if (pull_rows)
                    G1_TRY(launch_pull_rows(g, pull_edges, s));
                else
                    G1_TRY(launch_spmv_merge<MinPlusU8X>(n, g->m, g->t_row_off, g->t_col, g->t_w, g->dl_xb, g->yu,
                                                         g->t_ws, s));
                sssp_pull_apply_kernel<<<kMaxPushBlocks, kPushThreads, 0, s>>>(
                    n, g->row_off, g->yu, g->dstamp, it + 1, pq, db, static_cast<uint32_t>(lo), leaving ? limit : kInf);
                G1_LAUNCHED();
                // merge path: x 8n + n; SpMV 4(n+1) + 8m + n (x) + 4n (y); apply 4n (y) + 16n (dist word) +
                // 8n (degrees)
                prof_end(g, tok, s, bytes);
                ++pulls;
                fused = leaving;
            } else if (heavy_edges >= pull_min) {
                // x = the bucket's distances (the scan stamped them `it`), y = A^T (min,+) x, apply
                tok = prof_begin(g, "sssp_pull", s);
                sssp_pull_x_kernel<<<grid_for(n, kPushThreads, kMaxPushBlocks), kPushThreads, 0, s>>>(
                    n, g->dstamp, it, g->xu, pq, g->hist);
                G1_LAUNCHED();
                G1_TRY(launch_spmv_merge<MinPlusU32>(n, g->m, g->t_row_off, g->t_col, g->t_w, g->xu, g->yu, g->t_ws,
                                                     s));
                sssp_pull_apply_kernel<<<kMaxPushBlocks, kPushThreads, 0, s>>>(n, g->row_off, g->yu, g->dstamp, it + 1,
                                                                             pq, db);
                G1_LAUNCHED();
                prof_end(g, tok, s, 4.0 * (n + 1.0) + 8.0 * g->m + 44.0 * n);
                ++pulls;
            } else {
                SsspPushNoReturn pol{g->dstamp, it + 1, db};
                tok = prof_begin(g, "sssp_delta_heavy", s);
                push_advance_kernel<SsspPushNoReturn, HeavyRows><<<sssp_noreturn_blocks(), kPushThreads, 0, s>>>(
                    hrows, g->dl_col, g->dl_w, pq, pol, it, g->hist);
                G1_LAUNCHED();
                prof_end(g, tok, s, 16.0 * static_cast<uint32_t>(fr) + 12.0 * heavy_edges);
            }
            ++it;
        }
        if (limit == kInf) break;
        lo = limit;
        if (leaving) {
            // every vertex at or above lo with a finite distance, with all its edges, for the plain loop:
            // already enqueued by the pull step's apply, or from a scan
            if (fused) G1_TRY(read_frontier(g, it % 3, s, &fr));
            else G1_TRY(scan(FullRows(g->row_off), lo, kInf, &fr, &mn));
            if (static_cast<uint32_t>(fr) != 0) {
                uint32_t launched2 = it;
                G1_TRY(sssp_push_iterations(g, s, db, it, (fr >> 32) <= kSmallCap, &launched2, &dense, &pulls,
                                            fr >> 32));
                it = launched2;
            }
            break;
        }
        // the next bucket's light frontier; an empty bucket is skipped to the smallest distance left
        G1_TRY(scan(lrows, lo, lo + delta, &fr, &mn));
        if (mn == kInf) break;  // no finite distance at or above lo: done
        if (mn >= lo + delta) {
            lo = mn / delta * static_cast<uint64_t>(delta);
            G1_TRY(scan(lrows, lo, lo + delta, &fr, &mn));
        }
        first_edges = fr >> 32;
        small_first = first_edges <= kSmallCap;
    }
    if (g->out_inv) {
        G1_TRY(map_back<true>(g, g->dstamp, dist, s));
    } else {
        dstamp_extract_kernel<<<fill_blocks, 256, 0, s>>>(n, g->dstamp, dist);
        G1_LAUNCHED();
    }
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(collect_push_run(g, s, it, dist, st, true, true, g->out_inv != nullptr));
    st->pull_iterations += pulls;
    st->push_iterations -= std::min(st->push_iterations, pulls);
    st->dense_iterations = dense;
    st->frontier_sum = 1 + g->h_ctr[3];
    return G1_OK;
}

// ---------------------------------------------------------------- pull engines (SpMV)

g1_status run_minplus_pull(g1_graph* g, uint32_t src, uint32_t* d, bool weighted, int engine, g1_run_stats* st,
                           hipStream_t s) {
    G1_CHECK(g->has_t, G1_ERR_INVALID_ARG, "pull engine needs the graph's transpose");
    if (weighted) G1_CHECK(g->t_w != nullptr, G1_ERR_INVALID_ARG, "g1_sssp: graph has no weights");
    G1_TRY(ensure_pull_work(g, false));
    G1_TRY(ensure_push_work(g, s));  // stats slots
    if (engine == G1_ENGINE_BLAS_CSRVEC) G1_TRY(ensure_long_rows(g, false, s));
    const uint32_t n = g->n;
    const uint32_t cap = std::min<uint32_t>(persistent_cap(), kMaxPushBlocks);
    minplus_seed_kernel<<<grid_for(n, 256, cap), 256, 0, s>>>(n, src, d, g->xu);
    G1_LAUNCHED();
    clear_u64_kernel<<<1, 64, 0, s>>>(g->ctr + 6, 1);
    G1_LAUNCHED();
    const uint32_t* vals = weighted ? g->t_w : nullptr;
    const uint32_t lanes = csrvec_lanes(n, g->m);
    G1_HIP(hipEventRecord(g->ev[0], s));
    uint64_t prev = 0, frontier_sum = 1;
    uint32_t it = 0;
    const double spmv_bytes = 4.0 * (n + 1.0) + 4.0 * g->m * (weighted ? 2.0 : 1.0) + 8.0 * n;
    for (;;) {
        if (engine == G1_ENGINE_BLAS_CSRVEC) {
            const int tok = prof_begin(g, "spmv_csrvec_minplus", s);
            G1_TRY(launch_spmv_csrvec<MinPlusU32>(n, g->t_row_off, g->t_col, vals, g->xu, g->yu, lanes, s, g->t_long,
                                                  g->t_nlong, g->long_min));
            prof_end(g, tok, s, spmv_bytes);
        } else {
            const int tok = prof_begin(g, "spmv_merge_minplus", s);
            G1_TRY(launch_spmv_merge<MinPlusU32>(n, g->m, g->t_row_off, g->t_col, vals, g->xu, g->yu, g->t_ws, s));
            prof_end(g, tok, s, spmv_bytes);
        }
        const int tok = prof_begin(g, "apply_minplus", s);
        apply_minplus_kernel<<<grid_for(n, 256, cap), 256, 0, s>>>(n, g->yu, d, g->xu, g->slots);
        G1_LAUNCHED();
        prof_end(g, tok, s, 12.0 * n);
        sum_slots_kernel<<<1, 256, 0, s>>>(g->slots, kMaxPushBlocks, g->ctr + 6);
        G1_LAUNCHED();
        ++it;
        st->edges_sum += g->m;
        G1_TRY(read_ctr(g, s, 6, 1));
        const uint64_t c = g->h_ctr[6];
        const uint64_t delta = c - prev;
        prev = c;
        if (delta == 0 || it >= kMaxIterations) break;
        frontier_sum += delta;
    }
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(elapsed(g, st));
    G1_TRY(prof_flush(g));
    st->iterations = it;
    st->pull_iterations = it;
    st->frontier_sum = frontier_sum;
    return finish_stats(g, d, st, s);
}

// ---------------------------------------------------------------- PageRank / personalized PageRank

// out[perm[i]] = r[i]: ranks in the relabelled ids back to the graph's ids.
__global__ void unpermute_f32_kernel(uint32_t n, const uint32_t* __restrict__ perm, const float* __restrict__ r,
                                     float* __restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[perm[i]] = r[i];
}

// Plain PageRank relabels the vertices hub first unless G1_PR_RELABEL=0.
bool pr_relabel() {
    static const bool on = [] {
        const char* e = getenv("G1_PR_RELABEL");
        return !(e && e[0] == '0');
    }();
    return on;
}

g1_status run_pagerank(g1_graph* g, float alpha, uint32_t seed, float restart, uint32_t max_iter, double tol,
                       int mode, int engine, float* out, g1_run_stats* st, hipStream_t s) {
    G1_CHECK(g->has_t, G1_ERR_INVALID_ARG, "PageRank needs the graph's transpose");
    G1_TRY(ensure_pull_work(g, true));
    const uint32_t n = g->n;
    const uint32_t blocks = grid_for(n, 256, std::min<uint32_t>(kPartials, persistent_cap()));
    const uint32_t lanes = csrvec_lanes(n, g->m);
    const bool personalized = seed != kInf;
    // Plain PageRank runs over the vertices relabelled hub first (ensure_pr_relabel; G1_PR_RELABEL=0 or a
    // view handle keeps the graph's ids) and un-permutes the ranks at the end; personalized PageRank keeps
    // the graph's ids, the fold order the batched engines reproduce.
    const bool relabel = !personalized && pr_relabel() && n > 0 && !g->is_view;
    if (relabel) G1_TRY(ensure_pr_relabel(g, s));
    const uint32_t* t_ro = relabel ? g->pr_row_off : g->t_row_off;
    const uint32_t* t_ci = relabel ? g->pr_col : g->t_col;
    const float* vals = relabel ? g->pr_wf : g->t_wf;  // null for an unweighted graph: every weight is 1
    const float* osum = relabel ? g->pr_osum : g->osum;
    void* ws = relabel ? g->pr_ws : g->t_ws;
    if (engine == G1_ENGINE_BLAS_CSRVEC) G1_TRY(ensure_long_rows(g, relabel, s));
    const uint32_t* lrows = relabel ? g->pr_long : g->t_long;
    const uint32_t nlong = relabel ? g->pr_nlong : g->t_nlong;
    // alpha of the propagation: PageRank damping, or 1 - restart for the personalized variant.
    const float prop = personalized ? 1.0f - restart : alpha;
    const float r0 = n ? 1.0f / static_cast<float>(n) : 0.0f;
    pr_init_kernel<<<grid_for(n, 256, persistent_cap()), 256, 0, s>>>(n, osum, prop, g->iw, g->r[0],
                                                                      r0, g->xf, seed);
    G1_LAUNCHED();
    const float base_all = personalized ? 0.0f : static_cast<float>((1.0 - alpha) / static_cast<double>(n));
    const float base_seed = personalized ? restart : base_all;
    G1_HIP(hipEventRecord(g->ev[0], s));
    uint32_t it = 0;
    int cur = 0;  // r[cur] holds the latest iterate
    double delta = 0.0;
    if (mode == G1_PR_MAX_ABS) {
        // The reference's loop (apps/pagerank_blas.cu:63-78,93-123): ranks[1] starts at 0; before each
        // iteration stop if max|ranks[idx^1] - ranks[idx]| <= tol; the iteration writes ranks[idx]
        // in place (apply_each), then idx ^= 1. The result is ranks[idx] at exit.
        G1_TRY(g1_fill_f32(g->r[1], 0.0f, n, s));
        int idx = 0;
        for (;;) {
            max_abs_diff_kernel<<<blocks, 256, 0, s>>>(n, g->r[idx ^ 1], g->r[idx], g->partials);
            G1_LAUNCHED();
            reduce_partials_kernel<1><<<1, 256, 0, s>>>(g->partials, blocks, g->partials + kPartials);
            G1_LAUNCHED();
            G1_HIP(hipMemcpyAsync(g->h_scalar, g->partials + kPartials, sizeof(double), hipMemcpyDeviceToHost, s));
            G1_HIP(hipStreamSynchronize(s));
            delta = g->h_scalar[0];
            if (delta <= tol || it >= max_iter) break;
            if (engine == G1_ENGINE_BLAS_CSRVEC)
                G1_TRY(launch_spmv_csrvec<PlusTimesKahanF32>(n, t_ro, t_ci, vals, g->xf, g->yf, lanes, s, lrows, nlong,
                                                              g->long_min));
            else
                G1_TRY(launch_spmv_merge<PlusTimesKahanF32>(n, g->m, t_ro, t_ci, vals, g->xf, g->yf, ws, s));
            // In place on ranks[idx] as apply_each does; the partials of this call are unused.
            pr_apply_kernel<1><<<blocks, 256, 0, s>>>(n, g->yf, base_all, seed, base_seed, g->iw, g->r[idx ^ 1],
                                                      g->r[idx], g->xf, g->partials);
            G1_LAUNCHED();
            idx ^= 1;
            ++it;
        }
        cur = idx;
    } else {
        const double spmv_bytes = 4.0 * (n + 1.0) + 4.0 * g->m * (vals ? 2.0 : 1.0) + 8.0 * n;
        G1_TRY(ensure_rank_partials(g));
        const uint32_t nb = rank_blocks(n);
        G1_HIP(hipMemcpyAsync(g->seeds_dev, &seed, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        for (;;) {
            if (it >= max_iter) break;
            int tok;
            if (engine == G1_ENGINE_BLAS_CSRVEC) {
                tok = prof_begin(g, "spmv_csrvec_plustimes", s);
                G1_TRY(launch_spmv_csrvec<PlusTimesKahanF32>(n, t_ro, t_ci, vals, g->xf, g->yf, lanes, s, lrows, nlong,
                                                              g->long_min));
            } else {
                tok = prof_begin(g, "spmv_merge_plustimes", s);
                G1_TRY(launch_spmv_merge<PlusTimesKahanF32>(n, g->m, t_ro, t_ci, vals, g->xf, g->yf, ws, s));
            }
            prof_end(g, tok, s, spmv_bytes);
            tok = prof_begin(g, "pr_apply", s);
            rank_apply_kernel<<<nb, 256, 0, s>>>(n, g->yf, base_all, g->seeds_dev, base_seed, g->iw, g->r[cur],
                                                 g->r[cur ^ 1], g->xf, g->rpart);
            G1_LAUNCHED();
            prof_end(g, tok, s, 20.0 * n);
            reduce_partials_kernel<0><<<dim3(1, 1), 256, 0, s>>>(g->rpart, nb, g->rpart + static_cast<size_t>(nb) * kBatchCols);
            G1_LAUNCHED();
            cur ^= 1;
            ++it;
            G1_HIP(hipMemcpyAsync(g->h_scalar, g->rpart + static_cast<size_t>(nb) * kBatchCols, sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            G1_HIP(hipStreamSynchronize(s));
            delta = g->h_scalar[0];
            if (delta < tol) break;
        }
    }
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_TRY(elapsed(g, st));
    G1_TRY(prof_flush(g));
    if (relabel) {
        unpermute_f32_kernel<<<grid_for(n, 256, persistent_cap()), 256, 0, s>>>(n, g->pr_perm, g->r[cur], out);
        G1_LAUNCHED();
    } else {
        G1_HIP(hipMemcpyAsync(out, g->r[cur], sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    }
    G1_HIP(hipStreamSynchronize(s));
    st->iterations = it;
    st->pull_iterations = it;
    st->edges_sum = static_cast<uint64_t>(it) * g->m;
    st->last_delta = delta;
    st->reached = n;
    st->reached_edges = g->m;
    return G1_OK;
}

g1_status check_run(g1_graph* g, uint32_t src, const void* out, g1_run_stats* st, const char* what) {
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "%s: null graph", what);
    G1_CHECK(out != nullptr, G1_ERR_INVALID_ARG, "%s: null output", what);
    G1_CHECK(src < g->n, G1_ERR_INVALID_ARG, "%s: source %u out of range (n = %u)", what, src, g->n);
    if (st) memset(st, 0, sizeof(*st));
    return G1_OK;
}

}  // namespace

bool is_road_graph(const g1_graph* g) { return road_graph(g); }
g1_status prepare_road_work(g1_graph* g, hipStream_t s) { return ensure_road_work(g, s); }
uint64_t road_work_size(const g1_graph* g) { return road_work_bytes(g); }
uint32_t delta_width(const g1_graph* g) { return sssp_delta_width(g); }

g1_status road_queries(g1_graph* g, const uint32_t* kinds, const uint32_t* sources, uint32_t count, uint32_t** outs,
                       uint8_t* ok, uint32_t* iters, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g->mu);
    return road_batch_run(g, kinds, sources, count, outs, ok, iters, s);
}

}  // namespace g1

extern "C" {

g1_status g1_bfs(g1_graph* g, uint32_t src, uint32_t* depth, int32_t engine, g1_run_stats* stats,
                 g1_stream_t stream) {
    using namespace g1;
    g1_run_stats local;
    if (!stats) stats = &local;
    G1_TRY(check_run(g, src, depth, stats, "g1_bfs"));
    std::lock_guard<std::mutex> lock(g->mu);
    const hipStream_t s = as_stream(stream);
    switch (engine) {
        case G1_ENGINE_ADVANCE: return run_bfs_push(g, src, depth, false, stats, s);
        case G1_ENGINE_DIROPT: return run_bfs_push(g, src, depth, true, stats, s);
        case G1_ENGINE_BLAS_MERGE:
        case G1_ENGINE_BLAS_CSRVEC: return run_minplus_pull(g, src, depth, false, engine, stats, s);
        default: return set_error(G1_ERR_INVALID_ARG, "g1_bfs: unknown engine %d", engine);
    }
}

g1_status g1_sssp(g1_graph* g, uint32_t src, uint32_t* dist, int32_t engine, g1_run_stats* stats,
                  g1_stream_t stream) {
    using namespace g1;
    g1_run_stats local;
    if (!stats) stats = &local;
    G1_TRY(check_run(g, src, dist, stats, "g1_sssp"));
    std::lock_guard<std::mutex> lock(g->mu);
    const hipStream_t s = as_stream(stream);
    switch (engine) {
        case G1_ENGINE_ADVANCE: return run_sssp_push(g, src, dist, stats, s);
        case G1_ENGINE_DELTA: return run_sssp_delta(g, src, dist, stats, s);
        case G1_ENGINE_BLAS_MERGE:
        case G1_ENGINE_BLAS_CSRVEC: return run_minplus_pull(g, src, dist, true, engine, stats, s);
        default: return set_error(G1_ERR_INVALID_ARG, "g1_sssp: unknown engine %d", engine);
    }
}

g1_status g1_road_batch(g1_graph* g, const uint32_t* kinds, const uint32_t* sources, uint32_t count, uint32_t* out,
                        uint32_t* iterations, g1_stream_t stream) {
    using namespace g1;
    G1_CHECK(g != nullptr && (count == 0 || (sources != nullptr && out != nullptr)), G1_ERR_INVALID_ARG,
             "g1_road_batch: null argument");
    G1_CHECK(road_graph(g), G1_ERR_INVALID_ARG,
             "g1_road_batch: the graph has a vertex of out-degree %u (the road engine takes at most %u)",
             g->max_out_deg, kRoadMaxDeg);
    for (uint32_t i = 0; i < count; ++i) {
        G1_CHECK(sources[i] < g->n, G1_ERR_INVALID_ARG, "g1_road_batch: source %u out of range", sources[i]);
        G1_CHECK(!(kinds ? kinds[i] : 1u) || g->w != nullptr, G1_ERR_INVALID_ARG, "g1_road_batch: SSSP without weights");
    }
    std::lock_guard<std::mutex> lock(g->mu);
    const hipStream_t s = as_stream(stream);
    std::vector<uint32_t> k(count, 1u);
    if (kinds)
        for (uint32_t i = 0; i < count; ++i) k[i] = kinds[i] ? 1u : 0u;
    std::vector<uint32_t*> outs(count);
    for (uint32_t i = 0; i < count; ++i) outs[i] = out + static_cast<size_t>(i) * g->n;
    std::vector<uint8_t> ok(count, 0);
    std::vector<uint32_t> it(count, 0);
    G1_TRY(road_batch_run(g, k.data(), sources, count, outs.data(), ok.data(), it.data(), s));
    for (uint32_t i = 0; i < count; ++i) {
        if (ok[i]) continue;  // a far-list overflow: this query runs on the push engine
        g1_run_stats st;
        memset(&st, 0, sizeof(st));
        if (k[i]) G1_TRY(run_sssp_push(g, sources[i], outs[i], &st, s, false));
        else G1_TRY(run_bfs_push(g, sources[i], outs[i], false, &st, s));
        it[i] = st.iterations;
    }
    if (iterations)
        for (uint32_t i = 0; i < count; ++i) iterations[i] = it[i];
    return G1_OK;
}

g1_status g1_pagerank(g1_graph* g, float alpha, uint32_t max_iter, double tol, int32_t converge_mode,
                      int32_t engine, float* rank, g1_run_stats* stats, g1_stream_t stream) {
    using namespace g1;
    g1_run_stats local;
    if (!stats) stats = &local;
    G1_CHECK(g != nullptr && rank != nullptr, G1_ERR_INVALID_ARG, "g1_pagerank: null argument");
    G1_CHECK(converge_mode == G1_PR_L1_F64 || converge_mode == G1_PR_MAX_ABS, G1_ERR_INVALID_ARG,
             "g1_pagerank: unknown convergence mode %d", converge_mode);
    G1_CHECK(engine == G1_ENGINE_BLAS_MERGE || engine == G1_ENGINE_BLAS_CSRVEC, G1_ERR_INVALID_ARG,
             "g1_pagerank: engine must be BLAS_MERGE or BLAS_CSRVEC");
    memset(stats, 0, sizeof(*stats));
    std::lock_guard<std::mutex> lock(g->mu);
    return run_pagerank(g, alpha, kInf, 0.0f, max_iter, tol, converge_mode, engine, rank, stats, as_stream(stream));
}

g1_status g1_ppr(g1_graph* g, uint32_t seed_vertex, float restart, uint32_t max_iter, double tol, float* rank,
                 g1_run_stats* stats, g1_stream_t stream) {
    using namespace g1;
    g1_run_stats local;
    if (!stats) stats = &local;
    G1_TRY(check_run(g, seed_vertex, rank, stats, "g1_ppr"));
    G1_CHECK(restart > 0.0f && restart <= 1.0f, G1_ERR_INVALID_ARG, "g1_ppr: restart must be in (0, 1]");
    std::lock_guard<std::mutex> lock(g->mu);
    return run_pagerank(g, 0.0f, seed_vertex, restart, max_iter, tol, G1_PR_L1_F64, G1_ENGINE_BLAS_MERGE, rank, stats,
                        as_stream(stream));
}

}  // extern "C"
