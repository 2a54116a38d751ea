// g1_graph: device-resident CSR (+ transpose) and its work areas. Replaces graph::build +
// Graph::ToArch (include/GraphOne/graph/builder.h:9-46, graph.h:94-113): the host CSR is uploaded once
// and stays resident in HBM; the in-edge CSR is built from it on the device (convert.hip).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "g1/push.hpp"
#include "g1/sweep.hpp"
#include "g1/spmv_merge.hpp"
#include "g1/apply.hpp"
#include "graph.hpp"
#include "spmv_launch.hpp"

namespace g1 {

g1_status graph_alloc(g1_graph* g, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 4;
    G1_HIP(hipMalloc(p, bytes));
    g->owned.push_back(*p);
    g->device_bytes += bytes;
    return G1_OK;
}

template <typename T>
static g1_status alloc_n(g1_graph* g, T** p, size_t count) {
    void* v = nullptr;
    G1_TRY(graph_alloc(g, &v, count * sizeof(T)));
    *p = static_cast<T*>(v);
    return G1_OK;
}

g1_status ensure_push_work(g1_graph* g, hipStream_t s) {
    if (g->q[0]) return G1_OK;
    const size_t n = g->n ? g->n : 1;
    const size_t words = (n + 31) / 32 + 1;
    const uint64_t items = g->m > g->n ? g->m : g->n;
    const uint64_t cap = push_segment_capacity(n, items);
    G1_CHECK(cap * kSegments < (1ull << 34), G1_ERR_INVALID_ARG, "push queues too large");
    g->seg_cap = static_cast<uint32_t>(cap);
    for (int i = 0; i < 2; ++i) {
        G1_TRY(alloc_n(g, &g->q[i], cap * kSegments));
        G1_TRY(alloc_n(g, &g->qoff[i], cap * kSegments));
        G1_TRY(alloc_n(g, &g->fb[i], words));
    }
    G1_TRY(alloc_n(g, &g->stamp, n));
    G1_TRY(alloc_n(g, &g->fb3, 3 * words));
    G1_TRY(alloc_n(g, &g->dstamp, n));
    G1_TRY(alloc_n(g, &g->dbound, n + 256));
    G1_TRY(alloc_n(g, &g->hist, 2ull * kHistIters));
    G1_TRY(alloc_n(g, &g->qcnt, 3 * kSegments));
    G1_TRY(alloc_n(g, &g->slots, 4 * kMaxPushBlocks));  // four regions of per-workgroup stats
    G1_TRY(alloc_n(g, &g->overflow, 4));
    if (!g->h_hist)  // the host copy of the iteration history (pinned)
        G1_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->h_hist), 2ull * kHistIters * sizeof(unsigned long long)));
    // Cleared on the caller's stream, ahead of its first launch: a memset on the null stream is not
    // ordered before work on a non-blocking stream, and the allocation may hold an earlier buffer's data.
    const g1_stream_t gs = reinterpret_cast<g1_stream_t>(s);
    G1_TRY(g1_fill_u32(reinterpret_cast<uint32_t*>(g->slots), 0u, 8ull * kMaxPushBlocks, gs));
    G1_TRY(g1_fill_u32(g->overflow, 0u, 4, gs));
    return G1_OK;
}

// The light/heavy copy of the out-edge CSR for delta-stepping (split_rows_kernel), built on the device
// once per bucket width: 8 bytes per edge and 5 per vertex, plus the scan minima.
g1_status ensure_delta_split(g1_graph* g, uint32_t delta, hipStream_t s) {
    G1_CHECK(g->w != nullptr, G1_ERR_INVALID_ARG, "delta-stepping needs edge weights");
    if (g->dl_col && g->dl_delta == delta) return G1_OK;
    if (!g->dl_col) {
        G1_TRY(alloc_n(g, &g->dl_col, std::max<size_t>(g->m, 1)));
        G1_TRY(alloc_n(g, &g->dl_w, std::max<size_t>(g->m, 1)));
        G1_TRY(alloc_n(g, &g->dl_lc, std::max<size_t>(g->n, 1)));
        G1_TRY(alloc_n(g, &g->dl_mins, kMaxPushBlocks));
        G1_TRY(alloc_n(g, &g->dl_xb, (static_cast<size_t>(g->n) + 3) / 4 * 4 + 4));
    }
    if (g->n > 0) {
        const uint32_t waves_per_block = 256 / kWave;
        const uint32_t blocks = std::min<uint32_t>((g->n + waves_per_block - 1) / waves_per_block, 8192);
        split_rows_kernel<<<blocks, 256, 0, s>>>(g->n, g->row_off, g->col, g->w, delta, g->dl_col, g->dl_w, g->dl_lc);
        G1_HIP(hipGetLastError());
    }
    g->dl_delta = delta;
    return G1_OK;
}

// The rows of the in-edge CSR (or of PageRank's relabelled copy) longer than long_min nonzeros, for the
// CSR-vector engine's long-row kernel, longest first (ties by id); built once on the host. The in-edge
// list also counts its rows longer than huge_min (the delta engine's heavy pull).
g1_status ensure_long_rows(g1_graph* g, bool relabelled, hipStream_t s) {
    uint32_t** list = relabelled ? &g->pr_long : &g->t_long;
    uint32_t* count = relabelled ? &g->pr_nlong : &g->t_nlong;
    if (*list || !g->has_t) return G1_OK;
    const uint32_t* ro_dev = relabelled ? g->pr_row_off : g->t_row_off;
    G1_CHECK(ro_dev != nullptr, G1_ERR_INTERNAL, "long rows of a missing matrix");
    const uint32_t n = g->n;
    g->long_min = csrvec_long_per_lane() * csrvec_lanes(n, g->m);
    G1_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> ro(static_cast<size_t>(n) + 1), rows;
    G1_HIP(hipMemcpy(ro.data(), ro_dev, sizeof(uint32_t) * (n + 1ull), hipMemcpyDeviceToHost));
    for (uint32_t v = 0; v < n; ++v)
        if (ro[v + 1] - ro[v] > g->long_min) rows.push_back(v);
    std::stable_sort(rows.begin(), rows.end(),
                     [&](uint32_t a, uint32_t b) { return ro[a + 1] - ro[a] > ro[b + 1] - ro[b]; });
    if (!relabelled) {
        // G1_SSSP_DELTA_PULL_HUGE lowers the row length past which the heavy pull gives a row a workgroup
        // (default 4096; never below long_min, the list's own bound)
        const char* e = getenv("G1_SSSP_DELTA_PULL_HUGE");
        const long hv = e ? atol(e) : 0;
        const uint32_t huge = (hv > 0 && hv < static_cast<long>(kPullRowsWaveMax)) ? static_cast<uint32_t>(hv)
                                                                                   : kPullRowsWaveMax;
        g->huge_min = std::max(huge, g->long_min);
        g->t_nhuge = 0;
        while (g->t_nhuge < rows.size() && ro[rows[g->t_nhuge] + 1] - ro[rows[g->t_nhuge]] > g->huge_min)
            ++g->t_nhuge;
    }
    G1_TRY(alloc_n(g, list, rows.empty() ? 1 : rows.size()));
    if (!rows.empty())
        G1_HIP(hipMemcpy(*list, rows.data(), sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice));
    *count = static_cast<uint32_t>(rows.size());
    return G1_OK;
}

// The pieces of the out-edge rows longer than kSweepLong edges, for the dense SSSP sweep: counted per row,
// scanned and written on the device; the count is read back once (a stream synchronization, at the first
// dense iteration on this handle).
g1_status ensure_sweep_pieces(g1_graph* g, hipStream_t s) {
    if (g->sweep_pieces) return G1_OK;
    const uint32_t n = g->n;
    uint32_t *cnt = nullptr, *pos = nullptr, *part = nullptr;
    const size_t nn = n ? n : 1;
    G1_HIP(hipMalloc(reinterpret_cast<void**>(&cnt), nn * sizeof(uint32_t)));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&pos), nn * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&part), (nn / 4096 + 2) * sizeof(uint32_t));
    g1_status st = G1_OK;
    uint32_t last[2] = {0, 0};
    if (e != hipSuccess) {
        st = set_error(G1_ERR_OUT_OF_MEMORY, "ensure_sweep_pieces: %s", hipGetErrorString(e));
    } else if (n) {
        const uint32_t blocks = (n + 255) / 256;
        sweep_piece_count_kernel<<<blocks, 256, 0, s>>>(n, g->row_off, cnt);
        st = exclusive_scan_u32(cnt, pos, n, part, s);
        if (st == G1_OK) {
            e = hipMemcpyAsync(&last[0], pos + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipMemcpyAsync(&last[1], cnt + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) st = set_error(G1_ERR_HIP, "ensure_sweep_pieces: %s", hipGetErrorString(e));
        }
        const uint32_t total = last[0] + last[1];
        if (st == G1_OK) {
            void* p = nullptr;
            st = graph_alloc(g, &p, static_cast<size_t>(total ? total : 1) * sizeof(SweepPiece));
            if (st == G1_OK && total) {
                sweep_piece_write_kernel<<<blocks, 256, 0, s>>>(n, g->row_off, pos, static_cast<SweepPiece*>(p));
                e = hipStreamSynchronize(s);
                if (e != hipSuccess) st = set_error(G1_ERR_HIP, "ensure_sweep_pieces: %s", hipGetErrorString(e));
            }
            if (st == G1_OK) {
                g->sweep_pieces = p;
                g->n_sweep_pieces = total;
            }
        }
    }
    (void)hipFree(cnt);
    (void)hipFree(pos);
    (void)hipFree(part);
    return st;
}

g1_status ensure_pull_work(g1_graph* g, bool pagerank) {
    const size_t n = g->n ? g->n : 1;
    if (!g->xu) {
        G1_TRY(alloc_n(g, &g->xu, n));
        G1_TRY(alloc_n(g, &g->yu, n));
    }
    if (pagerank && !g->xf) {
        G1_TRY(alloc_n(g, &g->xf, n));
        G1_TRY(alloc_n(g, &g->yf, n));
        G1_TRY(alloc_n(g, &g->iw, n));
        G1_TRY(alloc_n(g, &g->r[0], n));
        G1_TRY(alloc_n(g, &g->r[1], n));
    }
    if (pagerank && !g->osum) {
        G1_TRY(alloc_n(g, &g->osum, n));
        if (g->n) {
            const uint32_t blocks = std::min<uint32_t>((g->n + 3) / 4, 16384);
            out_weight_sum_kernel<<<blocks, 256>>>(g->n, g->row_off, g->w, g->osum);
            G1_LAUNCHED();
            G1_HIP(hipDeviceSynchronize());
        }
    }
    if (pagerank && g->t_w && !g->t_wf) {
        // with the padding of the integer weights (zeros), converted on the device
        G1_TRY(alloc_n(g, &g->t_wf, static_cast<size_t>(g->m) + kCsrPad));
        G1_TRY(device_u32_to_f32(g->t_w, g->t_wf, static_cast<size_t>(g->m) + kCsrPad, nullptr));
        G1_HIP(hipDeviceSynchronize());
    }
    return G1_OK;
}

g1_status ensure_rank_partials(g1_graph* g) {
    if (g->rpart) return G1_OK;
    const size_t nb = rank_blocks(g->n ? g->n : 1);
    G1_TRY(alloc_n(g, &g->rpart, nb * kBatchCols + kBatchCols + 8));
    G1_TRY(alloc_n(g, &g->seeds_dev, kBatchCols));
    return G1_OK;
}

static hipEvent_t pool_event(g1_graph* g) {
    if (g->ev_used == g->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        g->ev_pool.push_back(e);
    }
    return g->ev_pool[g->ev_used++];
}

int prof_begin(g1_graph* g, const char* name, hipStream_t s) {
    if (!g->profile) return -1;
    if (g->prof_to) g = g->prof_to;
    int entry = -1;
    for (size_t i = 0; i < g->prof.size(); ++i)
        if (g->prof[i].name == name) entry = static_cast<int>(i);
    if (entry < 0) {
        g->prof.push_back(g1_graph::ProfEntry{name, 0, 0.0, 0.0});
        entry = static_cast<int>(g->prof.size() - 1);
    }
    hipEvent_t a = pool_event(g), b = pool_event(g);
    if (!a || !b) return -1;
    (void)hipEventRecord(a, s);
    g->pending.push_back(g1_graph::ProfPending{entry, a, b, 0.0});
    return static_cast<int>(g->pending.size() - 1);
}

void prof_end(g1_graph* g, int token, hipStream_t s, double bytes) {
    if (token < 0) return;
    if (g->prof_to) g = g->prof_to;
    (void)hipEventRecord(g->pending[token].b, s);
    g->pending[token].bytes = bytes;
}

void prof_set_bytes(g1_graph* g, int token, double bytes) {
    if (g->prof_to) g = g->prof_to;
    if (token >= 0) g->pending[token].bytes = bytes;
}

g1_status prof_flush(g1_graph* g) {
    if (g->prof_to) g = g->prof_to;
    for (auto& p : g->pending) {
        G1_HIP(hipEventSynchronize(p.b));
        float ms = 0.0f;
        G1_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto& e = g->prof[p.entry];
        e.launches += 1;
        e.ms += ms;
        e.bytes += p.bytes;
    }
    g->pending.clear();
    g->ev_used = 0;
    return G1_OK;
}

// What every handle holds besides its CSR: the counters, the partial sums, their pinned host mirrors and
// the run events.
static g1_status graph_common(g1_graph* g, const char* what) {
    G1_TRY(alloc_n(g, &g->ctr, 16));
    G1_TRY(alloc_n(g, &g->partials, kPartials + 8));
    G1_CHECK(hipHostMalloc(reinterpret_cast<void**>(&g->h_ctr), 256 * sizeof(unsigned long long)) == hipSuccess &&
                 hipHostMalloc(reinterpret_cast<void**>(&g->h_scalar), 8 * sizeof(double)) == hipSuccess &&
                 hipHostMalloc(reinterpret_cast<void**>(&g->h_mbox), 4 * sizeof(unsigned long long),
                               hipHostMallocCoherent) == hipSuccess,
             G1_ERR_HIP, "%s: pinned host allocation failed", what);
    memset(g->h_mbox, 0, 4 * sizeof(unsigned long long));
    G1_CHECK(hipEventCreate(&g->ev[0]) == hipSuccess && hipEventCreate(&g->ev[1]) == hipSuccess, G1_ERR_HIP,
             "%s: event creation failed", what);
    return G1_OK;
}
static uint64_t graph_common_bytes() { return 16 * 8 + (kPartials + 8) * 8; }

static void graph_free_all(g1_graph* g) {
    for (auto e : g->ev_pool) (void)hipEventDestroy(e);
    g->ev_pool.clear();
    for (void* p : g->owned) (void)hipFree(p);
    g->owned.clear();
    if (g->h_ctr) (void)hipHostFree(g->h_ctr);
    if (g->h_scalar) (void)hipHostFree(g->h_scalar);
    if (g->h_mbox) (void)hipHostFree(g->h_mbox);
    if (g->h_hist) (void)hipHostFree(g->h_hist);
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
}

// Whether a graph can have the hub-first copy: not a view (which shares the base's arrays and ids), not a
// copy itself, not a road-like graph (which runs on the road engine). G1_HUB=0 turns the copy off.
bool hub_eligible(const g1_graph* g) {
    if (g->is_view || g->out_inv || g->n == 0 || is_road_graph(g)) return false;
    const char* e = getenv("G1_HUB");
    return !(e && atoi(e) == 0);
}

// Bytes ensure_hub_copy allocates (here and in the copy), before the copy's engines add their work areas.
static uint64_t hub_base_bytes(const g1_graph* g) {
    const uint64_t n1 = static_cast<uint64_t>(g->n) + 1, mp = static_cast<uint64_t>(g->m) + kCsrPad;
    const uint64_t csr = 4 * n1 + 4 * mp * (g->w ? 2 : 1);
    uint64_t b = 4ull * g->n + csr + 4ull * g->n + graph_common_bytes();
    if (g->has_t && !g->t_alias) b += csr;
    if (g->has_t) b += merge_workspace_bytes(g->n, g->m);
    return b;
}

// The push engines' hub-first copy: a graph over the vertices grouped by out-degree, largest group first,
// so the vertices a run touches most (the hubs, reached early and relaxed often) share cache lines, with
// each row's columns ascending (build_hub_csr). It holds its own CSR, in-edge CSR (when this graph has
// one) and work areas; runs on it map the source in and their results back (out_inv), so a caller sees
// its own ids. On RMAT scale 22 (weights 1..255, hub source) delta-stepping SSSP took 0.65–0.66 ms on the
// copy against 0.73–0.74 ms on the graph's own ids, and the BFS levels 128–136 µs against 146–152 µs
// (before the 18–22 µs map back), with identical results.
g1_status ensure_hub_copy(g1_graph* g, hipStream_t s) {
    if (g->hub) return G1_OK;
    G1_CHECK(hub_eligible(g), G1_ERR_INTERNAL, "hub copy: graph not eligible");
    const uint32_t n = g->n, m = g->m;
    if (!g->hub_inv) G1_TRY(alloc_n(g, &g->hub_inv, n));
    g1_graph* h = new g1_graph();
    h->n = n;
    h->m = m;
    h->flags = g->flags;
    h->max_out_deg = g->max_out_deg;
    h->max_w = g->max_w;
    auto fail = [&](g1_status st) {
        graph_free_all(h);
        delete h;
        return st;
    };
    g1_status st;
    const size_t n1 = static_cast<size_t>(n) + 1, mp = static_cast<size_t>(m) + kCsrPad;
    const g1_stream_t gs = reinterpret_cast<g1_stream_t>(s);
    auto csr = [&](uint32_t** ro, uint32_t** c, uint32_t** w) -> g1_status {
        G1_TRY(alloc_n(h, ro, n1));
        G1_TRY(alloc_n(h, c, mp));
        G1_TRY(g1_fill_u32(*c + m, 0u, kCsrPad, gs));
        if (g->w) {
            G1_TRY(alloc_n(h, w, mp));
            G1_TRY(g1_fill_u32(*w + m, 0u, kCsrPad, gs));
        }
        return G1_OK;
    };
    if ((st = csr(&h->row_off, &h->col, &h->w)) != G1_OK) return fail(st);
    h->has_t = g->has_t;
    h->t_alias = g->t_alias;
    if (g->t_alias) {
        h->t_row_off = h->row_off;
        h->t_col = h->col;
        h->t_w = h->w;
    } else if (g->has_t && (st = csr(&h->t_row_off, &h->t_col, &h->t_w)) != G1_OK) {
        return fail(st);
    }
    const bool sym = g->t_alias;
    g->hub_inv_h.resize(n);
    if ((st = build_hub_csr(g, sym, g->hub_inv, g->hub_inv_h.data(), h->row_off, h->col, h->w, sym ? nullptr : h->t_row_off,
                            sym ? nullptr : h->t_col, sym ? nullptr : h->t_w, s)) != G1_OK)
        return fail(st);
    if (h->has_t) {
        h->t_ws_bytes = merge_workspace_bytes(n, m);
        if ((st = graph_alloc(h, &h->t_ws, h->t_ws_bytes)) != G1_OK) return fail(st);
        if ((st = launch_merge_partition(n, m, h->t_row_off, h->t_ws, s)) != G1_OK) return fail(st);
    }
    if ((st = alloc_n(h, &h->out_work, n)) != G1_OK) return fail(st);
    if ((st = graph_common(h, "hub copy")) != G1_OK) return fail(st);
    h->out_inv = g->hub_inv;
    h->out_row_off = g->row_off;
    h->prof_to = g;
    g->hub = h;
    return G1_OK;
}

}  // namespace g1

extern "C" {

g1_status g1_graph_create(g1_graph** out, const g1_host_csr* csr, uint32_t flags) {
    using namespace g1;
    G1_CHECK(out != nullptr && csr != nullptr, G1_ERR_INVALID_ARG, "g1_graph_create: null argument");
    *out = nullptr;
    G1_CHECK(csr->row_offsets != nullptr, G1_ERR_INVALID_ARG, "g1_graph_create: null row_offsets");
    G1_CHECK(csr->m == 0 || csr->col_indices != nullptr, G1_ERR_INVALID_ARG, "g1_graph_create: null col_indices");
    G1_CHECK(csr->row_offsets[0] == 0 && csr->row_offsets[csr->n] == csr->m, G1_ERR_INVALID_ARG,
             "g1_graph_create: row_offsets must start at 0 and end at m");
    G1_CHECK(csr->m < kInf, G1_ERR_INVALID_ARG, "g1_graph_create: too many edges");
    for (uint32_t v = 0; v < csr->n; ++v)
        G1_CHECK(csr->row_offsets[v] <= csr->row_offsets[v + 1], G1_ERR_INVALID_ARG,
                 "g1_graph_create: row_offsets decrease at row %u", v);
    for (uint32_t e = 0; e < csr->m; ++e)
        G1_CHECK(csr->col_indices[e] < csr->n, G1_ERR_INVALID_ARG,
                 "g1_graph_create: column %u of edge %u out of range", csr->col_indices[e], e);

    g1_graph* g = new g1_graph();
    g->n = csr->n;
    g->m = csr->m;
    g->flags = flags;
    for (uint32_t v = 0; v < csr->n; ++v)
        g->max_out_deg = std::max(g->max_out_deg, csr->row_offsets[v + 1] - csr->row_offsets[v]);
    if (csr->weights)
        for (uint32_t e = 0; e < csr->m; ++e) g->max_w = std::max(g->max_w, csr->weights[e]);
    auto fail = [&](g1_status s) {
        graph_free_all(g);
        delete g;
        return s;
    };
    g1_status s;
    const size_t n1 = static_cast<size_t>(csr->n) + 1;
    // Column and weight arrays carry kCsrPad zero entries past m: kernels that read a fixed run of
    // entries past a tile's end (masking the extra ones) need no bounds branch.
    const std::vector<uint32_t> zeros(kCsrPad, 0u);
    auto upload_padded = [&](uint32_t** dst, const uint32_t* src) -> g1_status {
        G1_TRY(alloc_n(g, dst, static_cast<size_t>(csr->m) + kCsrPad));
        G1_CHECK(!csr->m || hipMemcpy(*dst, src, size_t(csr->m) * 4, hipMemcpyHostToDevice) == hipSuccess, G1_ERR_HIP,
                 "g1_graph_create: upload failed");
        G1_CHECK(hipMemcpy(*dst + csr->m, zeros.data(), kCsrPad * 4, hipMemcpyHostToDevice) == hipSuccess, G1_ERR_HIP,
                 "g1_graph_create: upload failed");
        return G1_OK;
    };
    if ((s = alloc_n(g, &g->row_off, n1)) != G1_OK) return fail(s);
    if (hipMemcpy(g->row_off, csr->row_offsets, n1 * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(G1_ERR_HIP, "g1_graph_create: upload failed"));
    if ((s = upload_padded(&g->col, csr->col_indices)) != G1_OK) return fail(s);
    if (csr->weights && (s = upload_padded(&g->w, csr->weights)) != G1_OK) return fail(s);
    if (flags & G1_GRAPH_SYMMETRIC) {
        g->has_t = true;
        g->t_alias = true;
        g->t_row_off = g->row_off;
        g->t_col = g->col;
        g->t_w = g->w;
    } else if (flags & G1_GRAPH_TRANSPOSE) {
        // the in-edge CSR is built on the device (convert.hip) into padded arrays; its work areas are freed
        // before the graph is returned
        g->has_t = true;
        const size_t mp = static_cast<size_t>(csr->m) + kCsrPad;
        if ((s = alloc_n(g, &g->t_row_off, n1)) != G1_OK || (s = alloc_n(g, &g->t_col, mp)) != G1_OK ||
            (csr->weights && (s = alloc_n(g, &g->t_w, mp)) != G1_OK))
            return fail(s);
        if ((s = g1_fill_u32(g->t_col + csr->m, 0u, kCsrPad, nullptr)) != G1_OK ||
            (g->t_w && (s = g1_fill_u32(g->t_w + csr->m, 0u, kCsrPad, nullptr)) != G1_OK))
            return fail(s);
        void* tws = nullptr;
        if (hipMalloc(&tws, transpose_workspace_bytes(csr->n, csr->m, csr->weights != nullptr)) != hipSuccess)
            return fail(set_error(G1_ERR_OUT_OF_MEMORY, "g1_graph_create: transpose work areas"));
        s = device_transpose(csr->n, csr->m, g->row_off, g->col, g->w, g->t_row_off, g->t_col, g->t_w, tws, nullptr);
        const hipError_t sync = hipDeviceSynchronize();
        (void)hipFree(tws);
        if (s != G1_OK) return fail(s);
        if (sync != hipSuccess) return fail(set_error(G1_ERR_HIP, "g1_graph_create: transpose failed"));
    }
    if (g->has_t) {
        g->t_ws_bytes = merge_workspace_bytes(g->n, g->m);
        if ((s = graph_alloc(g, &g->t_ws, g->t_ws_bytes)) != G1_OK) return fail(s);
        if ((s = launch_merge_partition(g->n, g->m, g->t_row_off, g->t_ws, nullptr)) != G1_OK) return fail(s);
    }
    if ((s = graph_common(g, "g1_graph_create")) != G1_OK) return fail(s);
    if (hipDeviceSynchronize() != hipSuccess) return fail(set_error(G1_ERR_HIP, "g1_graph_create: sync failed"));
    *out = g;
    return G1_OK;
}

g1_status g1_graph_create_view(g1_graph* base, g1_graph** out) {
    using namespace g1;
    G1_CHECK(base != nullptr && out != nullptr, G1_ERR_INVALID_ARG, "g1_graph_create_view: null argument");
    *out = nullptr;
    {
        // the arrays a view shares are built in the base first (float weights, out-weight sums)
        std::lock_guard<std::mutex> lock(base->mu);
        if (base->has_t) G1_TRY(ensure_pull_work(base, true));
    }
    g1_graph* g = new g1_graph();
    g->n = base->n;
    g->m = base->m;
    g->flags = base->flags;
    g->max_out_deg = base->max_out_deg;
    g->max_w = base->max_w;
    g->row_off = base->row_off;
    g->col = base->col;
    g->w = base->w;
    g->t_row_off = base->t_row_off;
    g->t_col = base->t_col;
    g->t_w = base->t_w;
    g->t_wf = base->t_wf;
    g->has_t = base->has_t;
    g->is_view = true;
    g->t_alias = base->t_alias;
    g->t_ws_bytes = base->t_ws_bytes;  // own copy below: the workspace also holds per-call tile records
    g->t_starts = base->t_starts;
    g->osum = base->osum;
    g->best_k[0] = base->best_k[0];
    g->best_k[1] = base->best_k[1];
    auto fail = [&](g1_status s) {
        graph_free_all(g);
        delete g;
        return s;
    };
    g1_status s;
    if (base->t_ws) {
        // The merge-path workspace holds the partition (read-only) and the tile records every SpMV writes:
        // concurrent runs on two handles need their own records, so the view copies the whole workspace.
        if ((s = graph_alloc(g, &g->t_ws, g->t_ws_bytes)) != G1_OK) return fail(s);
        if (hipMemcpy(g->t_ws, base->t_ws, g->t_ws_bytes, hipMemcpyDeviceToDevice) != hipSuccess)
            return fail(set_error(G1_ERR_HIP, "g1_graph_create_view: workspace copy failed"));
    }
    if ((s = graph_common(g, "g1_graph_create_view")) != G1_OK) return fail(s);
    *out = g;
    return G1_OK;
}

g1_status g1_graph_destroy(g1_graph* g) {
    if (!g) return G1_OK;
    (void)hipDeviceSynchronize();
    if (g->hub) g1_graph_destroy(g->hub);
    g1::batch_release(g);
    g1::graph_free_all(g);
    delete g;
    return G1_OK;
}

g1_status g1_profile_enable(g1_graph* g, int32_t enable) {
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "g1_profile_enable: null graph");
    std::lock_guard<std::mutex> lock(g->mu);
    g->profile = enable != 0;
    return G1_OK;
}

g1_status g1_profile_read(g1_graph* g, g1_kernel_profile* out, uint32_t capacity, uint32_t* count) {
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "g1_profile_read: null graph");
    std::lock_guard<std::mutex> lock(g->mu);
    G1_TRY(g1::prof_flush(g));
    const uint32_t k = static_cast<uint32_t>(g->prof.size());
    if (count) *count = k;
    for (uint32_t i = 0; i < k && i < capacity && out; ++i) {
        memset(&out[i], 0, sizeof(out[i]));
        snprintf(out[i].name, sizeof(out[i].name), "%s", g->prof[i].name.c_str());
        out[i].launches = g->prof[i].launches;
        out[i].total_ms = g->prof[i].ms;
        out[i].algorithmic_bytes = g->prof[i].bytes;
    }
    g->prof.clear();
    return G1_OK;
}

// ------------------------------------------------------------------ work areas: size query and prepare

namespace g1 {

// Device bytes each ensure_* call would allocate now (0 for what the handle already holds). The long-row
// lists of the CSR-vector engine depend on the data: counted at their bound, 4 bytes per vertex.
static uint64_t push_work_bytes(const g1_graph* g) {
    if (g->q[0]) return 0;
    const uint64_t n1 = g->n ? g->n : 1;
    const uint64_t words = (n1 + 31) / 32 + 1;
    const uint64_t items = g->m > g->n ? g->m : g->n;
    const uint64_t cap = push_segment_capacity(n1, items);
    return 4 * cap * kSegments * 4 + 2 * words * 4 + n1 * 4 + 3 * words * 4 + n1 * 8 + (n1 + 256) +
           2ull * kHistIters * 8 + 3ull * kSegments * 8 + 4ull * kMaxPushBlocks * 8 + 4 * 4;
}
static uint64_t long_rows_bound(const g1_graph* g) { return 4ull * (g->n ? g->n : 1); }
static uint64_t pull_work_bytes(const g1_graph* g) {
    const uint64_t n1 = g->n ? g->n : 1;
    uint64_t b = g->xu ? 0 : 2 * n1 * 4;
    if (g->has_t && !g->t_long) b += long_rows_bound(g);
    return b;
}
static uint64_t rank_work_bytes(const g1_graph* g) {
    const uint64_t n1 = g->n ? g->n : 1;
    uint64_t b = g->xu ? 0 : 2 * n1 * 4;
    if (!g->xf) b += 5 * n1 * 4;
    if (!g->osum) b += n1 * 4;
    if (g->t_w && !g->t_wf) b += (static_cast<uint64_t>(g->m) + kCsrPad) * 4;
    if (!g->rpart) b += (static_cast<uint64_t>(rank_blocks(static_cast<uint32_t>(n1))) * kBatchCols + kBatchCols + 8) * 8 +
                        kBatchCols * 4;
    return b;
}
static uint64_t relabel_work_bytes(const g1_graph* g) {
    if (g->pr_perm || !g->has_t || g->is_view || g->n == 0) return 0;
    const uint64_t n1 = g->n, m = g->m;
    uint64_t b = n1 * 4 + (n1 + 1) * 4 + (m + kCsrPad) * 4 + n1 * 4 + merge_workspace_bytes(g->n, g->m);
    if (g->t_w) b += (m + kCsrPad) * 4;
    return b + long_rows_bound(g);
}
static uint64_t delta_work_bytes(const g1_graph* g) {
    if (g->dl_col || !g->w) return 0;
    const uint64_t m1 = g->m ? g->m : 1, n1 = g->n ? g->n : 1;
    return 2 * m1 * 4 + n1 * 4 + kMaxPushBlocks * 4 + (static_cast<uint64_t>(g->n) + 3) / 4 * 4 + 4;
}

}  // namespace g1

g1_status g1_graph_work_bytes(const g1_graph* g, int32_t work, uint64_t* bytes) {
    using namespace g1;
    G1_CHECK(g != nullptr && bytes != nullptr, G1_ERR_INVALID_ARG, "g1_graph_work_bytes: null argument");
    G1_CHECK((work & ~G1_WORK_ALL) == 0 && work != 0, G1_ERR_INVALID_ARG, "g1_graph_work_bytes: unknown work set %d",
             work);
    if (work & G1_WORK_RELABEL) work |= G1_WORK_RANK;  // the relabel needs the rank vectors' out-weight sums
    std::lock_guard<std::mutex> lock(const_cast<g1_graph*>(g)->mu);
    uint64_t b = 0;
    const bool push = (work & (G1_WORK_PUSH | G1_WORK_ROAD)) != 0;
    if (push) b += push_work_bytes(g);
    if (work & G1_WORK_ROAD) b += road_work_size(g);
    if (work & G1_WORK_RANK) b += rank_work_bytes(g);
    if ((work & G1_WORK_PULL) && !(work & G1_WORK_RANK)) b += pull_work_bytes(g);
    if ((work & G1_WORK_PULL) && (work & G1_WORK_RANK) && g->has_t && !g->t_long) b += long_rows_bound(g);
    if (work & G1_WORK_RELABEL) b += relabel_work_bytes(g);
    if (work & G1_WORK_DELTA) b += delta_work_bytes(g);
    if ((work & G1_WORK_HUB) && hub_eligible(g)) {
        // the copy and, in it, the work areas of the two engines that run there (push, pull, long rows,
        // delta split)
        g1_graph probe;
        const g1_graph* h = g->hub;
        if (!h) {
            b += hub_base_bytes(g);
            probe.n = g->n;
            probe.m = g->m;
            probe.w = g->w;
            probe.has_t = g->has_t;
            h = &probe;
        }
        b += push_work_bytes(h) + delta_work_bytes(h);
        if (h->has_t) b += pull_work_bytes(h);
    }
    *bytes = b;
    return G1_OK;
}

g1_status g1_graph_prepare(g1_graph* g, int32_t work, g1_stream_t stream) {
    using namespace g1;
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "g1_graph_prepare: null graph");
    G1_CHECK((work & ~G1_WORK_ALL) == 0 && work != 0, G1_ERR_INVALID_ARG, "g1_graph_prepare: unknown work set %d",
             work);
    if (work & G1_WORK_RELABEL) work |= G1_WORK_RANK;
    std::lock_guard<std::mutex> lock(g->mu);
    const hipStream_t s = as_stream(stream);
    if (work & (G1_WORK_PUSH | G1_WORK_ROAD)) G1_TRY(ensure_push_work(g, s));
    if (work & G1_WORK_ROAD) G1_TRY(prepare_road_work(g, s));
    if (work & (G1_WORK_PULL | G1_WORK_RANK)) G1_TRY(ensure_pull_work(g, (work & G1_WORK_RANK) != 0));
    if (work & G1_WORK_RANK) G1_TRY(ensure_rank_partials(g));
    if ((work & G1_WORK_PULL) && g->has_t) G1_TRY(ensure_long_rows(g, false, s));
    if ((work & G1_WORK_RELABEL) && g->has_t && !g->is_view && g->n > 0) {
        G1_TRY(ensure_pr_relabel(g, s));
        G1_TRY(ensure_long_rows(g, true, s));
    }
    if ((work & G1_WORK_DELTA) && g->w) G1_TRY(ensure_delta_split(g, delta_width(g), s));
    if ((work & G1_WORK_HUB) && hub_eligible(g)) {
        G1_TRY(ensure_hub_copy(g, s));
        g1_graph* h = g->hub;
        G1_TRY(ensure_push_work(h, s));
        if (h->has_t) {
            G1_TRY(ensure_pull_work(h, false));
            G1_TRY(ensure_long_rows(h, false, s));
        }
        if (h->w) G1_TRY(ensure_delta_split(h, delta_width(h), s));
    }
    G1_HIP(hipStreamSynchronize(s));
    return G1_OK;
}

g1_status g1_graph_info(const g1_graph* g, uint32_t* n, uint32_t* m, uint64_t* device_bytes) {
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "g1_graph_info: null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (device_bytes) {  // (the lock: a run may be building the hub-first copy)
        std::lock_guard<std::mutex> lock(const_cast<g1_graph*>(g)->mu);
        *device_bytes = g->device_bytes + (g->hub ? g->hub->device_bytes : 0);
    }
    return G1_OK;
}

g1_status g1_graph_csr_dev(const g1_graph* g, int32_t transpose, const uint32_t** row_offsets,
                           const uint32_t** col_indices, const uint32_t** weights) {
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "g1_graph_csr_dev: null graph");
    if (transpose) {
        G1_CHECK(g->has_t, G1_ERR_INVALID_ARG, "g1_graph_csr_dev: graph was created without its transpose");
        if (row_offsets) *row_offsets = g->t_row_off;
        if (col_indices) *col_indices = g->t_col;
        if (weights) *weights = g->t_w;
    } else {
        if (row_offsets) *row_offsets = g->row_off;
        if (col_indices) *col_indices = g->col;
        if (weights) *weights = g->w;
    }
    return G1_OK;
}

}  // extern "C"
