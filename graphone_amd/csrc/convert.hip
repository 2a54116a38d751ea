// Host side of the on-device CSR conversions (kernels in g1/convert.hpp): the transpose every graph with
// G1_GRAPH_TRANSPOSE builds at creation, its C-ABI export, and plain PageRank's hub-first relabelled copy of
// the in-edge CSR. Replaces mat::ToCsc (include/GraphOne/mat/convert.h:62-88), used by graph::build for the
// transpose view (include/GraphOne/graph/builder.h:9-46).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "g1/convert.hpp"
#include "g1/spmv_merge.hpp"
#include "graph.hpp"
#include "spmv_launch.hpp"

namespace g1 {
namespace {

uint32_t blocks_for(uint64_t threads) {
    const uint64_t b = (threads + 255) / 256;
    return static_cast<uint32_t>(b ? b : 1);
}

// Radix passes for keys below n: one per 8 bits of n - 1, at least one.
uint32_t radix_passes(uint32_t n) {
    uint32_t bits = 0;
    while (bits < 32 && n > 1 && ((n - 1) >> bits) != 0) ++bits;
    return std::max<uint32_t>(1u, (bits + 7) / 8);
}

uint32_t sort_tiles(uint32_t count) { return std::max<uint32_t>(1u, (count + kSortTile - 1) / kSortTile); }

size_t scan_partials_count(uint64_t count) { return std::max<uint64_t>(1, (count + kScanTile - 1) / kScanTile); }

// Bytes of the histogram and its scan partials for sorting `count` keys.
size_t sort_aux_bytes(uint32_t count) {
    const size_t hist = static_cast<size_t>(kSortBins) * sort_tiles(count);
    return (hist + scan_partials_count(hist)) * sizeof(uint32_t);
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

g1_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* partials, hipStream_t s) {
    if (count == 0) return G1_OK;
    const uint32_t nb = static_cast<uint32_t>(scan_partials_count(count));
    scan_partials_kernel<<<nb, 256, 0, s>>>(in, count, partials);
    G1_LAUNCHED();
    scan_carry_kernel<<<1, 1024, 0, s>>>(partials, nb);
    G1_LAUNCHED();
    scan_apply_kernel<<<nb, 256, 0, s>>>(in, count, partials, out);
    G1_LAUNCHED();
    return G1_OK;
}

// Stable LSD radix sort of `count` (key, id[, weight]) triples by key (keys below key_limit). Pass p reads
// in[p] and writes out[p]; the caller lays out the ping-pong so the last pass lands where it wants.
// aux holds sort_aux_bytes(count).
g1_status radix_sort_pairs(uint32_t count, uint32_t key_limit, const uint32_t* k0, const uint32_t* u0,
                           const uint32_t* w0, uint32_t* kbuf[2], uint32_t* ubuf[2], uint32_t* wbuf[2],
                           uint32_t* u_last, uint32_t* w_last, uint32_t* aux, const uint32_t** k_sorted,
                           hipStream_t s) {
    const uint32_t passes = radix_passes(key_limit);
    const uint32_t nb = sort_tiles(count);
    uint32_t* hist = aux;
    uint32_t* partials = aux + static_cast<size_t>(kSortBins) * nb;
    const uint32_t *kin = k0, *uin = u0, *win = w0;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = 8 * p;
        uint32_t* kout = kbuf[p & 1];
        uint32_t* uout = (p + 1 == passes) ? u_last : ubuf[p & 1];
        uint32_t* wout = w0 ? ((p + 1 == passes) ? w_last : wbuf[p & 1]) : nullptr;
        radix_hist_kernel<<<nb, kSortThreads, 0, s>>>(kin, count, shift, nb, hist);
        G1_LAUNCHED();
        G1_TRY(exclusive_scan_u32(hist, hist, kSortBins * nb, partials, s));
        if (w0) radix_scatter_kernel<true><<<nb, kSortThreads, 0, s>>>(kin, uin, win, count, shift, nb, hist, kout, uout, wout);
        else radix_scatter_kernel<false><<<nb, kSortThreads, 0, s>>>(kin, uin, nullptr, count, shift, nb, hist, kout, uout, nullptr);
        G1_LAUNCHED();
        kin = kout;
        uin = uout;
        win = wout;
    }
    *k_sorted = kin;
    return G1_OK;
}

// The transpose's work areas: the expanded rows, two key and one more id buffer for the passes' ping-pong,
// two weight buffers when weighted, the histogram and its scan partials. Depends on m only.
size_t transpose_workspace_bytes(uint32_t /*n*/, uint32_t m, bool weighted) {
    const size_t arr = align256(static_cast<size_t>(m ? m : 1) * sizeof(uint32_t));
    return arr * (weighted ? 6 : 4) + align256(sort_aux_bytes(m));
}

// In-edge CSR of (row_off, col, w): t_row_off[n + 1], t_col[m], t_w[m] (when w). See g1/convert.hpp.
g1_status device_transpose(uint32_t n, uint32_t m, const uint32_t* row_off, const uint32_t* col, const uint32_t* w,
                           uint32_t* t_row_off, uint32_t* t_col, uint32_t* t_w, void* ws, hipStream_t s) {
    if (m == 0) {
        G1_TRY(g1_fill_u32(t_row_off, 0u, static_cast<size_t>(n) + 1, reinterpret_cast<g1_stream_t>(s)));
        return G1_OK;
    }
    const size_t arr = align256(static_cast<size_t>(m) * sizeof(uint32_t));
    uint8_t* p = static_cast<uint8_t*>(ws);
    auto take = [&]() {
        uint32_t* a = reinterpret_cast<uint32_t*>(p);
        p += arr;
        return a;
    };
    uint32_t* expanded = take();
    uint32_t* kbuf[2] = {take(), take()};
    uint32_t* ubuf[2] = {take(), expanded};  // the expanded rows are dead after the first pass
    uint32_t* wbuf[2] = {nullptr, nullptr};
    if (w) {
        wbuf[0] = take();
        wbuf[1] = take();
    }
    uint32_t* aux = reinterpret_cast<uint32_t*>(p);
    csr_expand_rows_kernel<<<blocks_for((static_cast<uint64_t>(m) + kExpandPer - 1) / kExpandPer), 256, 0, s>>>(
        n, m, row_off, expanded);
    G1_LAUNCHED();
    const uint32_t* sorted = nullptr;
    G1_TRY(radix_sort_pairs(m, n, col, expanded, w, kbuf, ubuf, wbuf, t_col, t_w, aux, &sorted, s));
    row_starts_kernel<<<std::min<uint32_t>(blocks_for(static_cast<uint64_t>(n) + 1), 8192), 256, 0, s>>>(n, m, sorted,
                                                                                                       t_row_off);
    G1_LAUNCHED();
    return G1_OK;
}

// float copy of the in-edge weights (plus-times engines)
g1_status device_u32_to_f32(const uint32_t* in, float* out, size_t count, hipStream_t s) {
    if (count == 0) return G1_OK;
    u32_to_f32_kernel<<<std::min<uint32_t>(blocks_for(count), 8192), 256, 0, s>>>(count, in, out);
    G1_LAUNCHED();
    return G1_OK;
}

// The relabelled copy of the in-edge CSR for plain PageRank, built on the device: vertices ordered by
// out-degree, largest first (ties by id), so the sources gathered most often -- the hubs -- share cache
// lines; a row keeps its in-edges in their original order. Needs the out-weight sums (ensure_pull_work).
// Temporary arrays (about 20 bytes per vertex) are freed before it returns.
g1_status ensure_pr_relabel(g1_graph* g, hipStream_t s) {
    if (g->pr_perm || !g->has_t) return G1_OK;
    const uint32_t n = g->n, m = g->m;
    const size_t nn = std::max<size_t>(n, 1);
    G1_TRY(graph_alloc(g, reinterpret_cast<void**>(&g->pr_perm), nn * sizeof(uint32_t)));
    G1_TRY(graph_alloc(g, reinterpret_cast<void**>(&g->pr_row_off), (static_cast<size_t>(n) + 1) * sizeof(uint32_t)));
    G1_TRY(graph_alloc(g, reinterpret_cast<void**>(&g->pr_col), (static_cast<size_t>(m) + kCsrPad) * sizeof(uint32_t)));
    G1_TRY(graph_alloc(g, reinterpret_cast<void**>(&g->pr_osum), nn * sizeof(float)));
    if (g->t_wf)
        G1_TRY(graph_alloc(g, reinterpret_cast<void**>(&g->pr_wf), (static_cast<size_t>(m) + kCsrPad) * sizeof(float)));
    const g1_stream_t gs = reinterpret_cast<g1_stream_t>(s);
    G1_TRY(g1_fill_u32(g->pr_col + m, 0u, kCsrPad, gs));
    if (g->pr_wf) G1_TRY(g1_fill_f32(g->pr_wf + m, 0.0f, kCsrPad, gs));
    if (n > 0) {
        const size_t arr = align256(nn * sizeof(uint32_t));
        const size_t bytes = arr * 5 + align256((nn + 1) * sizeof(uint32_t)) + align256(sort_aux_bytes(n));
        void* tmp = nullptr;
        G1_HIP(hipMalloc(&tmp, bytes));
        uint8_t* p = static_cast<uint8_t*>(tmp);
        auto take = [&](size_t b) {
            uint32_t* a = reinterpret_cast<uint32_t*>(p);
            p += b;
            return a;
        };
        uint32_t* keys = take(arr);
        uint32_t* ids = take(arr);
        uint32_t* kbuf[2] = {take(arr), take(arr)};
        uint32_t* ubuf[2] = {take(arr), ids};
        uint32_t* wbuf[2] = {nullptr, nullptr};
        uint32_t* lens = take(align256((nn + 1) * sizeof(uint32_t)));
        uint32_t* aux = reinterpret_cast<uint32_t*>(p);
        const uint32_t maxd = g->max_out_deg;
        g1_status st = G1_OK;
        const uint32_t grid = std::min<uint32_t>(blocks_for(static_cast<uint64_t>(n) + 1), 8192);
        degree_keys_kernel<<<grid, 256, 0, s>>>(n, maxd, g->row_off, keys, ids);
        const uint32_t* sorted = nullptr;
        if (hipGetLastError() != hipSuccess) st = set_error(G1_ERR_HIP, "relabel: degree keys launch failed");
        if (st == G1_OK)
            st = radix_sort_pairs(n, maxd + 1, keys, ids, nullptr, kbuf, ubuf, wbuf, g->pr_perm, nullptr, aux, &sorted, s);
        if (st == G1_OK) {
            // inverse permutation (into keys, free now), new row lengths, out-weight sums in the new ids
            relabel_rows_kernel<<<grid, 256, 0, s>>>(n, g->pr_perm, g->t_row_off, g->osum, keys, lens, g->pr_osum);
            if (hipGetLastError() != hipSuccess) st = set_error(G1_ERR_HIP, "relabel: row launch failed");
        }
        if (st == G1_OK) st = exclusive_scan_u32(lens, g->pr_row_off, n + 1, aux, s);
        if (st == G1_OK && m > 0) {
            relabel_copy_kernel<float><<<blocks_for((static_cast<uint64_t>(m) + kExpandPer - 1) / kExpandPer), 256, 0, s>>>(
                n, m, g->pr_row_off, g->pr_perm, keys, g->t_row_off, g->t_col, g->t_wf, g->pr_col, g->pr_wf);
            if (hipGetLastError() != hipSuccess) st = set_error(G1_ERR_HIP, "relabel: copy launch failed");
        }
        const hipError_t sync = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        G1_TRY(st);
        G1_HIP(sync);
    } else {
        G1_TRY(g1_fill_u32(g->pr_row_off, 0u, 1, gs));
    }
    G1_TRY(graph_alloc(g, &g->pr_ws, merge_workspace_bytes(n, m)));
    G1_TRY(launch_merge_partition(n, m, g->pr_row_off, g->pr_ws, s));
    G1_HIP(hipStreamSynchronize(s));
    return G1_OK;
}

// The push engines' hub-first working copy of the out-edge CSR (ensure_hub_copy): vertices ordered by
// out-degree group, largest first (group k: degrees in [2^k - 1, 2^(k+1) - 1)), and by id within a group:
// vertex v is new id inv[v] (device) and inv_host[v] (host array of n). The hubs still come first and share
// cache lines, and a gather in the original ids (the results mapped back) reads a few dozen ascending
// streams rather than one random line per vertex, as an exact degree order gives. The rows are
// copied renamed, then transposed: the transpose is a stable sort by column, so each row of its output
// lists its entries by ascending source. Transposing that once more gives the out-edge CSR with each row's
// columns ascending (row_off, col, w); the intermediate is the copy's in-edge CSR (t_row_off, t_col, t_w)
// when those are given. For a symmetric graph the first transpose already is the sorted out-edge CSR:
// pass symmetric = true and it lands in (row_off, col, w). Temporary arrays are freed before it returns.
g1_status build_hub_csr(const g1_graph* g, bool symmetric, uint32_t* inv, uint32_t* inv_host, uint32_t* row_off, uint32_t* col, uint32_t* w,
                        uint32_t* t_row_off, uint32_t* t_col, uint32_t* t_w, hipStream_t s) {
    const uint32_t n = g->n, m = g->m;
    const bool weighted = g->w != nullptr;
    const bool keep_t = t_row_off != nullptr;
    const size_t arr = align256(std::max<size_t>(n, 1) * sizeof(uint32_t));
    const size_t n1 = align256((static_cast<size_t>(n) + 1) * sizeof(uint32_t));
    const size_t marr = align256((static_cast<size_t>(m) + kCsrPad) * sizeof(uint32_t));
    // perm sort: keys, ids, 2 key + 1 id ping-pong buffers, perm, lens, sort aux; the renamed rows (n + 1
    // offsets, columns, weights); the intermediate transpose when it is not kept; the transpose work areas
    const size_t sort_bytes = arr * 6 + n1 + align256(sort_aux_bytes(n));  // (the sixth: perm)
    const size_t rows_bytes = n1 + marr * (weighted ? 2 : 1);
    const size_t mid_bytes = (symmetric || keep_t) ? 0 : rows_bytes;
    const size_t bytes = sort_bytes + rows_bytes + mid_bytes + transpose_workspace_bytes(n, m, weighted);
    void* tmp = nullptr;
    G1_HIP(hipMalloc(&tmp, bytes));
    uint8_t* p = static_cast<uint8_t*>(tmp);
    auto take = [&](size_t b) {
        uint32_t* a = reinterpret_cast<uint32_t*>(p);
        p += b;
        return a;
    };
    uint32_t* keys = take(arr);
    uint32_t* ids = take(arr);
    uint32_t* kbuf[2] = {take(arr), take(arr)};
    uint32_t* ubuf[2] = {take(arr), ids};
    uint32_t* wbuf[2] = {nullptr, nullptr};
    uint32_t* perm = take(arr);
    uint32_t* lens = take(n1);
    uint32_t* aux = take(align256(sort_aux_bytes(n)));
    uint32_t* r_off = take(n1);
    uint32_t* r_col = take(marr);
    uint32_t* r_w = weighted ? take(marr) : nullptr;
    if (!symmetric && !keep_t) {
        t_row_off = take(n1);
        t_col = take(marr);
        t_w = weighted ? take(marr) : nullptr;
    }
    void* tws = p;
    if (symmetric) {  // the first transpose is the result
        t_row_off = row_off;
        t_col = col;
        t_w = w;
    }
    g1_status st = G1_OK;
    const uint32_t grid = std::min<uint32_t>(blocks_for(static_cast<uint64_t>(n) + 1), 8192);
    const uint32_t* sorted = nullptr;
    degree_group_keys_kernel<<<grid, 256, 0, s>>>(n, g->row_off, keys, ids);
    if (hipGetLastError() != hipSuccess) st = set_error(G1_ERR_HIP, "hub copy: degree keys launch failed");
    if (st == G1_OK) st = radix_sort_pairs(n, 33, keys, ids, nullptr, kbuf, ubuf, wbuf, perm, nullptr, aux, &sorted, s);
    if (st == G1_OK) {
        hub_rows_kernel<<<grid, 256, 0, s>>>(n, perm, g->row_off, inv, lens);
        if (hipGetLastError() != hipSuccess) st = set_error(G1_ERR_HIP, "hub copy: row launch failed");
    }
    if (st == G1_OK) st = exclusive_scan_u32(lens, r_off, n + 1, aux, s);
    if (st == G1_OK && m > 0) {
        relabel_copy_kernel<uint32_t><<<blocks_for((static_cast<uint64_t>(m) + kExpandPer - 1) / kExpandPer), 256, 0, s>>>(
            n, m, r_off, perm, inv, g->row_off, g->col, g->w, r_col, r_w);
        if (hipGetLastError() != hipSuccess) st = set_error(G1_ERR_HIP, "hub copy: copy launch failed");
    }
    if (st == G1_OK) st = device_transpose(n, m, r_off, r_col, r_w, t_row_off, t_col, t_w, tws, s);
    if (st == G1_OK && !symmetric) st = device_transpose(n, m, t_row_off, t_col, t_w, row_off, col, w, tws, s);
    if (st == G1_OK && hipMemcpyAsync(inv_host, inv, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        st = set_error(G1_ERR_HIP, "hub copy: map download failed");
    const hipError_t sync = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    G1_TRY(st);
    G1_HIP(sync);
    return G1_OK;
}

}  // namespace g1

extern "C" {

size_t g1_csr_transpose_workspace_bytes(uint32_t n, uint32_t m, int32_t weighted) {
    return g1::transpose_workspace_bytes(n, m, weighted != 0);
}

g1_status g1_csr_transpose(uint32_t n, uint32_t m, const uint32_t* row_offsets, const uint32_t* col_indices,
                           const uint32_t* weights, uint32_t* t_row_offsets, uint32_t* t_col_indices,
                           uint32_t* t_weights, void* workspace, size_t workspace_bytes, g1_stream_t stream) {
    using namespace g1;
    G1_CHECK(row_offsets && t_row_offsets, G1_ERR_INVALID_ARG, "g1_csr_transpose: null row offsets");
    G1_CHECK(m == 0 || (col_indices && t_col_indices), G1_ERR_INVALID_ARG, "g1_csr_transpose: null column indices");
    G1_CHECK((weights == nullptr) == (t_weights == nullptr), G1_ERR_INVALID_ARG,
             "g1_csr_transpose: weights and t_weights must both be given or both be null");
    G1_CHECK(m < kInf && n < kInf, G1_ERR_INVALID_ARG, "g1_csr_transpose: sizes must be below 2^32-1");
    const size_t need = transpose_workspace_bytes(n, m, weights != nullptr);
    G1_CHECK(m == 0 || (workspace && workspace_bytes >= need), G1_ERR_INVALID_ARG,
             "g1_csr_transpose: workspace of %zu bytes, %zu needed", workspace_bytes, need);
    return device_transpose(n, m, row_offsets, col_indices, weights, t_row_offsets, t_col_indices, t_weights,
                            workspace, as_stream(stream));
}

}  // extern "C"
