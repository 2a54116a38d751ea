// C-ABI exports of the step-level primitives (kernels in g1/step.hpp): the pieces a caller needs to run
// the reference's AdvanceGC / BlasGcEngine loops on its own buffers, one stream-ordered step at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "g1/apply.hpp"
#include "g1/convert.hpp"
#include "g1/semiring.hpp"
#include "g1/spmv_merge.hpp"
#include "g1/step.hpp"
#include "graph.hpp"
#include "internal.hpp"
#include "spmv_launch.hpp"

namespace g1 {
namespace {

constexpr uint32_t kStepBlocks = 2048;  // grid of the grid-stride step kernels

size_t al(size_t b) { return (b + 255) & ~size_t(255); }
uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Workspace layout (every region 256-byte aligned):
//   u32 offsets  : capacity + 2           (advance / density: the frontier's degree prefix)
//   u32 counts   : 2 * blocks + 2          (per-block counts and their scan: compaction, bitmap words, apply)
//   f64 partials : blocks + 2              (L1 partials)
//   u32 scan     : scan partials of the largest scan
// blocks = max(ceil(capacity / 256), ceil(n / 32), ceil(n / 256)).
struct StepWs {
    uint32_t* off;
    uint32_t* cnt;
    uint32_t* pos;
    double* part;
    uint32_t* scan;
};

uint64_t step_blocks(uint32_t n, uint32_t capacity) {
    return std::max<uint64_t>({cdiv(capacity, 256), cdiv(n, 32), cdiv(n, 256), 1});
}

size_t step_ws_bytes(uint32_t n, uint32_t capacity) {
    const uint64_t blocks = step_blocks(n, capacity);
    const uint64_t scan_items = std::max<uint64_t>(capacity + 1ull, blocks);
    return al((capacity + 2ull) * 4) + al((2 * blocks + 2) * 4) + al((blocks + 2) * 8) +
           al((cdiv(scan_items, kScanTile) + 2) * 4);
}

StepWs step_ws(void* ws, uint32_t n, uint32_t capacity) {
    const uint64_t blocks = step_blocks(n, capacity);
    char* p = static_cast<char*>(ws);
    StepWs w;
    w.off = reinterpret_cast<uint32_t*>(p);
    p += al((capacity + 2ull) * 4);
    w.cnt = reinterpret_cast<uint32_t*>(p);
    w.pos = w.cnt + blocks + 1;
    p += al((2 * blocks + 2) * 4);
    w.part = reinterpret_cast<double*>(p);
    p += al((blocks + 2) * 8);
    w.scan = reinterpret_cast<uint32_t*>(p);
    return w;
}

g1_status check_ws(const void* ws, size_t bytes, uint32_t n, uint32_t capacity, const char* what) {
    G1_CHECK(ws != nullptr, G1_ERR_INVALID_ARG, "%s: null workspace", what);
    G1_CHECK(bytes >= step_ws_bytes(n, capacity), G1_ERR_INVALID_ARG,
             "%s: workspace of %zu bytes, %zu needed (g1_step_workspace_bytes)", what, bytes,
             step_ws_bytes(n, capacity));
    return G1_OK;
}

uint32_t grid(uint64_t items) { return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(cdiv(items, 256), 1), kStepBlocks)); }

// off = exclusive prefix of the frontier's degrees, off[capacity] = E; *edges (may be null) = E and
// *out_count (may be null) = 0.
g1_status frontier_offsets(uint32_t n, const uint32_t* row_off, const uint32_t* f, const uint32_t* count,
                           uint32_t capacity, const StepWs& w, uint32_t* edges, uint32_t* out_count, hipStream_t s) {
    if (capacity) {
        step_degrees_kernel<<<grid(capacity), 256, 0, s>>>(row_off, f, count, capacity, n, w.off);
        G1_LAUNCHED();
        // the last degree is kept beside the scan (the scan is in place)
        G1_HIP(hipMemcpyAsync(w.off + capacity + 1, w.off + capacity - 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        G1_TRY(exclusive_scan_u32(w.off, w.off, capacity, w.scan, s));
    }
    step_total_kernel<<<1, 1, 0, s>>>(w.off + capacity + 1, w.off, capacity, edges, out_count);
    G1_LAUNCHED();
    return G1_OK;
}

template <bool kSssp>
g1_status advance(uint32_t n, const uint32_t* row_off, const uint32_t* col, const uint32_t* wts, const uint32_t* f_in,
                  const uint32_t* in_count, uint32_t capacity, uint32_t* dist, uint32_t* visited, uint32_t iter,
                  uint32_t* f_out, uint32_t* out_count, void* ws, size_t ws_bytes, hipStream_t s, const char* what) {
    G1_CHECK(row_off && dist && out_count && f_out, G1_ERR_INVALID_ARG, "%s: null argument", what);
    G1_CHECK(capacity == 0 || (f_in && in_count), G1_ERR_INVALID_ARG, "%s: null frontier", what);
    G1_CHECK(!kSssp || visited, G1_ERR_INVALID_ARG, "%s: null visited array", what);
    G1_CHECK(capacity < kInf, G1_ERR_INVALID_ARG, "%s: capacity must be below 2^32-1", what);
    G1_TRY(check_ws(ws, ws_bytes, n, capacity, what));
    const StepWs w = step_ws(ws, n, capacity);
    G1_TRY(frontier_offsets(n, row_off, f_in, in_count, capacity, w, nullptr, out_count, s));
    if (capacity == 0) return G1_OK;
    StepRelax<kSssp> rel{col, wts, dist, visited, iter, f_out, out_count};
    step_advance_kernel<StepRelax<kSssp>><<<kStepBlocks, 256, 0, s>>>(row_off, f_in, in_count, w.off, capacity, rel);
    G1_LAUNCHED();
    return G1_OK;
}

// SpMM workspace: the merge-path partition of the matrix, then the tile records of one 64-column pass
// (sized for the larger record, the compensated plus-times one).
size_t spmm_ws_bytes(uint32_t n_rows, uint32_t nnz) {
    const size_t t = merge_num_tiles(n_rows, nnz);
    return al((t + 1) * sizeof(MergeCoord)) + t * sizeof(SpmmTileRec<PlusTimesKahanF32, 64>) + 256;
}

template <typename S>
g1_status spmm(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, uint32_t k, const uint32_t* row_off, const uint32_t* col,
               const typename S::value_t* vals, const typename S::x_t* X, typename S::y_t* Y,
               const unsigned long long* mask, void* ws, size_t ws_bytes, hipStream_t s, const char* what) {
    G1_CHECK(n_rows == 0 || k == 0 || (row_off && Y), G1_ERR_INVALID_ARG, "%s: null row_offsets or Y", what);
    G1_CHECK(nnz == 0 || k == 0 || (col && X), G1_ERR_INVALID_ARG, "%s: null col_indices or X", what);
    G1_CHECK(n_cols > 0 || nnz == 0, G1_ERR_INVALID_ARG, "%s: nnz > 0 with zero columns", what);
    G1_CHECK(nnz < kInf, G1_ERR_INVALID_ARG, "%s: nnz must be below 2^32-1", what);
    if (n_rows == 0 || k == 0) return G1_OK;
    G1_CHECK(ws != nullptr && ws_bytes >= spmm_ws_bytes(n_rows, nnz), G1_ERR_INVALID_ARG,
             "%s: workspace of %zu bytes, %zu needed (g1_spmm_workspace_bytes)", what, ws_bytes,
             spmm_ws_bytes(n_rows, nnz));
    const uint32_t tiles = merge_num_tiles(n_rows, nnz);
    G1_TRY(launch_merge_partition(n_rows, nnz, row_off, ws, s));
    const MergeCoord* coords = merge_coords(ws);
    auto* recs = reinterpret_cast<SpmmTileRec<S, 64>*>(static_cast<char*>(ws) + al((tiles + 1ull) * sizeof(MergeCoord)));
    using T = typename S::x_t;
    const T neutral = std::is_same_v<S, MinPlusU32> ? static_cast<T>(kInf) : T(0);
    for (uint32_t c0 = 0; c0 < k; c0 += 64) {
        StridedColumnsPolicy<T> pol{X, Y, k, c0, std::min<uint32_t>(64u, k - c0), mask, neutral};
        spmm_tile_kernel<S, 64, StridedColumnsPolicy<T>><<<tiles, 256, 0, s>>>(n_rows, nnz, row_off, col, vals, coords,
                                                                             recs, pol, nullptr);
        G1_LAUNCHED();
        if (tiles > 1) {
            spmm_fixup_kernel<S, 64, StridedColumnsPolicy<T>><<<(tiles * 64u + 255) / 256, 256, 0, s>>>(
                tiles, coords, recs, pol, nullptr);
            G1_LAUNCHED();
        }
    }
    return G1_OK;
}

}  // namespace
}  // namespace g1

using namespace g1;

extern "C" {

size_t g1_step_workspace_bytes(uint32_t n, uint32_t capacity) { return step_ws_bytes(n, capacity); }

g1_status g1_frontier_density(uint32_t n, const uint32_t* row_offsets, const uint32_t* frontier, const uint32_t* count,
                              uint32_t capacity, uint32_t* edges, void* ws, size_t ws_bytes, g1_stream_t stream) {
    G1_CHECK(row_offsets && edges && (capacity == 0 || (frontier && count)), G1_ERR_INVALID_ARG,
             "g1_frontier_density: null argument");
    G1_TRY(check_ws(ws, ws_bytes, n, capacity, "g1_frontier_density"));
    return frontier_offsets(n, row_offsets, frontier, count, capacity, step_ws(ws, n, capacity), edges, nullptr,
                            as_stream(stream));
}

g1_status g1_advance_sssp_u32(uint32_t n, const uint32_t* row_offsets, const uint32_t* col_indices,
                              const uint32_t* weights, const uint32_t* frontier_in, const uint32_t* in_count,
                              uint32_t in_capacity, uint32_t* dist, uint32_t* visited, uint32_t iter,
                              uint32_t* frontier_out, uint32_t* out_count, void* ws, size_t ws_bytes,
                              g1_stream_t stream) {
    return advance<true>(n, row_offsets, col_indices, weights, frontier_in, in_count, in_capacity, dist, visited, iter,
                         frontier_out, out_count, ws, ws_bytes, as_stream(stream), "g1_advance_sssp_u32");
}

g1_status g1_advance_bfs(uint32_t n, const uint32_t* row_offsets, const uint32_t* col_indices,
                         const uint32_t* frontier_in, const uint32_t* in_count, uint32_t in_capacity, uint32_t* depth,
                         uint32_t level, uint32_t* frontier_out, uint32_t* out_count, void* ws, size_t ws_bytes,
                         g1_stream_t stream) {
    return advance<false>(n, row_offsets, col_indices, nullptr, frontier_in, in_count, in_capacity, depth, nullptr,
                          level, frontier_out, out_count, ws, ws_bytes, as_stream(stream), "g1_advance_bfs");
}

g1_status g1_filter_compact(const uint32_t* in, const uint32_t* in_count, uint32_t capacity, uint32_t* out,
                            uint32_t* out_count, void* ws, size_t ws_bytes, g1_stream_t stream) {
    G1_CHECK(out_count && (capacity == 0 || (in && in_count && out)), G1_ERR_INVALID_ARG,
             "g1_filter_compact: null argument");
    G1_TRY(check_ws(ws, ws_bytes, 0, capacity, "g1_filter_compact"));
    const hipStream_t s = as_stream(stream);
    if (capacity == 0) return g1_fill_u32(out_count, 0u, 1, stream);
    const StepWs w = step_ws(ws, 0, capacity);
    const uint32_t nb = static_cast<uint32_t>(cdiv(capacity, 256));
    compact_count_kernel<<<nb, 256, 0, s>>>(in, in_count, capacity, w.cnt);
    G1_LAUNCHED();
    G1_TRY(exclusive_scan_u32(w.cnt, w.pos, nb, w.scan, s));
    compact_write_kernel<<<nb, 256, 0, s>>>(in, in_count, capacity, w.pos, w.cnt, out, out_count);
    G1_LAUNCHED();
    return G1_OK;
}

g1_status g1_frontier_queue_to_bitmap(uint32_t n, const uint32_t* frontier, const uint32_t* count, uint32_t capacity,
                                      uint32_t* bitmap, g1_stream_t stream) {
    G1_CHECK(bitmap && (capacity == 0 || (frontier && count)), G1_ERR_INVALID_ARG,
             "g1_frontier_queue_to_bitmap: null argument");
    G1_TRY(g1_fill_u32(bitmap, 0u, cdiv(n, 32), stream));
    if (capacity == 0) return G1_OK;
    queue_to_bitmap_kernel<<<grid(capacity), 256, 0, as_stream(stream)>>>(frontier, count, capacity, n, bitmap);
    G1_LAUNCHED();
    return G1_OK;
}

g1_status g1_frontier_bitmap_to_queue(uint32_t n, const uint32_t* bitmap, uint32_t* frontier, uint32_t* count,
                                      void* ws, size_t ws_bytes, g1_stream_t stream) {
    G1_CHECK(count && (n == 0 || (bitmap && frontier)), G1_ERR_INVALID_ARG, "g1_frontier_bitmap_to_queue: null argument");
    G1_TRY(check_ws(ws, ws_bytes, n, 0, "g1_frontier_bitmap_to_queue"));
    if (n == 0) return g1_fill_u32(count, 0u, 1, stream);
    const hipStream_t s = as_stream(stream);
    const StepWs w = step_ws(ws, n, 0);
    const uint32_t words = static_cast<uint32_t>(cdiv(n, 32));
    bitmap_popc_kernel<<<grid(words), 256, 0, s>>>(bitmap, n, w.cnt);
    G1_LAUNCHED();
    G1_TRY(exclusive_scan_u32(w.cnt, w.pos, words, w.scan, s));
    bitmap_write_kernel<<<grid(words), 256, 0, s>>>(bitmap, n, w.pos, w.cnt, frontier, count);
    G1_LAUNCHED();
    return G1_OK;
}

g1_status g1_apply_minplus_u32(uint32_t n, const uint32_t* y, uint32_t* d, uint32_t* x, uint32_t* changed, void* ws,
                               size_t ws_bytes, g1_stream_t stream) {
    G1_CHECK(changed && (n == 0 || (y && d && x)), G1_ERR_INVALID_ARG, "g1_apply_minplus_u32: null argument");
    G1_TRY(check_ws(ws, ws_bytes, n, 0, "g1_apply_minplus_u32"));
    if (n == 0) return g1_fill_u32(changed, 0u, 1, stream);
    const hipStream_t s = as_stream(stream);
    const StepWs w = step_ws(ws, n, 0);
    const uint32_t nb = static_cast<uint32_t>(cdiv(n, 256));
    step_apply_minplus_kernel<<<nb, 256, 0, s>>>(n, y, d, x, w.cnt);
    G1_LAUNCHED();
    sum_counts_kernel<<<1, 256, 0, s>>>(w.cnt, nb, changed);
    G1_LAUNCHED();
    return G1_OK;
}

g1_status g1_pr_apply_f32(uint32_t n, const float* y, float base, const float* iw, const float* r_old, float* r_new,
                          float* x, double* l1, void* ws, size_t ws_bytes, g1_stream_t stream) {
    G1_CHECK(n == 0 || (y && iw && r_new && x), G1_ERR_INVALID_ARG, "g1_pr_apply_f32: null argument");
    G1_CHECK(!l1 || r_old || n == 0, G1_ERR_INVALID_ARG, "g1_pr_apply_f32: the L1 norm needs r_old");
    if (l1) G1_TRY(check_ws(ws, ws_bytes, n, 0, "g1_pr_apply_f32"));
    const hipStream_t s = as_stream(stream);
    if (n == 0) return l1 ? g1_fill_u32(reinterpret_cast<uint32_t*>(l1), 0u, 2, stream) : G1_OK;  // 0.0
    const uint32_t nb = static_cast<uint32_t>(cdiv(n, 256));
    double* part = l1 ? step_ws(ws, n, 0).part : nullptr;
    step_pr_apply_kernel<false><<<nb, 256, 0, s>>>(n, y, base, iw, l1 ? r_old : nullptr, r_new, x, part);
    G1_LAUNCHED();
    if (l1) {
        reduce_partials_kernel<0><<<dim3(1, 1), 256, 0, s>>>(part, nb, l1);
        G1_LAUNCHED();
    }
    return G1_OK;
}

g1_status g1_l1_delta_f64(uint32_t n, const float* a, const float* b, double* out, void* ws, size_t ws_bytes,
                          g1_stream_t stream) {
    G1_CHECK(out && (n == 0 || (a && b)), G1_ERR_INVALID_ARG, "g1_l1_delta_f64: null argument");
    G1_TRY(check_ws(ws, ws_bytes, n, 0, "g1_l1_delta_f64"));
    const hipStream_t s = as_stream(stream);
    const uint32_t nb = static_cast<uint32_t>(std::max<uint64_t>(cdiv(n, 256), 1));
    const StepWs w = step_ws(ws, n, 0);
    step_pr_apply_kernel<true><<<nb, 256, 0, s>>>(n, a, 0.0f, nullptr, b, nullptr, nullptr, w.part);
    G1_LAUNCHED();
    reduce_partials_kernel<0><<<dim3(1, 1), 256, 0, s>>>(w.part, nb, out);
    G1_LAUNCHED();
    return G1_OK;
}

size_t g1_spmm_workspace_bytes(uint32_t n_rows, uint32_t nnz) { return spmm_ws_bytes(n_rows, nnz); }

g1_status g1_spmm_minplus_u32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, uint32_t k, const uint32_t* row_offsets,
                              const uint32_t* col_indices, const uint32_t* values, const uint32_t* X, uint32_t* Y,
                              const uint64_t* col_mask, void* ws, size_t ws_bytes, g1_stream_t stream) {
    return spmm<MinPlusU32>(n_rows, n_cols, nnz, k, row_offsets, col_indices, values, X, Y,
                            reinterpret_cast<const unsigned long long*>(col_mask), ws, ws_bytes, as_stream(stream),
                            "g1_spmm_minplus_u32");
}

g1_status g1_spmm_plustimes_f32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, uint32_t k, const uint32_t* row_offsets,
                                const uint32_t* col_indices, const float* values, const float* X, float* Y,
                                const uint64_t* col_mask, void* ws, size_t ws_bytes, g1_stream_t stream) {
    return spmm<PlusTimesKahanF32>(n_rows, n_cols, nnz, k, row_offsets, col_indices, values, X, Y,
                                   reinterpret_cast<const unsigned long long*>(col_mask), ws, ws_bytes,
                                   as_stream(stream), "g1_spmm_plustimes_f32");
}

}  // extern "C"
