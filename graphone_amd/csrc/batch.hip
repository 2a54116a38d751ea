// Batched multi-source engines: k sources run together as k columns of one semiring SpMM over the
// in-edge CSR (g1/spmm.hpp), min-plus for SSSP/BFS and compensated plus-times for personalized PageRank.
//
// A batch of S sources is cut into ceil(S/k) column tiles of k in {8, 16, 32, 64}; k comes from the
// caller or from the autotuner (g1_batch_autotune), which times one tile of every width on the graph and
// keeps the fastest per source. Min-plus iterations are captured once per (graph, k) as a HIP graph of
// two iterations (the change masks swap back) and replayed until no column changes.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "g1/apply.hpp"
#include "g1/spmm.hpp"
#include "graph.hpp"
#include "session.hpp"
#include "spmv_launch.hpp"

// Merge tiles the k = 64 min-plus kernel may run on (kWaveTileItems below).
constexpr uint32_t kWaveTileCount = 4;

struct BatchWork {
    uint32_t K = 0;
    uint32_t* D = nullptr;                                 // n x K distances
    unsigned long long* mask[2] = {nullptr, nullptr};      // changed-column masks
    void* recs = nullptr;                                  // SpMM tile records
    unsigned long long* changed = nullptr;                 // [0], [1]: columns changed in iteration 1, 2 of a replay
    float* R[2] = {nullptr, nullptr};                      // PPR ranks, n x K
    float* X = nullptr;
    float* Y = nullptr;
    hipGraphExec_t minplus_exec[kWaveTileCount][2] = {};   // two captured iterations, [merge tile][weighted]
    int16_t* D16 = nullptr;                                // n x 64 distances, 16-bit storage (k = 64)
    uint32_t* overflow = nullptr;                          // set when a distance does not fit 16 bits
    hipGraphExec_t minplus16_exec[kWaveTileCount][2] = {};
    uint32_t* act = nullptr;                               // k = 64: bitmap of last iteration's changed vertices
};

namespace g1 {
namespace {

// state[0] = frozen, state[1] = 0 (the batched PPR's device-side stop-rule state)
__global__ void set_state_kernel(unsigned long long* state, unsigned long long frozen) {
    if (threadIdx.x == 0) {
        state[0] = frozen;
        state[1] = 0ull;
    }
}

// Min-plus SpMM kernel: at k = 64 the wavefront-per-tile kernel (unless G1_MINPLUS_WAVE=0); otherwise
// the chunked, barrier-free walk (default) or the round-based one of the plus-times path
// (G1_MINPLUS_ROUNDS=1). All give the same bits.
bool minplus_chunked() {
    static const bool chunked = [] {
        const char* e = getenv("G1_MINPLUS_ROUNDS");
        return !(e && e[0] == '1');
    }();
    return chunked;
}

// k = 64 batches keep distances in 16 bits while they fit (G1_MINPLUS_U16=0: always 32 bits).
bool minplus_u16() {
    static const bool on = [] {
        const char* e = getenv("G1_MINPLUS_U16");
        return !(e && e[0] == '0');
    }();
    return on;
}

bool minplus_wave() {
    static const bool wave = [] {
        const char* e = getenv("G1_MINPLUS_WAVE");
        return !(e && e[0] == '0');
    }();
    return wave;
}

int k_slot(uint32_t k) { return k == 8 ? 0 : k == 16 ? 1 : k == 32 ? 2 : k == 64 ? 3 : -1; }

template <typename T>
g1_status alloc_n(g1_graph* g, T** p, size_t count) {
    void* v = nullptr;
    G1_TRY(graph_alloc(g, &v, count * sizeof(T)));
    *p = static_cast<T*>(v);
    return G1_OK;
}

// out[j * n + v] = D16[v * 64 + j] widened to 32 bits (-1 -> G1_INF), through LDS like
// batch_columns_out_kernel.
__global__ void __launch_bounds__(256) batch_columns_out16_kernel(uint32_t n, const int16_t* __restrict__ D,
                                                                  uint32_t cols, uint32_t* __restrict__ out) {
    __shared__ uint32_t s[64][65];
    const uint32_t v0 = blockIdx.x * 64;
    for (uint32_t e = threadIdx.x; e < 64u * 64u; e += 256) {
        const uint32_t r = e / 64, c = e % 64;
        if (v0 + r < n) s[c][r] = widen_dist<int16_t>(D[static_cast<size_t>(v0 + r) * 64 + c]);
    }
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < 64u * cols; f += 256) {
        const uint32_t c = f / 64, r = f % 64;
        if (v0 + r < n) out[static_cast<size_t>(c) * n + v0 + r] = s[c][r];
    }
}

uint32_t fill_blocks(size_t items) {
    size_t b = (items + 255) / 256;
    if (b > 8192) b = 8192;
    return static_cast<uint32_t>(b ? b : 1);
}

// Words of the k = 64 activity bitmap (act_bits_kernel): two per 64 vertices.
size_t act_words(size_t n) { return 2 * ((n + 63) / 64) + 2; }

// The row-start bitmap of the in-edge CSR used by the k = 64 min-plus kernel; built once per graph
// (caller holds g->mu), finished before it returns so any stream may use it.
g1_status ensure_row_starts(g1_graph* g, hipStream_t s) {
    if (g->t_starts || !g->has_t) return G1_OK;
    const size_t words = (static_cast<size_t>(g->m) >> 6) + 2;
    G1_TRY(alloc_n(g, &g->t_starts, words));
    G1_TRY(g1_fill_u32(reinterpret_cast<uint32_t*>(g->t_starts), 0u, 2 * words, reinterpret_cast<g1_stream_t>(s)));
    if (g->n) {
        row_starts_kernel<<<fill_blocks(g->n), 256, 0, s>>>(g->n, g->t_row_off, g->t_starts);
        G1_LAUNCHED();
    }
    G1_HIP(hipStreamSynchronize(s));
    return G1_OK;
}

// Merge tiles (rows + nonzeros per wavefront) the autotuner tries for the k = 64 min-plus kernel. Min is
// order-free, so every partition gives the same bits; [kWaveTileOwn] is the graph's own partition (t_ws), shared
// with the single-source SpMV and the plus-times kernels, whose fold order it fixes.
constexpr uint32_t kWaveTileItems[kWaveTileCount] = {kMergeTile / 4, kMergeTile / 2, kMergeTile, kMergeTile * 2};
constexpr uint32_t kWaveTileOwn = 2;  // kWaveTileItems[kWaveTileOwn] == kMergeTile
static_assert(sizeof(g1_graph::wave_coords) / sizeof(void*) == kWaveTileCount, "one partition per merge tile");

// The merge partition of the in-edge CSR at kWaveTileItems[idx]; built once per graph and size (caller
// holds g->mu), finished before it returns.
g1_status ensure_wave_coords(g1_graph* g, uint32_t idx, hipStream_t s) {
    if (idx == kWaveTileOwn || g->wave_coords[idx] || g->n == 0) return G1_OK;
    const uint32_t tiles = merge_num_tiles_for(g->n, g->m, kWaveTileItems[idx]);
    MergeCoord* c = nullptr;
    G1_TRY(alloc_n(g, &c, static_cast<size_t>(tiles) + 1));
    merge_partition_kernel<<<(tiles + 1 + 255) / 256, 256, 0, s>>>(g->t_row_off, g->n, g->m, tiles, c,
                                                                   kWaveTileItems[idx]);
    G1_LAUNCHED();
    G1_HIP(hipStreamSynchronize(s));
    g->wave_coords[idx] = c;
    return G1_OK;
}

g1_status ensure_batch(g1_graph* g, uint32_t K, bool ppr, BatchWork** out, hipStream_t s) {
    const int slot = k_slot(K);
    G1_CHECK(slot >= 0, G1_ERR_INVALID_ARG, "column tile k must be 8, 16, 32 or 64 (got %u)", K);
    G1_CHECK(g->has_t, G1_ERR_INVALID_ARG, "batched engines need the graph's transpose");
    if (!g->batch[slot]) {
        BatchWork* b = new BatchWork();
        b->K = K;
        g->batch[slot] = b;
    }
    BatchWork* b = g->batch[slot];
    const size_t n = g->n ? g->n : 1;
    if (!b->recs) {
        // (min-plus k = 64 may run on the smallest merge tile: four times the graph's tiles)
        const size_t rb = std::max(static_cast<size_t>(merge_num_tiles_for(g->n, g->m, kWaveTileItems[0])) *
                                       sizeof(SpmmTileRec<MinPlusU32, 64>),
                                   spmm_records_bytes<PlusTimesKahanF32, 64>(g->n, g->m));
        void* v = nullptr;
        G1_TRY(graph_alloc(g, &v, rb + 256));
        b->recs = v;
        G1_TRY(alloc_n(g, &b->changed, 64));
    }
    if (!ppr && K == 64) G1_TRY(ensure_row_starts(g, s));
    if (!ppr && !b->D) {
        G1_TRY(alloc_n(g, &b->D, (n + 1) * K));  // row n: all unreached (read for inactive sources)
        G1_TRY(alloc_n(g, &b->mask[0], n));
        G1_TRY(alloc_n(g, &b->mask[1], n));
        if (K == 64) G1_TRY(alloc_n(g, &b->act, act_words(n)));
        if (K == 64 && minplus_u16() && (static_cast<uint64_t>(n) + 1) * 128 < (1ull << 32)) {
            G1_TRY(alloc_n(g, &b->D16, (n + 1) * 64));
            G1_TRY(alloc_n(g, &b->overflow, 64));
        }
    }
    if (ppr && !b->X) {
        G1_TRY(alloc_n(g, &b->R[0], n * K));
        G1_TRY(alloc_n(g, &b->R[1], n * K));
        G1_TRY(alloc_n(g, &b->X, n * K));
        G1_TRY(alloc_n(g, &b->Y, n * K));
        G1_TRY(ensure_rank_partials(g));
        G1_TRY(ensure_pull_work(g, true));  // iw
    }
    *out = b;
    return G1_OK;
}

// D = unreached (all ones in DT), mask = 0; then D[src_j][j] = 0 for the tile's valid columns and
// mask[v] = the columns whose source is v. The init kernel also clears the 16-bit overflow flag.
template <int K, typename DT>
__global__ void minplus_batch_init_kernel(uint32_t n, DT* __restrict__ D, unsigned long long* __restrict__ mask,
                                          uint32_t* __restrict__ overflow) {
    const size_t total = (static_cast<size_t>(n) + 1) * K;  // rows 0..n (row n stays unreached)
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t i = t0; i < total; i += stride) D[i] = static_cast<DT>(~DT(0));
    for (size_t v = t0; v < n; v += stride) mask[v] = 0ull;
    if (overflow && t0 == 0) *overflow = 0u;
}

template <int K, typename DT>
__global__ void minplus_batch_seed_kernel(DT* __restrict__ D, unsigned long long* __restrict__ mask,
                                          const uint32_t* __restrict__ srcs, uint32_t cols) {
    const uint32_t j = threadIdx.x;
    if (j < cols) {
        const uint32_t s = srcs[j];
        D[static_cast<size_t>(s) * K + j] = DT(0);
        atomicOr(&mask[s], 1ull << j);
    }
}

// R0 = e_seed per column and X0 = R0 * iw (pr_init_kernel's seed variant, per column); padding columns
// (seed kInf) start at zero.
template <int K>
__global__ void ppr_batch_init_kernel(uint32_t n, const float* __restrict__ iw, const uint32_t* __restrict__ seeds,
                                      float* __restrict__ R, float* __restrict__ X) {
    const size_t total = static_cast<size_t>(n) * K;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t v = static_cast<uint32_t>(i / K);
        const uint32_t j = static_cast<uint32_t>(i % K);
        const float rv = (v == seeds[j]) ? 1.0f : 0.0f;
        R[i] = rv;
        X[i] = rv * iw[v];
    }
}

__global__ void clear_flags_kernel(unsigned long long* __restrict__ f) {
    if (threadIdx.x < 2) f[threadIdx.x] = 0ull;
}

// out[j * n + v] = D[v * K + j] for the tile's valid columns j < cols: 64 rows of D pass through LDS so
// that both the row-major reads and the column-major writes are coalesced.
template <int K, typename T>
__global__ void __launch_bounds__(256) batch_columns_out_kernel(uint32_t n, const T* __restrict__ D, uint32_t cols,
                                                                T* __restrict__ out) {
    __shared__ T s[K][65];
    const uint32_t v0 = blockIdx.x * 64;
    for (uint32_t e = threadIdx.x; e < 64u * K; e += 256) {
        const uint32_t r = e / K, c = e % K;
        if (v0 + r < n) s[c][r] = D[static_cast<size_t>(v0 + r) * K + c];
    }
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < 64u * cols; f += 256) {
        const uint32_t c = f / 64, r = f % 64;
        if (v0 + r < n) out[static_cast<size_t>(c) * n + v0 + r] = s[c][r];
    }
}

// k = 64 plus-times SpMM on the wavefront kernel unless G1_PPR_WAVE=0 (same bits either way).
bool ppr_wave() {
    static const bool on = [] {
        const char* e = getenv("G1_PPR_WAVE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// One plus-times SpMM iteration: Y = A^T (x) X over the in-edges (k columns, Kahan sums in the fold order
// of the single-source SpMV), then the fixup of rows that cross tiles.
template <int K>
g1_status launch_plustimes(g1_graph* g, const float* X, float* Y, void* recs_v, hipStream_t s) {
    const uint32_t tiles = merge_num_tiles(g->n, g->m);
    PlusTimesBatchPolicy<K> pol{X, Y};
    auto* recs = static_cast<SpmmTileRec<PlusTimesKahanF32, K>*>(recs_v);
    const MergeCoord* coords = merge_coords(g->t_ws);
    const float* vals = g->t_wf;
    bool done = false;
    if constexpr (K == 64) {
        if (ppr_wave()) {
            const uint32_t blocks = (tiles + 3) / 4;
            if (vals)
                spmm_plustimes_wave_kernel<true><<<blocks, 256, 0, s>>>(g->n, tiles, g->t_row_off, g->t_col, vals,
                                                                        coords, recs, X, Y);
            else
                spmm_plustimes_wave_kernel<false><<<blocks, 256, 0, s>>>(g->n, tiles, g->t_row_off, g->t_col, vals,
                                                                         coords, recs, X, Y);
            G1_LAUNCHED();
            done = true;
        }
    }
    if (!done) {
        spmm_tile_kernel<PlusTimesKahanF32, K, PlusTimesBatchPolicy<K>><<<tiles, 256, 0, s>>>(
            g->n, g->m, g->t_row_off, g->t_col, vals, coords, recs, pol, nullptr);
        G1_LAUNCHED();
    }
    if (tiles > 1) {
        const uint32_t blocks = static_cast<uint32_t>((static_cast<uint64_t>(tiles) * K + 255) / 256);
        spmm_fixup_kernel<PlusTimesKahanF32, K, PlusTimesBatchPolicy<K>><<<blocks, 256, 0, s>>>(tiles, coords, recs,
                                                                                              pol, nullptr);
        G1_LAUNCHED();
    }
    return G1_OK;
}

// One min-plus SpMM iteration over the in-edges: D relaxed in place, change masks prev -> next, the
// columns that changed ORed into *flag; then the fixup of rows that cross tiles. DT is the distance
// storage (int16_t only at k = 64 with the wavefront kernel; overflow is then raised instead of storing
// a distance >= 0xFFFF).
template <int K, typename DT = uint32_t>
g1_status launch_minplus(g1_graph* g, DT* D, unsigned long long* mask_prev, unsigned long long* mask_next,
                         void* recs_v, bool weighted, unsigned long long* flag, hipStream_t s, uint32_t* act,
                         uint32_t* overflow = nullptr, uint32_t tile_idx = kWaveTileOwn) {
    uint32_t tiles = merge_num_tiles(g->n, g->m);
    auto* recs = static_cast<SpmmTileRec<MinPlusU32, K>*>(recs_v);
    const uint32_t* vals = weighted ? g->t_w : nullptr;
    const MergeCoord* coords = merge_coords(g->t_ws);
    uint32_t fix_blocks = static_cast<uint32_t>((static_cast<uint64_t>(tiles) * K + 255) / 256);
    if constexpr (K == 64) {
        if (sizeof(DT) == 2 || minplus_wave()) {
            // the wavefront kernel runs on any merge partition (kWaveTileItems[tile_idx])
            if (tile_idx != kWaveTileOwn) {
                G1_CHECK(tile_idx < kWaveTileCount && g->wave_coords[tile_idx] != nullptr, G1_ERR_INTERNAL,
                         "min-plus k = 64: merge partition %u missing", tile_idx);
                tiles = merge_num_tiles_for(g->n, g->m, kWaveTileItems[tile_idx]);
                coords = static_cast<const MergeCoord*>(g->wave_coords[tile_idx]);
                fix_blocks = static_cast<uint32_t>((static_cast<uint64_t>(tiles) * K + 255) / 256);
            }
            const uint32_t blocks = (tiles + 3) / 4;
            G1_CHECK(g->t_starts != nullptr && act != nullptr, G1_ERR_INTERNAL,
                     "min-plus k = 64: row-start or activity bitmap missing");
            // The wavefront kernel writes only non-zero change masks, and reads activity as a bitmap.
            G1_TRY(g1_fill_u32(reinterpret_cast<uint32_t*>(mask_next), 0u, 2ull * g->n, reinterpret_cast<g1_stream_t>(s)));
            act_bits_kernel<<<fill_blocks(static_cast<size_t>(g->n) / 4 + 64), 256, 0, s>>>(g->n, mask_prev, act);
            G1_LAUNCHED();
            if (weighted)
                spmm_minplus_wave_kernel<true, DT><<<blocks, 256, 0, s>>>(g->n, tiles, g->t_row_off, g->t_col, vals,
                                                                          g->t_starts, coords, recs, D, act,
                                                                          mask_next, flag, overflow);
            else
                spmm_minplus_wave_kernel<false, DT><<<blocks, 256, 0, s>>>(g->n, tiles, g->t_row_off, g->t_col, vals,
                                                                           g->t_starts, coords, recs, D, act,
                                                                           mask_next, flag, overflow);
            G1_LAUNCHED();
            if (tiles > 1) {
                if constexpr (sizeof(DT) == 2) {
                    MinPlusBatchPolicy16 pol{D, mask_prev, mask_next, overflow};
                    spmm_fixup_kernel<MinPlusU32, 64, MinPlusBatchPolicy16><<<fix_blocks, 256, 0, s>>>(
                        tiles, coords, recs, pol, flag);
                } else {
                    MinPlusBatchPolicy<64> pol{D, mask_prev, mask_next};
                    spmm_fixup_kernel<MinPlusU32, 64, MinPlusBatchPolicy<64>><<<fix_blocks, 256, 0, s>>>(
                        tiles, coords, recs, pol, flag);
                }
                G1_LAUNCHED();
            }
            return G1_OK;
        }
    }
    if constexpr (sizeof(DT) == 4) {
        MinPlusBatchPolicy<K> pol{D, mask_prev, mask_next};
        if (minplus_chunked())
            spmm_minplus_kernel<K, MinPlusBatchPolicy<K>><<<tiles, 256, 0, s>>>(
                g->n, g->m, g->t_row_off, g->t_col, vals, coords, recs, pol, flag);
        else
            spmm_tile_kernel<MinPlusU32, K, MinPlusBatchPolicy<K>><<<tiles, 256, 0, s>>>(
                g->n, g->m, g->t_row_off, g->t_col, vals, coords, recs, pol, flag);
        G1_LAUNCHED();
        if (tiles > 1) {
            spmm_fixup_kernel<MinPlusU32, K, MinPlusBatchPolicy<K>><<<fix_blocks, 256, 0, s>>>(tiles, coords, recs,
                                                                                             pol, flag);
            G1_LAUNCHED();
        }
    }
    return G1_OK;
}

// Replays the captured pair of iterations (mask[0] -> mask[1] -> mask[0]) until an iteration changes
// nothing; the count includes that last, unchanged iteration. The pair is captured once per (graph, k,
// storage, weighted) on a private stream; the pointers it holds never change.
template <int K, typename DT>
g1_status minplus_replay(g1_graph* g, BatchWork* b, DT* D, hipGraphExec_t& exec, bool weighted, uint32_t tile_idx,
                         uint32_t* iterations, hipStream_t s) {
    uint32_t* ovf = sizeof(DT) == 2 ? b->overflow : nullptr;
    if (!exec) {
        hipStream_t cs = nullptr;
        G1_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        g1_status st = G1_OK;
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            clear_flags_kernel<<<1, 64, 0, cs>>>(b->changed);
            st = launch_minplus<K, DT>(g, D, b->mask[0], b->mask[1], b->recs, weighted, b->changed, cs, b->act, ovf,
                                       tile_idx);
            if (st == G1_OK)
                st = launch_minplus<K, DT>(g, D, b->mask[1], b->mask[0], b->recs, weighted, b->changed + 1, cs, b->act,
                                           ovf, tile_idx);
            e = hipStreamEndCapture(cs, &graph);
        }
        (void)hipStreamDestroy(cs);
        G1_TRY(st);
        G1_HIP(e);
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        G1_HIP(e);
    }
    uint32_t it = 0;
    unsigned long long* hf = g->h_ctr + 12;
    for (;;) {
        G1_HIP(hipGraphLaunch(exec, s));
        G1_HIP(hipMemcpyAsync(hf, b->changed, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        G1_HIP(hipStreamSynchronize(s));
        if (hf[0] == 0) { it += 1; break; }
        it += 2;
        if (hf[1] == 0 || it >= kHistIters) break;
    }
    *iterations = it;
    return G1_OK;
}

template <int K>
g1_status minplus_tile_run(g1_graph* g, BatchWork* b, const uint32_t* srcs_host, uint32_t cols, bool weighted,
                           uint32_t* out, uint32_t* iterations, hipStream_t s) {
    const uint32_t n = g->n;
    G1_HIP(hipMemcpyAsync(g->seeds_dev, srcs_host, cols * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    // the autotuned merge tile (k = 64 wavefront kernel only; the other kernels use the graph's partition)
    const uint32_t tile = (K == 64 && g->minplus_tile < kWaveTileCount) ? g->minplus_tile : kWaveTileOwn;
    G1_TRY(ensure_wave_coords(g, tile, s));
    if constexpr (K == 64) {
        if (b->D16 && (!weighted || g->max_w <= 0xFFFF0000u)) {
            // 16-bit storage first; a distance that does not fit voids the run, which is redone in 32 bits.
            minplus_batch_init_kernel<64, int16_t><<<fill_blocks(static_cast<size_t>(n) * 64), 256, 0, s>>>(
                n, b->D16, b->mask[0], b->overflow);
            G1_LAUNCHED();
            minplus_batch_seed_kernel<64, int16_t><<<1, 64, 0, s>>>(b->D16, b->mask[0], g->seeds_dev, cols);
            G1_LAUNCHED();
            G1_TRY((minplus_replay<64, int16_t>(g, b, b->D16, b->minplus16_exec[tile][weighted ? 1 : 0], weighted,
                                                 tile, iterations, s)));
            uint32_t* hov = reinterpret_cast<uint32_t*>(g->h_ctr + 14);
            G1_HIP(hipMemcpyAsync(hov, b->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            G1_HIP(hipStreamSynchronize(s));
            if (*hov == 0) {
                batch_columns_out16_kernel<<<(n + 63) / 64, 256, 0, s>>>(n, b->D16, cols, out);
                G1_LAUNCHED();
                return G1_OK;
            }
        }
    }
    minplus_batch_init_kernel<K, uint32_t><<<fill_blocks(static_cast<size_t>(n) * K), 256, 0, s>>>(
        n, b->D, b->mask[0], nullptr);
    G1_LAUNCHED();
    minplus_batch_seed_kernel<K, uint32_t><<<1, 64, 0, s>>>(b->D, b->mask[0], g->seeds_dev, cols);
    G1_LAUNCHED();
    G1_TRY((minplus_replay<K, uint32_t>(g, b, b->D, b->minplus_exec[tile][weighted ? 1 : 0], weighted, tile,
                                        iterations, s)));
    batch_columns_out_kernel<K, uint32_t><<<(n + 63) / 64, 256, 0, s>>>(n, b->D, cols, out);
    G1_LAUNCHED();
    return G1_OK;
}

template <int K>
g1_status ppr_tile_run(g1_graph* g, BatchWork* b, const uint32_t* seeds_host, uint32_t cols, float restart,
                       uint32_t max_iter, double tol, float* out, uint32_t* iterations, hipStream_t s) {
    const uint32_t n = g->n;
    const uint32_t nb = rank_blocks(n);
    std::vector<uint32_t> seeds(K, kInf);
    for (uint32_t j = 0; j < cols; ++j) seeds[j] = seeds_host[j];
    G1_HIP(hipMemcpyAsync(g->seeds_dev, seeds.data(), K * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    // iw from the out-edges, then R0 = e_seed per column, X0 = R0 * iw
    pr_init_kernel<<<fill_blocks(n), 256, 0, s>>>(n, g->osum, 1.0f - restart, g->iw, g->r[0], 0.0f, g->xf, kInf);
    G1_LAUNCHED();
    ppr_batch_init_kernel<K><<<fill_blocks(static_cast<size_t>(n) * K), 256, 0, s>>>(n, g->iw, g->seeds_dev, b->R[0],
                                                                                      b->X);
    G1_LAUNCHED();
    // Device state of the stop rule: state[0] frozen columns (padding columns never run), state[1] the
    // iteration count; written with the seeds' stream order, by a kernel (no memset in the captured work).
    unsigned long long* state = b->changed + 4;
    unsigned long long frozen0 = 0;
    for (uint32_t j = cols; j < K; ++j) frozen0 |= 1ull << j;
    set_state_kernel<<<1, 64, 0, s>>>(state, frozen0);
    G1_LAUNCHED();
    const double* deltas = g->rpart + static_cast<size_t>(nb) * kBatchCols;
    // One iteration: SpMM, apply (frozen columns keep their ranks), the canonical L1 per column, the freeze.
    // An iteration after every column froze changes nothing (frozen columns copy R into the other buffer).
    auto iteration = [&](int cur, hipStream_t st) -> g1_status {
        G1_TRY(launch_plustimes<K>(g, b->X, b->Y, b->recs, st));
        rank_apply_batch_kernel<K><<<(nb + 256 / K - 1) / (256 / K), 256, 0, st>>>(
            n, b->Y, 0.0f, g->seeds_dev, restart, g->iw, b->R[cur], b->R[cur ^ 1], b->X, g->rpart, nb, 0ull, state);
        G1_LAUNCHED();
        reduce_partials_kernel<0><<<dim3(1, K), 256, 0, st>>>(g->rpart, nb, g->rpart + static_cast<size_t>(nb) * kBatchCols);
        G1_LAUNCHED();
        ppr_freeze_kernel<K><<<1, 64, 0, st>>>(deltas, tol, state);
        G1_LAUNCHED();
        return G1_OK;
    };
    // Two iterations (R[0] -> R[1] -> R[0]) captured as one HIP graph for this call (its arguments --
    // restart, tol -- are the call's), replayed while both fit in max_iter; the host reads the state (two
    // words) once per replay and stops when every column froze.
    hipGraphExec_t exec = nullptr;
    if (max_iter >= 2) {
        hipStream_t cs = nullptr;
        G1_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        g1_status st = G1_OK;
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            st = iteration(0, cs);
            if (st == G1_OK) st = iteration(1, cs);
            e = hipStreamEndCapture(cs, &graph);
        }
        (void)hipStreamDestroy(cs);
        G1_TRY(st);
        G1_HIP(e);
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        G1_HIP(e);
    }
    const unsigned long long all = (K == 64) ? ~0ull : ((1ull << K) - 1ull);
    unsigned long long* h = g->h_ctr + 12;
    h[0] = frozen0;
    h[1] = 0ull;
    uint32_t launched = 0;
    g1_status st = G1_OK;
    while (launched < max_iter) {
        if (max_iter - launched >= 2) {
            const hipError_t e = hipGraphLaunch(exec, s);
            if (e != hipSuccess) {
                st = set_error(G1_ERR_HIP, "ppr batch: graph launch: %s", hipGetErrorString(e));
                break;
            }
            launched += 2;
        } else {
            if ((st = iteration(0, s)) != G1_OK) break;  // an even count of iterations ran before: R[0] is current
            launched += 1;
        }
        hipError_t e = hipMemcpyAsync(h, state, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            st = set_error(G1_ERR_HIP, "ppr batch: state read: %s", hipGetErrorString(e));
            break;
        }
        if (h[0] == all) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    G1_TRY(st);
    // h[1] iterations ran with a running column: a frozen column kept its rank in both buffers since, and
    // the buffer written last is R[h[1] % 2] (no-op iterations after that copy ranks only).
    *iterations = static_cast<uint32_t>(h[1]);
    batch_columns_out_kernel<K, float><<<(n + 63) / 64, 256, 0, s>>>(n, b->R[h[1] % 2], cols, out);
    G1_LAUNCHED();
    return G1_OK;
}

template <int K>
g1_status sssp_batch_k(g1_graph* g, const uint32_t* srcs, uint32_t count, bool weighted, uint32_t* out,
                       g1_batch_stats* st, hipStream_t s) {
    BatchWork* b = nullptr;
    G1_TRY(ensure_batch(g, K, false, &b, s));
    G1_TRY(ensure_rank_partials(g));  // seeds_dev
    for (uint32_t c0 = 0; c0 < count; c0 += K) {
        const uint32_t cols = std::min<uint32_t>(K, count - c0);
        uint32_t iters = 0;
        G1_TRY(minplus_tile_run<K>(g, b, srcs + c0, cols, weighted, out + static_cast<size_t>(c0) * g->n, &iters, s));
        st->iterations_sum += iters;
        st->tiles += 1;
    }
    return G1_OK;
}

template <int K>
g1_status ppr_batch_k(g1_graph* g, const uint32_t* seeds, uint32_t count, float restart, uint32_t max_iter, double tol,
                      float* out, g1_batch_stats* st, hipStream_t s) {
    BatchWork* b = nullptr;
    G1_TRY(ensure_batch(g, K, true, &b, s));
    for (uint32_t c0 = 0; c0 < count; c0 += K) {
        const uint32_t cols = std::min<uint32_t>(K, count - c0);
        uint32_t iters = 0;
        G1_TRY(ppr_tile_run<K>(g, b, seeds + c0, cols, restart, max_iter, tol, out + static_cast<size_t>(c0) * g->n,
                               &iters, s));
        st->iterations_sum += iters;
        st->tiles += 1;
    }
    return G1_OK;
}

g1_status check_sources(g1_graph* g, const uint32_t* srcs, uint32_t count, const void* out, const char* what) {
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "%s: null graph", what);
    G1_CHECK(out != nullptr && (count == 0 || srcs != nullptr), G1_ERR_INVALID_ARG, "%s: null argument", what);
    for (uint32_t i = 0; i < count; ++i)
        G1_CHECK(srcs[i] < g->n, G1_ERR_INVALID_ARG, "%s: source %u (index %u) out of range (n = %u)", what, srcs[i], i,
                 g->n);
    return G1_OK;
}

g1_status timed(g1_graph* g, hipStream_t s, g1_batch_stats* st) {
    G1_HIP(hipEventRecord(g->ev[1], s));
    G1_HIP(hipEventSynchronize(g->ev[1]));
    float ms = 0.0f;
    G1_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
    st->elapsed_ms = ms;
    return G1_OK;
}

g1_status sssp_batch_dispatch(g1_graph* g, const uint32_t* srcs, uint32_t count, uint32_t k, bool weighted,
                              uint32_t* out, g1_batch_stats* st, hipStream_t s) {
    switch (k) {
        case 8: return sssp_batch_k<8>(g, srcs, count, weighted, out, st, s);
        case 16: return sssp_batch_k<16>(g, srcs, count, weighted, out, st, s);
        case 32: return sssp_batch_k<32>(g, srcs, count, weighted, out, st, s);
        case 64: return sssp_batch_k<64>(g, srcs, count, weighted, out, st, s);
        default: return set_error(G1_ERR_INVALID_ARG, "column tile k must be 8, 16, 32 or 64 (got %u)", k);
    }
}

g1_status ppr_batch_dispatch(g1_graph* g, const uint32_t* seeds, uint32_t count, uint32_t k, float restart,
                             uint32_t max_iter, double tol, float* out, g1_batch_stats* st, hipStream_t s) {
    switch (k) {
        case 8: return ppr_batch_k<8>(g, seeds, count, restart, max_iter, tol, out, st, s);
        case 16: return ppr_batch_k<16>(g, seeds, count, restart, max_iter, tol, out, st, s);
        case 32: return ppr_batch_k<32>(g, seeds, count, restart, max_iter, tol, out, st, s);
        case 64: return ppr_batch_k<64>(g, seeds, count, restart, max_iter, tol, out, st, s);
        default: return set_error(G1_ERR_INVALID_ARG, "column tile k must be 8, 16, 32 or 64 (got %u)", k);
    }
}

// Times one full column tile of every width on sources drawn from `srcs` (repeated as needed) and
// returns the width with the lowest device time per source. When min-plus picks k = 64, the merge tile of
// its wavefront kernel (kWaveTileItems) is timed the same way and kept in g->minplus_tile.
g1_status autotune(g1_graph* g, int kind, const uint32_t* srcs, uint32_t count, bool weighted, uint32_t* best,
                   hipStream_t s) {
    G1_CHECK(count > 0, G1_ERR_INVALID_ARG, "autotune needs at least one source");
    std::vector<uint32_t> sample(64);
    for (uint32_t i = 0; i < 64; ++i) sample[i] = srcs[i % count];
    void* out = nullptr;
    const size_t elem = kind == 1 ? sizeof(float) : sizeof(uint32_t);
    G1_HIP(hipMalloc(&out, elem * static_cast<size_t>(g->n) * 64));
    double best_per = 1e300;
    uint32_t best_k = 64;
    g1_status status = G1_OK;
    // one tile of k columns: an untimed first pass (allocation, graph capture), then the timed pass
    auto time_tile = [&](uint32_t k, double* ms) {
        g1_batch_stats st;
        for (int pass = 0; pass < 2 && status == G1_OK; ++pass) {
            memset(&st, 0, sizeof(st));
            if (hipEventRecord(g->ev[0], s) != hipSuccess) { status = set_error(G1_ERR_HIP, "event"); break; }
            status = kind == 1 ? ppr_batch_dispatch(g, sample.data(), k, k, 0.15f, 50, 1e-7, static_cast<float*>(out), &st, s)
                               : sssp_batch_dispatch(g, sample.data(), k, k, weighted,
                                                     static_cast<uint32_t*>(out), &st, s);
            if (status == G1_OK) status = timed(g, s, &st);
        }
        *ms = st.elapsed_ms;
    };
    if (kind == 0) g->minplus_tile = kWaveTileOwn;  // widths are compared on the graph's own partition
    double ms64 = 0.0;
    for (uint32_t k : {8u, 16u, 32u, 64u}) {
        double ms = 0.0;
        time_tile(k, &ms);
        if (status != G1_OK) break;
        if (k == 64) ms64 = ms;
        const double per = ms / k;
        if (per < best_per) {
            best_per = per;
            best_k = k;
        }
    }
    if (status == G1_OK && kind == 0 && best_k == 64) {
        uint32_t best_tile = kWaveTileOwn;
        double best_ms = ms64;
        for (uint32_t idx = 0; idx < kWaveTileCount; ++idx) {
            if (idx == kWaveTileOwn) continue;  // timed above
            g->minplus_tile = idx;
            double ms = 0.0;
            time_tile(64, &ms);
            if (status != G1_OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best_tile = idx;
            }
        }
        g->minplus_tile = best_tile;
    }
    if (status != G1_OK && kind == 0) g->minplus_tile = kWaveTileOwn;
    (void)hipFree(out);
    G1_TRY(status);
    *best = best_k;
    g->best_k[kind] = best_k;
    return G1_OK;
}

// ---------------------------------------------------------------- continuous-batching sessions

// Column col of a min-plus session starts over from `src` (kInf: the column is emptied).
template <int K>
__global__ void minplus_column_reset_kernel(uint32_t n, uint32_t* __restrict__ D, unsigned long long* __restrict__ mask,
                                            uint32_t col, uint32_t src) {
    const unsigned long long bit = 1ull << col;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        D[static_cast<size_t>(v) * K + col] = (v == src) ? 0u : kInf;
        mask[v] = (mask[v] & ~bit) | ((v == src) ? bit : 0ull);
    }
}

// Column col of a PPR session starts from e_seed: R = e_seed, X = R * iw (pr_init_kernel's seed variant).
template <int K>
__global__ void ppr_column_reset_kernel(uint32_t n, const float* __restrict__ iw, float* __restrict__ R,
                                        float* __restrict__ X, uint32_t col, uint32_t seed) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const float rv = (v == seed) ? 1.0f : 0.0f;
        R[static_cast<size_t>(v) * K + col] = rv;
        X[static_cast<size_t>(v) * K + col] = rv * iw[v];
    }
}

template <int K, typename T>
__global__ void column_extract_kernel(uint32_t n, const T* __restrict__ A, uint32_t col, T* __restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) out[v] = A[static_cast<size_t>(v) * K + col];
}

// iw[u] = alpha / osum[u] (0 without out-weight): the expression of pr_init_kernel, so the same bits.
__global__ void inv_weight_kernel(uint32_t n, const float* __restrict__ osum, float alpha, float* __restrict__ iw) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride) {
        const float sw = osum[u];
        iw[u] = (sw != 0.0f) ? alpha / sw : 0.0f;
    }
}

__global__ void set_u32_kernel(uint32_t* __restrict__ p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

}  // namespace

struct Session {
    g1_graph* g = nullptr;
    int kind = 0;
    uint32_t K = 0;
    float restart = 0.0f;
    hipStream_t s = nullptr;
    uint32_t* D = nullptr;
    unsigned long long* mask[2] = {nullptr, nullptr};
    int mcur = 0;  // mask[mcur]: the sources of the next iteration
    float* R[2] = {nullptr, nullptr};
    float* X = nullptr;
    float* Y = nullptr;
    float* iw = nullptr;
    int rcur = 0;  // R[rcur]: the latest iterate
    double* part = nullptr;
    double* deltas = nullptr;
    uint32_t* seeds_dev = nullptr;
    void* recs = nullptr;
    unsigned long long* changed = nullptr;
    unsigned long long* h_changed = nullptr;  // pinned
    double* h_deltas = nullptr;               // pinned
    uint64_t active = 0;
    uint64_t converged = 0;
    uint32_t iters[kBatchCols] = {};
    uint32_t max_iter[kBatchCols] = {};
    double tol[kBatchCols] = {};
    uint32_t* act = nullptr;  // k = 64 min-plus: activity bitmap
    std::vector<void*> owned;
};

namespace {

size_t records_bytes(const g1_graph* g, int kind, uint32_t k) {
    auto pick = [&](auto tag) -> size_t {
        constexpr int KK = decltype(tag)::value;
        return kind == kSessPpr ? spmm_records_bytes<PlusTimesKahanF32, KK>(g->n, g->m)
                                : spmm_records_bytes<MinPlusU32, KK>(g->n, g->m);
    };
    switch (k) {
        case 8: return pick(std::integral_constant<int, 8>());
        case 16: return pick(std::integral_constant<int, 16>());
        case 32: return pick(std::integral_constant<int, 32>());
        default: return pick(std::integral_constant<int, 64>());
    }
}

template <typename T>
g1_status sess_alloc(Session* ss, T** p, size_t count) {
    void* v = nullptr;
    G1_HIP(hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)));
    ss->owned.push_back(v);
    *p = static_cast<T*>(v);
    return G1_OK;
}

template <int K>
g1_status session_step_minplus(Session* ss, uint64_t* done) {
    g1_graph* g = ss->g;
    const hipStream_t s = ss->s;
    clear_flags_kernel<<<1, 64, 0, s>>>(ss->changed);
    G1_LAUNCHED();
    G1_TRY(launch_minplus<K>(g, ss->D, ss->mask[ss->mcur], ss->mask[ss->mcur ^ 1], ss->recs,
                             ss->kind == kSessMinPlusWeighted, ss->changed, s, ss->act));
    G1_HIP(hipMemcpyAsync(ss->h_changed, ss->changed, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    G1_HIP(hipStreamSynchronize(s));
    ss->mcur ^= 1;
    const uint64_t ch = ss->h_changed[0];
    uint64_t d = 0;
    for (uint32_t j = 0; j < K; ++j) {
        if (!((ss->active >> j) & 1ull) || ((ss->converged >> j) & 1ull)) continue;
        ++ss->iters[j];
        // Bellman-Ford settles within n rounds; the bound only guards the loop.
        if (!((ch >> j) & 1ull) || ss->iters[j] > g->n + 1) d |= 1ull << j;
    }
    ss->converged |= d;
    *done = d;
    return G1_OK;
}

template <int K>
g1_status session_step_ppr(Session* ss, uint64_t* done) {
    g1_graph* g = ss->g;
    const uint32_t n = g->n;
    const uint32_t nb = rank_blocks(n);
    const hipStream_t s = ss->s;
    G1_TRY(launch_plustimes<K>(g, ss->X, ss->Y, ss->recs, s));
    const unsigned long long live = ss->active & ~ss->converged;
    rank_apply_batch_kernel<K><<<(nb + 256 / K - 1) / (256 / K), 256, 0, s>>>(
        n, ss->Y, 0.0f, ss->seeds_dev, ss->restart, ss->iw, ss->R[ss->rcur], ss->R[ss->rcur ^ 1], ss->X, ss->part, nb,
        ~live, nullptr);
    G1_LAUNCHED();
    reduce_partials_kernel<0><<<dim3(1, K), 256, 0, s>>>(ss->part, nb, ss->deltas);
    G1_LAUNCHED();
    G1_HIP(hipMemcpyAsync(ss->h_deltas, ss->deltas, K * sizeof(double), hipMemcpyDeviceToHost, s));
    G1_HIP(hipStreamSynchronize(s));
    ss->rcur ^= 1;
    uint64_t d = 0;
    for (uint32_t j = 0; j < K; ++j) {
        if (!((live >> j) & 1ull)) continue;
        ++ss->iters[j];
        // g1_ppr's rule: stop after the iteration whose L1 delta is below tol, or at max_iter.
        if (ss->h_deltas[j] < ss->tol[j] || ss->iters[j] >= ss->max_iter[j]) d |= 1ull << j;
    }
    ss->converged |= d;
    *done = d;
    return G1_OK;
}

template <int K>
g1_status session_insert_k(Session* ss, uint32_t col, uint32_t src) {
    const uint32_t n = ss->g->n;
    const uint32_t blocks = fill_blocks(n);
    if (ss->kind == kSessPpr) {
        set_u32_kernel<<<1, 64, 0, ss->s>>>(ss->seeds_dev + col, src);
        G1_LAUNCHED();
        ppr_column_reset_kernel<K><<<blocks, 256, 0, ss->s>>>(n, ss->iw, ss->R[ss->rcur], ss->X, col, src);
    } else {
        minplus_column_reset_kernel<K><<<blocks, 256, 0, ss->s>>>(n, ss->D, ss->mask[ss->mcur], col, src);
    }
    G1_LAUNCHED();
    return G1_OK;
}

template <int K>
g1_status session_extract_k(Session* ss, uint32_t col, void* out) {
    const uint32_t n = ss->g->n;
    if (ss->kind == kSessPpr)
        column_extract_kernel<K, float><<<fill_blocks(n), 256, 0, ss->s>>>(n, ss->R[ss->rcur], col, static_cast<float*>(out));
    else
        column_extract_kernel<K, uint32_t><<<fill_blocks(n), 256, 0, ss->s>>>(n, ss->D, col, static_cast<uint32_t*>(out));
    G1_LAUNCHED();
    return G1_OK;
}

#define G1_SESSION_DISPATCH(fn, ss, ...)                                                              \
    switch ((ss)->K) {                                                                               \
        case 8: return fn<8>(ss, __VA_ARGS__);                                                       \
        case 16: return fn<16>(ss, __VA_ARGS__);                                                     \
        case 32: return fn<32>(ss, __VA_ARGS__);                                                     \
        default: return fn<64>(ss, __VA_ARGS__);                                                     \
    }

}  // namespace

size_t session_bytes(const g1_graph* g, int kind, uint32_t k) {
    const size_t n = g->n ? g->n : 1;
    size_t b = records_bytes(g, kind, k) + 64 * sizeof(unsigned long long);
    if (kind == kSessPpr)
        b += 4 * n * k * sizeof(float) + n * sizeof(float) +
             (static_cast<size_t>(rank_blocks(g->n)) * k + k) * sizeof(double) + k * sizeof(uint32_t);
    else
        b += (n + 1) * k * sizeof(uint32_t) + 2 * n * sizeof(unsigned long long);
    return b;
}

g1_status session_create(g1_graph* g, int kind, uint32_t k, float restart, hipStream_t stream, Session** out) {
    G1_CHECK(g != nullptr && out != nullptr, G1_ERR_INVALID_ARG, "session_create: null argument");
    G1_CHECK(k_slot(k) >= 0, G1_ERR_INVALID_ARG, "column width must be 8, 16, 32 or 64 (got %u)", k);
    G1_CHECK(g->has_t, G1_ERR_INVALID_ARG, "batched engines need the graph's transpose");
    G1_CHECK(kind != kSessMinPlusWeighted || g->t_w != nullptr, G1_ERR_INVALID_ARG, "graph has no weights");
    {
        std::lock_guard<std::mutex> lock(g->mu);  // lazily built graph arrays
        if (kind == kSessPpr) G1_TRY(ensure_pull_work(g, true));  // float weights, out-weight sums
        else if (k == 64) G1_TRY(ensure_row_starts(g, stream));
    }
    Session* ss = new Session();
    ss->g = g;
    ss->kind = kind;
    ss->K = k;
    ss->restart = restart;
    ss->s = stream;
    auto fail = [&](g1_status st) {
        session_destroy(ss);
        return st;
    };
    const size_t n = g->n ? g->n : 1;
    g1_status st = G1_OK;
    void* recs = nullptr;
    if ((st = sess_alloc(ss, reinterpret_cast<uint8_t**>(&recs), records_bytes(g, kind, k))) != G1_OK) return fail(st);
    ss->recs = recs;
    if ((st = sess_alloc(ss, &ss->changed, 64)) != G1_OK) return fail(st);
    if (hipHostMalloc(reinterpret_cast<void**>(&ss->h_changed), 64 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&ss->h_deltas), kBatchCols * sizeof(double)) != hipSuccess)
        return fail(set_error(G1_ERR_OUT_OF_MEMORY, "session_create: pinned host memory"));
    if (kind == kSessPpr) {
        for (int i = 0; i < 2 && st == G1_OK; ++i) st = sess_alloc(ss, &ss->R[i], n * k);
        if (st == G1_OK) st = sess_alloc(ss, &ss->X, n * k);
        if (st == G1_OK) st = sess_alloc(ss, &ss->Y, n * k);
        if (st == G1_OK) st = sess_alloc(ss, &ss->iw, n);
        if (st == G1_OK) st = sess_alloc(ss, &ss->part, static_cast<size_t>(rank_blocks(g->n)) * k);
        if (st == G1_OK) st = sess_alloc(ss, &ss->deltas, k);
        if (st == G1_OK) st = sess_alloc(ss, &ss->seeds_dev, k);
        if (st != G1_OK) return fail(st);
        inv_weight_kernel<<<fill_blocks(g->n), 256, 0, stream>>>(g->n, g->osum, 1.0f - restart, ss->iw);
        if (hipGetLastError() != hipSuccess) return fail(set_error(G1_ERR_HIP, "session_create: launch"));
    } else {
        if ((st = sess_alloc(ss, &ss->D, (n + 1) * k)) != G1_OK) return fail(st);  // row n: unreached
        if (k == 64 && (st = sess_alloc(ss, &ss->act, act_words(n))) != G1_OK) return fail(st);
        for (int i = 0; i < 2 && st == G1_OK; ++i) st = sess_alloc(ss, &ss->mask[i], n);
        if (st != G1_OK) return fail(st);
        // every column empty: D = INF, no source bits in either mask
        if ((st = g1_fill_u32(ss->D, kInf, (n + 1) * k, stream)) != G1_OK) return fail(st);
        for (int i = 0; i < 2 && st == G1_OK; ++i) st = g1_fill_u32(reinterpret_cast<uint32_t*>(ss->mask[i]), 0u, 2 * n, stream);
        if (st != G1_OK) return fail(st);
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(set_error(G1_ERR_HIP, "session_create: sync"));
    *out = ss;
    return G1_OK;
}

void session_destroy(Session* ss) {
    if (!ss) return;
    if (ss->s) (void)hipStreamSynchronize(ss->s);
    for (void* p : ss->owned) (void)hipFree(p);
    if (ss->h_changed) (void)hipHostFree(ss->h_changed);
    if (ss->h_deltas) (void)hipHostFree(ss->h_deltas);
    delete ss;
}

uint32_t session_width(const Session* ss) { return ss->K; }
uint64_t session_active(const Session* ss) { return ss->active; }

g1_status session_insert(Session* ss, uint32_t col, uint32_t source, uint32_t max_iter, double tol, bool* done_now) {
    G1_CHECK(col < ss->K && !((ss->active >> col) & 1ull), G1_ERR_INVALID_ARG, "session_insert: column %u busy", col);
    G1_CHECK(source < ss->g->n, G1_ERR_INVALID_ARG, "session_insert: source %u out of range", source);
    ss->active |= 1ull << col;
    ss->converged &= ~(1ull << col);
    ss->iters[col] = 0;
    ss->max_iter[col] = max_iter;
    ss->tol[col] = tol;
    const bool now = ss->kind == kSessPpr && max_iter == 0;
    if (now) ss->converged |= 1ull << col;
    if (done_now) *done_now = now;
    G1_SESSION_DISPATCH(session_insert_k, ss, col, source)
}

g1_status session_step(Session* ss, uint64_t* done, uint32_t* iters) {
    *done = 0;
    if (!(ss->active & ~ss->converged)) return G1_OK;
    g1_status st;
    if (ss->kind == kSessPpr) {
        switch (ss->K) {
            case 8: st = session_step_ppr<8>(ss, done); break;
            case 16: st = session_step_ppr<16>(ss, done); break;
            case 32: st = session_step_ppr<32>(ss, done); break;
            default: st = session_step_ppr<64>(ss, done); break;
        }
    } else {
        switch (ss->K) {
            case 8: st = session_step_minplus<8>(ss, done); break;
            case 16: st = session_step_minplus<16>(ss, done); break;
            case 32: st = session_step_minplus<32>(ss, done); break;
            default: st = session_step_minplus<64>(ss, done); break;
        }
    }
    if (iters)
        for (uint32_t j = 0; j < ss->K; ++j) iters[j] = ss->iters[j];
    return st;
}

g1_status session_extract(Session* ss, uint32_t col, void* out_dev) {
    G1_CHECK(col < ss->K && out_dev != nullptr, G1_ERR_INVALID_ARG, "session_extract: bad column or output");
    G1_SESSION_DISPATCH(session_extract_k, ss, col, out_dev)
}

void session_release(Session* ss, uint32_t col) {
    if (col >= ss->K || !((ss->active >> col) & 1ull)) return;
    // A min-plus column left before converging still has source bits: empty it so it stops propagating.
    if (ss->kind != kSessPpr && !((ss->converged >> col) & 1ull)) {
        switch (ss->K) {
            case 8: minplus_column_reset_kernel<8><<<fill_blocks(ss->g->n), 256, 0, ss->s>>>(ss->g->n, ss->D, ss->mask[ss->mcur], col, kInf); break;
            case 16: minplus_column_reset_kernel<16><<<fill_blocks(ss->g->n), 256, 0, ss->s>>>(ss->g->n, ss->D, ss->mask[ss->mcur], col, kInf); break;
            case 32: minplus_column_reset_kernel<32><<<fill_blocks(ss->g->n), 256, 0, ss->s>>>(ss->g->n, ss->D, ss->mask[ss->mcur], col, kInf); break;
            default: minplus_column_reset_kernel<64><<<fill_blocks(ss->g->n), 256, 0, ss->s>>>(ss->g->n, ss->D, ss->mask[ss->mcur], col, kInf); break;
        }
    }
    ss->active &= ~(1ull << col);
    ss->converged &= ~(1ull << col);
}

void batch_release(g1_graph* g) {
    for (BatchWork*& b : g->batch) {
        if (!b) continue;
        for (auto& row : b->minplus_exec)
            for (hipGraphExec_t& e : row)
                if (e) (void)hipGraphExecDestroy(e);
        for (auto& row : b->minplus16_exec)
            for (hipGraphExec_t& e : row)
                if (e) (void)hipGraphExecDestroy(e);
        delete b;  // device areas belong to the graph's allocation list
        b = nullptr;
    }
}

}  // namespace g1

extern "C" {

g1_status g1_batch_set_merge_tile(g1_graph* g, uint32_t tile_items) {
    using namespace g1;
    G1_CHECK(g != nullptr, G1_ERR_INVALID_ARG, "g1_batch_set_merge_tile: null graph");
    std::lock_guard<std::mutex> lock(g->mu);
    for (uint32_t i = 0; i < kWaveTileCount; ++i) {
        if (kWaveTileItems[i] == tile_items) {
            g->minplus_tile = i;
            return G1_OK;
        }
    }
    return set_error(G1_ERR_INVALID_ARG, "g1_batch_set_merge_tile: tile must be 512, 1024, 2048 or 4096 (got %u)",
                     tile_items);
}

g1_status g1_batch_autotune(g1_graph* g, int32_t kind, const uint32_t* sources, uint32_t count, int32_t weighted,
                            uint32_t* best_k, g1_stream_t stream) {
    using namespace g1;
    G1_CHECK(g != nullptr && best_k != nullptr, G1_ERR_INVALID_ARG, "g1_batch_autotune: null argument");
    G1_CHECK(kind == G1_BATCH_MINPLUS || kind == G1_BATCH_PPR, G1_ERR_INVALID_ARG, "g1_batch_autotune: unknown kind");
    G1_TRY(check_sources(g, sources, count, best_k, "g1_batch_autotune"));
    std::lock_guard<std::mutex> lock(g->mu);
        if (weighted && kind == G1_BATCH_MINPLUS)
        G1_CHECK(g->t_w != nullptr, G1_ERR_INVALID_ARG, "g1_batch_autotune: graph has no weights");
    return autotune(g, kind, sources, count, weighted != 0 && kind == G1_BATCH_MINPLUS, best_k, as_stream(stream));
}

g1_status g1_sssp_batch(g1_graph* g, const uint32_t* sources, uint32_t count, uint32_t k, int32_t weighted,
                        uint32_t* dist_dev, g1_batch_stats* stats, g1_stream_t stream) {
    using namespace g1;
    g1_batch_stats local;
    if (!stats) stats = &local;
    memset(stats, 0, sizeof(*stats));
    G1_TRY(check_sources(g, sources, count, dist_dev, "g1_sssp_batch"));
    if (weighted) G1_CHECK(g->t_w != nullptr, G1_ERR_INVALID_ARG, "g1_sssp_batch: graph has no weights");
    std::lock_guard<std::mutex> lock(g->mu);
    const hipStream_t s = as_stream(stream);
    if (k == 0) {
        if (!g->best_k[0]) {
            uint32_t best = 0;
            G1_TRY(autotune(g, G1_BATCH_MINPLUS, sources, count, weighted != 0, &best, s));
        }
        k = g->best_k[0];
    }
    stats->k = k;
    stats->merge_tile = k == 64 ? kWaveTileItems[g->minplus_tile < kWaveTileCount ? g->minplus_tile : kWaveTileOwn] : kMergeTile;
    G1_HIP(hipEventRecord(g->ev[0], s));
    G1_TRY(sssp_batch_dispatch(g, sources, count, k, weighted != 0, dist_dev, stats, s));
    return timed(g, s, stats);
}

g1_status g1_ppr_batch(g1_graph* g, const uint32_t* seeds, uint32_t count, float restart, uint32_t max_iter, double tol,
                       uint32_t k, float* rank_dev, g1_batch_stats* stats, g1_stream_t stream) {
    using namespace g1;
    g1_batch_stats local;
    if (!stats) stats = &local;
    memset(stats, 0, sizeof(*stats));
    G1_TRY(check_sources(g, seeds, count, rank_dev, "g1_ppr_batch"));
    G1_CHECK(restart > 0.0f && restart <= 1.0f, G1_ERR_INVALID_ARG, "g1_ppr_batch: restart must be in (0, 1]");
    std::lock_guard<std::mutex> lock(g->mu);
    const hipStream_t s = as_stream(stream);
    if (k == 0) {
        if (!g->best_k[1]) {
            uint32_t best = 0;
            G1_TRY(autotune(g, G1_BATCH_PPR, seeds, count, false, &best, s));
        }
        k = g->best_k[1];
    }
    stats->k = k;
    stats->merge_tile = kMergeTile;
    G1_HIP(hipEventRecord(g->ev[0], s));
    G1_TRY(ppr_batch_dispatch(g, seeds, count, k, restart, max_iter, tol, rank_dev, stats, s));
    return timed(g, s, stats);
}

}  // extern "C"
