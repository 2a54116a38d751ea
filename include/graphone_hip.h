/*
 * graphone_hip.h -- C-ABI of libgraphone_hip.so, the gfx950 (MI355X) graph-computation hot path.
 *
 * The reference (GraphOne) is a header-only C++/CUDA template library with no ABI of its own: its
 * engines call the SpMV dispatcher, LaunchKernel and Thrust wrappers directly. Every entry point below
 * replaces one of those internal boundaries; the comment on each names the reference interface it
 * stands in for (paths relative to the reference root).
 *
 * Conventions
 *   - Every function returns g1_status (0 = G1_OK). On failure g1_last_error_message() holds a
 *     thread-local description. Errors are always checked (the reference's checkCudaErrors is a no-op
 *     unless DEBUG_CUDA, include/GraphOne/archi/check/cuda.cuh:10-16).
 *   - Pointers named *_dev are device (HBM) pointers; *_host are host pointers. Callers own every
 *     buffer they pass. Functions never allocate device memory, except the graph object (which owns
 *     the resident CSR and its per-graph work areas) and the serving registry.
 *   - Vertex ids, edge offsets and distances are uint32 (vid_t/eid_t, include/GraphOne/type.hpp:8-9).
 *     G1_INF = 0xFFFFFFFF is the unreached distance and the invalid vertex
 *     (include/GraphOne/utils/limits.hpp:54-73).
 *   - All work is stream-ordered on the given stream (NULL = the default stream).
 */
#ifndef GRAPHONE_HIP_H
#define GRAPHONE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t g1_status;
#define G1_OK 0
#define G1_ERR_INVALID_ARG 1
#define G1_ERR_HIP 2
#define G1_ERR_OUT_OF_MEMORY 3
#define G1_ERR_BUDGET 4
#define G1_ERR_NOT_FOUND 5
#define G1_ERR_BUSY 6
#define G1_ERR_CANCELLED 7
#define G1_ERR_INTERNAL 8
#define G1_ERR_NOT_CONVERGED 9

#define G1_INF 0xFFFFFFFFu

typedef void* g1_stream_t; /* a hipStream_t; NULL is the default stream */
typedef void* g1_event_t;  /* a hipEvent_t */

/* ---------------------------------------------------------------------------------------------
 * Status and runtime.  Replaces archi::memalloc/memfree/memset/memcpy
 * (include/GraphOne/archi/mem/cuda.cuh:11-59), archi::fill (include/GraphOne/archi/thrust/functions.h:28-31)
 * and archi::LaunchSync (include/GraphOne/archi/kernel/cuda.cuh:93-95).
 * ------------------------------------------------------------------------------------------- */
const char* g1_status_string(g1_status status);
const char* g1_last_error_message(void);
int32_t g1_abi_version(void);

g1_status g1_device_info(int32_t* device_count, int32_t* cu_count, uint64_t* hbm_bytes,
                         char* name, size_t name_len);
g1_status g1_malloc(void** ptr_dev, size_t bytes);
g1_status g1_free(void* ptr_dev);
g1_status g1_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes, g1_stream_t stream);
g1_status g1_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes, g1_stream_t stream);
g1_status g1_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes, g1_stream_t stream);
/* Fill kernels (not hipMemset), so they are safe inside captured HIP graphs. */
g1_status g1_fill_u32(uint32_t* dst_dev, uint32_t value, size_t count, g1_stream_t stream);
g1_status g1_fill_f32(float* dst_dev, float value, size_t count, g1_stream_t stream);
g1_status g1_stream_create(g1_stream_t* stream);
g1_status g1_stream_destroy(g1_stream_t stream);
g1_status g1_stream_synchronize(g1_stream_t stream);
g1_status g1_event_create(g1_event_t* event);
g1_status g1_event_destroy(g1_event_t event);
g1_status g1_event_record(g1_event_t event, g1_stream_t stream);
g1_status g1_event_synchronize(g1_event_t event);
g1_status g1_event_elapsed_ms(float* ms, g1_event_t start, g1_event_t stop);

/* ---------------------------------------------------------------------------------------------
 * Semiring SpMV  y = A (x) x  on a CSR matrix A (n_rows x n_cols, nnz nonzeros).
 *   min-plus u32 : y[r] = min_j ( x[c_j] == INF ? INF : A[r,c_j] + x[c_j] ), empty row -> INF
 *                  (the SSSP semiring of apps/sssp_blas.cu:55-73 with initialize = default_result)
 *   plus-times f32: y[r] = sum_j A[r,c_j] * x[c_j] with compensated (Kahan-Babuska) accumulation,
 *                  empty row -> 0 (the PageRank semiring of apps/pagerank_blas.cu:82-111)
 * `values_dev` may be NULL: every nonzero then has value 1 (an unweighted graph).
 * Results are deterministic: the plus-times fold order is a function of the matrix only.
 *
 * Merge-path (Merrill & Garland): replaces SpmvDispatcher<SpmvCudaCsrMergeBased>
 * (include/GraphOne/archi/blas/spmv/cuda/cuda.cuh:64-114): GetBufferSize -> workspace_bytes,
 * SegmentSearch (run once per matrix) -> partition, SpmvMerge -> spmv_merge_*.
 * ------------------------------------------------------------------------------------------- */
size_t g1_spmv_merge_workspace_bytes(uint32_t n_rows, uint32_t nnz);
g1_status g1_spmv_merge_partition(uint32_t n_rows, uint32_t nnz, const uint32_t* row_offsets_dev,
                                  void* workspace_dev, size_t workspace_bytes, g1_stream_t stream);
g1_status g1_spmv_merge_minplus_u32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                                    const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                                    const uint32_t* values_dev, const uint32_t* x_dev, uint32_t* y_dev,
                                    void* workspace_dev, size_t workspace_bytes, g1_stream_t stream);
g1_status g1_spmv_merge_plustimes_f32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                                      const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                                      const float* values_dev, const float* x_dev, float* y_dev,
                                      void* workspace_dev, size_t workspace_bytes, g1_stream_t stream);

/* Merge-path min-plus SpMV that allocates nothing and takes no workspace. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Each call partitions the matrix and multiplies with
 * the library's own per-GPU scratch array of g1_spmv_merge_noalloc_max_workspace_bytes() bytes (8 MiB,
 * reserved when the library loads and held until the process ends); a matrix whose
 * g1_spmv_merge_workspace_bytes() exceeds it is rejected with G1_ERR_INVALID_ARG (up to about 4.2e8
 * n_rows + nnz fit). Results equal g1_spmv_merge_minplus_u32's. Calls on one GPU share the scratch:
 * calls that overlap in time on one GPU (other streams, other threads) race and must be serialized by
 * the caller. */
size_t g1_spmv_merge_noalloc_max_workspace_bytes(void);
g1_status g1_spmv_merge_minplus_u32_noalloc(uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                                            const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                                            const uint32_t* values_dev, const uint32_t* x_dev, uint32_t* y_dev,
                                            g1_stream_t stream);

/* CSR-vector: a vector of L lanes (L in 1..64, a power of two) per row inside a wave64, L = half of
 * nnz / n_rows rounded up to a power of two unless lanes > 0 forces it. Replaces SpmvDispatcher<SpmvCudaCsrVector>
 * (include/GraphOne/archi/blas/spmv/cuda/cuda.cuh:10-61, csr_vector.cuh:20-68), written from its
 * definition (the reference kernel does not compile and is never instantiated). */
uint32_t g1_spmv_csrvec_lanes(uint32_t n_rows, uint32_t nnz);
g1_status g1_spmv_csrvec_minplus_u32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                                     const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                                     const uint32_t* values_dev, const uint32_t* x_dev, uint32_t* y_dev,
                                     uint32_t lanes, g1_stream_t stream);
g1_status g1_spmv_csrvec_plustimes_f32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                                       const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                                       const float* values_dev, const float* x_dev, float* y_dev,
                                       uint32_t lanes, g1_stream_t stream);

/* ---------------------------------------------------------------------------------------------
 * Step-level primitives, for a caller that keeps the reference's engine loop and its own buffers and
 * routes each step through the library: AdvanceGC::Forward (include/GraphOne/engine/advance_gc.h:180-244)
 * and the BlasGcEngine Construct / Gather / Apply steps (include/GraphOne/engine/blas_gc.h:21-76,
 * 171-435). A queue frontier is an array of vertex ids with a capacity the caller allocated and a length
 * in device memory (*count_dev), so steps chain on a stream with no host read; G1_INF entries are
 * skipped (the reference's INVALID slots). A bitmap frontier holds bit v of word v / 32 (ceil(n / 32)
 * words). Every step that needs scratch takes a caller-owned workspace of at least
 * g1_step_workspace_bytes(n, capacity) bytes (n: vertices; capacity: the queue capacity the step reads),
 * reusable across steps on one stream.
 * ------------------------------------------------------------------------------------------- */
size_t g1_step_workspace_bytes(uint32_t n, uint32_t capacity);

/* *edges_dev = sum of the out-degrees of the valid entries of frontier[0..*count_dev): the E of
 * AdvanceGC::ResizeOuput (include/GraphOne/engine/advance_gc.h:187-210), the quantity a caller compares
 * with m to choose push or pull. */
g1_status g1_frontier_density(uint32_t n, const uint32_t* row_offsets_dev, const uint32_t* frontier_dev,
                              const uint32_t* count_dev, uint32_t capacity, uint32_t* edges_dev,
                              void* workspace_dev, size_t workspace_bytes, g1_stream_t stream);

/* One advance + filter of the SSSP functor (include/GraphOne/engine/advance_gc.h:180-244 with
 * apps/sssp_adv.cu:49-71): for every valid u of frontier_in[0..*in_count_dev) and every out-edge
 * (u, v, w): nd = dist[u] + w, old = atomicMin(&dist[v], nd); v joins frontier_out when nd < old and
 * visited[v] != iter, and visited[v] = iter marks it (the functor's filter). dist[u] is read live, as the
 * reference functor reads it. frontier_out (n entries suffice: a vertex joins once) gets the winners in
 * no particular order and no G1_INF entries; *out_count_dev is set by the call. weights_dev NULL: unit
 * weights. frontier_out must not overlap frontier_in. */
g1_status g1_advance_sssp_u32(uint32_t n, const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                              const uint32_t* weights_dev, const uint32_t* frontier_in_dev,
                              const uint32_t* in_count_dev, uint32_t in_capacity, uint32_t* dist_dev,
                              uint32_t* visited_dev, uint32_t iter, uint32_t* frontier_out_dev,
                              uint32_t* out_count_dev, void* workspace_dev, size_t workspace_bytes,
                              g1_stream_t stream);

/* One BFS advance: every out-neighbour v of a valid frontier entry with depth[v] == G1_INF gets
 * depth[v] = level (compare-and-swap: its first visitor only) and joins frontier_out. */
g1_status g1_advance_bfs(uint32_t n, const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                         const uint32_t* frontier_in_dev, const uint32_t* in_count_dev, uint32_t in_capacity,
                         uint32_t* depth_dev, uint32_t level, uint32_t* frontier_out_dev, uint32_t* out_count_dev,
                         void* workspace_dev, size_t workspace_bytes, g1_stream_t stream);

/* Stable compaction: the valid (non-G1_INF) entries of in[0..*in_count_dev), in order, to out;
 * *out_count_dev their number. Turns a reference-style output with INVALID slots
 * (include/GraphOne/engine/advance_gc.h:212-244) into a dense queue. */
g1_status g1_filter_compact(const uint32_t* in_dev, const uint32_t* in_count_dev, uint32_t capacity,
                            uint32_t* out_dev, uint32_t* out_count_dev, void* workspace_dev, size_t workspace_bytes,
                            g1_stream_t stream);

/* Queue -> bitmap (the bitmap is cleared first); bitmap -> queue in ascending vertex order, the order of
 * the sparse engine's stable copy_if (include/GraphOne/engine/blas_gc.h:239-243). */
g1_status g1_frontier_queue_to_bitmap(uint32_t n, const uint32_t* frontier_dev, const uint32_t* count_dev,
                                      uint32_t capacity, uint32_t* bitmap_dev, g1_stream_t stream);
g1_status g1_frontier_bitmap_to_queue(uint32_t n, const uint32_t* bitmap_dev, uint32_t* frontier_dev,
                                      uint32_t* count_dev, void* workspace_dev, size_t workspace_bytes,
                                      g1_stream_t stream);

/* Min-plus apply of the BLAS engines (apps/sssp_blas.cu:75-82 inside include/GraphOne/engine/blas_gc.h:43-55):
 * for every v, pred = y[v] != G1_INF && y[v] < d[v]; then d[v] = y[v] and x[v] = y[v] (the next
 * Construct), else x[v] = G1_INF; *changed_dev = the number of such v (the next frontier's size). */
g1_status g1_apply_minplus_u32(uint32_t n, const uint32_t* y_dev, uint32_t* d_dev, uint32_t* x_dev,
                               uint32_t* changed_dev, void* workspace_dev, size_t workspace_bytes,
                               g1_stream_t stream);

/* PageRank apply of the all-active engine (apps/pagerank_blas.cu:98-123 inside
 * include/GraphOne/engine/blas_gc.h:353-435): r_new[v] = base + y[v], x[v] = r_new[v] * iw[v]. With
 * l1_dev, also *l1_dev = sum |r_new - r_old| in FP64, in a fixed order (blocks of 256 vertices, four
 * sequential quarter sums each, blocks folded in order): the same bits on every run. */
g1_status g1_pr_apply_f32(uint32_t n, const float* y_dev, float base, const float* iw_dev, const float* r_old_dev,
                          float* r_new_dev, float* x_dev, double* l1_dev, void* workspace_dev,
                          size_t workspace_bytes, g1_stream_t stream);
/* *out_dev = sum |a - b| in FP64, in the fixed order of g1_pr_apply_f32. */
g1_status g1_l1_delta_f64(uint32_t n, const float* a_dev, const float* b_dev, double* out_dev, void* workspace_dev,
                          size_t workspace_bytes, g1_stream_t stream);

/* Semiring SpMM Y = A (x) X with k dense columns, both row-major (X: n_cols x k, Y: n_rows x k), replacing
 * the reference's SpMM boundary (include/GraphOne/archi/blas/spmm/cuda/cuda.cuh:65-116, the GNN aggregation
 * of include/GraphOne/engine/blas_gnn.h:18-122) and serving batched multi-source steps:
 *   min-plus u32 : Y[r][j] = min over the nonzeros (r, c) of (X[c][j] == G1_INF ? G1_INF : A[r,c] + X[c][j])
 *   plus-times f32: Y[r][j] = sum over the nonzeros of A[r,c] * X[c][j], left to right per row with the
 *                  compensated sum of the SpMV kernels
 * col_mask_dev (device, ceil(k / 64) words; NULL = every column): only the columns whose bit is set are
 * computed, the others keep their Y values. values_dev NULL: unit values. The rows and nonzeros are
 * cut into merge-path tiles of equal size (a row of any length spreads over tiles), 64 columns per pass;
 * the workspace (the partition and the tiles' carries) is g1_spmm_workspace_bytes(n_rows, nnz) bytes. */
size_t g1_spmm_workspace_bytes(uint32_t n_rows, uint32_t nnz);
g1_status g1_spmm_minplus_u32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, uint32_t k,
                              const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                              const uint32_t* values_dev, const uint32_t* x_dev, uint32_t* y_dev,
                              const uint64_t* col_mask_dev, void* workspace_dev, size_t workspace_bytes,
                              g1_stream_t stream);
g1_status g1_spmm_plustimes_f32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, uint32_t k,
                                const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                                const float* values_dev, const float* x_dev, float* y_dev,
                                const uint64_t* col_mask_dev, void* workspace_dev, size_t workspace_bytes,
                                g1_stream_t stream);

/* ---------------------------------------------------------------------------------------------
 * Host-side synthetic graphs (seeded, splitmix64; the random streams are spelled out at the top of
 * graphone_amd/csrc/gen.cpp).  They produce the CSR invariants of GraphLoader + EdgeCache::ToCsr
 * (include/GraphOne/loader/graph_loader.h:254-296, include/GraphOne/loader/edge_cache.h:50-105):
 * no self loops, no duplicate (src,dst), columns ascending within a row, n = number of vertex ids.
 * ------------------------------------------------------------------------------------------- */
typedef struct g1_host_csr {
    uint32_t n;
    uint32_t m;
    uint32_t* row_offsets; /* n + 1 */
    uint32_t* col_indices; /* m */
    uint32_t* weights;     /* m, or NULL */
} g1_host_csr;

/* R-MAT: 2^scale vertices, edge_factor * 2^scale raw directed edges, quadrant probabilities a,b,c
 * (d = 1-a-b-c); self loops and duplicates removed. undirected != 0 adds every reverse edge first. */
g1_status g1_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t undirected, g1_host_csr* out);
/* rows x cols 4-neighbour grid, vertex id r*cols + c; each undirected grid edge kept with probability
 * p_keep, stored in both directions. */
g1_status g1_gen_grid(uint32_t rows, uint32_t cols, double p_keep, uint64_t seed, g1_host_csr* out);
/* Builds a CSR from an edge list (src[i] -> dst[i]) with the EdgeCache rules: self loops dropped unless
 * keep_self, reverse edges added unless directed, (src,dst) duplicates dropped unless keep_dup (the
 * first occurrence in input order keeps its weight). weights may be NULL. */
g1_status g1_csr_from_edges(uint32_t n, uint64_t num_edges, const uint32_t* src, const uint32_t* dst,
                            const uint32_t* weights, int32_t directed, int32_t keep_self,
                            int32_t keep_dup, g1_host_csr* out);
/* u32 weights uniform in [lo, hi] from a hash of (seed, src, dst); symmetric != 0 hashes the unordered
 * pair so both directions of an undirected edge get the same weight. Allocates out->weights. */
g1_status g1_assign_weights_u32(g1_host_csr* csr, uint32_t lo, uint32_t hi, uint64_t seed,
                                int32_t symmetric);
/* Transpose (in-edge CSR): rows are destinations, columns ascending source ids; weights follow. */
g1_status g1_host_csr_transpose(const g1_host_csr* in, g1_host_csr* out);
void g1_host_csr_free(g1_host_csr* csr);

/* Transpose of a device CSR, on the device (replaces ToCsc, include/GraphOne/mat/convert.h:62-88: expand the
 * row ids, sort by column keeping the original order among equal columns, count the new row starts): a
 * stable LSD radix sort, so each in-edge row lists its sources in ascending order with their weights.
 * t_row_offsets holds n + 1 entries, t_col_indices and t_weights m. weights and t_weights are both NULL or
 * both given (any 4-byte payload travels as its bits). The workspace is
 * g1_csr_transpose_workspace_bytes(n, m, weighted) bytes of device memory. */
size_t g1_csr_transpose_workspace_bytes(uint32_t n, uint32_t m, int32_t weighted);
g1_status g1_csr_transpose(uint32_t n, uint32_t m, const uint32_t* row_offsets_dev, const uint32_t* col_indices_dev,
                           const uint32_t* weights_dev, uint32_t* t_row_offsets_dev, uint32_t* t_col_indices_dev,
                           uint32_t* t_weights_dev, void* workspace_dev, size_t workspace_bytes, g1_stream_t stream);

/* ---------------------------------------------------------------------------------------------
 * Device-resident graph.  Replaces graph::build<arch, views> + Graph/GraphArchRef
 * (include/GraphOne/graph/builder.h:9-46, graph.h:13-140, graph_arch_ref.h:19-116): the out-edge CSR
 * ("normal" view) and optionally its transpose (the view BlasViews uses, engine/blas_gc.h:438).
 * The graph also owns the work areas its engines reuse across runs; runs on one graph object must
 * therefore not overlap in time.
 * ------------------------------------------------------------------------------------------- */
typedef struct g1_graph g1_graph;
#define G1_GRAPH_TRANSPOSE 1u  /* also keep the in-edge CSR (needed by pull / BLAS engines) */
#define G1_GRAPH_SYMMETRIC 2u  /* the CSR is symmetric: the transpose aliases it            */

g1_status g1_graph_create(g1_graph** graph, const g1_host_csr* csr, uint32_t flags);
g1_status g1_graph_destroy(g1_graph* graph);
/* A second handle on the same resident graph with its own work areas: runs on a view and runs on the base
 * (or on other views) may proceed at the same time on different streams, where runs on one handle are
 * serialized. The view shares the base's CSR arrays (no copy), copies the merge-path workspace (partition
 * and per-call tile records, about 16 bytes per 2048 nonzeros) and owns its work areas; it must be
 * destroyed before the base.
 * No counterpart in the reference (one graph, one engine per process). */
g1_status g1_graph_create_view(g1_graph* base, g1_graph** view);
g1_status g1_graph_info(const g1_graph* graph, uint32_t* n, uint32_t* m, uint64_t* device_bytes);

/* Work areas. A handle allocates the device memory an engine works in on that engine's first run and
 * keeps it until it is destroyed (the reference's engines allocate theirs when constructed and grow the
 * output frontier per iteration: engine/advance_gc.h:177-188, engine/blas_gc.h:137-141,
 * frontier/sparse_dbl.h:66-90). No counterpart in the reference: g1_graph_work_bytes returns how
 * many bytes the given set would add now (0 for what the handle already holds; the CSR-vector engine's
 * long-row lists, whose size depends on the data, at their bound of 4 bytes per vertex), and
 * g1_graph_prepare allocates and builds them on `stream` (synchronously), so a caller can budget and
 * place the allocation. g1_graph_info's device_bytes is the handle's total.
 *   G1_WORK_PUSH     queues, counters, distance words and history of the push engines (every
 *                    BFS / SSSP push engine; also the stats of the SpMV engines)
 *   G1_WORK_PULL     SpMV x / y vectors (min-plus engines, the push engines' pull steps) and the
 *                    in-edge CSR's long-row list
 *   G1_WORK_RANK     PageRank / PPR vectors, out-weight sums, float weights, L1 partials
 *   G1_WORK_RELABEL  plain PageRank's hub-first relabelled in-edge copy (implies G1_WORK_RANK; none for a
 *                    view, which keeps the graph's ids)
 *   G1_WORK_DELTA    the delta-stepping engine's weight-split CSR copy (the engine also needs PUSH | PULL)
 *   G1_WORK_ROAD     the road engine's far lists and query records (implies G1_WORK_PUSH)
 *   G1_WORK_HUB      the hub-first working copy the delta-stepping SSSP and direction-optimizing BFS run
 *                    on: the vertices relabelled by out-degree, largest first, the CSR (and its transpose)
 *                    in those ids with each row's columns ascending, the map from the graph's ids (4 bytes
 *                    per vertex, and a host copy), and the copy's own push, pull and delta work areas.
 *                    Those runs map the source in and their results back, so callers see the graph's
 *                    ids. Built on the first such run when the graph has at least 2^20 vertices, or here
 *                    for any size; none for a view or a road-like graph. G1_HUB=0 turns it off.
 * Unknown bits are rejected with G1_ERR_INVALID_ARG. */
#define G1_WORK_PUSH 1
#define G1_WORK_PULL 2
#define G1_WORK_RANK 4
#define G1_WORK_RELABEL 8
#define G1_WORK_DELTA 16
#define G1_WORK_ROAD 32
#define G1_WORK_HUB 64
#define G1_WORK_ALL 127
g1_status g1_graph_work_bytes(const g1_graph* graph, int32_t work, uint64_t* bytes);
g1_status g1_graph_prepare(g1_graph* graph, int32_t work, g1_stream_t stream);
g1_status g1_graph_csr_dev(const g1_graph* graph, int32_t transpose, const uint32_t** row_offsets_dev,
                           const uint32_t** col_indices_dev, const uint32_t** weights_dev);

/* ---------------------------------------------------------------------------------------------
 * Whole-algorithm runs (the Run<F>(comp, frontier) loop, include/GraphOne/base/run.h:8-39, with the
 * functor fixed).  Engines:
 *   G1_ENGINE_ADVANCE      push advance + fused filter/compaction (engine/advance_gc.h:162-283)
 *   G1_ENGINE_BLAS_MERGE   pull semiring SpMV engine, merge-path kernel (engine/blas_gc.h:171-351)
 *   G1_ENGINE_BLAS_CSRVEC  the same engine over the CSR-vector kernel
 *   G1_ENGINE_DIROPT       BFS only: queue/bitmap switching push<->pull by frontier density
 *   G1_ENGINE_DELTA        SSSP only: delta-stepping over the advance engine's kernels (light/heavy
 *                          buckets of the edge weights, then the advance loop); the same distances as
 *                          G1_ENGINE_ADVANCE with fewer relaxations. No counterpart in the reference
 *                          (whose sssp_adv loop is G1_ENGINE_ADVANCE); its copy of the CSR split by
 *                          weight takes 8 bytes per edge and 4 per vertex, built on the first call.
 * ------------------------------------------------------------------------------------------- */
#define G1_ENGINE_ADVANCE 0
#define G1_ENGINE_BLAS_MERGE 1
#define G1_ENGINE_BLAS_CSRVEC 2
#define G1_ENGINE_DIROPT 3
#define G1_ENGINE_DELTA 4

typedef struct g1_run_stats {
    uint32_t iterations;       /* engine iterations (Forward calls)                         */
    uint32_t push_iterations;  /* iterations run top-down (advance)                          */
    uint32_t pull_iterations;  /* iterations run bottom-up / SpMV                            */
    uint32_t dense_iterations; /* SSSP push: iterations relaxed without the winner test      */
    uint64_t frontier_sum;     /* sum over iterations of |F_i| (vertices activated)          */
    uint64_t edges_sum;        /* sum over iterations of E(F_i) (push) or nnz scanned (pull) */
    uint64_t reached;          /* vertices with a finite result                              */
    uint64_t reached_edges;    /* sum of out-degrees of reached vertices (GTEPS numerator)   */
    double elapsed_ms;         /* device time of the run                                     */
    double last_delta;         /* PageRank: last L1 (or max) delta                           */
} g1_run_stats;

/* Per-kernel profile of the runs on one graph (HIP events around every engine launch, recorded on the
 * run's stream). While enabled, every launch adds its device time and its algorithmic bytes (the bytes
 * the kernel must move by definition of its work, not the bytes it happens to move) to the entry of
 * its kernel. g1_profile_read copies up to `capacity` entries and clears them. */
typedef struct g1_kernel_profile {
    char name[40];
    uint64_t launches;
    double total_ms;
    double algorithmic_bytes;
} g1_kernel_profile;
g1_status g1_profile_enable(g1_graph* graph, int32_t enable);
g1_status g1_profile_read(g1_graph* graph, g1_kernel_profile* out, uint32_t capacity, uint32_t* count);

/* BFS depths (hops), depth_dev[src] = 0, unreached = G1_INF. */
g1_status g1_bfs(g1_graph* graph, uint32_t src, uint32_t* depth_dev, int32_t engine,
                 g1_run_stats* stats, g1_stream_t stream);
/* SSSP distances over u32 weights (min-plus), dist_dev[src] = 0, unreached = G1_INF.
 * Relaxation is atomicMin on uint32 (apps/sssp_adv.cu:49-60). */
g1_status g1_sssp(g1_graph* graph, uint32_t src, uint32_t* dist_dev, int32_t engine,
                  g1_run_stats* stats, g1_stream_t stream);

/* PageRank (apps/pagerank_blas.cu:16-125): iw[u] = alpha / sum_out w (0 if none), r0 = 1/N,
 * r' = (1-alpha)/N + sum_{u->v} w * r[u] * iw[u]; no dangling redistribution.
 * FP32 values, Kahan row sums. converge_mode:
 *   G1_PR_L1_F64   stop when the FP64 L1 norm of r'-r is < tol or after max_iter iterations
 *   G1_PR_MAX_ABS  the reference's test: before each iteration stop if max|r_cur - r_pre| <= tol
 *                  (the first test compares r0 with a zero vector, as apps/pagerank_blas.cu:63-78)
 * rank_dev receives the final iterate (for G1_PR_MAX_ABS: the array the reference prints,
 * ranks[opt_idx], i.e. the iterate before the last). */
#define G1_PR_L1_F64 0
#define G1_PR_MAX_ABS 1
g1_status g1_pagerank(g1_graph* graph, float alpha, uint32_t max_iter, double tol,
                      int32_t converge_mode, int32_t engine, float* rank_dev, g1_run_stats* stats,
                      g1_stream_t stream);

/* Personalized PageRank from one seed: r0 = e_seed, r' = restart * e_seed + sum_{u->v} w*r[u]*iw[u]
 * with iw[u] = (1-restart)/sum_out w. Stops when FP64 L1(r'-r) < tol or after max_iter. */
g1_status g1_ppr(g1_graph* graph, uint32_t seed_vertex, float restart, uint32_t max_iter, double tol,
                 float* rank_dev, g1_run_stats* stats, g1_stream_t stream);

/* Many BFS (kinds[i] = 0) / SSSP (kinds[i] = 1; kinds NULL: all SSSP) queries on a low-degree graph (every
 * out-degree <= 16: grids, road networks), one workgroup per query, 64 per launch: each runs to the end
 * inside its workgroup (SSSP near-far, bucket width 16x the heaviest edge), which suits graphs of
 * thousands of levels. out_dev holds count x n u32, source-major; iterations (host, may be NULL) gets
 * each query's iteration count. Results equal g1_bfs / g1_sssp bit for bit. No counterpart in the
 * reference. */
g1_status g1_road_batch(g1_graph* graph, const uint32_t* kinds, const uint32_t* sources, uint32_t count,
                        uint32_t* out_dev, uint32_t* iterations, g1_stream_t stream);

/* ---- batched multi-source engines ------------------------------------------------------------
 * No single counterpart in the reference: they run many single-source queries of the reference's apps
 * (apps/sssp_blas.cu:60-80 for min-plus, apps/pagerank_blas.cu:93-123 for the PageRank iteration)
 * together, as the columns of one semiring SpMM over the in-edge CSR. Sources are cut into column tiles
 * of k in {8, 16, 32, 64}; k = 0 picks the autotuned width of the graph (g1_batch_autotune on first use).
 * Column c of the output (out_dev + c * n) is exactly what the single-source call returns for
 * sources[c]: bit-identical distances, and bit-identical PPR ranks (same fold order of every row sum,
 * same per-256-vertex FP64 L1 tree, same stop rule per column). The graph needs its transpose. */
typedef struct g1_batch_stats {
    uint32_t k;               /* column tile width used                               */
    uint32_t tiles;           /* column tiles run                                     */
    uint64_t iterations_sum;  /* SpMM iterations summed over the tiles                */
    double elapsed_ms;        /* device time of the whole call (autotuning excluded)  */
    uint32_t merge_tile;      /* rows + nonzeros per merge tile of the SpMM kernel      */
    uint32_t reserved;
} g1_batch_stats;

#define G1_BATCH_MINPLUS 0
#define G1_BATCH_PPR 1

/* Times one tile of every width on up to 64 of `sources` (repeated when fewer) and stores the width with
 * the lowest time per source as the graph's default for `kind`. weighted selects the min-plus variant
 * timed (ignored for G1_BATCH_PPR). When min-plus picks k = 64, the merge tile of its kernel (512, 1024,
 * 2048 or 4096 rows + nonzeros; every choice gives the same bits) is timed too and kept for the graph's later
 * batches (g1_batch_stats.merge_tile reports it). */
g1_status g1_batch_autotune(g1_graph* graph, int32_t kind, const uint32_t* sources, uint32_t count,
                            int32_t weighted, uint32_t* best_k, g1_stream_t stream);

/* Sets the merge tile of the graph's k = 64 min-plus batches (512, 1024, 2048 or 4096 rows + nonzeros;
 * the autotuner's choice until then, 2048 without one). No counterpart in the reference: it extends
 * g1_batch_autotune, for callers that tune once and reuse the result. */
g1_status g1_batch_set_merge_tile(g1_graph* graph, uint32_t tile_items);

/* Multi-source SSSP (weighted != 0) or BFS depths (weighted == 0): dist_dev holds count x n u32,
 * source-major. sources are host memory. */
g1_status g1_sssp_batch(g1_graph* graph, const uint32_t* sources, uint32_t count, uint32_t k,
                        int32_t weighted, uint32_t* dist_dev, g1_batch_stats* stats, g1_stream_t stream);

/* Multi-seed personalized PageRank (g1_ppr per column): rank_dev holds count x n floats, seed-major. */
g1_status g1_ppr_batch(g1_graph* graph, const uint32_t* seeds, uint32_t count, float restart,
                       uint32_t max_iter, double tol, uint32_t k, float* rank_dev, g1_batch_stats* stats,
                       g1_stream_t stream);

/* ---- query server ------------------------------------------------------------------------------
 * No counterpart in the reference (its apps run one query per process, apps/sssp_adv.cu:74-124): a
 * persistent request loop over graphs resident in HBM. Per graph, a session thread and `single_workers`
 * single-source threads, each with its own HIP stream and its own view of the graph (g1_graph_create_view),
 * so queries on one graph and on different graphs overlap on the device. BFS runs on the single-source
 * direction-optimizing engine; SSSP and personalized PageRank join a continuous-batching column session
 * (a multi-source SpMM whose columns are freed and refilled one query at a time while the others keep
 * iterating) when one is open or when enough of them wait (the session opens if its arrays fit the budget;
 * SSSP only on graphs whose single-source runs took at most 64 iterations, where a full SpMM per iteration
 * pays), and run on the single-source engines otherwise; on low-degree graphs (g1_road_batch) a
 * single-source thread runs up to 32 waiting BFS / SSSP queries in one launch. Every result is the one the
 * single-source call returns (bit-identical), whichever path and whichever neighbours it ran with.
 * Admission: a query is admitted when its result vector fits in the HBM budget, counting the registered
 * graphs, open sessions and results not yet fetched; otherwise it waits, in order, for memory to be freed
 * (fetch, cancel, a session closing after half a second idle). */
typedef struct g1_server g1_server;

typedef struct g1_server_config {
    uint64_t hbm_budget_bytes; /* 0: 90% of the device's free memory at creation                   */
    uint32_t batch_k;          /* widest session (8, 16, 32, 64); 0: 64. A session opens at the power
                                  of two >= twice the queries waiting, capped here                   */
    uint32_t batch_min_sssp;   /* open an SSSP session when this many SSSP queries wait; 0: 32      */
    uint32_t batch_min_ppr;    /* open a PPR session when this many PPR queries wait; 0: 2           */
    float ppr_restart;         /* restart probability of every PPR query; 0: 0.15                   */
    uint32_t single_workers;   /* single-source runs in flight per graph (graph views); 0: 4         */
    uint32_t reserved[2];
} g1_server_config;

#define G1_QUERY_BFS 0
#define G1_QUERY_SSSP 1
#define G1_QUERY_PPR 2

typedef struct g1_query {
    uint32_t graph_id;
    int32_t kind;       /* G1_QUERY_*                                   */
    uint32_t source;    /* source (BFS/SSSP) or seed vertex (PPR)       */
    uint32_t max_iter;  /* PPR only                                     */
    double tol;         /* PPR only: FP64 L1 stop threshold             */
} g1_query;

#define G1_Q_QUEUED 0
#define G1_Q_RUNNING 1
#define G1_Q_DONE 2
#define G1_Q_FAILED 3
#define G1_Q_CANCELLED 4
#define G1_Q_FETCHED 5

#define G1_PATH_SINGLE 0  /* single-source engine           */
#define G1_PATH_BATCHED 1 /* a column of a batching session */

typedef struct g1_query_status {
    int32_t state;          /* G1_Q_*                                                          */
    int32_t status;         /* g1_status of a failed query                                     */
    int32_t path;           /* G1_PATH_*                                                       */
    uint32_t iterations;    /* engine iterations the query ran                                 */
    double t_submit_ms;     /* server clock (ms since the server started)                      */
    double t_admit_ms;
    double t_done_ms;
    uint64_t result_bytes;  /* n * 4: u32 distances / depths, or f32 ranks                     */
} g1_query_status;

typedef struct g1_server_stats {
    uint64_t submitted, completed, failed, cancelled;
    uint64_t queued, running;
    uint64_t admission_waits;  /* queries that waited for HBM at the head of their graph's queue */
    uint64_t batched, single;  /* completed queries per path                                    */
    uint64_t session_steps;    /* batched iterations run                                        */
    uint64_t hbm_budget, hbm_used, hbm_peak;
} g1_server_stats;

g1_status g1_server_create(const g1_server_config* config, g1_server** out);
/* Cancels what is still queued or running, joins the workers, frees every result and session. */
g1_status g1_server_destroy(g1_server* server);
/* The graph stays owned by the caller and must outlive the server; it needs its transpose. */
g1_status g1_server_add_graph(g1_server* server, g1_graph* graph, uint32_t* graph_id);
/* Unregisters a resident graph: new submits on it fail with G1_ERR_NOT_FOUND; the call blocks until its
 * queued and running queries reach a final state (their results stay fetchable), then stops the graph's
 * threads, frees its views and returns its bytes to the budget. The graph itself stays the caller's. Not
 * to be called concurrently with g1_server_destroy. */
g1_status g1_server_remove_graph(g1_server* server, uint32_t graph_id);
/* Sets the HBM budget (bytes > 0); a larger budget admits waiting queries at once, a smaller one holds new
 * admissions until usage falls below it. */
g1_status g1_server_set_budget(g1_server* server, uint64_t hbm_budget_bytes);
g1_status g1_server_submit(g1_server* server, const g1_query* query, uint64_t* query_id);
g1_status g1_server_poll(g1_server* server, uint64_t query_id, g1_query_status* status);
/* Waits until the query is done, failed or cancelled, or timeout_ms passes (G1_ERR_BUSY then). */
g1_status g1_server_wait(g1_server* server, uint64_t query_id, double timeout_ms, g1_query_status* status);
/* Copies a done query's result (n values) to dst and frees it on the device. For a failed query it
 * returns the engine's status, and g1_last_error_message() names the engine's error. */
g1_status g1_server_fetch(g1_server* server, uint64_t query_id, void* dst, uint64_t dst_bytes, int32_t dst_on_device);
g1_status g1_server_cancel(g1_server* server, uint64_t query_id);
g1_status g1_server_get_stats(g1_server* server, g1_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHONE_HIP_H */
