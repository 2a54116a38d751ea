// The device-resident graph object behind g1_graph and its reusable work areas.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

struct g1_graph {
    uint32_t n = 0;
    uint32_t m = 0;
    uint32_t flags = 0;
    uint32_t max_out_deg = 0;  // largest out-degree (set at creation)
    uint32_t max_w = 0;        // largest edge weight (set at creation; 0 without weights)
    // out-edge CSR ("normal" view)
    uint32_t* row_off = nullptr;
    uint32_t* col = nullptr;
    uint32_t* w = nullptr;
    // in-edge CSR ("transpose" view); aliases the out-edge CSR for a symmetric graph
    uint32_t* t_row_off = nullptr;
    uint32_t* t_col = nullptr;
    uint32_t* t_w = nullptr;
    float* t_wf = nullptr;  // float copy of t_w for plus-times engines (lazy)
    bool has_t = false;
    bool is_view = false;  // a handle of g1_graph_create_view (shares the base's CSR arrays)
    bool t_alias = false;
    void* t_ws = nullptr;  // merge-path partition of the in-edge CSR
    size_t t_ws_bytes = 0;
    unsigned long long* t_starts = nullptr;  // bitmap of in-edge CSR row starts (lazy, batched min-plus)

    // plain PageRank over the vertices relabelled hub first (lazy, ensure_pr_relabel): new id i is old id
    // pr_perm[i]; the in-edge CSR, its weights, the out-weight sums and the merge partition in the new ids
    uint32_t* pr_perm = nullptr;
    uint32_t* pr_row_off = nullptr;
    uint32_t* pr_col = nullptr;
    float* pr_wf = nullptr;
    float* pr_osum = nullptr;
    void* pr_ws = nullptr;
    // the push engines' hub-first working copy (lazy, ensure_hub_copy): a graph of its own over the vertices
    // in groups of out-degree (powers of two), largest first, with each row's columns ascending; vertex v
    // here is vertex hub_inv[v] there (device, and the host copy for sources). The copy's runs write their
    // results in this graph's ids: out_inv is set in the copy (hub_inv) and out_row_off to this graph's
    // row offsets (for the reached-edge stats), out_work holds the copy's BFS depths before they are
    // mapped, and its profile entries go to prof_to.
    g1_graph* hub = nullptr;
    uint32_t* hub_inv = nullptr;
    std::vector<uint32_t> hub_inv_h;
    const uint32_t* out_inv = nullptr;
    const uint32_t* out_row_off = nullptr;
    uint32_t* out_work = nullptr;
    g1_graph* prof_to = nullptr;
    // CSR-vector long rows (lazy, ensure_long_rows): rows of the in-edge CSR, and of the relabelled one,
    // longer than long_min nonzeros, longest first; the in-edge list's first t_nhuge rows are longer than
    // huge_min (the delta engine's heavy pull gives those a workgroup each)
    uint32_t* t_long = nullptr;
    uint32_t t_nlong = 0;
    uint32_t t_nhuge = 0;
    uint32_t huge_min = 0;
    uint32_t* pr_long = nullptr;
    uint32_t pr_nlong = 0;
    uint32_t long_min = 0;

    // push work areas
    uint32_t* q[2] = {nullptr, nullptr};
    uint32_t* qoff[2] = {nullptr, nullptr};
    uint32_t* stamp = nullptr;
    uint32_t seg_cap = 0;                       // entries per queue segment
    unsigned long long* qcnt = nullptr;         // 3 rotating sets of kSegments queue counters
    unsigned long long* slots = nullptr;        // per-workgroup stats slots (4 regions of kMaxPushBlocks)
    uint32_t* overflow = nullptr;
    uint32_t* fb[2] = {nullptr, nullptr};
    uint32_t* fb3 = nullptr;                    // three rotating frontier bitmaps (device-driven BFS)
    unsigned long long* dstamp = nullptr;       // SSSP (dist << 32 | stamp) words
    uint8_t* dbound = nullptr;                  // SSSP distance upper bounds, one byte per vertex (DistBound)
    void* sweep_pieces = nullptr;               // SweepPiece records of the long out-edge rows (lazy)
    uint32_t n_sweep_pieces = 0;
    unsigned long long* hist = nullptr;         // per-iteration history of the push engines
    // the delta-stepping SSSP engine's copy of the out-edge CSR with each row's light edges (weight below
    // dl_delta) first (lazy, ensure_delta_split): columns, weights, light counts; per-workgroup minima of
    // its distance scans
    uint32_t* dl_col = nullptr;
    uint32_t* dl_w = nullptr;
    uint32_t* dl_lc = nullptr;
    uint32_t* dl_mins = nullptr;
    uint8_t* dl_xb = nullptr;  // the heavy pull step's byte x (MinPlusU8X), n rounded up to 4
    uint32_t dl_delta = 0;
    uint32_t* road_far[2] = {nullptr, nullptr};  // far lists of the road-graph kernel (road_far_cap each)
    uint32_t road_far_cap = 0;
    void* road_q = nullptr;                     // RoadQuery records (kRoadSlots)
    uint32_t** road_outs = nullptr;             // per-record output vectors (kRoadSlots)
    // per-query state of batched road-graph runs (slot i: dstamp, 2 near buffers, 2 far lists)
    std::vector<void*> road_slots;
    unsigned long long* ctr = nullptr;  // [0..2] frontier counters, [3] activated, [4..5] reached stats, [6] pull count
    // pull work areas
    uint32_t* xu = nullptr;
    uint32_t* yu = nullptr;
    float* xf = nullptr;
    float* yf = nullptr;
    float* iw = nullptr;
    float* osum = nullptr;  // exact out-weight sum (or out-degree) of every vertex, rounded to float
    float* r[2] = {nullptr, nullptr};
    double* partials = nullptr;  // per-block partial sums, then the folded scalar at [kPartials]
    double* rpart = nullptr;     // canonical L1 partials: kBatchCols columns x ceil(n/256) blocks (+ sums)
    uint32_t* seeds_dev = nullptr;  // seed vertex per column (kBatchCols)
    struct BatchWork* batch[4] = {nullptr, nullptr, nullptr, nullptr};  // per column count 8, 16, 32, 64
    uint32_t best_k[2] = {0, 0};                                        // autotuned k: [0] min-plus, [1] PPR
    uint32_t minplus_tile = 2;  // merge tile of the k = 64 min-plus kernel: index into kWaveTileItems (2: 2048)
    void* wave_coords[4] = {};  // merge partitions of the in-edge CSR per kWaveTileItems ([2]: t_ws)
    unsigned long long* h_ctr = nullptr;  // pinned host mirror of ctr (16) and a counter set (at [16])
    double* h_scalar = nullptr;           // pinned host scalar
    unsigned long long* h_hist = nullptr;  // pinned copy of the push engines' iteration history (lazy)
    unsigned long long* h_mbox = nullptr;  // coherent pinned mailbox: [0] post sequence, [1] nf, [2] E
    unsigned long long post_seq = 0;       // last sequence number posted to h_mbox
    hipEvent_t ev[2] = {nullptr, nullptr};

    // per-kernel profile (g1_profile_enable / g1_profile_read)
    struct ProfEntry {
        std::string name;
        uint64_t launches = 0;
        double ms = 0.0;
        double bytes = 0.0;
    };
    struct ProfPending {
        int entry;
        hipEvent_t a, b;
        double bytes;
    };
    bool profile = false;
    std::vector<ProfEntry> prof;
    std::vector<ProfPending> pending;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;

    uint64_t device_bytes = 0;
    std::vector<void*> owned;  // every device allocation, freed by g1_graph_destroy
    std::mutex mu;
};

namespace g1 {

constexpr uint32_t kPartials = 4096;     // max blocks of the apply kernels (and partial slots)
constexpr uint32_t kHistIters = 1u << 17;  // iterations a push run may take
constexpr uint32_t kBatchCols = 64;        // widest column tile of the batched engines
constexpr uint32_t kCsrPad = 64;           // zero entries past m in the column and weight arrays
inline uint32_t rank_blocks(uint32_t n) { return (n + 255) / 256; }
g1_status ensure_rank_partials(g1_graph* g);
void batch_release(g1_graph* g);  // captured graphs and host records of the batched engines

g1_status graph_alloc(g1_graph* g, void** p, size_t bytes);

// Low-degree graphs (grids, road networks) run BFS / SSSP on the single-workgroup road engine;
// road_queries runs count such queries (kinds[i]: 1 SSSP, 0 BFS) together, ok[i] = 0 when query i must be
// run again on the push engine.
bool is_road_graph(const g1_graph* g);
g1_status road_queries(g1_graph* g, const uint32_t* kinds, const uint32_t* sources, uint32_t count, uint32_t** outs,
                       uint8_t* ok, uint32_t* iters, hipStream_t s);
g1_status ensure_push_work(g1_graph* g, hipStream_t s);  // work areas cleared on stream s
g1_status ensure_pull_work(g1_graph* g, bool pagerank);
g1_status ensure_sweep_pieces(g1_graph* g, hipStream_t s);  // pieces of the long rows of the out-edge CSR
g1_status ensure_pr_relabel(g1_graph* g, hipStream_t s);  // after ensure_pull_work(g, true)
// The push engines' hub-first copy (graph.hip; 8 bytes per vertex here, the copy's CSR and work areas in
// it) and its CSR build (convert.hip). Used by the delta-stepping SSSP and direction-optimizing BFS runs
// when hub_wanted (engines.hip).
g1_status ensure_hub_copy(g1_graph* g, hipStream_t s);
g1_status build_hub_csr(const g1_graph* g, bool symmetric, uint32_t* inv, uint32_t* inv_host, uint32_t* row_off, uint32_t* col, uint32_t* w,
                        uint32_t* t_row_off, uint32_t* t_col, uint32_t* t_w, hipStream_t s);
bool hub_eligible(const g1_graph* g);
g1_status ensure_delta_split(g1_graph* g, uint32_t delta, hipStream_t s);  // 8m + 4n bytes, rebuilt when delta changes

// On-device CSR conversions (convert.hip): the transpose (workspace of transpose_workspace_bytes), an
// exclusive scan of u32 (partials: one entry per 4096 items), u32 -> f32.
size_t transpose_workspace_bytes(uint32_t n, uint32_t m, bool weighted);
g1_status device_transpose(uint32_t n, uint32_t m, const uint32_t* row_off, const uint32_t* col, const uint32_t* w,
                           uint32_t* t_row_off, uint32_t* t_col, uint32_t* t_w, void* ws, hipStream_t s);
g1_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* partials, hipStream_t s);
g1_status device_u32_to_f32(const uint32_t* in, float* out, size_t count, hipStream_t s);
g1_status ensure_long_rows(g1_graph* g, bool relabelled, hipStream_t s);
// engines.hip: the road engine's work areas (allocate / bytes they would add) and the delta engine's width
g1_status prepare_road_work(g1_graph* g, hipStream_t s);
uint64_t road_work_size(const g1_graph* g);
uint32_t delta_width(const g1_graph* g);

// Profiling: prof_begin records a start event (returns -1 when profiling is off), prof_end the stop
// event with the launch's algorithmic bytes; prof_flush (after the stream is synchronized) adds the
// elapsed times to the per-kernel entries.
int prof_begin(g1_graph* g, const char* name, hipStream_t s);
void prof_end(g1_graph* g, int token, hipStream_t s, double bytes);
void prof_set_bytes(g1_graph* g, int token, double bytes);
g1_status prof_flush(g1_graph* g);

}  // namespace g1
