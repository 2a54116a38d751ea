// The device-resident graph object behind g1_graph and its reusable work areas.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "internal.hpp"

struct g1_graph {
    uint32_t n = 0;
    uint32_t m = 0;
    uint32_t flags = 0;
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
    bool t_alias = false;
    void* t_ws = nullptr;  // merge-path partition of the in-edge CSR
    size_t t_ws_bytes = 0;

    // push work areas
    uint32_t* q[2] = {nullptr, nullptr};
    uint32_t* qoff[2] = {nullptr, nullptr};
    uint32_t* stamp = nullptr;
    uint32_t* fb[2] = {nullptr, nullptr};
    unsigned long long* ctr = nullptr;  // [0..2] frontier counters, [3] activated, [4..5] reached stats, [6] pull count
    // pull work areas
    uint32_t* xu = nullptr;
    uint32_t* yu = nullptr;
    float* xf = nullptr;
    float* yf = nullptr;
    float* iw = nullptr;
    float* r[2] = {nullptr, nullptr};
    double* partials = nullptr;  // per-block partial sums, then the folded scalar at [kPartials]
    unsigned long long* h_ctr = nullptr;  // pinned host mirror of ctr
    double* h_scalar = nullptr;           // pinned host scalar
    hipEvent_t ev[2] = {nullptr, nullptr};

    uint64_t device_bytes = 0;
    std::vector<void*> owned;  // every device allocation, freed by g1_graph_destroy
    std::mutex mu;
};

namespace g1 {

constexpr uint32_t kPartials = 4096;  // max blocks of the apply kernels (and partial slots)

g1_status graph_alloc(g1_graph* g, void** p, size_t bytes);
g1_status ensure_push_work(g1_graph* g);
g1_status ensure_pull_work(g1_graph* g, bool pagerank);

}  // namespace g1
