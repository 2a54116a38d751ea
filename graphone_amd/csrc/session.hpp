// Continuous-batching column sessions: the batched engines of batch.hip opened up so that queries can
// join and leave a running SpMM one column at a time (the serving loop of serve.cpp drives them).
//
// A session belongs to one graph and one kind (min-plus SSSP/BFS, or personalized PageRank at one
// restart value) and owns every mutable device array it uses; it only reads the graph's CSR arrays, so
// sessions of one graph, and the graph's single-source engines, can run at the same time on different
// streams. A session is not thread-safe: one host thread drives it.
//
// Column arithmetic never depends on the other columns, so a query's result is the same whichever
// columns run beside it: bit-identical to the single-source run (g1_sssp / g1_bfs / g1_ppr).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "graph.hpp"

namespace g1 {

enum SessionKind : int { kSessMinPlusWeighted = 0, kSessMinPlusUnit = 1, kSessPpr = 2 };

struct Session;

// Device bytes a session of this graph, kind and width allocates.
size_t session_bytes(const g1_graph* g, int kind, uint32_t k);
// k in {8, 16, 32, 64}; restart is used by kSessPpr only.
g1_status session_create(g1_graph* g, int kind, uint32_t k, float restart, hipStream_t stream, Session** out);
void session_destroy(Session* s);

uint32_t session_width(const Session* s);
uint64_t session_active(const Session* s);  // bit j: column j holds a query
// Starts a query in a free column (source vertex; PPR: tol and max_iter). *done_now is set when the
// query needs no iteration (PPR with max_iter 0): extract it at once.
g1_status session_insert(Session* s, uint32_t col, uint32_t source, uint32_t max_iter, double tol, bool* done_now);
// One iteration over every active column. *done receives the columns that finished with it (min-plus:
// no distance of the column changed; PPR: the column's L1 delta < tol or its max_iter reached) and
// iters[j] the iterations column j has run. The columns stay allocated until session_release.
g1_status session_step(Session* s, uint64_t* done, uint32_t* iters);
// Copies column col (n values, u32 distances or f32 ranks) to out_dev, in stream order.
g1_status session_extract(Session* s, uint32_t col, void* out_dev);
void session_release(Session* s, uint32_t col);

}  // namespace g1
