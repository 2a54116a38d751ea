// Query server: graphs resident in HBM, a request loop per graph, HBM-budget admission and continuous
// batching of SSSP / personalized-PageRank queries into column sessions (session.hpp).
//
// The reference has no serving layer (each app loads one graph and answers one query per process,
// apps/sssp_adv.cu:74-124, apps/pagerank_blas.cu:127-175); the engines under it are the same ones the
// single-source C-ABI calls use (g1_bfs, g1_sssp, g1_ppr) and the batched SpMM of batch.hip.
//
// Threads: callers submit, poll, wait, fetch and cancel from any thread (one server mutex guards the
// query table and the budget); one worker thread per registered graph owns that graph's HIP stream and
// sessions, and never holds the server mutex while it calls into HIP.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "graph.hpp"
#include "internal.hpp"
#include "session.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Query {
    uint64_t id = 0;
    g1_query q{};
    int32_t state = G1_Q_QUEUED;
    g1_status status = G1_OK;
    int32_t path = G1_PATH_SINGLE;
    uint32_t iterations = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t bytes = 0;        // result bytes (n * 4)
    void* result = nullptr;    // device buffer, owned by the server until fetched
    bool cancel = false;       // cancel requested while running
    bool reserved = false;     // its bytes are counted in hbm_used
    std::string error;         // the engine's error message when the query failed
};

struct GraphSlot {
    uint32_t id = 0;
    g1_graph* g = nullptr;
    uint32_t n = 0;
    // guarded by the server mutex
    std::deque<Query*> queue;     // submitted, not admitted (FIFO)
    std::deque<Query*> admitted;  // admitted, not yet routed by the session thread
    std::deque<Query*> singles;   // routed to a single-source engine, not started
    uint64_t sssp_runs = 0, sssp_iters = 0;  // single-source SSSP iterations seen on this graph
    // threads: one session thread, `single_workers` single-source threads with a view each
    std::thread worker;
    std::vector<std::thread> helpers;
    std::vector<g1_graph*> views;
    uint64_t bytes = 0;     // the graph's reservation in the budget
    bool draining = false;  // g1_server_remove_graph: no new queries; the graph's queries finish
    bool closing = false;   // its threads exit
    bool removed = false;   // threads joined, views destroyed, bytes returned
};

}  // namespace

struct g1_server {
    g1_server_config cfg{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work;  // new submissions / cancellations / budget freed / stop
    std::condition_variable cv_done;  // a query reached a final state
    bool stop = false;
    std::vector<std::unique_ptr<GraphSlot>> graphs;
    std::unordered_map<uint64_t, std::unique_ptr<Query>> queries;
    uint64_t next_id = 1;
    uint64_t hbm_used = 0, hbm_peak = 0;  // reserved: graphs, sessions, live results
    // Result buffers returned by fetch, kept for reuse by size (hipFree may wait for the whole device).
    std::unordered_map<uint64_t, std::vector<void*>> pool;
    uint64_t pooled = 0;
    g1_server_stats st{};

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
    void reserve(uint64_t b) {
        hbm_used += b;
        hbm_peak = std::max(hbm_peak, hbm_used);
    }
    void unreserve(uint64_t b) { hbm_used -= std::min(hbm_used, b); }
};

namespace {

// A device buffer of `bytes` for a result (worker thread): reused from the pool when one is free.
// Pooled buffers beyond the budget are freed first, so reserved + pooled memory stays within it.
void* result_buffer(g1_server* sv, uint64_t bytes) {
    std::vector<void*> drop;
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        auto it = sv->pool.find(bytes);
        if (it != sv->pool.end() && !it->second.empty()) {
            p = it->second.back();
            it->second.pop_back();
            sv->pooled -= bytes;
            return p;
        }
        for (auto& kv : sv->pool) {
            while (!kv.second.empty() && sv->hbm_used + sv->pooled > sv->cfg.hbm_budget_bytes) {
                drop.push_back(kv.second.back());
                kv.second.pop_back();
                sv->pooled -= kv.first;
            }
        }
    }
    for (void* d : drop) (void)hipFree(d);
    if (hipMalloc(&p, std::max<uint64_t>(bytes, 4)) != hipSuccess) return nullptr;
    return p;
}

// Lock held: returns a result buffer to the pool.
void recycle(g1_server* sv, void* p, uint64_t bytes) {
    if (!p) return;
    sv->pool[bytes].push_back(p);
    sv->pooled += bytes;
}

bool final_state(int32_t s) { return s == G1_Q_DONE || s == G1_Q_FAILED || s == G1_Q_CANCELLED || s == G1_Q_FETCHED; }

int session_kind_of(int32_t qkind) { return qkind == G1_QUERY_PPR ? g1::kSessPpr : g1::kSessMinPlusWeighted; }

// ---- worker side ---------------------------------------------------------------------------------

// Marks q final under the server lock and releases what it holds when it did not complete. Called on
// the thread whose engine call returned `status`, so a failure's message is that thread's last error.
void finish(g1_server* sv, Query* q, g1_status status, uint32_t iterations, bool cancel = false) {
    const std::string message = status != G1_OK ? std::string(g1_last_error_message()) : std::string();
    std::lock_guard<std::mutex> lock(sv->mu);
    if (cancel) q->cancel = true;
    q->iterations = iterations;
    q->t_done = sv->now_ms();
    if (q->cancel) {
        q->state = G1_Q_CANCELLED;
        ++sv->st.cancelled;
    } else if (status != G1_OK) {
        q->state = G1_Q_FAILED;
        q->status = status;
        q->error = message;
        ++sv->st.failed;
    } else {
        q->state = G1_Q_DONE;
        ++sv->st.completed;
        if (q->path == G1_PATH_BATCHED) ++sv->st.batched; else ++sv->st.single;
    }
    if (sv->st.running) --sv->st.running;
    if (q->state != G1_Q_DONE) {
        recycle(sv, q->result, q->bytes);
        q->result = nullptr;
        if (q->reserved) sv->unreserve(q->bytes);
        q->reserved = false;
        sv->cv_work.notify_all();
    }
    sv->cv_done.notify_all();
}

// Road-graph queries one single-source thread runs together in one launch.
constexpr size_t kRoadBatch = 32;

// A single-source thread: takes routed queries of its graph and runs them on its own view and stream.
void single_loop(g1_server* sv, GraphSlot* slot, g1_graph* view) {
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) stream = nullptr;
    const g1_stream_t s = reinterpret_cast<g1_stream_t>(stream);
    for (;;) {
        Query* q = nullptr;
        std::vector<Query*> road;  // BFS / SSSP on a low-degree graph: several run in one launch
        {
            std::unique_lock<std::mutex> lock(sv->mu);
            sv->cv_work.wait(lock, [&] { return sv->stop || slot->closing || !slot->singles.empty(); });
            if (sv->stop || slot->closing) break;
            q = slot->singles.front();
            slot->singles.pop_front();
            if (q->cancel) {
                lock.unlock();
                finish(sv, q, G1_OK, 0);
                continue;
            }
            if (g1::is_road_graph(view) && q->q.kind != G1_QUERY_PPR) {
                road.push_back(q);
                for (auto it = slot->singles.begin(); it != slot->singles.end() && road.size() < kRoadBatch;) {
                    if ((*it)->q.kind != G1_QUERY_PPR && !(*it)->cancel) {
                        road.push_back(*it);
                        it = slot->singles.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
        }
        if (!road.empty()) {
            const uint32_t c = static_cast<uint32_t>(road.size());
            std::vector<uint32_t> kinds(c), srcs(c), iters(c, 0);
            std::vector<uint32_t*> outs(c);
            std::vector<uint8_t> ok(c, 0);
            for (uint32_t i = 0; i < c; ++i) {
                kinds[i] = road[i]->q.kind == G1_QUERY_SSSP ? 1u : 0u;
                srcs[i] = road[i]->q.source;
                outs[i] = static_cast<uint32_t*>(road[i]->result);
            }
            g1_status st = g1::road_queries(view, kinds.data(), srcs.data(), c, outs.data(), ok.data(), iters.data(),
                                            stream);
            for (uint32_t i = 0; i < c; ++i) {
                g1_status qs = st;
                uint32_t its = iters[i];
                if (qs == G1_OK && !ok[i]) {  // rerun on the push engine
                    g1_run_stats rs;
                    memset(&rs, 0, sizeof(rs));
                    qs = kinds[i] ? g1_sssp(view, srcs[i], outs[i], G1_ENGINE_ADVANCE, &rs, s)
                                  : g1_bfs(view, srcs[i], outs[i], G1_ENGINE_ADVANCE, &rs, s);
                    if (qs == G1_OK && hipStreamSynchronize(stream) != hipSuccess) qs = G1_ERR_HIP;
                    its = rs.iterations;
                }
                finish(sv, road[i], qs, its);
            }
            continue;
        }
        g1_run_stats rs;
        memset(&rs, 0, sizeof(rs));
        g1_status st;
        switch (q->q.kind) {
            case G1_QUERY_BFS:
                st = g1_bfs(view, q->q.source, static_cast<uint32_t*>(q->result), G1_ENGINE_DIROPT, &rs, s);
                break;
            case G1_QUERY_SSSP:
                st = g1_sssp(view, q->q.source, static_cast<uint32_t*>(q->result), G1_ENGINE_ADVANCE, &rs, s);
                break;
            default:
                st = g1_ppr(view, q->q.source, sv->cfg.ppr_restart, q->q.max_iter, q->q.tol,
                            static_cast<float*>(q->result), &rs, s);
                break;
        }
        if (st == G1_OK && hipStreamSynchronize(stream) != hipSuccess) st = G1_ERR_HIP;
        if (st == G1_OK && q->q.kind == G1_QUERY_SSSP) {
            std::lock_guard<std::mutex> lock(sv->mu);
            ++slot->sssp_runs;
            slot->sssp_iters += rs.iterations;
        }
        finish(sv, q, st, rs.iterations);
    }
    if (stream) (void)hipStreamDestroy(stream);
}

// The session thread of a graph: routes admitted queries and drives the column sessions.
struct Worker {
    g1_server* sv;
    GraphSlot* slot;
    hipStream_t stream = nullptr;
    g1::Session* sess[3] = {nullptr, nullptr, nullptr};  // [G1_QUERY_SSSP], [G1_QUERY_PPR] used
    Query* col_query[3][g1::kBatchCols] = {};
    std::deque<Query*> backlog[3];  // admitted, path not chosen yet, per kind
    std::deque<Query*> pending[3];  // routed to a session, waiting for a free column
    uint64_t sess_bytes[3] = {0, 0, 0};
    Clock::time_point idle_since[3];

    void finish(Query* q, g1_status status, uint32_t iterations, bool cancel = false) {
        ::finish(sv, q, status, iterations, cancel);
    }

    // One session iteration of kind k: fill free columns, step, hand out the finished columns.
    void session_round(int k) {
        g1::Session* ss = sess[k];
        const uint32_t width = g1::session_width(ss);
        uint64_t active = g1::session_active(ss);
        uint64_t done_now = 0;
        // cancelled columns leave at once
        for (uint32_t j = 0; j < width; ++j) {
            Query* q = col_query[k][j];
            if (!q) continue;
            bool c;
            {
                std::lock_guard<std::mutex> lock(sv->mu);
                c = q->cancel;
            }
            if (c) {
                g1::session_release(ss, j);
                col_query[k][j] = nullptr;
                finish(q, G1_OK, 0);
            }
        }
        active = g1::session_active(ss);
        for (uint32_t j = 0; j < width && !pending[k].empty(); ++j) {
            if ((active >> j) & 1ull) continue;
            Query* q = pending[k].front();
            pending[k].pop_front();
            bool now = false;
            const g1_status st = g1::session_insert(ss, j, q->q.source, q->q.max_iter, q->q.tol, &now);
            if (st != G1_OK) {
                finish(q, st, 0);
                continue;
            }
            col_query[k][j] = q;
            active |= 1ull << j;
            if (now) done_now |= 1ull << j;
        }
        uint64_t done = 0;
        uint32_t iters[g1::kBatchCols] = {};
        g1_status st = g1::session_step(ss, &done, iters);
        {
            std::lock_guard<std::mutex> lock(sv->mu);
            ++sv->st.session_steps;
        }
        done |= done_now;
        if (st != G1_OK) done = active;  // a failed step fails every column in flight
        for (uint32_t j = 0; j < width; ++j) {
            if (!((done >> j) & 1ull) || !col_query[k][j]) continue;
            Query* q = col_query[k][j];
            g1_status qs = st;
            if (qs == G1_OK) qs = g1::session_extract(ss, j, q->result);
            if (qs == G1_OK && hipStreamSynchronize(stream) != hipSuccess) qs = G1_ERR_HIP;
            g1::session_release(ss, j);
            col_query[k][j] = nullptr;
            finish(q, qs, iters[j]);
        }
    }

    bool busy() const {
        for (int k = 0; k < 3; ++k) {
            if (!pending[k].empty()) return true;
            if (sess[k] && g1::session_active(sess[k])) return true;
        }
        return false;
    }

    // A session idle for half a second is closed and its memory returned to the budget.
    void close_idle_sessions() {
        const auto now = Clock::now();
        for (int k = 1; k < 3; ++k) {
            if (!sess[k]) continue;
            if (g1::session_active(sess[k]) || !pending[k].empty() || !backlog[k].empty()) {
                idle_since[k] = now;
                continue;
            }
            if (now - idle_since[k] < std::chrono::milliseconds(500)) continue;
            g1::session_destroy(sess[k]);
            sess[k] = nullptr;
            std::lock_guard<std::mutex> lock(sv->mu);
            sv->unreserve(sess_bytes[k]);
            sess_bytes[k] = 0;
            sv->cv_work.notify_all();
        }
    }

    // Path choice: BFS always single-source. SSSP / PPR: an open session takes every query of its kind;
    // with no session, when the queries of the kind waiting (fresh plus routed but not started) reach the
    // threshold, a session opens if its memory fits the budget (SSSP only on low-diameter graphs: at most
    // 64 iterations per single-source run seen on the graph) and takes them all; otherwise they go to the
    // single-source threads.
    void route(int k) {
        if (k == G1_QUERY_BFS || (!sess[k] && backlog[k].empty())) {
            to_singles(k);
            return;
        }
        if (!sess[k]) {
            const uint32_t threshold = k == G1_QUERY_PPR ? sv->cfg.batch_min_ppr : sv->cfg.batch_min_sssp;
            bool open = false;
            uint32_t width = 8;
            uint64_t bytes = 0;
            {
                std::lock_guard<std::mutex> lock(sv->mu);
                size_t waiting = backlog[k].size();
                for (Query* q : slot->singles) waiting += (q->q.kind == k);
                // a column costs a full row of every gathered vector: the width grows with the backlog (room
                // for as many arrivals again), up to batch_k
                while (width < sv->cfg.batch_k && width < 2 * waiting) width *= 2;
                bytes = g1::session_bytes(slot->g, session_kind_of(k), width);
                const bool shallow = k == G1_QUERY_PPR ||
                                     (slot->sssp_runs > 0 && slot->sssp_iters <= 64 * slot->sssp_runs);
                if (waiting >= threshold && shallow && sv->hbm_used + bytes <= sv->cfg.hbm_budget_bytes) {
                    sv->reserve(bytes);
                    open = true;
                    // take back the routed-but-not-started queries of this kind
                    for (auto it = slot->singles.begin(); it != slot->singles.end();) {
                        if ((*it)->q.kind == k) {
                            backlog[k].push_back(*it);
                            it = slot->singles.erase(it);
                        } else {
                            ++it;
                        }
                    }
                }
            }
            if (open) {
                g1::Session* ss = nullptr;
                const g1_status st = g1::session_create(slot->g, session_kind_of(k), width, sv->cfg.ppr_restart,
                                                        stream, &ss);
                if (st == G1_OK) {
                    sess[k] = ss;
                    sess_bytes[k] = bytes;
                    idle_since[k] = Clock::now();
                } else {
                    std::lock_guard<std::mutex> lock(sv->mu);
                    sv->unreserve(bytes);
                }
            }
        }
        if (!sess[k]) {
            to_singles(k);
            return;
        }
        std::lock_guard<std::mutex> lock(sv->mu);
        while (!backlog[k].empty()) {
            Query* q = backlog[k].front();
            backlog[k].pop_front();
            q->path = G1_PATH_BATCHED;
            pending[k].push_back(q);
        }
    }

    void to_singles(int k) {
        if (backlog[k].empty()) return;
        std::lock_guard<std::mutex> lock(sv->mu);
        while (!backlog[k].empty()) {
            slot->singles.push_back(backlog[k].front());
            backlog[k].pop_front();
        }
        sv->cv_work.notify_all();
    }

    void loop() {
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) stream = nullptr;
        for (;;) {
            std::vector<Query*> fresh;
            {
                std::unique_lock<std::mutex> lock(sv->mu);
                if (!busy())
                    sv->cv_work.wait_for(lock, std::chrono::milliseconds(20),
                                         [&] { return sv->stop || slot->closing || !slot->admitted.empty(); });
                if (sv->stop || slot->closing) break;
                while (!slot->admitted.empty()) {
                    fresh.push_back(slot->admitted.front());
                    slot->admitted.pop_front();
                }
            }
            // result buffers (their bytes were reserved at admission)
            for (Query* q : fresh) {
                void* p = result_buffer(sv, q->bytes);
                if (!p) {
                    finish(q, G1_ERR_OUT_OF_MEMORY, 0);
                    continue;
                }
                q->result = p;
                backlog[q->q.kind].push_back(q);
            }
            // route, then one iteration of each open session
            for (int k = 0; k < 3; ++k) route(k);
            for (int k = 1; k < 3; ++k)
                if (sess[k]) session_round(k);
            close_idle_sessions();
        }
        // shutdown (or the graph's removal): everything still held is cancelled, sessions return their memory
        for (int k = 1; k < 3; ++k) {
            for (uint32_t j = 0; j < g1::kBatchCols; ++j)
                if (col_query[k][j]) finish(col_query[k][j], G1_OK, 0, true);
            for (Query* q : pending[k]) finish(q, G1_OK, 0, true);
            if (sess[k]) {
                g1::session_destroy(sess[k]);
                std::lock_guard<std::mutex> lock(sv->mu);
                sv->unreserve(sess_bytes[k]);
            }
        }
        for (int k = 0; k < 3; ++k)
            for (Query* q : backlog[k]) finish(q, G1_OK, 0, true);
        if (stream) (void)hipStreamDestroy(stream);
        std::vector<Query*> left;
        {
            std::lock_guard<std::mutex> lock(sv->mu);
            left.assign(slot->singles.begin(), slot->singles.end());
            slot->singles.clear();
        }
        for (Query* q : left) finish(q, G1_OK, 0, true);
    }
};

// Admission (server lock held): moves queries from the head of the graph's queue to its worker while
// their memory fits. A query that does not fit stops its graph's queue (FIFO) until memory is freed.
void admit(g1_server* sv, GraphSlot* gs) {
    while (!gs->queue.empty()) {
        Query* q = gs->queue.front();
        if (q->cancel) {
            gs->queue.pop_front();
            if (sv->st.queued) --sv->st.queued;
            q->state = G1_Q_CANCELLED;
            q->t_done = sv->now_ms();
            ++sv->st.cancelled;
            sv->cv_done.notify_all();
            continue;
        }
        if (sv->hbm_used + q->bytes > sv->cfg.hbm_budget_bytes) {
            if (!q->reserved && q->t_admit == 0) {
                q->t_admit = -1;  // counted once
                ++sv->st.admission_waits;
            }
            return;
        }
        gs->queue.pop_front();
        sv->reserve(q->bytes);
        q->reserved = true;
        q->path = G1_PATH_SINGLE;
        q->state = G1_Q_RUNNING;
        q->t_admit = sv->now_ms();
        if (sv->st.queued) --sv->st.queued;
        ++sv->st.running;
        gs->admitted.push_back(q);
    }
}

void admit_all(g1_server* sv) {
    for (auto& gs : sv->graphs) admit(sv, gs.get());
    sv->cv_work.notify_all();
}

Query* find(g1_server* sv, uint64_t id) {
    auto it = sv->queries.find(id);
    return it == sv->queries.end() ? nullptr : it->second.get();
}

void fill_status(const Query* q, g1_query_status* st) {
    memset(st, 0, sizeof(*st));
    st->state = q->state;
    st->status = q->status;
    st->path = q->path;
    st->iterations = q->iterations;
    st->t_submit_ms = q->t_submit;
    st->t_admit_ms = q->t_admit > 0 ? q->t_admit : 0.0;
    st->t_done_ms = q->t_done;
    st->result_bytes = q->bytes;
}

}  // namespace

extern "C" {

g1_status g1_server_create(const g1_server_config* config, g1_server** out) {
    G1_CHECK(out != nullptr, G1_ERR_INVALID_ARG, "g1_server_create: null output");
    g1_server_config c;
    memset(&c, 0, sizeof(c));
    if (config) c = *config;
    if (c.batch_k == 0) c.batch_k = 64;
    G1_CHECK(c.batch_k == 8 || c.batch_k == 16 || c.batch_k == 32 || c.batch_k == 64, G1_ERR_INVALID_ARG,
             "g1_server_create: batch_k must be 8, 16, 32 or 64 (got %u)", c.batch_k);
    if (c.batch_min_sssp == 0) c.batch_min_sssp = 32;
    if (c.batch_min_ppr == 0) c.batch_min_ppr = 2;
    if (c.ppr_restart == 0.0f) c.ppr_restart = 0.15f;
    if (c.single_workers == 0) c.single_workers = 4;
    G1_CHECK(c.single_workers <= 32, G1_ERR_INVALID_ARG, "g1_server_create: at most 32 single-source workers");
    G1_CHECK(c.ppr_restart > 0.0f && c.ppr_restart <= 1.0f, G1_ERR_INVALID_ARG,
             "g1_server_create: ppr_restart must be in (0, 1]");
    if (c.hbm_budget_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        G1_HIP(hipMemGetInfo(&free_b, &total_b));
        c.hbm_budget_bytes = static_cast<uint64_t>(static_cast<double>(free_b) * 0.9);
    }
    g1_server* sv = new g1_server();
    sv->cfg = c;
    sv->t0 = Clock::now();
    sv->st.hbm_budget = c.hbm_budget_bytes;
    *out = sv;
    return G1_OK;
}

g1_status g1_server_destroy(g1_server* sv) {
    if (!sv) return G1_OK;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        sv->stop = true;
        for (auto& gs : sv->graphs) {
            for (Query* q : gs->admitted) q->cancel = true;
        }
        sv->cv_work.notify_all();
        sv->cv_done.notify_all();
    }
    for (auto& gs : sv->graphs) {
        if (gs->worker.joinable()) gs->worker.join();
        for (auto& h : gs->helpers)
            if (h.joinable()) h.join();
        for (g1_graph* v : gs->views) g1_graph_destroy(v);
    }
    for (auto& kv : sv->queries)
        if (kv.second->result) (void)hipFree(kv.second->result);
    for (auto& kv : sv->pool)
        for (void* p : kv.second) (void)hipFree(p);
    delete sv;
    return G1_OK;
}

g1_status g1_server_add_graph(g1_server* sv, g1_graph* graph, uint32_t* graph_id) {
    G1_CHECK(sv != nullptr && graph != nullptr, G1_ERR_INVALID_ARG, "g1_server_add_graph: null argument");
    uint32_t n = 0, m = 0;
    uint64_t bytes = 0;
    G1_TRY(g1_graph_info(graph, &n, &m, &bytes));
    G1_CHECK(graph->has_t, G1_ERR_INVALID_ARG, "g1_server_add_graph: the graph needs its transpose");
    {
        // the budget is checked and the graph's bytes reserved under the lock ...
        std::lock_guard<std::mutex> lock(sv->mu);
        G1_CHECK(!sv->stop, G1_ERR_CANCELLED, "g1_server_add_graph: server is shutting down");
        G1_CHECK(sv->hbm_used + bytes <= sv->cfg.hbm_budget_bytes, G1_ERR_BUDGET,
                 "g1_server_add_graph: graph (%llu bytes) exceeds the HBM budget",
                 static_cast<unsigned long long>(bytes));
        sv->reserve(bytes);
    }
    // ... the views are created without it (they allocate, copy and synchronize the device, which must not
    // stall submit/poll/wait or the workers of the graphs already resident) ...
    auto gs = std::make_unique<GraphSlot>();
    gs->g = graph;
    gs->n = n;
    gs->bytes = bytes;
    for (uint32_t i = 0; i < sv->cfg.single_workers; ++i) {
        g1_graph* v = nullptr;
        const g1_status st = g1_graph_create_view(graph, &v);
        if (st != G1_OK) {
            for (g1_graph* x : gs->views) g1_graph_destroy(x);
            std::lock_guard<std::mutex> lock(sv->mu);
            sv->unreserve(bytes);
            return st;
        }
        gs->views.push_back(v);
    }
    // ... and the slot is published under the lock again.
    std::lock_guard<std::mutex> lock(sv->mu);
    gs->id = static_cast<uint32_t>(sv->graphs.size());
    GraphSlot* raw = gs.get();
    sv->graphs.push_back(std::move(gs));
    raw->worker = std::thread([sv, raw] {
        Worker w{sv, raw};
        w.loop();
    });
    for (g1_graph* v : raw->views) raw->helpers.emplace_back([sv, raw, v] { single_loop(sv, raw, v); });
    if (graph_id) *graph_id = raw->id;
    return G1_OK;
}

g1_status g1_server_remove_graph(g1_server* sv, uint32_t graph_id) {
    G1_CHECK(sv != nullptr, G1_ERR_INVALID_ARG, "g1_server_remove_graph: null server");
    GraphSlot* gs = nullptr;
    {
        std::unique_lock<std::mutex> lock(sv->mu);
        G1_CHECK(graph_id < sv->graphs.size() && !sv->graphs[graph_id]->draining, G1_ERR_NOT_FOUND,
                 "g1_server_remove_graph: unknown or removed graph %u", graph_id);
        gs = sv->graphs[graph_id].get();
        gs->draining = true;  // submit refuses the graph from now on
        // the graph's queued and running queries finish first (their results stay fetchable)
        auto live = [&] {
            for (const auto& kv : sv->queries) {
                const Query* q = kv.second.get();
                if (q->q.graph_id == graph_id && (q->state == G1_Q_QUEUED || q->state == G1_Q_RUNNING)) return true;
            }
            return false;
        };
        sv->cv_done.wait(lock, [&] { return sv->stop || !live(); });
        G1_CHECK(!sv->stop, G1_ERR_CANCELLED, "g1_server_remove_graph: server is shutting down");
        gs->closing = true;
        sv->cv_work.notify_all();
    }
    // the graph's threads exit (without the lock: they take it on their way out), then its views go
    if (gs->worker.joinable()) gs->worker.join();
    for (auto& h : gs->helpers)
        if (h.joinable()) h.join();
    for (g1_graph* v : gs->views) g1_graph_destroy(v);
    gs->views.clear();
    std::lock_guard<std::mutex> lock(sv->mu);
    sv->unreserve(gs->bytes);
    gs->bytes = 0;
    gs->removed = true;
    for (auto& other : sv->graphs)
        if (!other->draining) admit(sv, other.get());  // the freed bytes may admit waiting queries
    sv->cv_work.notify_all();
    return G1_OK;
}

g1_status g1_server_set_budget(g1_server* sv, uint64_t hbm_budget_bytes) {
    G1_CHECK(sv != nullptr && hbm_budget_bytes > 0, G1_ERR_INVALID_ARG, "g1_server_set_budget: null server or zero budget");
    std::lock_guard<std::mutex> lock(sv->mu);
    sv->cfg.hbm_budget_bytes = hbm_budget_bytes;
    sv->st.hbm_budget = hbm_budget_bytes;
    for (auto& gs : sv->graphs)
        if (!gs->draining) admit(sv, gs.get());
    sv->cv_work.notify_all();
    return G1_OK;
}

g1_status g1_server_submit(g1_server* sv, const g1_query* query, uint64_t* query_id) {
    G1_CHECK(sv != nullptr && query != nullptr && query_id != nullptr, G1_ERR_INVALID_ARG,
             "g1_server_submit: null argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    G1_CHECK(!sv->stop, G1_ERR_CANCELLED, "g1_server_submit: server is shutting down");
    G1_CHECK(query->graph_id < sv->graphs.size() && !sv->graphs[query->graph_id]->draining, G1_ERR_NOT_FOUND,
             "g1_server_submit: unknown or removed graph %u", query->graph_id);
    G1_CHECK(query->kind == G1_QUERY_BFS || query->kind == G1_QUERY_SSSP || query->kind == G1_QUERY_PPR,
             G1_ERR_INVALID_ARG, "g1_server_submit: unknown query kind %d", query->kind);
    GraphSlot* gs = sv->graphs[query->graph_id].get();
    G1_CHECK(query->source < gs->n, G1_ERR_INVALID_ARG, "g1_server_submit: source %u out of range (n = %u)",
             query->source, gs->n);
    G1_CHECK(query->kind != G1_QUERY_SSSP || gs->g->w != nullptr, G1_ERR_INVALID_ARG,
             "g1_server_submit: SSSP on a graph without weights");
    const uint64_t bytes = static_cast<uint64_t>(gs->n) * 4;
    G1_CHECK(sv->graphs.size() && bytes <= sv->cfg.hbm_budget_bytes, G1_ERR_BUDGET,
             "g1_server_submit: the result alone exceeds the HBM budget");
    auto q = std::make_unique<Query>();
    q->id = sv->next_id++;
    q->q = *query;
    q->bytes = bytes;
    q->t_submit = sv->now_ms();
    Query* raw = q.get();
    sv->queries.emplace(raw->id, std::move(q));
    gs->queue.push_back(raw);
    ++sv->st.submitted;
    ++sv->st.queued;
    *query_id = raw->id;
    admit(sv, gs);
    sv->cv_work.notify_all();
    return G1_OK;
}

g1_status g1_server_poll(g1_server* sv, uint64_t query_id, g1_query_status* status) {
    G1_CHECK(sv != nullptr && status != nullptr, G1_ERR_INVALID_ARG, "g1_server_poll: null argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    Query* q = find(sv, query_id);
    G1_CHECK(q != nullptr, G1_ERR_NOT_FOUND, "g1_server_poll: unknown query %llu",
             static_cast<unsigned long long>(query_id));
    fill_status(q, status);
    return G1_OK;
}

g1_status g1_server_wait(g1_server* sv, uint64_t query_id, double timeout_ms, g1_query_status* status) {
    G1_CHECK(sv != nullptr, G1_ERR_INVALID_ARG, "g1_server_wait: null server");
    std::unique_lock<std::mutex> lock(sv->mu);
    Query* q = find(sv, query_id);
    G1_CHECK(q != nullptr, G1_ERR_NOT_FOUND, "g1_server_wait: unknown query %llu",
             static_cast<unsigned long long>(query_id));
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                             std::chrono::duration<double, std::milli>(std::max(timeout_ms, 0.0)));
    while (!final_state(q->state)) {
        if (sv->cv_done.wait_until(lock, deadline) == std::cv_status::timeout && !final_state(q->state)) {
            if (status) fill_status(q, status);
            return g1::set_error(G1_ERR_BUSY, "g1_server_wait: query %llu not finished after %.1f ms",
                                 static_cast<unsigned long long>(query_id), timeout_ms);
        }
    }
    if (status) fill_status(q, status);
    return G1_OK;
}

g1_status g1_server_fetch(g1_server* sv, uint64_t query_id, void* dst, uint64_t dst_bytes, int32_t dst_on_device) {
    G1_CHECK(sv != nullptr && dst != nullptr, G1_ERR_INVALID_ARG, "g1_server_fetch: null argument");
    void* src = nullptr;
    uint64_t bytes = 0;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        Query* q = find(sv, query_id);
        G1_CHECK(q != nullptr, G1_ERR_NOT_FOUND, "g1_server_fetch: unknown query %llu",
                 static_cast<unsigned long long>(query_id));
        G1_CHECK(q->state != G1_Q_FAILED, q->status, "g1_server_fetch: query %llu failed: %s",
                 static_cast<unsigned long long>(query_id), q->error.c_str());
        G1_CHECK(q->state == G1_Q_DONE, q->state == G1_Q_FETCHED ? G1_ERR_INVALID_ARG : G1_ERR_BUSY,
                 "g1_server_fetch: query %llu is not done (state %d)", static_cast<unsigned long long>(query_id),
                 q->state);
        G1_CHECK(dst_bytes >= q->bytes, G1_ERR_INVALID_ARG, "g1_server_fetch: destination holds %llu of %llu bytes",
                 static_cast<unsigned long long>(dst_bytes), static_cast<unsigned long long>(q->bytes));
        src = q->result;
        bytes = q->bytes;
        q->result = nullptr;
        q->state = G1_Q_FETCHED;
    }
    const hipError_t e = hipMemcpy(dst, src, bytes, dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost);
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        recycle(sv, src, bytes);
        Query* q = find(sv, query_id);
        if (q->reserved) sv->unreserve(q->bytes);
        q->reserved = false;
        admit_all(sv);
    }
    G1_CHECK(e == hipSuccess, G1_ERR_HIP, "g1_server_fetch: copy failed: %s", hipGetErrorString(e));
    return G1_OK;
}

g1_status g1_server_cancel(g1_server* sv, uint64_t query_id) {
    G1_CHECK(sv != nullptr, G1_ERR_INVALID_ARG, "g1_server_cancel: null server");
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        Query* q = find(sv, query_id);
        G1_CHECK(q != nullptr, G1_ERR_NOT_FOUND, "g1_server_cancel: unknown query %llu",
                 static_cast<unsigned long long>(query_id));
        if (q->state == G1_Q_QUEUED) {
            q->cancel = true;
            GraphSlot* gs = sv->graphs[q->q.graph_id].get();
            auto it = std::find(gs->queue.begin(), gs->queue.end(), q);
            if (it != gs->queue.end()) {
                gs->queue.erase(it);
                if (sv->st.queued) --sv->st.queued;
            }
            q->state = G1_Q_CANCELLED;
            q->t_done = sv->now_ms();
            ++sv->st.cancelled;
            sv->cv_done.notify_all();
            admit_all(sv);
        } else if (q->state == G1_Q_RUNNING) {
            q->cancel = true;  // the worker drops it at its next step
        } else if (q->state == G1_Q_DONE) {
            recycle(sv, q->result, q->bytes);  // a finished, unfetched result is discarded
            q->result = nullptr;
            q->state = G1_Q_CANCELLED;
            if (q->reserved) sv->unreserve(q->bytes);
            q->reserved = false;
            ++sv->st.cancelled;
            admit_all(sv);
        }
    }
    return G1_OK;
}

g1_status g1_server_get_stats(g1_server* sv, g1_server_stats* stats) {
    G1_CHECK(sv != nullptr && stats != nullptr, G1_ERR_INVALID_ARG, "g1_server_get_stats: null argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    *stats = sv->st;
    stats->hbm_budget = sv->cfg.hbm_budget_bytes;
    stats->hbm_used = sv->hbm_used;
    stats->hbm_peak = sv->hbm_peak;
    return G1_OK;
}

}  // extern "C"
