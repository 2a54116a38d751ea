/*
 * g1_oracle.c -- CPU restatement of GraphOne's graph-computation hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker of the GPU library: only tests/, the smoke test
 * of __graft_entry__.py and the cpu_baseline leg of bench.py may load it. The product never calls it.
 *
 * Every function names the reference lines (paths relative to the reference root) whose semantics it
 * restates. How each one is pinned (checked against closed forms, textbook algorithms, library
 * routines, brute force, or the reference's own compiled naive SpMV) is listed in oracle/README.md and
 * tested in tests/test_oracle.py.
 *
 * Arrays follow the reference's app-level types: vid_t/eid_t = uint32 (include/GraphOne/type.hpp:8-9);
 * INF = 0xFFFFFFFF for distances, the invalid vertex of include/GraphOne/utils/limits.hpp:54-73.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 0xFFFFFFFFu

/* Host threads of the OpenMP loops (orc_pagerank_f64 and orc_sssp_advance_omp): every thread of the
 * machine by default. The results do not depend on it (see each function). */
void orc_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs()); }
int orc_get_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------------------------------------
 * Semiring SpMV, the naive row fold of include/GraphOne/archi/blas/spmv/cpu/naive.hpp:5-20:
 *   y[r] = reduce(...reduce(initialize(), combine(A[r,j0], x[c_j0]))..., combine(A[r,jk], x[c_jk]))
 * with columns taken in storage order. vals == NULL means every nonzero is 1.
 * ------------------------------------------------------------------------------------------- */

/* min-plus over uint32, the SSSP functor of apps/sssp_blas.cu:55-73 (initialize = default_result = INF,
 * combine = (x == INF) ? INF : w + x, reduce = min). */
void orc_spmv_minplus_u32(uint32_t n_rows, const uint32_t* ap, const uint32_t* aj, const uint32_t* ax,
                          const uint32_t* x, uint32_t* y) {
    for (uint32_t r = 0; r < n_rows; ++r) {
        uint32_t sum = ORC_INF;
        for (uint32_t k = ap[r]; k < ap[r + 1]; ++k) {
            const uint32_t xv = x[aj[k]];
            const uint32_t w = ax ? ax[k] : 1u;
            const uint32_t p = (xv == ORC_INF) ? ORC_INF : w + xv;
            sum = p < sum ? p : sum;
        }
        y[r] = sum;
    }
}

/* plus-times, the PageRank functor of apps/pagerank_blas.cu:82-111 (combine = info * weight,
 * reduce = +, initialize = 0), folded in float exactly as the naive loop does. */
void orc_spmv_plustimes_f32(uint32_t n_rows, const uint32_t* ap, const uint32_t* aj, const float* ax,
                            const float* x, float* y) {
    for (uint32_t r = 0; r < n_rows; ++r) {
        float sum = 0.0f;
        for (uint32_t k = ap[r]; k < ap[r + 1]; ++k) {
            const float w = ax ? ax[k] : 1.0f;
            sum = sum + x[aj[k]] * w;
        }
        y[r] = sum;
    }
}

/* The same fold in double precision (the FP64 oracle of the plus-times kernels). */
void orc_spmv_plustimes_f64(uint32_t n_rows, const uint32_t* ap, const uint32_t* aj, const double* ax,
                            const double* x, double* y) {
    for (uint32_t r = 0; r < n_rows; ++r) {
        double sum = 0.0;
        for (uint32_t k = ap[r]; k < ap[r + 1]; ++k) {
            const double w = ax ? ax[k] : 1.0;
            sum = sum + x[aj[k]] * w;
        }
        y[r] = sum;
    }
}

/* ---------------------------------------------------------------------------------------------
 * Push engine SSSP, the loop of base/run.h:21-34 over AdvanceGC::Forward (engine/advance_gc.h:180-244)
 * with SSSPFunctor (apps/sssp_adv.cu:17-72), restated sequentially:
 *   Init: d = INF, d[src] = 0, frontier = {src}, visited = 0, iter = 0
 *   while frontier non-empty: ++iter (BeforeEngine);
 *     advance: for each frontier vertex u (in order) and out-edge (u,v,w) in storage order:
 *       nd = d[u] + w (d[u] read live); old = d[v]; d[v] = min(old, nd); emit v iff nd < old
 *     filter: keep the first emission of v in this iteration (visited[v] == iter drops it)
 *   The result d is unique (shortest-path distances), whatever order the GPU relaxes in.
 * stats (may be NULL): [0] iterations, [1] sum |F_i| (frontier vertices), [2] sum E(F_i).
 * ------------------------------------------------------------------------------------------- */
int orc_sssp_advance(uint32_t n, const uint32_t* rp, const uint32_t* ci, const uint32_t* w, uint32_t src,
                     uint32_t* d, uint64_t* stats) {
    uint32_t* fin = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* fout = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* visited = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    if (!fin || !fout || !visited) { free(fin); free(fout); free(visited); return -1; }
    for (uint32_t v = 0; v < n; ++v) d[v] = ORC_INF;
    d[src] = 0;
    uint32_t nin = 1, iter = 0;
    uint64_t fsum = 0, esum = 0;
    fin[0] = src;
    while (nin > 0) {
        ++iter;
        uint32_t nout = 0;
        fsum += nin;
        for (uint32_t i = 0; i < nin; ++i) {
            const uint32_t u = fin[i];
            esum += rp[u + 1] - rp[u];
            for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) {
                const uint32_t v = ci[k];
                const uint32_t nd = d[u] + (w ? w[k] : 1u);
                const uint32_t old = d[v];
                if (nd < old) {
                    d[v] = nd;
                    if (visited[v] != iter) {
                        visited[v] = iter;
                        fout[nout++] = v;
                    }
                }
            }
        }
        uint32_t* t = fin; fin = fout; fout = t;
        nin = nout;
    }
    if (stats) { stats[0] = iter; stats[1] = fsum; stats[2] = esum; }
    free(fin); free(fout); free(visited);
    return 0;
}

/* The same advance/filter loop on every host thread (this build's own variant, timed as the all-core CPU
 * baseline; the reference has no CPU engine): the frontier's vertices are split among the threads, a
 * relaxation is an atomic min on d[v] (compare-and-swap loop) that reports an improvement when it lowered
 * the value, and the visited filter claims v for this iteration with an atomic exchange of its stamp.
 * Frontier order varies with the threads; the final distances are the unique shortest-path distances.
 * stats as orc_sssp_advance. */
int orc_sssp_advance_omp(uint32_t n, const uint32_t* rp, const uint32_t* ci, const uint32_t* w, uint32_t src,
                         uint32_t* d, uint64_t* stats) {
    uint32_t* fin = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* fout = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* visited = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    if (!fin || !fout || !visited) { free(fin); free(fout); free(visited); return -1; }
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)n; ++v) d[v] = ORC_INF;
    d[src] = 0;
    uint32_t nin = 1, iter = 0;
    uint64_t fsum = 0, esum = 0;
    fin[0] = src;
    while (nin > 0) {
        ++iter;
        uint32_t nout = 0;
        fsum += nin;
        uint64_t e_it = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : e_it)
        for (int64_t i = 0; i < (int64_t)nin; ++i) {
            const uint32_t u = fin[i];
            const uint32_t du = __atomic_load_n(&d[u], __ATOMIC_RELAXED);
            e_it += rp[u + 1] - rp[u];
            for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) {
                const uint32_t v = ci[k];
                const uint32_t nd = du + (w ? w[k] : 1u);
                uint32_t old = __atomic_load_n(&d[v], __ATOMIC_RELAXED);
                int improved = 0;
                while (nd < old) {
                    if (__atomic_compare_exchange_n(&d[v], &old, nd, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                        improved = 1;
                        break;
                    }
                }
                if (improved && __atomic_exchange_n(&visited[v], iter, __ATOMIC_RELAXED) != iter)
                    fout[__atomic_fetch_add(&nout, 1u, __ATOMIC_RELAXED)] = v;
            }
        }
        esum += e_it;
        uint32_t* t = fin; fin = fout; fout = t;
        nin = nout;
    }
    if (stats) { stats[0] = iter; stats[1] = fsum; stats[2] = esum; }
    free(fin); free(fout); free(visited);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Pull (BLAS) engine SSSP/BFS over the transposed CSR, BlasGcEngine<SPARSE> (engine/blas_gc.h:171-247)
 * with the functor of apps/sssp_blas.cu:48-84 (BFS: unit weights, tw == NULL):
 *   Construct: x = INF; x[v] = d[v] for v in F            (blas_gc.h:199-217)
 *   Gather:    y = A^T (min,+) x                          (naive fold, as above)
 *   Apply:     flag[v] = y[v] != INF && y[v] < d[v] (then d[v] = y[v]); F' = {v : flag} ascending
 *              (blas_gc.h:224-245)
 * until F is empty. stats: [0] iterations, [1] sum |F_i|.
 * ------------------------------------------------------------------------------------------- */
int orc_minplus_blas(uint32_t n, const uint32_t* tp, const uint32_t* tj, const uint32_t* tw, uint32_t src,
                     uint32_t* d, uint64_t* stats) {
    uint32_t* x = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* y = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* f = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!x || !y || !f) { free(x); free(y); free(f); return -1; }
    for (uint32_t v = 0; v < n; ++v) d[v] = ORC_INF;
    d[src] = 0;
    uint32_t nf = 1, iter = 0;
    uint64_t fsum = 0;
    f[0] = src;
    while (nf > 0) {
        ++iter;
        fsum += nf;
        for (uint32_t v = 0; v < n; ++v) x[v] = ORC_INF;
        for (uint32_t i = 0; i < nf; ++i) x[f[i]] = d[f[i]];
        orc_spmv_minplus_u32(n, tp, tj, tw, x, y);
        nf = 0;
        for (uint32_t v = 0; v < n; ++v) {
            if (y[v] != ORC_INF && y[v] < d[v]) {
                d[v] = y[v];
                f[nf++] = v;
            }
        }
    }
    if (stats) { stats[0] = iter; stats[1] = fsum; }
    free(x); free(y); free(f);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Textbook references used to pin the restatements.
 * ------------------------------------------------------------------------------------------- */

/* Queue BFS: depth in hops, INF if unreachable. */
int orc_bfs_queue(uint32_t n, const uint32_t* rp, const uint32_t* ci, uint32_t src, uint32_t* depth) {
    uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!q) return -1;
    for (uint32_t v = 0; v < n; ++v) depth[v] = ORC_INF;
    depth[src] = 0;
    uint32_t head = 0, tail = 0;
    q[tail++] = src;
    while (head < tail) {
        const uint32_t u = q[head++];
        for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) {
            const uint32_t v = ci[k];
            if (depth[v] == ORC_INF) {
                depth[v] = depth[u] + 1;
                q[tail++] = v;
            }
        }
    }
    free(q);
    return 0;
}

/* Dijkstra with a binary heap (lazy deletion). */
typedef struct { uint64_t key; uint32_t v; } orc_heap_item;

static void heap_push(orc_heap_item* h, size_t* sz, uint64_t key, uint32_t v) {
    size_t i = (*sz)++;
    h[i].key = key; h[i].v = v;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (h[p].key <= h[i].key) break;
        orc_heap_item t = h[p]; h[p] = h[i]; h[i] = t;
        i = p;
    }
}

static orc_heap_item heap_pop(orc_heap_item* h, size_t* sz) {
    orc_heap_item top = h[0];
    h[0] = h[--(*sz)];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < *sz && h[l].key < h[m].key) m = l;
        if (r < *sz && h[r].key < h[m].key) m = r;
        if (m == i) break;
        orc_heap_item t = h[m]; h[m] = h[i]; h[i] = t;
        i = m;
    }
    return top;
}

int orc_dijkstra(uint32_t n, const uint32_t* rp, const uint32_t* ci, const uint32_t* w, uint32_t src, uint32_t* d) {
    const size_t m = rp[n];
    orc_heap_item* h = (orc_heap_item*)malloc(sizeof(orc_heap_item) * (m + 2));
    if (!h) return -1;
    uint64_t* dd = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!dd) { free(h); return -1; }
    for (uint32_t v = 0; v < n; ++v) dd[v] = UINT64_MAX;
    size_t sz = 0;
    dd[src] = 0;
    heap_push(h, &sz, 0, src);
    while (sz > 0) {
        orc_heap_item it = heap_pop(h, &sz);
        if (it.key != dd[it.v]) continue;
        const uint32_t u = it.v;
        for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) {
            const uint64_t nd = dd[u] + (w ? w[k] : 1u);
            if (nd < dd[ci[k]]) {
                dd[ci[k]] = nd;
                heap_push(h, &sz, nd, ci[k]);
            }
        }
    }
    for (uint32_t v = 0; v < n; ++v) d[v] = dd[v] == UINT64_MAX ? ORC_INF : (uint32_t)dd[v];
    free(h); free(dd);
    return 0;
}

/* Floyd-Warshall on a dense n x n matrix (n small): dist[i*n+j]. Brute-force pin for Dijkstra. */
void orc_floyd_warshall(uint32_t n, const uint32_t* rp, const uint32_t* ci, const uint32_t* w, uint64_t* dist) {
    for (uint64_t i = 0; i < (uint64_t)n * n; ++i) dist[i] = UINT64_MAX;
    for (uint32_t i = 0; i < n; ++i) dist[(uint64_t)i * n + i] = 0;
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) {
            const uint64_t c = w ? w[k] : 1u;
            uint64_t* e = &dist[(uint64_t)u * n + ci[k]];
            if (c < *e) *e = c;
        }
    for (uint32_t k = 0; k < n; ++k)
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t a = dist[(uint64_t)i * n + k];
            if (a == UINT64_MAX) continue;
            for (uint32_t j = 0; j < n; ++j) {
                const uint64_t b = dist[(uint64_t)k * n + j];
                if (b == UINT64_MAX) continue;
                if (a + b < dist[(uint64_t)i * n + j]) dist[(uint64_t)i * n + j] = a + b;
            }
        }
}

/* ---------------------------------------------------------------------------------------------
 * PageRank in FP64 with the formula of apps/pagerank_blas.cu:16-125:
 *   iw[u] = alpha / sum_out w(u)  (0 when u has no out-edge)           (:41-53)
 *   r0 = 1/N; r'[v] = (1 - alpha)/N + sum_{u->v} w(u,v) * r[u] * iw[u]   (:82-111)
 *   no dangling redistribution.
 * Stopping: mode 0 -- after max_iter iterations or once the L1 norm of r' - r is < tol;
 *           mode 1 -- the reference's test (:63-78): before each iteration stop when
 *           max|ranks[idx^1] - ranks[idx]| <= tol (ranks[1] starts at 0), the iteration updates
 *           ranks[idx] in place then idx ^= 1; the result is ranks[idx].
 * Gathers run over the in-edge CSR (tp, tj, tw); out-weights come from (rp, w). Weights may be NULL.
 * Personalized variant (seed != INF): r0 = e_seed, r'[v] = restart*[v == seed] + sum ..., with
 * iw[u] = (1 - restart) / sum_out w(u).
 * Returns the number of iterations; *last_delta receives the last norm.
 * ------------------------------------------------------------------------------------------- */
int orc_pagerank_f64(uint32_t n, const uint32_t* rp, const uint32_t* w, const uint32_t* tp, const uint32_t* tj,
                     const uint32_t* tw, double alpha, uint32_t seed, double restart, uint32_t max_iter,
                     double tol, int mode, double* out, double* last_delta) {
    double* iw = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* r0 = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* r1 = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* x = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (!iw || !r0 || !r1 || !x) { free(iw); free(r0); free(r1); free(x); return -1; }
    const int personalized = seed != ORC_INF;
    const double prop = personalized ? 1.0 - restart : alpha;
    for (uint32_t u = 0; u < n; ++u) {
        double s = 0.0;
        for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) s += w ? (double)w[k] : 1.0;
        iw[u] = s != 0.0 ? prop / s : 0.0;
        r0[u] = personalized ? (u == seed ? 1.0 : 0.0) : 1.0 / (double)n;
        r1[u] = 0.0;
    }
    const double base = personalized ? 0.0 : (1.0 - alpha) / (double)n;
    double* ranks[2] = {r0, r1};
    uint32_t it = 0;
    double delta = 0.0;
    int cur = 0;
    if (mode == 1) {
        /* x is built once from ranks[0] (the first-iteration Construct, engine/blas_gc.h:388-400), then
         * by apply_construct from the rank just written (blas_gc.h:64-74 with construct_each of
         * apps/pagerank_blas.cu:94-100). */
        int idx = 0;
        for (uint32_t u = 0; u < n; ++u) x[u] = ranks[0][u] * iw[u];
        double* y = (double*)malloc(sizeof(double) * (n ? n : 1));
        if (!y) { free(iw); free(r0); free(r1); free(x); return -1; }
        for (;;) {
            delta = 0.0;
            for (uint32_t v = 0; v < n; ++v) {
                const double dl = fabs(ranks[idx ^ 1][v] - ranks[idx][v]);
                if (dl > delta) delta = dl;
            }
            if (delta <= tol || it >= max_iter) break;
            for (uint32_t v = 0; v < n; ++v) {
                double s = 0.0;
                for (uint32_t k = tp[v]; k < tp[v + 1]; ++k) s += x[tj[k]] * (tw ? (double)tw[k] : 1.0);
                y[v] = s;
            }
            for (uint32_t v = 0; v < n; ++v) {
                ranks[idx][v] = base + y[v] + ((personalized && v == seed) ? restart : 0.0);
                x[v] = ranks[idx][v] * iw[v];
            }
            idx ^= 1;
            ++it;
        }
        free(y);
        cur = idx;
    } else {
        /* Rows are independent and each row sums in storage order, so the iterates do not depend on the
         * thread count; the L1 norm is summed per block of 65536 vertices, then over the blocks in order. */
        const uint32_t kBlock = 65536;
        const uint32_t nblk = (n + kBlock - 1) / kBlock;
        double* part = (double*)malloc(sizeof(double) * (nblk ? nblk : 1));
        if (!part) { free(iw); free(r0); free(r1); free(x); return -1; }
        while (it < max_iter) {
            const double* rc = ranks[cur];
            double* rn_out = ranks[cur ^ 1];
#pragma omp parallel for schedule(static)
            for (int64_t u = 0; u < (int64_t)n; ++u) x[u] = rc[u] * iw[u];
#pragma omp parallel for schedule(dynamic, 1)
            for (int64_t b = 0; b < (int64_t)nblk; ++b) {
                const uint32_t v0 = (uint32_t)b * kBlock;
                const uint32_t v1 = v0 + kBlock < n ? v0 + kBlock : n;
                double dsum = 0.0;
                for (uint32_t v = v0; v < v1; ++v) {
                    double s = 0.0;
                    for (uint32_t k = tp[v]; k < tp[v + 1]; ++k) s += x[tj[k]] * (tw ? (double)tw[k] : 1.0);
                    const double rn = base + s + ((personalized && v == seed) ? restart : 0.0);
                    dsum += fabs(rn - rc[v]);
                    rn_out[v] = rn;
                }
                part[b] = dsum;
            }
            delta = 0.0;
            for (uint32_t b = 0; b < nblk; ++b) delta += part[b];
            cur ^= 1;
            ++it;
            if (delta < tol) break;
        }
        free(part);
    }
    memcpy(out, ranks[cur], sizeof(double) * n);
    if (last_delta) *last_delta = delta;
    free(iw); free(r0); free(r1); free(x);
    return (int)it;
}
