// Host-side synthetic graph generators and CSR construction (multi-threaded, deterministic).
//
// CSR invariants follow the reference loader (include/GraphOne/loader/graph_loader.h:254-296 and
// include/GraphOne/loader/edge_cache.h:50-105): self loops dropped, reverse edges added for undirected
// input, (src,dst) duplicates removed, columns ascending within each row, u32 offsets.
//
// Random streams (all splitmix64 finalizer based, so any thread layout gives the same graph):
//   mix64(z)      : z ^= z>>30; z *= 0xBF58476D1CE4E5B9; z ^= z>>27; z *= 0x94D049BB133111EB; z ^= z>>31
//   next(state)   : state += 0x9E3779B97F4A7C15; return mix64(state)
//   unit(r)       : (r >> 11) * 2^-53, a double in [0, 1)
//   R-MAT edge e  : state = mix64(seed ^ mix64(e + 0x632BE59BD9B4E019)); for each of the `scale` levels,
//                   u = unit(next(state)); quadrant (src bit, dst bit) = (0,0) if u < a, (0,1) if
//                   u < a+b, (1,0) if u < a+b+c, else (1,1); bits are appended from the most
//                   significant end.
//   pair hash     : h(seed, p, q) = mix64(seed ^ mix64((uint64(p) << 32) | q))
//   grid edge     : kept iff unit(h(seed, lo, hi)) < p_keep, (lo, hi) the ordered endpoint ids
//   u32 weight    : lo + (((h >> 32) * (hi - lo + 1)) >> 32) with h = h(seed, src, dst), or
//                   h(seed, min, max) when symmetric
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <thread>
#include <vector>

#include "graphone_hip.h"
#include "internal.hpp"

namespace g1 {
namespace {

inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
inline uint64_t next64(uint64_t& state) {
    state += 0x9E3779B97F4A7C15ull;
    return mix64(state);
}
inline double unit64(uint64_t r) { return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0); }
inline uint64_t pair_hash(uint64_t seed, uint32_t p, uint32_t q) {
    return mix64(seed ^ mix64((static_cast<uint64_t>(p) << 32) | q));
}
inline uint32_t weight_from_hash(uint64_t h, uint32_t lo, uint32_t hi) {
    const uint64_t span = static_cast<uint64_t>(hi) - lo + 1;
    return lo + static_cast<uint32_t>(((h >> 32) * span) >> 32);
}

int host_threads() {
    int t = static_cast<int>(std::thread::hardware_concurrency());
    if (const char* env = std::getenv("G1_HOST_THREADS")) t = std::atoi(env);
    if (t < 1) t = 1;
    if (t > 64) t = 64;
    return t;
}

template <typename F>
void parallel_for(size_t n, int threads, F&& f) {
    if (n == 0) return;
    if (threads <= 1 || n < 4096) {
        f(0, size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const size_t lo = t * chunk;
        if (lo >= n) break;
        const size_t hi = std::min(n, lo + chunk);
        pool.emplace_back([&, t, lo, hi] { f(t, lo, hi); });
    }
    for (auto& th : pool) th.join();
}

// Stable LSD radix sort of 64-bit keys by their low key_bits bits (higher bits travel with the key and do
// not take part in the order). Result ends in `a`.
void radix_sort_u64(std::vector<uint64_t>& a, int key_bits, int threads) {
    const size_t n = a.size();
    if (n < 2) return;
    constexpr int kBits = 11;
    constexpr size_t kR = size_t(1) << kBits;
    std::vector<uint64_t> tmp(n);
    const int passes = (key_bits + kBits - 1) / kBits;
    const int T = std::max(1, std::min<int>(threads, static_cast<int>((n + 65535) / 65536)));
    const size_t chunk = (n + T - 1) / T;
    std::vector<size_t> hist(static_cast<size_t>(T) * kR);
    uint64_t* src = a.data();
    uint64_t* dst = tmp.data();
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kBits;
        const uint64_t dmask = (size_t(1) << std::min(kBits, key_bits - shift)) - 1;
        std::fill(hist.begin(), hist.end(), 0);
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t) {
            pool.emplace_back([&, t] {
                const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
                size_t* h = &hist[t * kR];
                for (size_t i = lo; i < hi; ++i) ++h[(src[i] >> shift) & dmask];
            });
        }
        for (auto& th : pool) th.join();
        pool.clear();
        size_t run = 0;
        for (size_t d = 0; d < kR; ++d) {
            for (int t = 0; t < T; ++t) {
                const size_t c = hist[t * kR + d];
                hist[t * kR + d] = run;
                run += c;
            }
        }
        for (int t = 0; t < T; ++t) {
            pool.emplace_back([&, t] {
                const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
                size_t* h = &hist[t * kR];
                for (size_t i = lo; i < hi; ++i) {
                    const uint64_t k = src[i];
                    dst[h[(k >> shift) & dmask]++] = k;
                }
            });
        }
        for (auto& th : pool) th.join();
        std::swap(src, dst);
    }
    if (src != a.data()) a.swap(tmp);
}

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

g1_status alloc_csr(g1_host_csr* out, uint32_t n, uint64_t m, bool weights) {
    G1_CHECK(m < 0xFFFFFFFFull, G1_ERR_INVALID_ARG, "graph has %llu edges; at most 2^32-2 are supported",
             static_cast<unsigned long long>(m));
    out->n = n;
    out->m = static_cast<uint32_t>(m);
    out->row_offsets = static_cast<uint32_t*>(malloc(sizeof(uint32_t) * (static_cast<size_t>(n) + 1)));
    out->col_indices = static_cast<uint32_t*>(malloc(sizeof(uint32_t) * (m ? m : 1)));
    out->weights = weights ? static_cast<uint32_t*>(malloc(sizeof(uint32_t) * (m ? m : 1))) : nullptr;
    if (!out->row_offsets || !out->col_indices || (weights && !out->weights)) {
        g1_host_csr_free(out);
        return set_error(G1_ERR_OUT_OF_MEMORY, "host allocation of a CSR with %llu edges failed",
                         static_cast<unsigned long long>(m));
    }
    return G1_OK;
}

// Sorted unique keys (src << shift | dst) -> CSR.
g1_status csr_from_sorted_keys(const std::vector<uint64_t>& keys, int shift, uint32_t n, g1_host_csr* out) {
    G1_TRY(alloc_csr(out, n, keys.size(), false));
    const uint64_t mask = (shift >= 64) ? ~0ull : ((1ull << shift) - 1);
    std::fill(out->row_offsets, out->row_offsets + static_cast<size_t>(n) + 1, 0u);
    for (size_t i = 0; i < keys.size(); ++i) {
        out->row_offsets[(keys[i] >> shift) + 1]++;
        out->col_indices[i] = static_cast<uint32_t>(keys[i] & mask);
    }
    for (uint32_t r = 0; r < n; ++r) out->row_offsets[r + 1] += out->row_offsets[r];
    return G1_OK;
}

void unique_inplace(std::vector<uint64_t>& keys) {
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
}

}  // namespace
}  // namespace g1

extern "C" {

void g1_host_csr_free(g1_host_csr* csr) {
    if (!csr) return;
    free(csr->row_offsets);
    free(csr->col_indices);
    free(csr->weights);
    csr->row_offsets = nullptr;
    csr->col_indices = nullptr;
    csr->weights = nullptr;
    csr->n = 0;
    csr->m = 0;
}

g1_status g1_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                      int32_t undirected, g1_host_csr* out) {
    using namespace g1;
    G1_CHECK(out != nullptr, G1_ERR_INVALID_ARG, "g1_gen_rmat: null output");
    memset(out, 0, sizeof(*out));
    G1_CHECK(scale >= 1 && scale <= 30, G1_ERR_INVALID_ARG, "g1_gen_rmat: scale must be in 1..30");
    G1_CHECK(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, G1_ERR_INVALID_ARG,
             "g1_gen_rmat: probabilities must be non-negative with a+b+c <= 1");
    const uint64_t n = 1ull << scale;
    const uint64_t m_raw = static_cast<uint64_t>(edge_factor) << scale;
    G1_CHECK(m_raw * (undirected ? 2 : 1) < 0xFFFFFFFFull, G1_ERR_INVALID_ARG,
             "g1_gen_rmat: too many edges for u32 offsets");
    const int threads = host_threads();
    const double ab = a + b, abc = a + b + c;
    // Generate keys src << scale | dst; self loops become the sentinel ~0 and are dropped after sort.
    const uint64_t kDrop = ~0ull;
    std::vector<uint64_t> keys(m_raw * (undirected ? 2 : 1));
    parallel_for(m_raw, threads, [&](int, size_t lo, size_t hi) {
        for (size_t e = lo; e < hi; ++e) {
            uint64_t state = mix64(seed ^ mix64(e + 0x632BE59BD9B4E019ull));
            uint64_t s = 0, d = 0;
            for (uint32_t l = 0; l < scale; ++l) {
                const double u = unit64(next64(state));
                uint64_t sb, db;
                if (u < a) { sb = 0; db = 0; }
                else if (u < ab) { sb = 0; db = 1; }
                else if (u < abc) { sb = 1; db = 0; }
                else { sb = 1; db = 1; }
                s = (s << 1) | sb;
                d = (d << 1) | db;
            }
            if (s == d) {
                keys[e] = kDrop;
                if (undirected) keys[m_raw + e] = kDrop;
            } else {
                keys[e] = (s << scale) | d;
                if (undirected) keys[m_raw + e] = (d << scale) | s;
            }
        }
    });
    radix_sort_u64(keys, 2 * static_cast<int>(scale) + 1, threads);  // the +1 bit orders kDrop last
    while (!keys.empty() && keys.back() == kDrop) keys.pop_back();
    unique_inplace(keys);
    return csr_from_sorted_keys(keys, static_cast<int>(scale), static_cast<uint32_t>(n), out);
}

g1_status g1_gen_grid(uint32_t rows, uint32_t cols, double p_keep, uint64_t seed, g1_host_csr* out) {
    using namespace g1;
    G1_CHECK(out != nullptr, G1_ERR_INVALID_ARG, "g1_gen_grid: null output");
    memset(out, 0, sizeof(*out));
    G1_CHECK(rows >= 1 && cols >= 1, G1_ERR_INVALID_ARG, "g1_gen_grid: empty grid");
    const uint64_t n64 = static_cast<uint64_t>(rows) * cols;
    G1_CHECK(n64 < 0xFFFFFFFFull, G1_ERR_INVALID_ARG, "g1_gen_grid: too many vertices");
    const uint32_t n = static_cast<uint32_t>(n64);
    auto keep = [&](uint32_t p, uint32_t q) {
        const uint32_t lo = p < q ? p : q, hi = p < q ? q : p;
        return unit64(pair_hash(seed, lo, hi)) < p_keep;
    };
    // Neighbour order (up, left, right, down) is ascending in vertex id.
    std::vector<uint32_t> deg(n);
    const int threads = host_threads();
    parallel_for(n, threads, [&](int, size_t lo, size_t hi) {
        for (size_t v = lo; v < hi; ++v) {
            const uint32_t r = static_cast<uint32_t>(v / cols), c = static_cast<uint32_t>(v % cols);
            const uint32_t id = static_cast<uint32_t>(v);
            uint32_t dcount = 0;
            if (r > 0 && keep(id, id - cols)) ++dcount;
            if (c > 0 && keep(id, id - 1)) ++dcount;
            if (c + 1 < cols && keep(id, id + 1)) ++dcount;
            if (r + 1 < rows && keep(id, id + cols)) ++dcount;
            deg[v] = dcount;
        }
    });
    uint64_t m = 0;
    for (uint32_t v = 0; v < n; ++v) m += deg[v];
    G1_TRY(alloc_csr(out, n, m, false));
    out->row_offsets[0] = 0;
    for (uint32_t v = 0; v < n; ++v) out->row_offsets[v + 1] = out->row_offsets[v] + deg[v];
    parallel_for(n, threads, [&](int, size_t lo, size_t hi) {
        for (size_t v = lo; v < hi; ++v) {
            const uint32_t r = static_cast<uint32_t>(v / cols), c = static_cast<uint32_t>(v % cols);
            const uint32_t id = static_cast<uint32_t>(v);
            uint32_t* o = out->col_indices + out->row_offsets[v];
            if (r > 0 && keep(id, id - cols)) *o++ = id - cols;
            if (c > 0 && keep(id, id - 1)) *o++ = id - 1;
            if (c + 1 < cols && keep(id, id + 1)) *o++ = id + 1;
            if (r + 1 < rows && keep(id, id + cols)) *o++ = id + cols;
        }
    });
    return G1_OK;
}

g1_status g1_csr_from_edges(uint32_t n, uint64_t num_edges, const uint32_t* src, const uint32_t* dst,
                            const uint32_t* weights, int32_t directed, int32_t keep_self, int32_t keep_dup,
                            g1_host_csr* out) {
    using namespace g1;
    G1_CHECK(out != nullptr, G1_ERR_INVALID_ARG, "g1_csr_from_edges: null output");
    memset(out, 0, sizeof(*out));
    G1_CHECK(num_edges == 0 || (src && dst), G1_ERR_INVALID_ARG, "g1_csr_from_edges: null edge arrays");
    // Entries: (key = src << 32 | dst, input position) sorted stably by key; the first occurrence of a
    // duplicate pair keeps its weight.
    std::vector<uint64_t> keys;
    std::vector<uint64_t> pos;  // position in the expanded list: 2*i (as given) or 2*i+1 (reversed)
    keys.reserve(num_edges * (directed ? 1 : 2));
    pos.reserve(keys.capacity());
    for (uint64_t i = 0; i < num_edges; ++i) {
        G1_CHECK(src[i] < n && dst[i] < n, G1_ERR_INVALID_ARG,
                 "g1_csr_from_edges: edge %llu (%u -> %u) out of range for %u vertices",
                 static_cast<unsigned long long>(i), src[i], dst[i], n);
        if (src[i] == dst[i] && !keep_self) continue;
        keys.push_back((static_cast<uint64_t>(src[i]) << 32) | dst[i]);
        pos.push_back(2 * i);
        if (!directed && src[i] != dst[i]) {
            keys.push_back((static_cast<uint64_t>(dst[i]) << 32) | src[i]);
            pos.push_back(2 * i + 1);
        }
    }
    std::vector<uint64_t> order(keys.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return keys[x] < keys[y]; });
    std::vector<uint64_t> kept;
    kept.reserve(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        if (!keep_dup && i > 0 && keys[order[i]] == keys[order[i - 1]]) continue;
        kept.push_back(order[i]);
    }
    G1_TRY(alloc_csr(out, n, kept.size(), weights != nullptr));
    std::fill(out->row_offsets, out->row_offsets + static_cast<size_t>(n) + 1, 0u);
    for (size_t i = 0; i < kept.size(); ++i) {
        const uint64_t k = keys[kept[i]];
        out->row_offsets[(k >> 32) + 1]++;
        out->col_indices[i] = static_cast<uint32_t>(k & 0xFFFFFFFFull);
        if (weights) out->weights[i] = weights[pos[kept[i]] >> 1];
    }
    for (uint32_t r = 0; r < n; ++r) out->row_offsets[r + 1] += out->row_offsets[r];
    return G1_OK;
}

g1_status g1_assign_weights_u32(g1_host_csr* csr, uint32_t lo, uint32_t hi, uint64_t seed, int32_t symmetric) {
    using namespace g1;
    G1_CHECK(csr != nullptr && csr->row_offsets != nullptr, G1_ERR_INVALID_ARG, "g1_assign_weights_u32: null csr");
    G1_CHECK(lo <= hi, G1_ERR_INVALID_ARG, "g1_assign_weights_u32: lo > hi");
    if (!csr->weights) {
        csr->weights = static_cast<uint32_t*>(malloc(sizeof(uint32_t) * (csr->m ? csr->m : 1)));
        G1_CHECK(csr->weights != nullptr, G1_ERR_OUT_OF_MEMORY, "g1_assign_weights_u32: host allocation failed");
    }
    parallel_for(csr->n, host_threads(), [&](int, size_t vlo, size_t vhi) {
        for (size_t v = vlo; v < vhi; ++v) {
            const uint32_t s = static_cast<uint32_t>(v);
            for (uint32_t e = csr->row_offsets[v]; e < csr->row_offsets[v + 1]; ++e) {
                const uint32_t d = csr->col_indices[e];
                const uint64_t h = symmetric ? pair_hash(seed, s < d ? s : d, s < d ? d : s) : pair_hash(seed, s, d);
                csr->weights[e] = weight_from_hash(h, lo, hi);
            }
        }
    });
    return G1_OK;
}

g1_status g1_host_csr_transpose(const g1_host_csr* in, g1_host_csr* out) {
    using namespace g1;
    G1_CHECK(in != nullptr && out != nullptr && in->row_offsets != nullptr, G1_ERR_INVALID_ARG,
             "g1_host_csr_transpose: null argument");
    memset(out, 0, sizeof(*out));
    const uint32_t n = in->n;
    const uint32_t m = in->m;
    G1_TRY(alloc_csr(out, n, m, in->weights != nullptr));
    const int threads = host_threads();
    // A stable sort of the edges by destination, the edges listed by ascending source: every in-edge row
    // ends up with its sources ascending and the weights follow their edges (the device transpose of
    // convert.hip gives the same bytes). Keys: (edge index << vb) | destination, sorted on the low vb bits.
    const int vb = std::max(1, bits_for(n > 0 ? n - 1 : 0));
    G1_CHECK(m == 0 || bits_for(m - 1) + vb <= 64, G1_ERR_INVALID_ARG, "g1_host_csr_transpose: graph too large");
    std::vector<uint64_t> keys(m);
    std::vector<uint32_t> rowof(m);
    parallel_for(n, threads, [&](int, size_t lo, size_t hi) {
        for (size_t u = lo; u < hi; ++u)
            for (uint32_t e = in->row_offsets[u]; e < in->row_offsets[u + 1]; ++e) {
                rowof[e] = static_cast<uint32_t>(u);
                keys[e] = (static_cast<uint64_t>(e) << vb) | in->col_indices[e];
            }
    });
    radix_sort_u64(keys, vb, threads);
    const uint64_t vmask = (1ull << vb) - 1;
    parallel_for(m, threads, [&](int, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const uint64_t e = keys[i] >> vb;
            out->col_indices[i] = rowof[e];
            if (out->weights) out->weights[i] = in->weights[e];
            // row starts: every vertex in (previous destination, this destination] starts at i
            const uint32_t v = static_cast<uint32_t>(keys[i] & vmask);
            const int64_t prev = i == 0 ? -1 : static_cast<int64_t>(keys[i - 1] & vmask);
            for (int64_t x = prev + 1; x <= static_cast<int64_t>(v); ++x) out->row_offsets[x] = static_cast<uint32_t>(i);
        }
    });
    const int64_t last = m == 0 ? -1 : static_cast<int64_t>(keys[m - 1] & vmask);
    for (int64_t x = last + 1; x <= static_cast<int64_t>(n); ++x) out->row_offsets[x] = m;
    return G1_OK;
}

}  // extern "C"
