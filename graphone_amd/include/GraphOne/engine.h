// Engines and the algorithm API of the graph_one C++ API.
//
//   ComponentX / make_component       include/GraphOne/base/component.h:7-56
//   Run<Functor>(comp, frontier)      include/GraphOne/base/run.h:8-39
//   AdvanceFunctor, AdvanceGC         include/GraphOne/engine/advance_gc.h:162-303
//   BlasFunctor, BlasGcEngine<kind>   include/GraphOne/engine/blas_gc.h:99-474
//
// User functors are __device__ statics inlined into the templated gfx950 kernels of g1/push.hpp and
// g1/spmv_merge.hpp (instantiated in the caller's translation unit by hipcc); memory and the merge-path
// partition go through the C-ABI. The advance engine fuses advance, filter and compaction (the output
// frontier holds exactly the kept vertices, in no particular order, instead of the reference's
// INVALID-padded list); the BLAS engines keep the reference's Construct / Gather / Apply structure, with
// Apply fused with the next Construct where the reference also fuses it (apply_construct_kernel).
#pragma once

#include <stdio.h>
#include <stdlib.h>

#include <type_traits>
#include <utility>
#include <vector>

#include "GraphOne/frontier.h"
#include "GraphOne/graph.h"
#include "GraphOne/spmv.h"
#include "g1/push.hpp"
#include "g1/step.hpp"

namespace graph_one {

template <typename graph_t, typename hstatus_t, typename dstatus_t>
struct ComponentX {
    using graph_type = graph_t;
    using hstatus_type = hstatus_t;
    using dstatus_type = dstatus_t;
    using vertex_type = typename graph_t::vertex_type;
    using edge_type = typename graph_t::edge_type;
    using dstatus_ref_t = std::conditional_t<std::is_same_v<dstatus_t, empty_t>, empty_t, dstatus_t&>;

    ComponentX(const graph_t& g, hstatus_t& h, dstatus_ref_t d) : graph(g), h_status(h), d_status(d) {}
    virtual ~ComponentX() = default;
    virtual void Init() {}
    virtual bool IsConvergent() { return false; }
    virtual void BeforeEngine() {}
    virtual void AfterEngine() {}
    virtual void Final() {}

    const graph_t& graph;
    hstatus_t& h_status;
    dstatus_ref_t d_status;
};

template <template <typename> class comp_t, typename graph_t, typename hstatus_t, typename dstatus_t>
comp_t<graph_t> make_component(const graph_t& g, hstatus_t& h, dstatus_t& d) {
    return comp_t<graph_t>{g, h, d};
}
template <template <typename> class comp_t, typename graph_t, typename hstatus_t>
comp_t<graph_t> make_component(const graph_t& g, hstatus_t& h) {
    return comp_t<graph_t>{g, h};
}

namespace engine {

template <typename comp_t, typename frontier_t>
class GcEngine {
public:
    using graph_type = typename comp_t::graph_type;
    using hstatus_type = typename comp_t::hstatus_type;
    using dstatus_type = typename comp_t::dstatus_type;
    GcEngine(comp_t& comp, frontier_t& frontier)
        : graph_(comp.graph), h_status_(comp.h_status), d_status_(comp.d_status), frontier_(frontier) {}
    virtual ~GcEngine() = default;
    virtual void Forward() = 0;

protected:
    const graph_type& graph_;
    hstatus_type& h_status_;
    dstatus_type& d_status_;
    frontier_t& frontier_;
};

// ------------------------------------------------------------------------------------------ kernels

// deg[i] = out-degree of list[i] for i < n (0 for an invalid vertex) and deg[n] = 0, so that an exclusive
// scan of n + 1 entries leaves the total in deg[n].
template <typename vertex_t>
__global__ void list_degrees_kernel(const vertex_t* __restrict__ list, uint32_t n, uint32_t n_vertices,
                                    const uint32_t* __restrict__ row_off, uint32_t* __restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        uint32_t d = 0;
        if (i < n) {
            const uint32_t v = static_cast<uint32_t>(list[i]);
            if (v < n_vertices) d = row_off[v + 1] - row_off[v];
        }
        out[i] = d;
    }
}

constexpr uint32_t kScanBlock = 2048;  // items per block of the three-pass scan (256 threads x 8)

static __global__ void scan_block_sums_kernel(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t s[4];
    const uint32_t b0 = blockIdx.x * kScanBlock;
    uint32_t acc = 0;
    for (uint32_t k = threadIdx.x; k < kScanBlock; k += 256)
        if (b0 + k < n) acc += in[b0 + k];
    acc = g1::wave_sum(acc);
    if (g1::lane_id() == 0) s[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// Exclusive scan of sums[0..count) in place by one block; total written to sums[count].
static __global__ void scan_sums_kernel(uint32_t* __restrict__ sums, uint32_t count) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < count ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = g1::block_exclusive_sum<256>(v, tot, lds);
        if (i < count) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sums[count] = carry;
}

// In-place exclusive scan of a block's items with its offset; items per thread are contiguous.
static __global__ void scan_apply_kernel(uint32_t* __restrict__ data, uint32_t n, const uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[4];
    const uint32_t b0 = blockIdx.x * kScanBlock + threadIdx.x * 8;
    uint32_t v[8];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = (b0 + k < n) ? data[b0 + k] : 0u;
        local += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + g1::block_exclusive_sum<256>(local, tot, lds);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (b0 + k < n) data[b0 + k] = run;
        run += v[k];
    }
}

// Exclusive scan of data[0..n) in place, on the device only; the total is left in scratch[blocks].
// scratch holds >= n/kScanBlock + 2 entries.
inline void exclusive_scan_u32_dev(uint32_t* data, uint32_t n, uint32_t* scratch) {
    if (n == 0) return;
    const uint32_t blocks = (n + kScanBlock - 1) / kScanBlock;
    scan_block_sums_kernel<<<blocks, 256>>>(data, n, scratch);
    scan_sums_kernel<<<1, 256>>>(scratch, blocks);
    scan_apply_kernel<<<blocks, 256>>>(data, n, scratch);
    archi::launch_check("exclusive scan");
}

// The same, returning the total to the host (one device-to-host read).
inline uint32_t exclusive_scan_u32(uint32_t* data, uint32_t n, uint32_t* scratch) {
    if (n == 0) return 0;
    exclusive_scan_u32_dev(data, n, scratch);
    uint32_t total = 0;
    archi::memcpy<arch_t::cpu, arch_t::hip, uint32_t>(&total, scratch + (n + kScanBlock - 1) / kScanBlock, 1);
    return total;
}

// Clears the advance engine's packed output counter ((edges << 32) | count) and overflow flag.
static __global__ void clear_counters_kernel(unsigned long long* packed, uint32_t* overflow) {
    *packed = 0ull;
    *overflow = 0u;
}

// Appends each lane's kept vertices (bit u of `keep`: slot u's vertex v[u], out-degree dg[u]) to the output
// list with one reservation per wave on the packed counter (count in the low word, out-edges in the high
// word); entries past `capacity` are dropped and flagged.
template <int N>
__device__ __forceinline__ void append_kept(uint32_t keep, const uint32_t (&v)[N], const uint32_t (&dg)[N],
                                            uint32_t* out, uint32_t capacity, unsigned long long* packed,
                                            uint32_t* overflow) {
    unsigned long long mine = 0ull;
#pragma unroll
    for (int u = 0; u < N; ++u)
        if (keep & (1u << u)) mine += (static_cast<unsigned long long>(dg[u]) << 32) | 1ull;
    const unsigned long long incl = g1::wave_inclusive_sum(mine);
    const unsigned long long total = __shfl(incl, g1::kWave - 1);
    if (total == 0ull) return;
    unsigned long long base = 0ull;
    if (g1::lane_id() == 0) base = atomicAdd(packed, total);
    base = __shfl(base, 0);
    uint32_t at = static_cast<uint32_t>(base + incl - mine);
#pragma unroll
    for (int u = 0; u < N; ++u) {
        if (keep & (1u << u)) {
            if (at < capacity) out[at] = v[u];
            else atomicExch(overflow, 1u);
            ++at;
        }
    }
}

// Push relaxation of a user AdvanceFunctor for g1::step_advance_kernel: the source value is the source id;
// the functor's advance() decides each edge and its filter() (when present) drops the kept vertex. Kept
// vertices are appended with their out-degrees (the next iteration's edge count).
template <typename F, typename GRef, typename D, bool kHasFilter>
struct FunctorPush {
    using vertex_t = typename GRef::vertex_type;
    using edge_t = typename GRef::edge_type;
    GRef g;
    D d;
    const uint32_t* row_off;
    const uint32_t* col;
    uint32_t* out;
    uint32_t capacity;
    unsigned long long* packed;
    uint32_t* overflow;
    __device__ __forceinline__ uint32_t source(uint32_t u) const { return u; }
    __device__ __forceinline__ void operator()(const uint32_t (&e)[g1::kSweepUnroll], const uint32_t (&src)[g1::kSweepUnroll],
                                               const bool (&ok)[g1::kSweepUnroll]) const {
        D dl = d;
        uint32_t v[g1::kSweepUnroll], dg[g1::kSweepUnroll];
        uint32_t keep = 0;
#pragma unroll
        for (int u = 0; u < g1::kSweepUnroll; ++u) v[u] = ok[u] ? col[e[u]] : 0u;
#pragma unroll
        for (int u = 0; u < g1::kSweepUnroll; ++u) {
            dg[u] = 0;
            if (!ok[u]) continue;
            bool k = F::advance(static_cast<vertex_t>(src[u]), static_cast<vertex_t>(v[u]), static_cast<edge_t>(e[u]),
                                g.get_out_edge_weight(static_cast<edge_t>(e[u])), dl);
            if constexpr (kHasFilter) {
                if (k) k = !F::filter(static_cast<vertex_t>(v[u]), dl);
            }
            if (k) {
                keep |= 1u << u;
                dg[u] = row_off[v[u] + 1] - row_off[v[u]];
            }
        }
        append_kept(keep, v, dg, out, capacity, packed, overflow);
    }
};

// Pull step of the adaptive frontier (DblBufFrontier): groups of 64 vertices per wave; every in-edge
// (u, v) of the group whose source is in the input frontier (bitmap fb) calls the functor's advance(u, v,
// in-edge index, weight) -- the relaxations the push step would make, grouped by target -- and v joins the
// output once if any of them kept it and the filter (when present) does not drop it. The in-edges are
// dealt out 64 per step (g1::wave_expand); a 64-flag LDS array per wave collects the keeps.
template <typename F, typename GRef, typename D, bool kHasFilter, typename weight_t>
__global__ void __launch_bounds__(256)
functor_pull_kernel(uint32_t n, const uint32_t* __restrict__ t_row_off, const uint32_t* __restrict__ t_col,
                    const weight_t* __restrict__ t_w, const uint32_t* __restrict__ fb, GRef g, D d,
                    const uint32_t* __restrict__ row_off, uint32_t* out, uint32_t capacity,
                    unsigned long long* packed, uint32_t* overflow) {
    using vertex_t = typename GRef::vertex_type;
    using edge_t = typename GRef::edge_type;
    __shared__ uint32_t s_mark[256];
    __shared__ uint32_t s_keep[256];
    const uint32_t lane = g1::lane_id();
    const uint32_t wave = threadIdx.x / g1::kWave;
    uint32_t* keepf = s_keep + wave * g1::kWave;
    const uint32_t groups = (n + g1::kWave - 1) / g1::kWave;
    const uint32_t waves = gridDim.x * (blockDim.x / g1::kWave);
    D dl = d;
    for (uint32_t grp = blockIdx.x * (blockDim.x / g1::kWave) + wave; grp < groups; grp += waves) {  // wave-uniform
        const uint32_t v0 = grp * g1::kWave;
        const uint32_t v = v0 + lane;
        uint32_t rs = 0, deg = 0;
        if (v < n) {
            rs = t_row_off[v];
            deg = t_row_off[v + 1] - rs;
        }
        keepf[lane] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        auto relax = [&](const uint32_t (&e)[g1::kSweepUnroll], const uint32_t (&tv)[g1::kSweepUnroll],
                         const bool (&ok)[g1::kSweepUnroll]) {
            uint32_t u[g1::kSweepUnroll];
            bool in[g1::kSweepUnroll];
#pragma unroll
            for (int k = 0; k < g1::kSweepUnroll; ++k) u[k] = ok[k] ? t_col[e[k]] : 0u;
#pragma unroll
            for (int k = 0; k < g1::kSweepUnroll; ++k) in[k] = ok[k] && ((fb[u[k] >> 5] >> (u[k] & 31u)) & 1u);
#pragma unroll
            for (int k = 0; k < g1::kSweepUnroll; ++k) {
                if (!in[k]) continue;
                weight_t w{};
                if constexpr (!std::is_same_v<weight_t, empty_t>) w = t_w ? t_w[e[k]] : weight_t(1);
                if (F::advance(static_cast<vertex_t>(u[k]), static_cast<vertex_t>(tv[k]), static_cast<edge_t>(e[k]), w, dl))
                    keepf[tv[k] - v0] = 1u;
            }
        };
        g1::wave_expand(rs, deg, v, s_mark + wave * g1::kWave, relax);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool k = v < n && keepf[lane] != 0u;
        if constexpr (kHasFilter) {
            if (k) k = !F::filter(static_cast<vertex_t>(v), dl);
        }
        uint32_t vv[1] = {v}, dg[1] = {k ? row_off[v + 1] - row_off[v] : 0u};
        append_kept<1>(k ? 1u : 0u, vv, dg, out, capacity, packed, overflow);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// x[v] = construct_each(v) for the listed vertices (the sparse Construct's transform + scatter).
template <typename F, typename D, typename vertex_t, typename info_t>
__global__ void construct_list_kernel(const vertex_t* __restrict__ list, uint32_t n, D d, info_t* __restrict__ x) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const vertex_t v = list[i];
        x[v] = F::construct_each(v, d);
    }
}

// x[v] = construct_each(v) for every vertex (the all-active engine's first Construct).
template <typename F, typename D, typename info_t>
__global__ void construct_all_kernel(uint32_t n, D d, info_t* __restrict__ x) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) x[v] = F::construct_each(v, d);
}

// flags[v] = res != default_result && apply_each(v, res) (the sparse Apply's transform); the flags are
// then scanned into output positions so the next frontier lists the vertices in ascending order, as the
// reference's stable copy_if does (blas_gc.h:224-245).
template <typename F, typename D, typename gather_t>
__global__ void apply_flags_kernel(const gather_t* __restrict__ y, uint32_t n, D d, uint32_t* __restrict__ flags) {
    D dl = d;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const gather_t res = y[v];
        flags[v] = (res != F::default_result() && F::apply_each(v, res, dl)) ? 1u : 0u;
    }
}

template <typename vertex_t>
__global__ void compact_flags_kernel(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ flags_copy,
                                     uint32_t n, vertex_t* __restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        if (flags_copy[v]) out[pos[v]] = static_cast<vertex_t>(v);
}

// apply_construct_kernel of blas_gc.h:21-76: pred = (counted ? res != default_result && : ) apply_each;
// x[v] = pred ? construct_each(v) : default_info; with counted, per-block counts go to block_counts.
template <typename F, typename D, typename gather_t, typename info_t, bool kCounted>
__global__ void apply_construct_kernel(const gather_t* __restrict__ y, uint32_t n, D d, info_t* __restrict__ x,
                                       uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t s[4];
    D dl = d;
    uint32_t cnt = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const gather_t res = y[v];
        bool pred;
        if constexpr (kCounted) pred = res != F::default_result() && F::apply_each(v, res, dl);
        else pred = F::apply_each(v, res, dl);
        x[v] = pred ? F::construct_each(v, dl) : F::default_info();
        cnt += pred ? 1u : 0u;
    }
    if constexpr (kCounted) {
        cnt = g1::wave_sum(cnt);
        if (g1::lane_id() == 0) s[threadIdx.x / 64] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) block_counts[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
    }
}

// A whole advance step of a small frontier (at most kSmallIn entries, its edge count known) in one
// workgroup of 1024 threads: the entries' degrees are scanned in LDS, each edge slot finds its entry with
// a binary search over the LDS offsets, and the functor relaxes it (the push relaxation of FunctorPush).
// The packed output counter is cleared here first, so the step is one launch.
constexpr uint32_t kSmallIn = 2048;
constexpr uint32_t kSmallThreads = 1024;

template <typename Rel, typename vertex_t>
__global__ void __launch_bounds__(kSmallThreads)
small_advance_kernel(const vertex_t* __restrict__ f, uint32_t nin, uint32_t n_vertices,
                     const uint32_t* __restrict__ row_off, Rel rel) {
    __shared__ uint32_t s_off[kSmallIn + 1];
    __shared__ uint32_t s_rs[kSmallIn];
    __shared__ uint32_t s_src[kSmallIn];
    __shared__ uint32_t s_red[kSmallThreads / g1::kWave];
    if (threadIdx.x == 0) {
        rel.packed[0] = 0ull;
        *rel.overflow = 0u;
    }
    // degrees, two entries per thread, then the block scan of the pairs
    uint32_t d[2], a[2], u[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t i = threadIdx.x * 2 + k;
        d[k] = 0;
        a[k] = 0;
        u[k] = g1::kInf;
        if (i < nin) {
            u[k] = static_cast<uint32_t>(f[i]);
            if (u[k] < n_vertices) {
                a[k] = row_off[u[k]];
                d[k] = row_off[u[k] + 1] - a[k];
            }
        }
    }
    uint32_t total = 0;
    const uint32_t ex = g1::block_exclusive_sum<kSmallThreads>(d[0] + d[1], total, s_red);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t i = threadIdx.x * 2 + k;
        if (i < kSmallIn) {
            s_off[i] = ex + (k ? d[0] : 0u);
            s_rs[i] = a[k];
            s_src[i] = u[k];
        }
    }
    if (threadIdx.x == 0) s_off[kSmallIn] = total;
    __syncthreads();
    // edge slots in rounds of kSweepUnroll x 1024; every thread of a wave takes part in each round
    for (uint32_t base = 0; base < total; base += kSmallThreads * g1::kSweepUnroll) {
        uint32_t e[g1::kSweepUnroll], src[g1::kSweepUnroll];
        bool ok[g1::kSweepUnroll];
#pragma unroll
        for (int k = 0; k < g1::kSweepUnroll; ++k) {
            const uint32_t slot = base + k * kSmallThreads + threadIdx.x;
            ok[k] = slot < total;
            e[k] = 0;
            src[k] = 0;
            if (ok[k]) {
                // the last entry whose first slot is <= slot (entries without edges share offsets)
                uint32_t lo = 0, hi = kSmallIn;
                while (lo + 1 < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_off[mid] <= slot) lo = mid; else hi = mid;
                }
                e[k] = s_rs[lo] + (slot - s_off[lo]);
                src[k] = s_src[lo];
            }
        }
        rel(e, src, ok);
    }
}

// ------------------------------------------------------------------------------------------ AdvanceGC

// AdvanceGC::Forward (advance_gc.h:180-244): ResizeOutput, Advance and Filter as one device-driven step.
// The input list's degrees are scanned on the device into edge offsets; its edge count E comes from the
// previous step's read-back (or, for a list the engine did not write, one read after the scan) and sizes
// the output; the functor runs inside g1::step_advance_kernel (edge windows of 4096 slots, each dealt 64
// per step over the entries that overlap it) and appends each kept vertex with its out-degree. The one
// host read per iteration is the output's (count, out-edges) pair, which Run's convergence test needs.
// On a DblBufFrontier an iteration whose E reaches pull_fraction() x m runs as a pull over the in-edges.
template <typename functor_t, typename comp_t, typename frontier_t>
class AdvanceGC : public GcEngine<comp_t, frontier_t> {
public:
    using base_t = GcEngine<comp_t, frontier_t>;
    using graph_t = typename base_t::graph_type;
    using dstatus_t = typename base_t::dstatus_type;
    using vertex_t = typename graph_t::vertex_type;
    using edge_t = typename graph_t::edge_type;
    using weight_t = typename graph_t::weight_type;
    using index_t = typename frontier_t::index_type;
    using gref_t = decltype(std::declval<const graph_t&>().ToArch());
    constexpr static arch_t arch = graph_t::arch_value;
    static_assert(frontier_t::kind == SPARSE_BASED, "AdvanceGC runs on a sparse frontier");
    static_assert(arch == arch_t::hip, "AdvanceGC runs on the device");
    static_assert(sizeof(vertex_t) == 4 && sizeof(edge_t) == 4, "32-bit vertex and edge ids");

    struct HasFilter {
        template <typename T>
        static std::true_type test(decltype(&T::filter));
        template <typename T>
        static std::false_type test(...);
        constexpr static bool value = decltype(test<functor_t>(0))::value;
    };

    AdvanceGC(comp_t& comp, frontier_t& frontier) : base_t(comp, frontier), scratch_(64), ctr_(2) {}

    ~AdvanceGC() override {
        if constexpr (frontier_t::adaptive) {
            if (getenv("G1_DBL_TRACE"))
                fprintf(stderr, "[graph_one] DblBufFrontier: %u push steps, %u pull steps\n", pushes_, pulls_);
        }
    }

    void Forward() override {
        auto& fr = this->frontier_;
        const auto gref = this->graph_.ToArch();
        const uint32_t nin = fr.input_size();
        const uint32_t nv = static_cast<uint32_t>(this->graph_.num_vertices());
        const uint32_t m = static_cast<uint32_t>(this->graph_.num_edges());
        const uint32_t* row_off = reinterpret_cast<const uint32_t*>(gref.row_offsets);
        if (nin == 0) {
            fr.reset_output(0);
            fr.set_output_edges(0);
            return;
        }
        // A small frontier whose edge count is known runs as one launch (small_advance_kernel).
        if (fr.input_edges_valid() && nin <= kSmallIn) {
            bool pull = false;
            if constexpr (frontier_t::adaptive) {
                const double f = frontier_t::pull_fraction();
                pull = f > 0.0 && m > 0 && static_cast<double>(fr.input_edges()) >= f * static_cast<double>(m);
            }
            if (!pull) {
                const uint32_t cap = std::max<uint32_t>(fr.input_edges(), 1u);
                fr.reserve_output_capacity(cap);
                FunctorPush<functor_t, gref_t, dstatus_t, HasFilter::value> rel{
                    gref, this->d_status_, row_off, reinterpret_cast<const uint32_t*>(gref.col_indices),
                    reinterpret_cast<uint32_t*>(fr.output().data()), cap, packed(), overflow()};
                small_advance_kernel<decltype(rel), vertex_t><<<1, kSmallThreads>>>(fr.input().data(), nin, nv, row_off,
                                                                                   rel);
                archi::launch_check("small advance");
                ++pushes_;
                read_output();
                return;
            }
        }
        // edge offsets of the input list: off_[0..nin], off_[nin] = E
        grow(off_, nin + 1);
        grow(scratch_, (nin + 1) / kScanBlock + 4);
        list_degrees_kernel<vertex_t><<<archi::grid_blocks(nin + 1), 256>>>(fr.input().data(), nin, nv, row_off,
                                                                          off_.data());
        archi::launch_check("advance degrees");
        exclusive_scan_u32_dev(off_.data(), nin + 1, scratch_.data());
        uint32_t E;
        if (fr.input_edges_valid()) {
            E = fr.input_edges();
        } else {
            archi::memcpy<arch_t::cpu, arch_t::hip, uint32_t>(&E, off_.data() + nin, 1);
        }
        bool pull = false;
        if constexpr (frontier_t::adaptive) {
            const double f = frontier_t::pull_fraction();
            pull = f > 0.0 && m > 0 && static_cast<double>(E) >= f * static_cast<double>(m);
        }
        // each kept edge (push) or vertex (pull) appends one entry
        const uint32_t cap = std::max<uint32_t>(pull ? nv : E, 1u);
        fr.reserve_output_capacity(cap);
        clear_counters_kernel<<<1, 1>>>(packed(), overflow());
        if (pull) {
            run_pull(gref, nin, nv, m, fr.output().data(), cap);
            ++pulls_;
        } else if (E > 0) {
            FunctorPush<functor_t, gref_t, dstatus_t, HasFilter::value> rel{
                gref, this->d_status_, row_off, reinterpret_cast<const uint32_t*>(gref.col_indices),
                reinterpret_cast<uint32_t*>(fr.output().data()), cap, packed(), overflow()};
            const uint32_t windows = (E + g1::kStepWindow - 1) / g1::kStepWindow;
            const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(2048u, (windows + 3) / 4));
            g1::step_advance_kernel<decltype(rel)><<<blocks, 256>>>(row_off, reinterpret_cast<const uint32_t*>(fr.input().data()),
                                                                    nullptr, off_.data(), nin, rel);
            archi::launch_check("advance");
            ++pushes_;
        }
        read_output();
    }

private:
    // The one read of the iteration: (count, out-edges) of the output and the overflow flag.
    void read_output() {
        auto& fr = this->frontier_;
        unsigned long long h[2] = {0ull, 0ull};
        archi::memcpy<arch_t::cpu, arch_t::hip, unsigned long long>(h, ctr_.data(), 2);
        const unsigned long long pk = h[0];
        if (static_cast<uint32_t>(h[1])) LOG_ERROR("advance: output list overflow");
        fr.reset_output(static_cast<index_t>(static_cast<uint32_t>(pk)));
        fr.set_output_edges(static_cast<uint32_t>(pk >> 32));
    }

    // The pull step: the input as a bitmap, then functor_pull_kernel over the in-edge CSR (the graph's
    // transpose view, or one built here on the device at the first pull).
    void run_pull(const gref_t& gref, uint32_t nin, uint32_t nv, uint32_t m, vertex_t* out, uint32_t cap) {
        auto& fr = this->frontier_;
        const uint32_t words = (nv + 31) / 32;
        grow(fb_, words + 1);
        archi::fill<arch_t::hip>(fb_.begin(), fb_.end(), 0u);
        g1::queue_to_bitmap_kernel<<<archi::grid_blocks(nin), 256>>>(reinterpret_cast<const uint32_t*>(fr.input().data()),
                                                                    nullptr, nin, nv, fb_.data());
        archi::launch_check("pull bitmap");
        const uint32_t* t_ro;
        const uint32_t* t_ci;
        const weight_t* t_w;
        if constexpr (graph_t::both_mat) {
            t_ro = reinterpret_cast<const uint32_t*>(gref.trans_row_offsets);
            t_ci = reinterpret_cast<const uint32_t*>(gref.trans_col_indices);
            t_w = gref.trans_values;
        } else {
            ensure_transpose(gref, nv, m);
            t_ro = t_ro_.data();
            t_ci = t_ci_.data();
            t_w = has_w_ ? reinterpret_cast<const weight_t*>(t_w_.data()) : nullptr;
        }
        const uint32_t groups = (nv + 63) / 64;
        const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(2048u, (groups + 3) / 4));
        functor_pull_kernel<functor_t, gref_t, dstatus_t, HasFilter::value, weight_t><<<blocks, 256>>>(
            nv, t_ro, t_ci, t_w, fb_.data(), gref, this->d_status_, reinterpret_cast<const uint32_t*>(gref.row_offsets),
            reinterpret_cast<uint32_t*>(out), cap, packed(), overflow());
        archi::launch_check("pull");
    }

    // The in-edge CSR of a graph kept in its out-edge view only, through the library's device transpose
    // (4-byte weights travel as their bit patterns).
    void ensure_transpose(const gref_t& gref, uint32_t nv, uint32_t m) {
        if (t_ro_.size()) return;
        static_assert(sizeof(weight_t) == 4 || std::is_same_v<weight_t, empty_t>, "4-byte weights");
        has_w_ = !std::is_same_v<weight_t, empty_t> && gref.values != nullptr;
        t_ro_ = Buffer<arch_t::hip, uint32_t, uint32_t>(nv + 1);
        t_ci_ = Buffer<arch_t::hip, uint32_t, uint32_t>(std::max<uint32_t>(m, 1));
        if (has_w_) t_w_ = Buffer<arch_t::hip, uint32_t, uint32_t>(std::max<uint32_t>(m, 1));
        const size_t bytes = g1_csr_transpose_workspace_bytes(nv, m, has_w_ ? 1 : 0);
        Buffer<arch_t::hip, uint8_t, size_t> ws(std::max<size_t>(bytes, 4));
        archi::check(g1_csr_transpose(nv, m, reinterpret_cast<const uint32_t*>(gref.row_offsets),
                                             reinterpret_cast<const uint32_t*>(gref.col_indices),
                                             has_w_ ? reinterpret_cast<const uint32_t*>(gref.values) : nullptr,
                                             t_ro_.data(), t_ci_.data(), has_w_ ? t_w_.data() : nullptr, ws.data(),
                                             ws.size(), nullptr),
                            "transpose");
        archi::LaunchSync<arch_t::hip>();
    }

    template <typename T>
    static void grow(Buffer<arch_t::hip, T, uint32_t>& b, uint32_t n) {
        if (b.size() < n) b.resize(n);
    }

    Buffer<arch_t::hip, uint32_t, uint32_t> off_;
    Buffer<arch_t::hip, uint32_t, uint32_t> scratch_;
    // [0]: the output's packed counter (edges << 32 | count); [1]: the overflow flag (low word)
    Buffer<arch_t::hip, unsigned long long, uint32_t> ctr_;
    unsigned long long* packed() { return ctr_.data(); }
    uint32_t* overflow() { return reinterpret_cast<uint32_t*>(ctr_.data() + 1); }
    Buffer<arch_t::hip, uint32_t, uint32_t> fb_;
    Buffer<arch_t::hip, uint32_t, uint32_t> t_ro_, t_ci_, t_w_;
    bool has_w_ = false;
    uint32_t pushes_ = 0, pulls_ = 0;
};

// ------------------------------------------------------------------------------------------ BLAS engines

template <typename functor_t, typename comp_t, typename frontier_t>
class BlasGcEngineBase : public GcEngine<comp_t, frontier_t> {
public:
    using base_t = GcEngine<comp_t, frontier_t>;
    using graph_t = typename base_t::graph_type;
    using dstatus_t = typename base_t::dstatus_type;
    using vertex_t = typename graph_t::vertex_type;
    using edge_t = typename graph_t::edge_type;
    using weight_t = typename graph_t::weight_type;
    using info_t = typename functor_t::info_type;
    using gather_t = typename functor_t::gather_type;
    constexpr static arch_t arch = graph_t::arch_value;

    // The functor as an SpMV semiring (blas_gc.h:113-128).
    struct SpmvFunctor {
        __host__ __device__ __forceinline__ static gather_t initialize() { return functor_t::default_result(); }
        __device__ __forceinline__ static gather_t combine(const weight_t& w, const info_t& x) {
            return functor_t::gather_combine(w, x);
        }
        __device__ __forceinline__ static gather_t reduce(const gather_t& a, const gather_t& b) {
            return functor_t::gather_reduce(a, b);
        }
    };
    using spmv_t = blas::SpmvDispatcher<blas::SpmvHipMergeBased, SpmvFunctor, vertex_t, edge_t, weight_t, info_t, gather_t>;

    BlasGcEngineBase(comp_t& comp, frontier_t& frontier)
        : base_t(comp, frontier), ipt_vec_(comp.graph.num_vertices()), opt_vec_(comp.graph.num_vertices()),
          spmv_(make_spmv(comp.graph, ipt_vec_.data(), opt_vec_.data())) {}

protected:
    static spmv_t make_spmv(const graph_t& g, info_t* x, gather_t* y) {
        const auto r = g.ToArch();
        // The transpose when both views are kept, otherwise the single CSR (the transpose under BlasViews).
        if constexpr (graph_t::both_mat) {
            return blas::MakeCsrSpMV<arch, blas::SpmvHipMergeBased, SpmvFunctor>(
                r.num_vertices, r.num_vertices, r.num_edges, r.trans_row_offsets, r.trans_col_indices, r.trans_values,
                static_cast<const info_t*>(x), y);
        } else {
            return blas::MakeCsrSpMV<arch, blas::SpmvHipMergeBased, SpmvFunctor>(
                r.num_vertices, r.num_vertices, r.num_edges, r.row_offsets, r.col_indices, r.values,
                static_cast<const info_t*>(x), y);
        }
    }

    DenseVec<arch, info_t, vertex_t> ipt_vec_;
    DenseVec<arch, gather_t, vertex_t> opt_vec_;
    spmv_t spmv_;
};

template <typename functor_t, typename comp_t, typename frontier_t, typename = void>
class BlasGcEngine;

// Sparse frontier (blas_gc.h:171-247).
template <typename functor_t, typename comp_t, typename frontier_t>
class BlasGcEngine<functor_t, comp_t, frontier_t, std::enable_if_t<frontier_t::kind == FrontierKind::SPARSE_BASED>>
    : public BlasGcEngineBase<functor_t, comp_t, frontier_t> {
public:
    using base_t = BlasGcEngineBase<functor_t, comp_t, frontier_t>;
    using typename base_t::dstatus_t;
    using typename base_t::gather_t;
    using typename base_t::info_t;
    using typename base_t::vertex_t;

    BlasGcEngine(comp_t& comp, frontier_t& frontier)
        : base_t(comp, frontier), flags_(comp.graph.num_vertices() + 1), pos_(comp.graph.num_vertices() + 1),
          scratch_(comp.graph.num_vertices() / kScanBlock + 4) {}

    // One host read per iteration: the output's size (the scan total), which Run's convergence test needs;
    // every other step stays on the stream.
    void Forward() override {
        auto& fr = this->frontier_;
        const uint32_t n = this->ipt_vec_.size();
        // Construct: x = default_info; x[F] = construct_each(F)
        if (n) {
            archi::fill_value_kernel<info_t><<<archi::grid_blocks(n), 256>>>(this->ipt_vec_.data(),
                                                                         functor_t::default_info(), n);
            archi::launch_check("construct fill");
        }
        const uint32_t nin = fr.input_size();
        if (nin) {
            construct_list_kernel<functor_t><<<archi::grid_blocks(nin), 256>>>(fr.input().data(), nin, this->d_status_,
                                                                               this->ipt_vec_.data());
            archi::launch_check("construct");
        }
        // Gather: y = A^T (x) x
        this->spmv_(false);
        // Apply: flags, then ordered compaction into the output list
        apply_flags_kernel<functor_t><<<archi::grid_blocks(n), 256>>>(this->opt_vec_.data(), n, this->d_status_,
                                                                      flags_.data());
        archi::launch_check("apply");
        archi::check(g1_memcpy_d2d(pos_.data(), flags_.data(), sizeof(uint32_t) * n, nullptr), "flags copy");
        const uint32_t nout = exclusive_scan_u32(pos_.data(), n, scratch_.data());
        fr.reset_output(nout);
        if (nout) {
            compact_flags_kernel<vertex_t><<<archi::grid_blocks(n), 256>>>(pos_.data(), flags_.data(), n,
                                                                           fr.output().data());
            archi::launch_check("compact");
        }
    }

private:
    Buffer<arch_t::hip, uint32_t, uint32_t> flags_;
    Buffer<arch_t::hip, uint32_t, uint32_t> pos_;
    Buffer<arch_t::hip, uint32_t, uint32_t> scratch_;
};

// Dense frontier (blas_gc.h:250-351): x from the seeds on the first iteration, then apply_construct.
template <typename functor_t, typename comp_t, typename frontier_t>
class BlasGcEngine<functor_t, comp_t, frontier_t, std::enable_if_t<frontier_t::kind == FrontierKind::DENSE_BASED>>
    : public BlasGcEngineBase<functor_t, comp_t, frontier_t> {
public:
    using base_t = BlasGcEngineBase<functor_t, comp_t, frontier_t>;
    using typename base_t::gather_t;
    using typename base_t::info_t;

    BlasGcEngine(comp_t& comp, frontier_t& frontier) : base_t(comp, frontier), counts_(kBlocks) {}

    void Forward() override {
        auto& fr = this->frontier_;
        auto& init = fr.init_buf();
        if (!init.empty()) {
            archi::fill<base_t::arch>(this->ipt_vec_.begin(), this->ipt_vec_.end(), functor_t::default_info());
            construct_list_kernel<functor_t><<<archi::grid_blocks(init.size()), 256>>>(
                init.data(), init.size(), this->d_status_, this->ipt_vec_.data());
            archi::launch_check("construct");
            init = std::remove_reference_t<decltype(init)>();
        }
        this->spmv_(false);
        const uint32_t n = this->opt_vec_.size();
        apply_construct_kernel<functor_t, typename base_t::dstatus_t, gather_t, info_t, true><<<kBlocks, 256>>>(
            this->opt_vec_.data(), n, this->d_status_, this->ipt_vec_.data(), counts_.data());
        archi::launch_check("apply_construct");
        std::vector<uint32_t> h(kBlocks);
        archi::memcpy<arch_t::cpu, arch_t::hip, uint32_t>(h.data(), counts_.data(), kBlocks);
        uint32_t c = 0;
        for (uint32_t v : h) c += v;
        fr.output_size() = c;
    }

private:
    static constexpr uint32_t kBlocks = 1024;
    Buffer<arch_t::hip, uint32_t, uint32_t> counts_;
};

// All-active frontier (blas_gc.h:353-435): x from every vertex on the first iteration, then apply on
// every vertex without the default_result test.
template <typename functor_t, typename comp_t, typename frontier_t>
class BlasGcEngine<functor_t, comp_t, frontier_t, std::enable_if_t<frontier_t::kind == FrontierKind::ALL_ACTIVE>>
    : public BlasGcEngineBase<functor_t, comp_t, frontier_t> {
public:
    using base_t = BlasGcEngineBase<functor_t, comp_t, frontier_t>;
    using typename base_t::gather_t;
    using typename base_t::info_t;

    BlasGcEngine(comp_t& comp, frontier_t& frontier) : base_t(comp, frontier) {}

    void Forward() override {
        const uint32_t n = this->opt_vec_.size();
        if (first_) {
            first_ = false;
            construct_all_kernel<functor_t><<<archi::grid_blocks(n), 256>>>(n, this->d_status_, this->ipt_vec_.data());
            archi::launch_check("construct");
        }
        this->spmv_(false);
        apply_construct_kernel<functor_t, typename base_t::dstatus_t, gather_t, info_t, false><<<1024, 256>>>(
            this->opt_vec_.data(), n, this->d_status_, this->ipt_vec_.data(), nullptr);
        archi::launch_check("apply_construct");
        archi::LaunchSync<base_t::arch>();
    }

private:
    bool first_ = true;
};

}  // namespace engine

constexpr graph_view_t AdvanceViews = graph_view_t::csr | graph_view_t::normal;
constexpr graph_view_t BlasViews = graph_view_t::csr | graph_view_t::transpose;

template <typename vertex_t, typename edge_t, typename weight_t, typename dstatus_t, typename vprop_t = empty_t>
struct AdvanceFunctor {
    template <typename functor_t, typename comp_t, typename frontier_t>
    using engine_type = engine::AdvanceGC<functor_t, comp_t, frontier_t>;

    __device__ __forceinline__ static bool advance(const vertex_t&, const vertex_t&, const edge_t&, const weight_t&,
                                                   dstatus_t&) {
        return false;
    }
};

template <typename vertex_t, typename weight_t, typename dstatus_t, typename info_t, typename gather_t>
struct BlasFunctor {
    template <typename functor_t, typename comp_t, typename frontier_t>
    using engine_type = engine::BlasGcEngine<functor_t, comp_t, frontier_t>;
    using info_type = info_t;
    using gather_type = gather_t;

    __host__ __device__ __forceinline__ static info_t default_info() { return info_t{0}; }
    __host__ __device__ __forceinline__ static gather_t default_result() { return gather_t{0}; }
    __device__ __forceinline__ static info_t construct_each(const vertex_t&, const dstatus_t&) { return info_t{}; }
    __device__ __forceinline__ static gather_t gather_combine(const weight_t&, const info_t&) { return gather_t{}; }
    __device__ __forceinline__ static gather_t gather_reduce(const gather_t& a, const gather_t&) { return a; }
    __device__ __forceinline__ static bool apply_each(const vertex_t&, const gather_t&, dstatus_t&) { return false; }
};

// The engine loop (base/run.h:8-39): Init; while neither the component nor the frontier has converged:
// BeforeEngine (component, then frontier), Forward, AfterEngine (frontier, then component); Final.
template <typename functor_t, typename comp_t, typename frontier_t>
void Run(comp_t& comp, frontier_t& frontier) {
    using engine_t = typename functor_t::template engine_type<functor_t, comp_t, frontier_t>;
    engine_t engine(comp, frontier);
    comp.Init();
    while (!comp.IsConvergent() && !frontier.IsConvergent()) {
        comp.BeforeEngine();
        frontier.BeforeEngine();
        engine.Forward();
        frontier.AfterEngine();
        comp.AfterEngine();
    }
    comp.Final();
}

}  // namespace graph_one
