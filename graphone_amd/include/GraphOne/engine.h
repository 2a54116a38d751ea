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

#include <type_traits>
#include <vector>

#include "GraphOne/frontier.h"
#include "GraphOne/graph.h"
#include "GraphOne/spmv.h"
#include "g1/push.hpp"

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

// qoff[i] = out-degree of list[i] (0 for an invalid vertex); one pass, then an exclusive scan in place.
template <vstart_t vstart, typename vertex_t>
__global__ void list_degrees_kernel(const vertex_t* __restrict__ list, uint32_t n, const uint32_t* __restrict__ row_off,
                                    uint32_t* __restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const vertex_t v = list[i];
        out[i] = utils::is_vertex_valid<vstart>(v) ? row_off[v + 1] - row_off[v] : 0u;
    }
}

static __global__ void nonzero_flags_kernel(const uint32_t* __restrict__ deg, uint32_t n, uint32_t* __restrict__ flags) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) flags[i] = deg[i] ? 1u : 0u;
}

template <typename vertex_t>
__global__ void compact_nonzero_kernel(const vertex_t* __restrict__ list, const uint32_t* __restrict__ deg,
                                       const uint32_t* __restrict__ pos, uint32_t n, uint32_t* __restrict__ out,
                                       uint32_t* __restrict__ out_deg) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (deg[i]) {
            out[pos[i]] = static_cast<uint32_t>(list[i]);
            out_deg[pos[i]] = deg[i];
        }
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

// Exclusive scan of data[0..n) in place; returns the total. scratch holds >= n/kScanBlock + 2 entries.
inline uint32_t exclusive_scan_u32(uint32_t* data, uint32_t n, uint32_t* scratch) {
    if (n == 0) return 0;
    const uint32_t blocks = (n + kScanBlock - 1) / kScanBlock;
    scan_block_sums_kernel<<<blocks, 256>>>(data, n, scratch);
    scan_sums_kernel<<<1, 256>>>(scratch, blocks);
    scan_apply_kernel<<<blocks, 256>>>(data, n, scratch);
    archi::launch_check("exclusive scan");
    uint32_t total = 0;
    archi::memcpy<arch_t::cpu, arch_t::hip, uint32_t>(&total, scratch + blocks, 1);
    return total;
}

// Gathers the kSegments queue segments into one contiguous list with global edge offsets.
template <typename vertex_t>
__global__ void gather_segments_kernel(const uint32_t* __restrict__ q, const uint32_t* __restrict__ qoff,
                                       const unsigned long long* __restrict__ cnt, uint32_t cap,
                                       vertex_t* __restrict__ out, uint32_t* __restrict__ out_off) {
    __shared__ uint32_t vb[g1::kSegments + 1], eb[g1::kSegments + 1];
    if (threadIdx.x < 64) g1::load_segment_bases(cnt, vb, eb);
    __syncthreads();
    const uint32_t nf = vb[g1::kSegments];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += stride) {
        const uint32_t s = g1::search_last_le(vb, g1::kSegments, i);
        const size_t at = static_cast<size_t>(s) * cap + (i - vb[s]);
        out[i] = static_cast<vertex_t>(q[at]);
        out_off[i] = eb[s] + qoff[at];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out_off[nf] = eb[g1::kSegments];
}

// Advance policy around a user AdvanceFunctor: the source value is the source id itself; the functor's
// advance() decides the edge and its filter() (when present) drops the vertex.
template <typename F, typename GRef, typename D, bool kHasFilter>
struct FunctorAdvance {
    using src_t = uint32_t;
    static constexpr bool kWeighted = false;
    static constexpr bool kQueueLeaves = true;  // the output frontier lists every kept vertex
    static constexpr bool kEnqueue = true;
    GRef g;
    D d;
    __device__ __forceinline__ uint32_t source_value(uint32_t u) const { return u; }
    template <int N>
    __device__ __forceinline__ void relax_many(const uint32_t (&)[N], const uint32_t (&u)[N], const uint32_t (&v)[N],
                                               const uint32_t (&)[N], const uint32_t (&e)[N], bool (&won)[N]) const {
        D dl = d;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            won[r] = false;
            if (v[r] == g1::kInf) continue;
            using vertex_t = typename GRef::vertex_type;
            using edge_t = typename GRef::edge_type;
            bool keep = F::advance(static_cast<vertex_t>(u[r]), static_cast<vertex_t>(v[r]), static_cast<edge_t>(e[r]),
                                   g.get_out_edge_weight(static_cast<edge_t>(e[r])), dl);
            if constexpr (kHasFilter) {
                if (keep) keep = !F::filter(static_cast<vertex_t>(v[r]), dl);
            }
            won[r] = keep;
        }
    }
};

template <typename P>
__global__ void __launch_bounds__(g1::kPushThreads)
functor_advance_kernel(const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ col, g1::PushQueues pq,
                       P pol) {
    __shared__ uint32_t s_vb[g1::kSegments + 1];
    __shared__ uint32_t s_eb[g1::kSegments + 1];
    if (!g1::push_prologue(pq, s_vb, s_eb)) return;
    g1::advance_tiles(row_off, col, nullptr, pq, pol, s_vb, s_eb);
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

// ------------------------------------------------------------------------------------------ AdvanceGC

template <typename functor_t, typename comp_t, typename frontier_t>
class AdvanceGC : public GcEngine<comp_t, frontier_t> {
public:
    using base_t = GcEngine<comp_t, frontier_t>;
    using graph_t = typename base_t::graph_type;
    using dstatus_t = typename base_t::dstatus_type;
    using vertex_t = typename graph_t::vertex_type;
    using edge_t = typename graph_t::edge_type;
    using index_t = typename frontier_t::index_type;
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

    AdvanceGC(comp_t& comp, frontier_t& frontier)
        : base_t(comp, frontier), scratch_(4096), cnt_(3 * g1::kSegments), stats_(2 * g1::kMaxPushBlocks), overflow_(4) {}

    void Forward() override {
        auto& fr = this->frontier_;
        const auto gref = this->graph_.ToArch();
        const uint32_t nin = fr.input_size();
        // Prepare: the input entries that own edges, in order, with their edge offsets.
        uint32_t k = 0, E = 0;
        if (nin > 0) {
            grow(deg_, nin + 1);
            grow(pos_, nin + 1);
            grow(list_, nin + 1);
            grow(off_, nin + 1);
            ensure_scratch(nin);
            list_degrees_kernel<graph_t::vstart_value, vertex_t><<<archi::grid_blocks(nin), 256>>>(
                fr.input().data(), nin, reinterpret_cast<const uint32_t*>(gref.row_offsets), deg_.data());
            nonzero_flags_kernel<<<archi::grid_blocks(nin), 256>>>(deg_.data(), nin, pos_.data());
            archi::launch_check("prepare");
            k = exclusive_scan_u32(pos_.data(), nin, scratch_.data());
            compact_nonzero_kernel<vertex_t><<<archi::grid_blocks(nin), 256>>>(fr.input().data(), deg_.data(), pos_.data(),
                                                                               nin, list_.data(), off_.data());
            archi::launch_check("prepare compact");
            E = exclusive_scan_u32(off_.data(), k, scratch_.data());
        }
        // Every edge emits at most one vertex: size the output segments for E edges.
        const uint64_t cap = g1::push_segment_capacity(~0ull, E > 0 ? E : 1);
        grow(segq_, static_cast<uint32_t>(cap * g1::kSegments));
        grow(segoff_, static_cast<uint32_t>(cap * g1::kSegments));
        // The prepared list is segment 0 of the input: counter (E << 32) | k, every other segment empty.
        std::vector<unsigned long long> c(3 * g1::kSegments, 0ull);
        c[0] = (static_cast<unsigned long long>(E) << 32) | k;
        archi::memcpy<arch_t::hip, arch_t::cpu, unsigned long long>(cnt_.data(), c.data(), c.size());
        if (E > 0) {
            g1::PushQueues pq;
            pq.q_in = list_.data();
            pq.qoff_in = off_.data();
            pq.cnt_in = cnt_.data();
            pq.q_out = segq_.data();
            pq.qoff_out = segoff_.data();
            pq.cnt_out = cnt_.data() + g1::kSegments;
            pq.cnt_clear = cnt_.data() + 2 * g1::kSegments;
            pq.stats = stats_.data();
            pq.overflow = overflow_.data();
            pq.cap = static_cast<uint32_t>(cap);  // only input segment 0 is used, so the stride is the output's
            const uint32_t blocks = std::min<uint32_t>(g1::kMaxPushBlocks, (E + g1::kPushTile - 1) / g1::kPushTile);
            FunctorAdvance<functor_t, decltype(gref), dstatus_t, HasFilter::value> pol{gref, this->d_status_};
            functor_advance_kernel<decltype(pol)><<<blocks, g1::kPushThreads>>>(
                reinterpret_cast<const uint32_t*>(gref.row_offsets), reinterpret_cast<const uint32_t*>(gref.col_indices),
                pq, pol);
            archi::launch_check("advance");
        }
        // Gather the output segments into the frontier's output list.
        std::vector<unsigned long long> oc(g1::kSegments);
        archi::memcpy<arch_t::cpu, arch_t::hip, unsigned long long>(oc.data(), cnt_.data() + g1::kSegments, g1::kSegments);
        uint64_t nout = 0;
        for (auto v : oc) nout += static_cast<uint32_t>(v);
        uint32_t of = 0;
        archi::memcpy<arch_t::cpu, arch_t::hip, uint32_t>(&of, overflow_.data(), 1);
        if (of) LOG_ERROR("advance: frontier segment overflow");
        fr.reset_output(static_cast<index_t>(nout));
        if (nout > 0) {
            grow(off_, static_cast<uint32_t>(nout) + 1);
            gather_segments_kernel<vertex_t><<<archi::grid_blocks(nout), 256>>>(
                segq_.data(), segoff_.data(), cnt_.data() + g1::kSegments, static_cast<uint32_t>(cap),
                fr.output().data(), off_.data());
            archi::launch_check("gather segments");
        }
        archi::LaunchSync<arch>();
    }

private:
    static void grow(Buffer<arch_t::hip, uint32_t, uint32_t>& b, uint32_t n) {
        if (b.size() < n) b.resize(n);
    }
    void ensure_scratch(uint32_t n) { grow(scratch_, n / kScanBlock + 4); }

    Buffer<arch_t::hip, uint32_t, uint32_t> deg_;
    Buffer<arch_t::hip, uint32_t, uint32_t> pos_;
    Buffer<arch_t::hip, uint32_t, uint32_t> list_;
    Buffer<arch_t::hip, uint32_t, uint32_t> off_;
    Buffer<arch_t::hip, uint32_t, uint32_t> scratch_;
    Buffer<arch_t::hip, unsigned long long, uint32_t> cnt_;
    Buffer<arch_t::hip, unsigned long long, uint32_t> stats_;
    Buffer<arch_t::hip, uint32_t, uint32_t> overflow_;
    Buffer<arch_t::hip, uint32_t, uint32_t> segq_;
    Buffer<arch_t::hip, uint32_t, uint32_t> segoff_;
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

    void Forward() override {
        auto& fr = this->frontier_;
        const uint32_t n = this->ipt_vec_.size();
        // Construct: x = default_info; x[F] = construct_each(F)
        archi::fill<base_t::arch>(this->ipt_vec_.begin(), this->ipt_vec_.end(), functor_t::default_info());
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
        archi::memcpy<arch_t::hip, arch_t::hip, uint32_t>(pos_.data(), flags_.data(), n);
        const uint32_t nout = exclusive_scan_u32(pos_.data(), n, scratch_.data());
        fr.reset_output(nout);
        if (nout) {
            compact_flags_kernel<vertex_t><<<archi::grid_blocks(n), 256>>>(pos_.data(), flags_.data(), n,
                                                                           fr.output().data());
            archi::launch_check("compact");
        }
        archi::LaunchSync<base_t::arch>();
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
