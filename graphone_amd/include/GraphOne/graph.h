// Graph containers of the graph_one C++ API: views, the resident Graph and its device view
// GraphArchRef, and graph::build.
// Reference: include/GraphOne/graph/graph_view.hpp:10-89, graph.h:13-140, graph_arch_ref.h:19-116,
// builder.h:9-46 (same template parameters, accessors and view rules: the "normal" view is the
// out-edge CSR, "transpose" the in-edge CSR; with both views the transpose is kept beside it).
#pragma once

#include <cstdint>
#include <type_traits>

#include "GraphOne/loader.h"

namespace graph_one {

enum class graph_view_t : uint8_t {
    normal = 1 << 0,
    transpose = 1 << 1,
    coo = 1 << 2,
    csr = 1 << 3
};

constexpr inline graph_view_t operator|(graph_view_t a, graph_view_t b) {
    return static_cast<graph_view_t>(static_cast<uint8_t>(a) | static_cast<uint8_t>(b));
}
constexpr inline bool has_view(graph_view_t target, graph_view_t v) {
    return (static_cast<uint8_t>(target) & static_cast<uint8_t>(v)) == static_cast<uint8_t>(v);
}
template <typename... Ts>
constexpr inline bool has_all_views(graph_view_t target, Ts... vs) {
    return (has_view(target, vs) && ...);
}
template <typename... Ts>
constexpr inline bool has_some_views(graph_view_t target, Ts... vs) {
    return (has_view(target, vs) || ...);
}

// Device view: plain pointers and sizes, passed by value into kernels.
template <graph_view_t views, arch_t arch, vstart_t v_start, typename vprop_t, typename vertex_t, typename edge_t,
          typename weight_t>
struct GraphArchRef {
    using vprop_type = vprop_t;
    using vertex_type = vertex_t;
    using edge_type = edge_t;
    using weight_type = weight_t;
    constexpr static arch_t arch_value = arch;
    constexpr static vstart_t vstart_value = v_start;
    constexpr static bool both_mat = has_all_views(views, graph_view_t::normal, graph_view_t::transpose);

    __host__ __device__ __forceinline__ vertex_t get_out_degree(vertex_t v) const {
        static_assert(has_view(views, graph_view_t::normal), "out-degree needs the normal view");
        return static_cast<vertex_t>(row_offsets[v + 1] - row_offsets[v]);
    }
    __host__ __device__ __forceinline__ vertex_t get_in_degree(vertex_t v) const {
        static_assert(has_view(views, graph_view_t::transpose), "in-degree needs the transpose view");
        if constexpr (both_mat) return static_cast<vertex_t>(trans_row_offsets[v + 1] - trans_row_offsets[v]);
        else return static_cast<vertex_t>(row_offsets[v + 1] - row_offsets[v]);
    }
    __host__ __device__ __forceinline__ edge_t get_starting_out_edge(vertex_t v) const {
        static_assert(has_view(views, graph_view_t::normal), "out-edges need the normal view");
        return row_offsets[v];
    }
    __host__ __device__ __forceinline__ vertex_t get_dst_vertex_by_out_edge(edge_t e) const { return col_indices[e]; }
    __host__ __device__ __forceinline__ weight_t get_out_edge_weight(edge_t e) const { return values[e]; }

    vertex_t num_vertices;
    edge_t num_edges;
    const edge_t* row_offsets;
    const vertex_t* col_indices;
    const weight_t* values;
    const edge_t* trans_row_offsets;
    const vertex_t* trans_col_indices;
    const weight_t* trans_values;
    vprop_t vprops{};
};

template <arch_t arch, graph_view_t views, vstart_t v_start, typename vprop_t, typename vertex_t, typename edge_t,
          typename weight_t>
class Graph {
public:
    using vprop_type = vprop_t;
    using vertex_type = vertex_t;
    using edge_type = edge_t;
    using weight_type = weight_t;
    constexpr static arch_t arch_value = arch;
    constexpr static vstart_t vstart_value = v_start;
    constexpr static graph_view_t views_value = views;
    constexpr static bool both_mat = has_all_views(views, graph_view_t::normal, graph_view_t::transpose);
    using graph_mat_t = CsrMat<arch, weight_t, vertex_t, edge_t, v_start>;
    using arch_ref_t = GraphArchRef<views, arch, v_start, vprop_t, vertex_t, edge_t, weight_t>;
    static_assert(has_some_views(views, graph_view_t::normal, graph_view_t::transpose), "a graph needs a view");
    static_assert(has_view(views, graph_view_t::csr), "only CSR graphs are supported");

    Graph(graph_mat_t&& mat, vprop_t&& vprops) : mat_(std::move(mat)), vprops_(std::move(vprops)) {}
    Graph(graph_mat_t&& mat, graph_mat_t&& tmat, vprop_t&& vprops)
        : mat_(std::move(mat)), tmat_(std::move(tmat)), vprops_(std::move(vprops)) {}

    vertex_t num_vertices() const { return mat_.n_rows; }
    edge_t num_edges() const { return static_cast<edge_t>(mat_.nnz); }
    const vprop_t& vprops() const { return vprops_; }
    const graph_mat_t& graph_mat() const { return mat_; }
    const graph_mat_t& graph_transpose_mat() const { return tmat_; }

    arch_ref_t ToArch() const {
        arch_ref_t r;
        r.num_vertices = mat_.n_rows;
        r.num_edges = static_cast<edge_t>(mat_.nnz);
        r.row_offsets = mat_.row_offsets.data();
        r.col_indices = mat_.col_indices.data();
        r.values = mat_.values.data();
        r.trans_row_offsets = both_mat ? tmat_.row_offsets.data() : nullptr;
        r.trans_col_indices = both_mat ? tmat_.col_indices.data() : nullptr;
        r.trans_values = both_mat ? tmat_.values.data() : nullptr;
        return r;
    }

    std::string ToString() const { return "Graph{ mat_:" + mat_.ToString() + " }"; }

protected:
    graph_mat_t mat_;
    graph_mat_t tmat_;
    vprop_t vprops_;
};

namespace graph {

// Builds the views requested from an edge cache: normal + transpose keeps both CSRs; a single view
// keeps one (for "transpose" only, the in-edge CSR becomes the graph's matrix, as builder.h:35-46).
template <arch_t arch, graph_view_t views, vstart_t v_start, typename weight_t, typename vertex_t, typename edge_t>
Graph<arch, views, v_start, empty_t, vertex_t, edge_t, weight_t> build(EdgeCache<v_start, weight_t, vertex_t, edge_t>& cache) {
    using G = Graph<arch, views, v_start, empty_t, vertex_t, edge_t, weight_t>;
    if constexpr (has_all_views(views, graph_view_t::normal, graph_view_t::transpose)) {
        return G(cache.template ToCsr<arch, false>(), cache.template ToCsr<arch, true>(), empty_t{});
    } else if constexpr (has_view(views, graph_view_t::normal)) {
        return G(cache.template ToCsr<arch, false>(), empty_t{});
    } else {
        return G(cache.template ToCsr<arch, true>(), empty_t{});
    }
}

}  // namespace graph
}  // namespace graph_one
