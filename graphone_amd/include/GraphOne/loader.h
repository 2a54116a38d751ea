// Edge-list loading and CSR construction of the graph_one C++ API.
//
// LoadEdgeOpts, EdgeUnit, EdgeCache and GraphLoader keep the reference's names and rules
// (include/GraphOne/loader/opts.hpp:7-34, edge_cache.h:15-105, graph_loader.h:20-305):
//   - text lines "src dst [weight]" separated by any of opts.line_sep, lines starting with
//     opts.comment_prefix skipped, the file name must end with opts.file_ext;
//   - a missing weight reads as 1;
//   - self loops dropped unless keep_self_loop; ids shifted down by one for FROM_1_TO_0 (edges that
//     touch vertex 0 are then dropped); ids renumbered in order of appearance when reorder is on;
//   - undirected input (the default) adds every reverse edge;
//   - unless keep_duplicate_edges, edges are sorted by (src, dst) and duplicates removed (the first
//     occurrence in sorted order survives);
//   - n = largest id + 1; ToCsr places each row's edges in cache order (counting sort), so the
//     transpose's rows are ordered by source.
// CsrMat keeps the reference's members (include/GraphOne/mat/csr.h:11-68).
#pragma once

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "GraphOne/buffer.h"

namespace graph_one {

struct LoadEdgeOpts {
    std::string file_ext = ".adj";
    std::string comment_prefix = "%";
    std::string line_sep = " \t";
    bool is_directed = false;
    bool keep_self_loop = false;
    bool keep_duplicate_edges = false;

    std::string ToString() const {
        return "LoadEdgeOpts{ file_ext: " + file_ext + ", comment_prefix: " + comment_prefix +
               ", is_directed: " + utils::ToString(is_directed) + ", keep_self_loop: " +
               utils::ToString(keep_self_loop) + ", keep_duplicate_edges: " + utils::ToString(keep_duplicate_edges) + " }";
    }
};

// The apps' common command-line options (include/GraphOne/utils/cli.hpp:18-32).
inline void add_common_args(CLI::App& app, std::string& input_graph, std::string& output_path, bool& reorder_vid,
                            LoadEdgeOpts& opts) {
    app.add_option("-i,--input_graph", input_graph, "input graph dataset file")->required();
    app.add_option("-o,--output", output_path, "output path to save result");
    app.add_flag("-r,--reorder", reorder_vid, "reorder vertex id in graph dataset to compress (default 'false')");
    app.add_option("--file_ext", opts.file_ext, "graph dataset file extension (default '.adj')");
    app.add_option("--comment_prefix", opts.comment_prefix, "graph dataset file comment line prefix (default '%')");
    app.add_option("--line_sep", opts.line_sep, "graph dataset file line data delimiter (default ' \\t')");
    app.add_flag("-d,--is_directed", opts.is_directed, "directed graph dataset (default 'false')");
    app.add_flag("--keep_self", opts.keep_self_loop, "keep self loop edges in graph (default 'false')");
    app.add_flag("--keep_dup", opts.keep_duplicate_edges, "keep duplicate edges in graph (default 'false')");
}

template <arch_t arch, typename value_t, typename index_t = uint32_t, typename offset_t = uint64_t,
          vstart_t v_start = vstart_t::FROM_0_TO_0>
struct CsrMat {
    constexpr static arch_t arch_value = arch;
    constexpr static vstart_t vstart_value = v_start;
    using value_type = value_t;
    using index_type = index_t;
    using offset_type = offset_t;

    CsrMat() = default;
    CsrMat(index_t rows, index_t cols, offset_t nz, Buffer<arch, offset_t, index_t>&& ro,
           Buffer<arch, index_t, offset_t>&& ci, Buffer<arch, value_t, offset_t>&& vals)
        : n_rows(rows), n_cols(cols), nnz(nz), row_offsets(std::move(ro)), col_indices(std::move(ci)),
          values(std::move(vals)) {}

    std::string ToString() const {
        return "CsrMat{ n_rows:" + utils::NumToString(n_rows) + ", n_cols:" + utils::NumToString(n_cols) +
               ", nnz:" + utils::NumToString(nnz) + " }";
    }

    index_t n_rows{0};
    index_t n_cols{0};
    offset_t nnz{0};
    Buffer<arch, offset_t, index_t> row_offsets;
    Buffer<arch, index_t, offset_t> col_indices;
    Buffer<arch, value_t, offset_t> values;
};

template <typename edata_t, typename index_t>
struct EdgeUnit {
    index_t src;
    index_t dst;
    edata_t edata;
    bool operator<(const EdgeUnit& r) const { return src < r.src || (src == r.src && dst < r.dst); }
    bool operator==(const EdgeUnit& r) const { return src == r.src && dst == r.dst; }
};

template <vstart_t v_start, typename edata_t, typename index_t, typename offset_t>
class EdgeCache : public std::vector<EdgeUnit<edata_t, index_t>> {
    using base_t = std::vector<EdgeUnit<edata_t, index_t>>;

public:
    using vertex_type = index_t;
    using edge_type = offset_t;
    using weight_type = edata_t;
    constexpr static vstart_t vstart_value = v_start;

    void push_back(const EdgeUnit<edata_t, index_t>& e) {
        base_t::push_back(e);
        max_vid_ = std::max(max_vid_, std::max(e.src, e.dst));
    }
    index_t max_vid() const { return max_vid_; }
    offset_t num_edges() const { return static_cast<offset_t>(base_t::size()); }
    index_t num_vertices() const { return base_t::empty() ? 0 : max_vid_ + 1; }

    template <arch_t arch, bool is_transpose = false>
    CsrMat<arch, edata_t, index_t, offset_t, v_start> ToCsr() const {
        const index_t n = num_vertices();
        const offset_t nnz = num_edges();
        std::vector<offset_t> ro(static_cast<size_t>(n) + 1, 0);
        std::vector<index_t> ci(nnz);
        std::vector<edata_t> vals(nnz);
        for (const auto& e : *this) ++ro[(is_transpose ? e.dst : e.src) + 1];
        for (index_t r = 0; r < n; ++r) ro[r + 1] += ro[r];
        std::vector<offset_t> cur(ro.begin(), ro.end() - 1);
        for (const auto& e : *this) {
            const index_t row = is_transpose ? e.dst : e.src;
            const offset_t at = cur[row]++;
            ci[at] = is_transpose ? e.src : e.dst;
            vals[at] = e.edata;
        }
        return CsrMat<arch, edata_t, index_t, offset_t, v_start>(
            n, n, nnz, Buffer<arch, offset_t, index_t>(ro), Buffer<arch, index_t, offset_t>(ci),
            Buffer<arch, edata_t, offset_t>(vals));
    }

protected:
    index_t max_vid_{0};
};

template <vstart_t v_start = vstart_t::FROM_0_TO_0, typename index_t = vid_t, typename offset_t = eid_t>
class GraphLoader {
public:
    explicit GraphLoader(bool reorder_vid = false)
        : reorder_vid_(reorder_vid), new_vid_(v_start == vstart_t::FROM_0_TO_0 ? 0 : 1) {}

    template <typename edata_t>
    EdgeCache<v_start, edata_t, index_t, offset_t> LoadEdgesFromTxt(const std::string& path,
                                                                    const LoadEdgeOpts& opts = LoadEdgeOpts()) {
        if (!utils::StrEndWith(path, opts.file_ext))
            LOG_ERROR("file extension does not match, it should \"", opts.file_ext, "\"");
        std::ifstream fin(path);
        if (!fin) LOG_ERROR("cannot open graph adj file: ", path);
        EdgeCache<v_start, edata_t, index_t, offset_t> cache;
        std::string line;
        std::vector<char> buf;
        while (std::getline(fin, line)) {
            line = utils::StrStrip(line);
            if (line.empty() || utils::StrStartWith(line, opts.comment_prefix)) continue;
            buf.assign(line.begin(), line.end());
            buf.push_back('\0');
            char* save = nullptr;
            char* tok = strtok_r(buf.data(), opts.line_sep.c_str(), &save);
            if (!tok) continue;
            EdgeUnit<edata_t, index_t> e;
            e.src = static_cast<index_t>(std::strtoul(tok, nullptr, 10));
            tok = strtok_r(nullptr, opts.line_sep.c_str(), &save);
            if (!tok) {
                LOG_WARNING("can not extract destination from (", line, ")");
                continue;
            }
            e.dst = static_cast<index_t>(std::strtoul(tok, nullptr, 10));
            if constexpr (std::is_arithmetic_v<edata_t>) {
                tok = strtok_r(nullptr, opts.line_sep.c_str(), &save);
                if (tok) {
                    if constexpr (std::is_integral_v<edata_t>) e.edata = static_cast<edata_t>(std::strtoul(tok, nullptr, 10));
                    else e.edata = static_cast<edata_t>(std::strtod(tok, nullptr));
                } else {
                    e.edata = edata_t(1);
                }
            } else {
                e.edata = edata_t{};
            }
            if (!before_edge(e, opts)) continue;
            cache.push_back(e);
            if (!opts.is_directed && e.src != e.dst) {
                std::swap(e.src, e.dst);
                cache.push_back(e);
            }
        }
        if (!opts.keep_duplicate_edges) {
            std::stable_sort(cache.begin(), cache.end());
            cache.erase(std::unique(cache.begin(), cache.end()), cache.end());
        }
        return cache;
    }

    // True if vid is a vertex of the loaded graph; with reordering, vid is replaced by its new id.
    bool ReorderedVid(index_t& vid) const {
        if (!reorder_vid_) return true;
        auto it = vid_map_.find(vid);
        if (it == vid_map_.end()) return false;
        vid = it->second;
        return true;
    }

    const std::unordered_map<index_t, index_t>& get_vid_map() const { return vid_map_; }

private:
    template <typename edata_t>
    bool before_edge(EdgeUnit<edata_t, index_t>& e, const LoadEdgeOpts& opts) {
        if (!opts.keep_self_loop && e.src == e.dst) return false;
        if constexpr (v_start == vstart_t::FROM_1_TO_0) {
            if (e.src == 0 || e.dst == 0) {
                LOG_WARNING("edge (", e.src, ", ", e.dst, ") names vertex 0 in a file numbered from 1");
                return false;
            }
            --e.src;
            --e.dst;
        }
        if (reorder_vid_) {
            e.src = reorder(e.src);
            e.dst = reorder(e.dst);
        }
        return true;
    }
    index_t reorder(index_t v) {
        auto it = vid_map_.find(v);
        if (it != vid_map_.end()) return it->second;
        vid_map_[v] = new_vid_;
        return new_vid_++;
    }

    bool reorder_vid_;
    index_t new_vid_;
    std::unordered_map<index_t, index_t> vid_map_;
};

}  // namespace graph_one
