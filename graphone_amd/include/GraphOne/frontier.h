// Frontiers of the graph_one C++ API.
// Reference: include/GraphOne/frontier/base.h:7-35 (FrontierKind, BaseFrontier), sparse_dbl.h:15-181
// (SpDblFrontier), dense_dbl.h:14-67 (DenDblFrontier), all_active.h:12-49 (AllActiveFrontier); the
// README's DblBufFrontier (README.md:102) is the adaptive double buffer: a sparse frontier whose advance
// steps switch to a pull over the in-edges when the frontier's out-edges pass a share of the graph's.
//
// SpDblFrontier keeps the reference's interface (double-buffered vertex lists, input/output sizes on the
// host, swap after every engine step, convergence when the input is empty). Each list also carries, when
// the engine that wrote it knows it, the sum of its vertices' out-degrees (`edges`, valid flag beside it):
// the advance engine sizes its output with it, so an iteration reads back only the output's size and
// edge count, together.
#pragma once

#include <stdlib.h>

#include <initializer_list>
#include <string>
#include <vector>

#include "GraphOne/buffer.h"

namespace graph_one {

enum FrontierKind : uint8_t {
    SPARSE_BASED = 1 << 1,
    DENSE_BASED = 1 << 2,
    ALL_ACTIVE = (1 << 1) | (1 << 2)
};

template <FrontierKind kind_, bool has_output_>
struct BaseFrontier {
    constexpr static FrontierKind kind = kind_;
    constexpr static bool has_output = has_output_;
    virtual ~BaseFrontier() = default;
    virtual void BeforeEngine() {}
    virtual void AfterEngine() {}
    virtual bool IsConvergent() = 0;
};

template <arch_t arch, typename vertex_t, typename index_t = vertex_t>
class SpDblFrontier : public BaseFrontier<SPARSE_BASED, true> {
public:
    using vertex_type = vertex_t;
    using index_type = index_t;
    constexpr static bool adaptive = false;  // advance steps always push

    // Device view (reference: sparse_dbl.h:28-64): get(i) reads the input list, set(i, v) writes the output.
    struct arch_ref_t {
        constexpr static bool has_output = true;
        using vertex_type = vertex_t;
        using index_type = index_t;
        __host__ __device__ __forceinline__ vertex_t get(index_t i) const { return frontiers[input_selector][i]; }
        __host__ __device__ __forceinline__ void set(index_t i, vertex_t v) { frontiers[input_selector ^ 1][i] = v; }
        __host__ __device__ __forceinline__ const vertex_t* input() const { return frontiers[input_selector]; }
        __host__ __device__ __forceinline__ vertex_t* output() { return frontiers[input_selector ^ 1]; }
        int input_selector;
        index_t input_size;
        index_t output_size;
        vertex_t* frontiers[2];
    };

    SpDblFrontier(index_t size, std::initializer_list<vertex_t> vids) : sel_(0), in_size_(0), out_size_(0) {
        const index_t sz = static_cast<index_t>(vids.size());
        const index_t cap = std::max(size, sz);
        reserve_input(cap);
        reserve_output(cap);
        reset_input(sz);
        lists_[sel_].copy_from(std::vector<vertex_t>(vids));
    }
    virtual ~SpDblFrontier() = default;

    void reserve_input(index_t size) { grow(sel_, size, false); }
    void reserve_output(index_t size) { grow(sel_ ^ 1, size, false); }
    void reset_input(index_t size) {
        grow(sel_, size, true);
        in_size_ = size;
        edges_valid_[sel_] = false;
    }
    void reset_output(index_t size) {
        grow(sel_ ^ 1, size, true);
        out_size_ = size;
        edges_valid_[sel_ ^ 1] = false;
    }
    index_t input_size() const { return in_size_; }
    index_t output_size() const { return out_size_; }
    void swap_inout() {
        sel_ ^= 1;
        std::swap(in_size_, out_size_);
    }
    const Buffer<arch, vertex_t, index_t>& input() const { return lists_[sel_]; }
    Buffer<arch, vertex_t, index_t>& input_mut() { return lists_[sel_]; }
    Buffer<arch, vertex_t, index_t>& output() { return lists_[sel_ ^ 1]; }

    // Out-edge sums of the lists (the advance engine's output sizing).
    bool input_edges_valid() const { return edges_valid_[sel_]; }
    uint32_t input_edges() const { return edges_[sel_]; }
    void set_output_edges(uint32_t e) {
        edges_[sel_ ^ 1] = e;
        edges_valid_[sel_ ^ 1] = true;
    }
    // A list of at least `size` entries that outlives the swap (the output, grown without changing its size).
    void reserve_output_capacity(index_t size) { grow(sel_ ^ 1, size, true); }
    index_t output_capacity() const { return lists_[sel_ ^ 1].size(); }

    arch_ref_t ToArch() {
        arch_ref_t r;
        r.input_selector = sel_;
        r.input_size = in_size_;
        r.output_size = out_size_;
        r.frontiers[0] = lists_[0].data();
        r.frontiers[1] = lists_[1].data();
        return r;
    }

    void AfterEngine() override { swap_inout(); }
    bool IsConvergent() override { return in_size_ == 0; }

    std::string ToString() const {
        std::vector<vertex_t> h(in_size_);
        archi::memcpy<arch_t::cpu, arch, vertex_t>(h.data(), lists_[sel_].data(), in_size_);
        std::string s = "SpDblFrontier{ in_size_:" + utils::NumToString(in_size_) +
                        ", out_size_:" + utils::NumToString(out_size_) + ", input:[";
        for (index_t i = 0; i < in_size_; ++i) s += (i ? "," : "") + utils::NumToString(h[i]);
        return s + "] }";
    }

protected:
    // Grows buffer b to at least `size` (twice the request when `doubling`, as reset_* does in the
    // reference), keeping its contents.
    void grow(int b, index_t size, bool doubling) {
        if (size <= lists_[b].size()) return;
        const index_t cap = doubling ? size * 2 : size;
        Buffer<arch, vertex_t, index_t> nl(cap);
        if (lists_[b].size()) nl.copy_from(lists_[b]);
        lists_[b] = std::move(nl);
    }

    int sel_;
    index_t in_size_;
    index_t out_size_;
    Buffer<arch, vertex_t, index_t> lists_[2];
    bool edges_valid_[2] = {false, false};
    uint32_t edges_[2] = {0, 0};
};

// The frontier name of the reference README (README.md:102): the sparse double buffer, adaptive. The
// advance engine runs an iteration as a pull over the in-edges (every vertex with an in-edge from the
// frontier calls the functor's advance for those edges and joins the output once) when the frontier's
// out-edges reach pull_fraction() of the graph's edges, and as a push otherwise. The in-edge CSR is the
// graph's transpose view when it has one, otherwise built on the device at the first pull.
template <arch_t arch, typename vertex_t, typename index_t = vertex_t>
class DblBufFrontier : public SpDblFrontier<arch, vertex_t, index_t> {
public:
    using SpDblFrontier<arch, vertex_t, index_t>::SpDblFrontier;
    constexpr static bool adaptive = true;
    // G1_DBL_PULL (a fraction in (0, 1]; 0 disables pulling), default 0.25.
    static double pull_fraction() {
        static const double f = [] {
            const char* e = getenv("G1_DBL_PULL");
            return e ? atof(e) : 0.25;
        }();
        return f;
    }
};

template <arch_t arch, typename vertex_t, typename index_t = vertex_t>
class DenDblFrontier : public BaseFrontier<DENSE_BASED, true> {
public:
    using vertex_type = vertex_t;
    using index_type = index_t;

    DenDblFrontier(index_t, std::initializer_list<vertex_t> vids)
        : in_size_(static_cast<index_t>(vids.size())), out_size_(0),
          init_buf_(std::vector<vertex_t>(vids)) {}

    index_t input_size() const { return in_size_; }
    index_t& output_size() { return out_size_; }
    void swap_inout() { std::swap(in_size_, out_size_); }
    Buffer<arch, vertex_t, index_t>& init_buf() { return init_buf_; }

    void AfterEngine() override { swap_inout(); }
    bool IsConvergent() override { return in_size_ == 0; }
    std::string ToString() const {
        return "DenDblFrontier{ in_size_:" + utils::NumToString(in_size_) + ", out_size_:" +
               utils::NumToString(out_size_) + " }";
    }

protected:
    index_t in_size_;
    index_t out_size_;
    Buffer<arch, vertex_t, index_t> init_buf_;
};

template <arch_t arch, typename vertex_t, typename index_t = vertex_t>
class AllActiveFrontier : public BaseFrontier<ALL_ACTIVE, false> {
public:
    using vertex_type = vertex_t;
    using index_type = index_t;

    AllActiveFrontier(index_t size, std::initializer_list<vertex_t>) : size_(size) {}
    index_t input_size() const { return size_; }
    index_t output_size() const { return size_; }
    index_t size() const { return size_; }
    bool IsConvergent() override { return false; }
    std::string ToString() const { return "AllActiveFrontier{ size_: " + utils::NumToString(size_) + " }"; }

protected:
    index_t size_;
};

}  // namespace graph_one
