// Semiring SpMV dispatch of the graph_one C++ API (namespace graph_one::blas).
// Reference: include/GraphOne/archi/blas/spmv/def.hpp:8-73 (SpmvCsrParams, MakeCsrSpMV, SpmvDispatcher),
// cpu/naive.hpp:5-20 + cpu/cpu.hpp:8-32 (the host fold), archi/blas/spmv/cuda/cuda.cuh:10-114 (CSR-vector and merge-based
// dispatchers). The functor contract is the same: static initialize(), combine(nonzero, x), reduce(a, b).
// The device kinds are SpmvHipMergeBased and SpmvHipCsrVector; their kernels are the gfx950 templates of
// g1/spmv_merge.hpp instantiated with the caller's functor, and the once-per-matrix merge-path partition
// runs through the C-ABI (g1_spmv_merge_partition).
#pragma once

#include "GraphOne/archi.h"
#include "g1/spmv_merge.hpp"

namespace graph_one::blas {

template <typename index_t, typename offset_t, typename mat_value_t, typename vec_x_value_t, typename vec_y_value_t>
struct SpmvCsrParams {
    index_t n_rows;
    index_t n_cols;
    offset_t nnz;
    const offset_t* row_offsets;
    const index_t* col_indices;
    const mat_value_t* values;
    const vec_x_value_t* vector_x;
    vec_y_value_t* vector_y;
};

struct SpmvCpuNaive {
    constexpr static arch_t arch_type = arch_t::cpu;
};
struct SpmvHipMergeBased {
    constexpr static arch_t arch_type = arch_t::hip;
};
struct SpmvHipCsrVector {
    constexpr static arch_t arch_type = arch_t::hip;
};

// The host fold, row by row, columns in storage order.
template <typename functor_t, typename index_t, typename offset_t, typename mat_value_t, typename vecx_value_t,
          typename vecy_value_t>
void SpMV_cpu_naive(index_t n_rows, const offset_t* ap, const index_t* aj, const mat_value_t* ax,
                    const vecx_value_t* x, vecy_value_t* y) {
    for (index_t r = 0; r < n_rows; ++r) {
        vecy_value_t sum = functor_t::initialize();
        for (offset_t k = ap[r]; k < ap[r + 1]; ++k) sum = functor_t::reduce(sum, functor_t::combine(ax[k], x[aj[k]]));
        y[r] = sum;
    }
}

template <typename kind, typename functor_t, typename index_t, typename offset_t, typename mat_value_t,
          typename vec_x_value_t, typename vec_y_value_t>
struct SpmvDispatcher;

template <typename functor_t, typename index_t, typename offset_t, typename mat_value_t, typename vec_x_value_t,
          typename vec_y_value_t>
struct SpmvDispatcher<SpmvCpuNaive, functor_t, index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t> {
    using params_t = SpmvCsrParams<index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t>;
    explicit SpmvDispatcher(const params_t& p) : params(p) {}
    void operator()(bool = true) {
        SpMV_cpu_naive<functor_t>(params.n_rows, params.row_offsets, params.col_indices, params.values,
                                  params.vector_x, params.vector_y);
    }
    params_t params;
};

template <typename functor_t, typename index_t, typename offset_t, typename mat_value_t, typename vec_x_value_t,
          typename vec_y_value_t>
struct SpmvDispatcher<SpmvHipMergeBased, functor_t, index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t> {
    static_assert(sizeof(index_t) == 4 && sizeof(offset_t) == 4, "the gfx950 SpMV kernels take 32-bit indices");
    using params_t = SpmvCsrParams<index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t>;
    using semiring_t = g1::FunctorSemiring<functor_t, mat_value_t, vec_x_value_t, vec_y_value_t>;

    explicit SpmvDispatcher(const params_t& p) : params(p) {
        ws_bytes = g1_spmv_merge_workspace_bytes(p.n_rows, p.nnz);
        archi::check(g1_malloc(&ws, ws_bytes), "spmv workspace");
        archi::check(g1_spmv_merge_partition(p.n_rows, p.nnz, reinterpret_cast<const uint32_t*>(p.row_offsets), ws,
                                             ws_bytes, nullptr),
                     "merge partition");
    }
    SpmvDispatcher(const SpmvDispatcher&) = delete;
    SpmvDispatcher(SpmvDispatcher&& o) noexcept : params(o.params), ws(o.ws), ws_bytes(o.ws_bytes) { o.ws = nullptr; }
    ~SpmvDispatcher() {
        if (ws) g1_free(ws);
    }

    void operator()(bool sync = true) {
        if (params.n_rows == 0) return;
        const uint32_t n = params.n_rows, nnz = static_cast<uint32_t>(params.nnz);
        const uint32_t tiles = g1::merge_num_tiles(n, nnz);
        auto* coords = g1::merge_coords(ws);
        auto* recs = g1::merge_records<semiring_t>(ws, n, nnz);
        const auto* ro = reinterpret_cast<const uint32_t*>(params.row_offsets);
        const auto* ci = reinterpret_cast<const uint32_t*>(params.col_indices);
        g1::spmv_merge_tile_kernel<semiring_t, true><<<tiles, g1::kMergeThreads>>>(
            n, nnz, ro, ci, params.values, params.vector_x, params.vector_y, coords, recs);
        archi::launch_check("spmv merge tile");
        if (tiles > 1) {
            g1::spmv_merge_fixup_kernel<semiring_t><<<(tiles + 255) / 256, 256>>>(tiles, coords, recs, params.vector_y);
            archi::launch_check("spmv merge fixup");
        }
        if (sync) archi::LaunchSync<arch_t::hip>();
    }

    params_t params;
    void* ws = nullptr;
    size_t ws_bytes = 0;
};

template <typename functor_t, typename index_t, typename offset_t, typename mat_value_t, typename vec_x_value_t,
          typename vec_y_value_t>
struct SpmvDispatcher<SpmvHipCsrVector, functor_t, index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t> {
    using params_t = SpmvCsrParams<index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t>;
    using semiring_t = g1::FunctorSemiring<functor_t, mat_value_t, vec_x_value_t, vec_y_value_t>;

    explicit SpmvDispatcher(const params_t& p) : params(p), lanes(g1_spmv_csrvec_lanes(p.n_rows, p.nnz)) {}

    template <int L>
    void launch() {
        const uint64_t threads = static_cast<uint64_t>(params.n_rows) * L;
        g1::spmv_csrvec_kernel<semiring_t, true, L><<<static_cast<unsigned>((threads + 255) / 256), 256>>>(
            params.n_rows, reinterpret_cast<const uint32_t*>(params.row_offsets),
            reinterpret_cast<const uint32_t*>(params.col_indices), params.values, params.vector_x, params.vector_y);
    }
    void operator()(bool sync = true) {
        if (params.n_rows == 0) return;
        switch (lanes) {
            case 1: launch<1>(); break;
            case 2: launch<2>(); break;
            case 4: launch<4>(); break;
            case 8: launch<8>(); break;
            case 16: launch<16>(); break;
            case 32: launch<32>(); break;
            default: launch<64>(); break;
        }
        archi::launch_check("spmv csr-vector");
        if (sync) archi::LaunchSync<arch_t::hip>();
    }
    params_t params;
    uint32_t lanes;
};

template <arch_t arch, typename kind, typename functor_t, typename index_t, typename offset_t, typename mat_value_t,
          typename vec_x_value_t, typename vec_y_value_t>
SpmvDispatcher<kind, functor_t, index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t>
MakeCsrSpMV(index_t n_rows, index_t n_cols, offset_t nnz, const offset_t* row_offsets, const index_t* col_indices,
            const mat_value_t* values, const vec_x_value_t* x, vec_y_value_t* y) {
    static_assert(arch == kind::arch_type, "SpMV kind does not run on this arch");
    SpmvCsrParams<index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t> p{n_rows, n_cols, nnz, row_offsets,
                                                                                 col_indices, values, x, y};
    return SpmvDispatcher<kind, functor_t, index_t, offset_t, mat_value_t, vec_x_value_t, vec_y_value_t>(p);
}

}  // namespace graph_one::blas
