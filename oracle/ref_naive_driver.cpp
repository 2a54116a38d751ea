// Driver that instantiates the reference's own CPU SpMV, SpMV_cpu_navie
// (include/GraphOne/archi/blas/spmv/cpu/naive.hpp:5-20), for the two semirings of the hot path and
// exports it with a C ABI. Compiled by oracle/Makefile against the reference tree into
// oracle/_ref/libref_naive.so; used only by the tests, to pin the oracle's restated fold.
//
// The functor bodies are the ones the reference apps give their BLAS functors: SSSP
// (apps/sssp_blas.cu:55-73, dist_t instantiated as uint32 here) and PageRank (apps/pagerank_blas.cu:82-111).
#include <stdint.h>

#include <algorithm>
#include <limits>

#include "GraphOne/archi/blas/spmv/cpu/naive.hpp"

namespace {

struct MinPlusU32Functor {
    static uint32_t initialize() { return std::numeric_limits<uint32_t>::max(); }
    static uint32_t combine(const uint32_t& w, const uint32_t& info) {
        return (info == std::numeric_limits<uint32_t>::max()) ? info : w + info;
    }
    static uint32_t reduce(const uint32_t& a, const uint32_t& b) { return std::min(a, b); }
};

struct PlusTimesF32Functor {
    static float initialize() { return 0.0f; }
    static float combine(const float& w, const float& info) { return info * w; }
    static float reduce(const float& a, const float& b) { return a + b; }
};

}  // namespace

extern "C" {

void ref_spmv_minplus_u32(uint32_t n_rows, const uint32_t* ap, const uint32_t* aj, const uint32_t* ax,
                          const uint32_t* x, uint32_t* y) {
    graph_one::blas::SpMV_cpu_navie<MinPlusU32Functor>(n_rows, ap, aj, ax, x, y);
}

void ref_spmv_plustimes_f32(uint32_t n_rows, const uint32_t* ap, const uint32_t* aj, const float* ax,
                            const float* x, float* y) {
    graph_one::blas::SpMV_cpu_navie<PlusTimesF32Functor>(n_rows, ap, aj, ax, x, y);
}

}  // extern "C"
