// Host-side launchers of the SpMV kernels, shared by the C-ABI exports and the engines.
#pragma once

#include <stdlib.h>

#include <hip/hip_runtime.h>

#include "g1/spmv_merge.hpp"
#include "internal.hpp"

namespace g1 {

inline uint32_t csrvec_lanes(uint32_t n_rows, uint32_t nnz) {
    // Half the average row length, rounded up to a power of two, at most one wave per row: on skewed graphs
    // most rows are far shorter than the average, and rows past 16 lanes' worth run one workgroup each.
    // Measured on RMAT22 (CSR-vector BFS / SSSP) and RMAT24 (PageRank per iteration), average 15 nonzeros:
    // 16 lanes 5.1 / 12.5 / 2.34 ms, 8 lanes 3.8 / 9.1 / 2.01 ms, 4 lanes 3.6 / 9.2 / 2.06 ms. G1_CSRVEC_LANES
    // (a power of two up to 64) overrides it.
    static const uint32_t forced = [] {
        const char* e = getenv("G1_CSRVEC_LANES");
        const int v = e ? atoi(e) : 0;
        return (v > 0 && v <= 64 && (v & (v - 1)) == 0) ? static_cast<uint32_t>(v) : 0u;
    }();
    if (forced) return forced;
    const uint32_t avg = n_rows ? nnz / n_rows : 0;
    const uint32_t half = (avg + 1) / 2;
    uint32_t lanes = 1;
    while (lanes < half && lanes < 64) lanes <<= 1;
    return lanes;
}

inline g1_status launch_merge_partition(uint32_t n_rows, uint32_t nnz, const uint32_t* row_offsets, void* ws,
                                        hipStream_t stream) {
    if (n_rows == 0) return G1_OK;
    const uint32_t tiles = merge_num_tiles(n_rows, nnz);
    const uint32_t blocks = (tiles + 1 + 255) / 256;
    merge_partition_kernel<<<dim3(blocks), dim3(256), 0, stream>>>(row_offsets, n_rows, nnz, tiles,
                                                                   merge_coords(ws));
    G1_LAUNCHED();
    return G1_OK;
}

template <typename S>
g1_status launch_spmv_merge(uint32_t n_rows, uint32_t nnz, const uint32_t* row_offsets, const uint32_t* col,
                            const typename S::value_t* vals, const typename S::x_t* x, typename S::y_t* y,
                            void* ws, hipStream_t stream) {
    if (n_rows == 0) return G1_OK;
    const uint32_t tiles = merge_num_tiles(n_rows, nnz);
    MergeCoord* coords = merge_coords(ws);
    MergeTileRec<S>* recs = merge_records<S>(ws, n_rows, nnz);
    if (vals) {
        spmv_merge_tile_kernel<S, true><<<dim3(tiles), dim3(kMergeThreads), 0, stream>>>(
            n_rows, nnz, row_offsets, col, vals, x, y, coords, recs);
    } else {
        spmv_merge_tile_kernel<S, false><<<dim3(tiles), dim3(kMergeThreads), 0, stream>>>(
            n_rows, nnz, row_offsets, col, vals, x, y, coords, recs);
    }
    G1_LAUNCHED();
    if (tiles > 1) {
        spmv_merge_fixup_kernel<S><<<dim3((tiles + 255) / 256), dim3(256), 0, stream>>>(tiles, coords, recs, y);
        G1_LAUNCHED();
    }
    return G1_OK;
}

template <typename S, bool V, int L>
inline void csrvec_launch_one(uint32_t n_rows, const uint32_t* row_offsets, const uint32_t* col,
                              const typename S::value_t* vals, const typename S::x_t* x, typename S::y_t* y,
                              uint32_t long_min, const uint32_t* long_rows, uint32_t n_long, hipStream_t stream) {
    const uint64_t threads = static_cast<uint64_t>(n_rows) * L;
    const uint32_t blocks = static_cast<uint32_t>((threads + 255) / 256) + n_long;
    spmv_csrvec_kernel<S, V, L><<<dim3(blocks), dim3(256), 0, stream>>>(n_rows, row_offsets, col, vals, x, y,
                                                                        long_min, long_rows, n_long);
}

template <typename S, bool V>
inline void csrvec_dispatch(uint32_t lanes, uint32_t n_rows, const uint32_t* row_offsets, const uint32_t* col,
                            const typename S::value_t* vals, const typename S::x_t* x, typename S::y_t* y,
                            uint32_t long_min, const uint32_t* long_rows, uint32_t n_long, hipStream_t stream) {
    switch (lanes) {
        case 1: csrvec_launch_one<S, V, 1>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
        case 2: csrvec_launch_one<S, V, 2>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
        case 4: csrvec_launch_one<S, V, 4>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
        case 8: csrvec_launch_one<S, V, 8>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
        case 16: csrvec_launch_one<S, V, 16>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
        case 32: csrvec_launch_one<S, V, 32>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
        default: csrvec_launch_one<S, V, 64>(n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream); break;
    }
}

// Rows longer than this many nonzeros per lane go to the long-row workgroups (when the caller has their
// list): 16, or G1_CSRVEC_LONG (1..1024). At 8 lanes, CSR-vector SSSP on RMAT22 takes 9.0 ms with 16 (9.1
// with 8, 12.5 with 32, 18.5 with 64); RMAT24 PageRank 2.02 ms per iteration (2.18, 2.00, 1.99).
inline uint32_t csrvec_long_per_lane() {
    static const uint32_t v = [] {
        const char* e = getenv("G1_CSRVEC_LONG");
        const int x = e ? atoi(e) : 0;
        return (x > 0 && x <= 1024) ? static_cast<uint32_t>(x) : 16u;
    }();
    return v;
}

// CSR-vector SpMV; with a list of the matrix's rows longer than long_min (ensure_long_rows), those rows run
// one workgroup each at the start of the same launch.
template <typename S>
g1_status launch_spmv_csrvec(uint32_t n_rows, const uint32_t* row_offsets, const uint32_t* col,
                             const typename S::value_t* vals, const typename S::x_t* x, typename S::y_t* y,
                             uint32_t lanes, hipStream_t stream, const uint32_t* long_rows = nullptr,
                             uint32_t n_long = 0, uint32_t long_min = kInf) {
    if (n_rows == 0) return G1_OK;
    if (!long_rows || n_long == 0) {
        long_min = kInf;
        n_long = 0;
    }
    if (vals) csrvec_dispatch<S, true>(lanes, n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream);
    else csrvec_dispatch<S, false>(lanes, n_rows, row_offsets, col, vals, x, y, long_min, long_rows, n_long, stream);
    G1_LAUNCHED();
    return G1_OK;
}

}  // namespace g1
