// C-ABI exports of the semiring SpMV kernels (merge-path and CSR-vector) for the two fixed semirings
// of the hot path: min-plus over uint32 (SSSP/BFS) and compensated plus-times over float (PageRank).
#include <hip/hip_runtime.h>

#include <mutex>

#include "g1/spmv_merge.hpp"
#include "internal.hpp"
#include "spmv_launch.hpp"

namespace g1 {

template <typename S>
g1_status spmv_merge_exported(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, const uint32_t* row_offsets,
                              const uint32_t* col, const typename S::value_t* vals,
                              const typename S::x_t* x, typename S::y_t* y, void* ws, size_t ws_bytes,
                              g1_stream_t stream) {
    G1_CHECK(n_rows == 0 || (row_offsets && y), G1_ERR_INVALID_ARG, "spmv_merge: null row_offsets or y");
    G1_CHECK(nnz == 0 || (col && x), G1_ERR_INVALID_ARG, "spmv_merge: null col_indices or x");
    G1_CHECK(n_cols > 0 || nnz == 0, G1_ERR_INVALID_ARG, "spmv_merge: nnz > 0 with zero columns");
    G1_CHECK(ws != nullptr || n_rows == 0, G1_ERR_INVALID_ARG, "spmv_merge: null workspace");
    G1_CHECK(ws_bytes >= merge_workspace_bytes(n_rows, nnz), G1_ERR_INVALID_ARG,
             "spmv_merge: workspace of %zu bytes, %zu needed", ws_bytes, merge_workspace_bytes(n_rows, nnz));
    G1_CHECK(nnz < kInf, G1_ERR_INVALID_ARG, "spmv_merge: nnz must be below 2^32-1");
    return launch_spmv_merge<S>(n_rows, nnz, row_offsets, col, vals, x, y, ws, as_stream(stream));
}

template <typename S>
g1_status spmv_csrvec_exported(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, const uint32_t* row_offsets,
                               const uint32_t* col, const typename S::value_t* vals,
                               const typename S::x_t* x, typename S::y_t* y, uint32_t lanes,
                               g1_stream_t stream) {
    G1_CHECK(n_rows == 0 || (row_offsets && y), G1_ERR_INVALID_ARG, "spmv_csrvec: null row_offsets or y");
    G1_CHECK(nnz == 0 || (col && x), G1_ERR_INVALID_ARG, "spmv_csrvec: null col_indices or x");
    G1_CHECK(n_cols > 0 || nnz == 0, G1_ERR_INVALID_ARG, "spmv_csrvec: nnz > 0 with zero columns");
    if (lanes == 0) lanes = csrvec_lanes(n_rows, nnz);
    G1_CHECK(lanes <= 64 && (lanes & (lanes - 1)) == 0, G1_ERR_INVALID_ARG,
             "spmv_csrvec: lanes must be a power of two in 1..64, got %u", lanes);
    return launch_spmv_csrvec<S>(n_rows, row_offsets, col, vals, x, y, lanes, as_stream(stream));
}

// Scratch of g1_spmv_merge_minplus_u32_noalloc: the merge-path coordinates and tile records of one
// multiply, in a device array the library reserves when it is loaded (kNoAllocScratchBytes on every GPU it
// runs on, held until the process ends), for callers that cannot allocate. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol): its device address comes
// from hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kNoAllocMaxDevices];  // address of `scratch` on each device, once looked up

g1_status noalloc_scratch(void** out) {
    int dev = 0;
    G1_HIP(hipGetDevice(&dev));
    G1_CHECK(dev >= 0 && dev < kNoAllocMaxDevices, G1_ERR_INVALID_ARG,
             "spmv_merge_noalloc: device %d beyond the %d devices the scratch table covers", dev,
             kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        G1_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        G1_CHECK(p != nullptr, G1_ERR_HIP, "spmv_merge_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return G1_OK;
}

}  // namespace g1

extern "C" {

size_t g1_spmv_merge_noalloc_max_workspace_bytes(void) { return g1::kNoAllocScratchBytes; }

g1_status g1_spmv_merge_minplus_u32_noalloc(uint32_t n_rows, uint32_t n_cols, uint32_t nnz,
                                            const uint32_t* row_offsets, const uint32_t* col, const uint32_t* vals,
                                            const uint32_t* x, uint32_t* y, g1_stream_t stream) {
    const size_t need = g1::merge_workspace_bytes(n_rows, nnz);
    G1_CHECK(need <= g1::kNoAllocScratchBytes, G1_ERR_INVALID_ARG,
             "spmv_merge_noalloc: %u rows + %u nonzeros need a %zu-byte workspace, above the %zu-byte scratch "
             "(n_rows + nnz up to about %zu); use g1_spmv_merge_minplus_u32 with a workspace",
             n_rows, nnz, need, g1::kNoAllocScratchBytes,
             g1::kNoAllocScratchBytes / (g1::kMergeRecBytes + sizeof(g1::MergeCoord)) * g1::kMergeTile);
    G1_CHECK(n_rows == 0 || row_offsets, G1_ERR_INVALID_ARG, "spmv_merge_noalloc: null row_offsets");
    G1_CHECK(nnz < g1::kInf, G1_ERR_INVALID_ARG, "spmv_merge_noalloc: nnz must be below 2^32-1");
    if (n_rows == 0) return G1_OK;
    void* ws = nullptr;
    G1_TRY(g1::noalloc_scratch(&ws));
    // the partition overwrites every coordinate and the multiply every record it later reads, so the
    // previous call's contents need no clearing
    G1_TRY(g1::launch_merge_partition(n_rows, nnz, row_offsets, ws, g1::as_stream(stream)));
    return g1::spmv_merge_exported<g1::MinPlusU32>(n_rows, n_cols, nnz, row_offsets, col, vals, x, y, ws,
                                                   g1::kNoAllocScratchBytes, stream);
}

size_t g1_spmv_merge_workspace_bytes(uint32_t n_rows, uint32_t nnz) {
    return g1::merge_workspace_bytes(n_rows, nnz);
}

g1_status g1_spmv_merge_partition(uint32_t n_rows, uint32_t nnz, const uint32_t* row_offsets, void* ws,
                                  size_t ws_bytes, g1_stream_t stream) {
    G1_CHECK(ws_bytes >= g1::merge_workspace_bytes(n_rows, nnz), G1_ERR_INVALID_ARG,
             "spmv_merge_partition: workspace of %zu bytes, %zu needed", ws_bytes,
             g1::merge_workspace_bytes(n_rows, nnz));
    G1_CHECK(n_rows == 0 || (row_offsets && ws), G1_ERR_INVALID_ARG, "spmv_merge_partition: null pointer");
    G1_CHECK(nnz < g1::kInf, G1_ERR_INVALID_ARG, "spmv_merge_partition: nnz must be below 2^32-1");
    return g1::launch_merge_partition(n_rows, nnz, row_offsets, ws, g1::as_stream(stream));
}

g1_status g1_spmv_merge_minplus_u32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, const uint32_t* row_offsets,
                                    const uint32_t* col, const uint32_t* vals, const uint32_t* x, uint32_t* y,
                                    void* ws, size_t ws_bytes, g1_stream_t stream) {
    return g1::spmv_merge_exported<g1::MinPlusU32>(n_rows, n_cols, nnz, row_offsets, col, vals, x, y, ws,
                                                   ws_bytes, stream);
}

g1_status g1_spmv_merge_plustimes_f32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, const uint32_t* row_offsets,
                                      const uint32_t* col, const float* vals, const float* x, float* y,
                                      void* ws, size_t ws_bytes, g1_stream_t stream) {
    return g1::spmv_merge_exported<g1::PlusTimesKahanF32>(n_rows, n_cols, nnz, row_offsets, col, vals, x, y,
                                                          ws, ws_bytes, stream);
}

uint32_t g1_spmv_csrvec_lanes(uint32_t n_rows, uint32_t nnz) { return g1::csrvec_lanes(n_rows, nnz); }

g1_status g1_spmv_csrvec_minplus_u32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, const uint32_t* row_offsets,
                                     const uint32_t* col, const uint32_t* vals, const uint32_t* x, uint32_t* y,
                                     uint32_t lanes, g1_stream_t stream) {
    return g1::spmv_csrvec_exported<g1::MinPlusU32>(n_rows, n_cols, nnz, row_offsets, col, vals, x, y, lanes,
                                                    stream);
}

g1_status g1_spmv_csrvec_plustimes_f32(uint32_t n_rows, uint32_t n_cols, uint32_t nnz, const uint32_t* row_offsets,
                                       const uint32_t* col, const float* vals, const float* x, float* y,
                                       uint32_t lanes, g1_stream_t stream) {
    return g1::spmv_csrvec_exported<g1::PlusTimesKahanF32>(n_rows, n_cols, nnz, row_offsets, col, vals, x, y,
                                                           lanes, stream);
}

}  // extern "C"
