// Kernels of the step-level C-ABI (csrc/step.hip): one advance + filter over a plain frontier queue, the
// frontier conversions, the BLAS engines' apply steps and a semiring SpMM with caller-owned dense
// operands. They serve a caller that keeps its own engine loop and buffers, as the reference's
// AdvanceGC (engine/advance_gc.h:162-283) and BlasGcEngine (engine/blas_gc.h:99-435) do.
//
// Queue frontiers are arrays of vertex ids whose length lives in device memory (so steps chain on a
// stream without host reads); G1_INF entries are skipped, as the reference's INVALID slots are.
// Bitmap frontiers hold bit v of word v / 32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"
#include "g1/spmm.hpp"
#include "g1/sweep.hpp"

namespace g1 {

constexpr uint32_t kStepWindow = 4096;  // edge slots per wave work item of the step advance

// off[i] = out-degree of frontier entry i (0 for an invalid entry or i >= count), i < capacity.
static __global__ void __launch_bounds__(256)
step_degrees_kernel(const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ f,
                    const uint32_t* __restrict__ count, uint32_t capacity, uint32_t n, uint32_t* __restrict__ off) {
    const uint32_t cnt = *count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < capacity; i += stride) {
        uint32_t d = 0;
        if (i < cnt) {
            const uint32_t u = f[i];
            if (u < n) d = row_off[u + 1] - row_off[u];
        }
        off[i] = d;
    }
}

// After the exclusive scan of the degrees: *edges = E (the frontier's out-edges), off[capacity] = E.
static __global__ void step_total_kernel(const uint32_t* __restrict__ deg_last, uint32_t* __restrict__ off,
                                         uint32_t capacity, uint32_t* __restrict__ edges,
                                         uint32_t* __restrict__ out_count) {
    const uint32_t e = capacity ? off[capacity - 1] + *deg_last : 0u;
    off[capacity] = e;
    if (edges) *edges = e;
    if (out_count) *out_count = 0u;
}

// The relaxation of one advance step: SSSP (dist[v] lowered by an atomic min; the first improver of the
// step claims v through visited[v] = iter) or BFS (dist[v] from G1_INF to `iter` by compare-and-swap).
template <bool kSssp>
struct StepRelax {
    const uint32_t* col;
    const uint32_t* wts;  // may be null: unit weights
    uint32_t* dist;
    uint32_t* visited;
    uint32_t iter;
    uint32_t* out;
    uint32_t* out_count;
    // the value a source carries to its edges: its live distance (SSSP), nothing (BFS)
    __device__ __forceinline__ uint32_t source(uint32_t u) const {
        return kSssp ? __hip_atomic_load(&dist[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    }
    __device__ __forceinline__ void operator()(const uint32_t (&e)[kSweepUnroll], const uint32_t (&du)[kSweepUnroll],
                                               const bool (&ok)[kSweepUnroll]) const {
        uint32_t v[kSweepUnroll], w[kSweepUnroll], cur[kSweepUnroll];
#pragma unroll
        for (int u = 0; u < kSweepUnroll; ++u) {
            v[u] = ok[u] ? col[e[u]] : 0u;
            w[u] = (ok[u] && wts) ? wts[e[u]] : 1u;
        }
#pragma unroll
        for (int u = 0; u < kSweepUnroll; ++u) cur[u] = ok[u] ? dist[v[u]] : 0u;
        uint32_t won = 0;  // bit u: slot u claimed its target
#pragma unroll
        for (int u = 0; u < kSweepUnroll; ++u) {
            if (!ok[u]) continue;
            if (kSssp) {
                const uint32_t nd = du[u] + w[u];
                if (nd < cur[u] && nd < atomicMin(&dist[v[u]], nd) && atomicExch(&visited[v[u]], iter) != iter)
                    won |= 1u << u;
            } else {
                if (cur[u] == kInf && atomicCAS(&dist[v[u]], kInf, iter) == kInf) won |= 1u << u;
            }
        }
        // one queue reservation per wave for the whole batch
        const uint32_t mine = static_cast<uint32_t>(__popc(won));
        const uint32_t incl = wave_inclusive_sum(mine);
        const uint32_t total = static_cast<uint32_t>(__shfl(static_cast<int>(incl), kWave - 1));
        if (total == 0) return;
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(out_count, total);
        base = static_cast<uint32_t>(__shfl(static_cast<int>(base), 0));
        uint32_t at = base + incl - mine;
#pragma unroll
        for (int u = 0; u < kSweepUnroll; ++u)
            if (won & (1u << u)) out[at++] = v[u];
    }
};

// One advance step over the queue frontier f (off: the exclusive prefix of its degrees, off[cap] = E).
// The E edge slots are cut into windows of kStepWindow; a wave takes a window, finds the entry owning
// its first slot with a 64-way search over off, and deals the window's edges 64 at a time over the
// entries that overlap it (wave_expand), 64 entries per round. Source values (rel.source(u)) are read
// when a round stages its entries: live, as the reference functor reads d[src]. The relaxation `rel`
// is called with up to kSweepUnroll x 64 edges per call (edge index, source value, valid).
template <typename Relax>
static __global__ void __launch_bounds__(256)
step_advance_kernel(const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ f,
                    const uint32_t* __restrict__ count, const uint32_t* __restrict__ off, uint32_t capacity,
                    Relax rel) {
    __shared__ uint32_t s_mark[256];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t cnt = (count && *count < capacity) ? *count : capacity;  // null count: capacity entries
    const uint32_t E = off[capacity];
    if (cnt == 0 || E == 0) return;
    const uint32_t windows = (E + kStepWindow - 1) / kStepWindow;
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    uint32_t* mark = s_mark + wave * kWave;
    for (uint32_t wi = blockIdx.x * (blockDim.x / kWave) + wave; wi < windows; wi += waves) {  // wave-uniform
        const uint32_t a = wi * kStepWindow;
        const uint32_t b = (E - a < kStepWindow) ? E : a + kStepWindow;
        uint32_t i0 = wave_search_last_le(off, cnt, a);
        for (;;) {  // wave-uniform: rounds of 64 entries until the window is covered
            const uint32_t i = i0 + lane;
            uint32_t rs = 0, deg = 0, du = 0;
            uint32_t lo = b;
            if (i < cnt) {
                lo = off[i];
                const uint32_t hi = off[i + 1];
                const uint32_t s0 = lo > a ? lo : a;
                const uint32_t s1 = hi < b ? hi : b;
                if (s0 < s1) {
                    const uint32_t u = f[i];
                    rs = row_off[u] + (s0 - lo);
                    deg = s1 - s0;
                    du = rel.source(u);
                }
            }
            wave_expand(rs, deg, du, mark, rel);
            // the next round starts at entry i0 + 64 if it still begins inside the window
            const uint32_t next = i0 + kWave;
            if (next >= cnt) break;
            const uint32_t next_lo = off[next];
            if (next_lo >= b) break;
            i0 = next;
        }
    }
}

// Queue -> bitmap (the bitmap is cleared by the caller's earlier fill); a null count means capacity entries.
static __global__ void __launch_bounds__(256)
queue_to_bitmap_kernel(const uint32_t* __restrict__ f, const uint32_t* __restrict__ count, uint32_t capacity,
                       uint32_t n, uint32_t* __restrict__ bitmap) {
    const uint32_t cnt = (count && *count < capacity) ? *count : capacity;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += stride) {
        const uint32_t v = f[i];
        if (v < n) atomicOr(&bitmap[v >> 5], 1u << (v & 31u));
    }
}

// Bitmap -> ascending queue, pass 1: set bits per word (words beyond the last vertex masked off).
static __global__ void __launch_bounds__(256)
bitmap_popc_kernel(const uint32_t* __restrict__ bitmap, uint32_t n, uint32_t* __restrict__ cnt) {
    const uint32_t words = (n + 31) / 32;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        uint32_t w = bitmap[i];
        if (i == words - 1 && (n & 31u)) w &= (1u << (n & 31u)) - 1u;
        cnt[i] = static_cast<uint32_t>(__popc(w));
    }
}

// pass 2: each word writes its vertices at its exclusive offset, in ascending order; the last word
// writes the count.
static __global__ void __launch_bounds__(256)
bitmap_write_kernel(const uint32_t* __restrict__ bitmap, uint32_t n, const uint32_t* __restrict__ pos,
                    const uint32_t* __restrict__ cnt, uint32_t* __restrict__ f, uint32_t* __restrict__ count) {
    const uint32_t words = (n + 31) / 32;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        uint32_t w = bitmap[i];
        if (i == words - 1 && (n & 31u)) w &= (1u << (n & 31u)) - 1u;
        uint32_t at = pos[i];
        while (w) {
            const uint32_t b = static_cast<uint32_t>(__ffs(static_cast<int>(w)) - 1);
            f[at++] = i * 32 + b;
            w &= w - 1u;
        }
        if (i == words - 1) *count = pos[i] + cnt[i];
    }
}

// Min-plus apply of the BLAS engines (apps/sssp_blas.cu:75-82 with blas_gc.h:43-55): pred = y[v] != INF
// && y[v] < d[v]; d[v] = y[v] and x[v] = y[v] when pred, x[v] = INF otherwise; per-block counts of pred.
static __global__ void __launch_bounds__(256)
step_apply_minplus_kernel(uint32_t n, const uint32_t* __restrict__ y, uint32_t* __restrict__ d,
                          uint32_t* __restrict__ x, uint32_t* __restrict__ block_cnt) {
    __shared__ uint32_t s_red[256 / kWave];
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0;
    if (v < n) {
        const uint32_t res = y[v];
        uint32_t xv = kInf;
        if (res != kInf && res < d[v]) {
            d[v] = res;
            xv = res;
            c = 1;
        }
        x[v] = xv;
    }
    const uint32_t w = wave_sum(c);
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) block_cnt[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// Sums per-block u32 counts into *out in one workgroup (fixed order).
static __global__ void __launch_bounds__(256)
sum_counts_kernel(const uint32_t* __restrict__ cnt, uint32_t nb, uint32_t* __restrict__ out) {
    __shared__ uint32_t s[256 / kWave];
    uint32_t acc = 0;
    for (uint32_t i = threadIdx.x; i < nb; i += 256) acc += cnt[i];
    acc = wave_sum(acc);
    if (lane_id() == 0) s[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *out = s[0] + s[1] + s[2] + s[3];
}

// PageRank apply (apps/pagerank_blas.cu:98-123, all-active engine blas_gc.h:353-435): r_new = base + y,
// x = r_new * iw; with r_old, per-block canonical FP64 partials of |r_new - r_old| (four sequential
// quarter sums of 64 vertices, ((q0 + q1) + q2) + q3), folded by reduce_partials_kernel<0>.
// l1_only: r_new and x are not written and y is read as the second vector (the L1 norm of y - r_old).
template <bool kL1Only>
static __global__ void __launch_bounds__(256)
step_pr_apply_kernel(uint32_t n, const float* __restrict__ y, float base, const float* __restrict__ iw,
                     const float* __restrict__ r_old, float* __restrict__ r_new, float* __restrict__ x,
                     double* __restrict__ partials) {
    __shared__ double s_d[256];
    __shared__ double s_q[256 / kWave];
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    double d = 0.0;
    if (v < n) {
        const float rn = kL1Only ? y[v] : base + y[v];
        if (!kL1Only) {
            r_new[v] = rn;
            x[v] = rn * iw[v];
        }
        if (r_old) d = fabs(static_cast<double>(rn) - static_cast<double>(r_old[v]));
    }
    if (!partials) return;
    s_d[threadIdx.x] = d;
    __syncthreads();
    if (lane_id() == 0) {
        const uint32_t w0 = (threadIdx.x / kWave) * kWave;
        double q = 0.0;
        for (int k = 0; k < kWave; ++k) q += s_d[w0 + k];
        s_q[threadIdx.x / kWave] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = ((s_q[0] + s_q[1]) + s_q[2]) + s_q[3];
}

// Semiring SpMM, Y = A (x) X with k dense columns stored row-major (X: n_cols x k, Y: n_rows x k), on the
// merge-path tile kernels of the batched engines (spmm_tile_kernel + spmm_fixup_kernel at 64 columns per
// pass): every tile holds the same number of rows plus nonzeros, so a hub row is spread over many tiles
// instead of one wave. This policy reads and writes the 64-column chunk starting at column c0 of the
// caller's arrays (row stride ld = k); lanes past the chunk's kc columns, and columns whose bit is clear in
// the device mask (bit j of word j / 64; null: every column), read nothing and write nothing, so their Y
// values stay. Plus-times folds in the single-source merge-path SpMV's order (Kahan-Babuska), min-plus is
// order-free.
template <typename T>
struct StridedColumnsPolicy {
    using aux_t = uint32_t;
    static constexpr bool kHasAux = false;
    static constexpr bool kRowOld = false;
    const T* X;
    T* Y;
    uint32_t ld;  // k
    uint32_t c0;  // first column of the chunk
    uint32_t kc;  // columns in the chunk (<= 64)
    const unsigned long long* mask;
    T neutral;    // what a skipped column reads (the semiring's absorbing x: INF for min-plus, 0 for plus-times)
    __device__ __forceinline__ bool on(uint32_t j) const {
        return j < kc && (!mask || ((mask[(c0 + j) >> 6] >> ((c0 + j) & 63u)) & 1ull));
    }
    __device__ __forceinline__ aux_t aux(uint32_t) const { return 0u; }
    __device__ __forceinline__ T x(uint32_t u, aux_t, uint32_t j, bool has) const {
        return (has && on(j)) ? X[static_cast<size_t>(u) * ld + c0 + j] : neutral;
    }
    __device__ __forceinline__ T row_old(uint32_t, uint32_t) const { return neutral; }
    __device__ __forceinline__ uint64_t row_done(uint32_t row, uint32_t j, T y, T) const {
        if (on(j)) Y[static_cast<size_t>(row) * ld + c0 + j] = y;
        return 0ull;
    }
};

// Stable compaction of the valid (non-G1_INF) entries of in[0..count), pass 1: per 256-entry block counts.
static __global__ void __launch_bounds__(256)
compact_count_kernel(const uint32_t* __restrict__ in, const uint32_t* __restrict__ count, uint32_t capacity,
                     uint32_t* __restrict__ block_cnt) {
    __shared__ uint32_t s_red[256 / kWave];
    const uint32_t cnt = *count < capacity ? *count : capacity;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool ok = i < cnt && in[i] != kInf;
    const uint32_t c = block_count<256>(ok, s_red);
    if (threadIdx.x == 0) block_cnt[blockIdx.x] = c;
}

// pass 2: each block writes its valid entries in order at its scanned offset; the last block writes the
// total to *out_count.
static __global__ void __launch_bounds__(256)
compact_write_kernel(const uint32_t* __restrict__ in, const uint32_t* __restrict__ count, uint32_t capacity,
                     const uint32_t* __restrict__ pos, const uint32_t* __restrict__ block_cnt,
                     uint32_t* __restrict__ out, uint32_t* __restrict__ out_count) {
    __shared__ uint32_t s_scan[256 / kWave];
    const uint32_t cnt = *count < capacity ? *count : capacity;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t v = i < cnt ? in[i] : kInf;
    uint32_t tot = 0;
    const uint32_t ex = block_exclusive_sum<256, uint32_t>(v != kInf ? 1u : 0u, tot, s_scan);
    if (v != kInf) out[pos[blockIdx.x] + ex] = v;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *out_count = pos[blockIdx.x] + block_cnt[blockIdx.x];
}

}  // namespace g1
