// Apply/construct kernels of the pull (BLAS) engine, engine/blas_gc.h:21-76 and :171-435 redesigned:
// the reference's Construct (fill + transform + scatter) and Apply (transform + copy_if) passes are
// fused into one kernel that reads y once and writes the next x, the per-vertex state and a count.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "g1/device.hpp"

namespace g1 {

// Min-plus apply (apps/sssp_blas.cu:75-82 + blas_gc.h:43-55):
//   pred = y[v] != INF && y[v] < d[v];  if pred: d[v] = y[v];  x[v] = pred ? d[v] : INF
// The number of applied vertices goes to this block's slot (slots[blockIdx.x] += count), never to a
// shared counter: device-scope atomics on one address serialize across the whole GPU.
static __global__ void __launch_bounds__(256) apply_minplus_kernel(uint32_t n, const uint32_t* __restrict__ y,
                                                                   uint32_t* __restrict__ d, uint32_t* __restrict__ x,
                                                                   unsigned long long* __restrict__ slots) {
    __shared__ uint32_t s_red[256 / kWave];
    uint32_t local = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t res = y[v];
        uint32_t xv = kInf;
        if (res != kInf) {
            const uint32_t dv = d[v];
            if (res < dv) {
                d[v] = res;
                xv = res;
                ++local;
            }
        }
        x[v] = xv;
    }
    const uint32_t w = wave_sum(local);
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = s_red[0] + s_red[1] + s_red[2] + s_red[3];
        if (t) slots[blockIdx.x] += t;
    }
}

// x = INF everywhere except x[src] = 0 (the sparse engine's Construct for the frontier {src}),
// d likewise (the component's Init: fill(dists, INF); dists.set(src, 0)).
static __global__ void __launch_bounds__(256) minplus_seed_kernel(uint32_t n, uint32_t src, uint32_t* __restrict__ d,
                                                           uint32_t* __restrict__ x) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t val = (v == src) ? 0u : kInf;
        d[v] = val;
        if (x) x[v] = val;
    }
}

// iw[u] = alpha / sum_out w(u) (0 when u has no out-edge), summed in float left to right as
// apps/pagerank_blas.cu:41-53 does; r[u] = r0; x[u] = r[u] * iw[u].
// Exact out-weight sum of every vertex (apps/pagerank_blas.cu:41-53 sums the out-edge weights; here the
// u32 weights are summed exactly in 64 bits and rounded once to float). One wave per row.
static __global__ void __launch_bounds__(256) out_weight_sum_kernel(uint32_t n, const uint32_t* __restrict__ row_off,
                                                             const uint32_t* __restrict__ wts,
                                                             float* __restrict__ osum) {
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += waves) {
        const uint32_t b = row_off[u], e = row_off[u + 1];
        unsigned long long s = 0;
        if (wts) {
            for (uint32_t k = b + lane_id(); k < e; k += kWave) s += wts[k];
#pragma unroll
            for (int off = kWave / 2; off >= 1; off >>= 1) s += __shfl_down(s, off);
        } else {
            s = e - b;
        }
        if (lane_id() == 0) osum[u] = static_cast<float>(s);
    }
}

// iw[u] = alpha / osum[u] (0 for a vertex without out-weight), r0 and x0 = r0 * iw; with a seed vertex,
// r0 = e_seed (personalized PageRank).
static __global__ void __launch_bounds__(256) pr_init_kernel(uint32_t n, const float* __restrict__ osum, float alpha,
                                                      float* __restrict__ iw, float* __restrict__ r, float r0,
                                                      float* __restrict__ x, uint32_t seed_vertex) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride) {
        const float sw = osum[u];
        const float inv = (sw != 0.0f) ? alpha / sw : 0.0f;
        iw[u] = inv;
        const float rv = (seed_vertex == kInf) ? r0 : (u == seed_vertex ? 1.0f : 0.0f);
        r[u] = rv;
        x[u] = rv * inv;
    }
}

// r_new[v] = base(v) + y[v]; x[v] = r_new[v] * iw[v]; per-block FP64 partial of |r_new - r_old|
// (mode 0: sum, the L1 norm) or max |r_new - r_old| in float (mode 1, the reference's test).
// base(v) = base_all, plus base_seed when v == seed (personalized PageRank).
template <int kMode>
__global__ void __launch_bounds__(256) pr_apply_kernel(uint32_t n, const float* __restrict__ y, float base_all,
                                                       uint32_t seed, float base_seed, const float* __restrict__ iw,
                                                       const float* __restrict__ r_old, float* __restrict__ r_new,
                                                       float* __restrict__ x, double* __restrict__ partials) {
    __shared__ double s_part[256 / kWave];
    double acc = 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const float b = (v == seed) ? base_seed : base_all;
        const float rn = b + y[v];
        r_new[v] = rn;
        x[v] = rn * iw[v];
        if (kMode == 0) {
            acc += fabs(static_cast<double>(rn) - static_cast<double>(r_old[v]));
        } else {
            const double dlt = static_cast<double>(fabsf(rn - r_old[v]));
            acc = dlt > acc ? dlt : acc;
        }
    }
    // Fixed-shape reduction: lanes, then waves in order.
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) {
        const double o = __shfl_down(acc, off);
        if (kMode == 0) acc += o; else acc = o > acc ? o : acc;
    }
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = s_part[0];
        for (int w = 1; w < 256 / kWave; ++w) {
            if (kMode == 0) t += s_part[w]; else t = s_part[w] > t ? s_part[w] : t;
        }
        partials[blockIdx.x] = t;
    }
}

// Canonical FP64 L1 norm of r' - r, shared by the single-source and the batched PageRank so that a
// batched column stops exactly when its single-source run stops: for every block b of 256 vertices
// [256b, 256b + 256) and every column, the four quarters of 64 vertices are each summed sequentially
// in vertex order (starting from 0.0), and partials[j * nblocks + b] = ((q0 + q1) + q2) + q3.
//   r'[v] = (v == seed ? base_seed : base_all) + y[v]
//   x[v]  = r'[v] * iw[v]
// Single column: thread t owns vertex 256b + t; lane 0 of each wave folds its quarter from LDS.
static __global__ void __launch_bounds__(256) rank_apply_kernel(uint32_t n, const float* __restrict__ y, float base_all,
                                                         const uint32_t* __restrict__ seeds, float base_seed,
                                                         const float* __restrict__ iw,
                                                         const float* __restrict__ r_old, float* __restrict__ r_new,
                                                         float* __restrict__ x, double* __restrict__ partials) {
    __shared__ double s_d[256];
    __shared__ double s_q[256 / kWave];
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    double d = 0.0;
    if (v < n) {
        const float b = (v == seeds[0]) ? base_seed : base_all;
        const float rn = b + y[v];
        r_new[v] = rn;
        x[v] = rn * iw[v];
        d = fabs(static_cast<double>(rn) - static_cast<double>(r_old[v]));
    }
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

// K columns, row-major n x K: a group of K lanes (lane j = column j) walks one block of 256 vertices in
// vertex order with coalesced K-wide accesses; a workgroup covers 256 / K blocks. Column j with bit j of
// `frozen` set keeps its rank (r' = r) and contributes nothing (it has converged). With frozen_dev the
// mask is read from device memory instead (a captured iteration whose mask the device updates).
template <int K>
__global__ void __launch_bounds__(256) rank_apply_batch_kernel(uint32_t n, const float* __restrict__ Y,
                                                               float base_all, const uint32_t* __restrict__ seeds,
                                                               float base_seed, const float* __restrict__ iw,
                                                               const float* __restrict__ R_old,
                                                               float* __restrict__ R_new, float* __restrict__ X,
                                                               double* __restrict__ partials, uint32_t nblocks,
                                                               unsigned long long frozen,
                                                               const unsigned long long* __restrict__ frozen_dev) {
    constexpr uint32_t G = 256 / K;
    const uint32_t j = threadIdx.x % K;
    const uint32_t blk = blockIdx.x * G + threadIdx.x / K;
    if (blk >= nblocks) return;
    const bool fz = ((frozen_dev ? *frozen_dev : frozen) >> j) & 1ull;
    const uint32_t seed = seeds[j];
    double q[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        double acc = 0.0;
        const uint32_t v0 = blk * 256 + w * kWave;
#pragma unroll 8
        for (uint32_t k = 0; k < static_cast<uint32_t>(kWave); ++k) {
            const uint32_t v = v0 + k;
            if (v >= n) break;
            const size_t at = static_cast<size_t>(v) * K + j;
            const float ro = R_old[at];
            if (fz) {
                R_new[at] = ro;
            } else {
                const float b = (v == seed) ? base_seed : base_all;
                const float rn = b + Y[at];
                R_new[at] = rn;
                X[at] = rn * iw[v];
                acc += fabs(static_cast<double>(rn) - static_cast<double>(ro));
            }
        }
        q[w] = acc;
    }
    partials[static_cast<size_t>(j) * nblocks + blk] = ((q[0] + q[1]) + q[2]) + q[3];
}

// The stop rule of the batched personalized PageRank, per column, on the device: a column still running
// whose FP64 L1 delta (deltas[j]) fell below tol freezes; state[0] is the frozen mask, state[1] counts the
// iterations that began with a running column. One wave.
template <int K>
__global__ void __launch_bounds__(64) ppr_freeze_kernel(const double* __restrict__ deltas, double tol,
                                                        unsigned long long* __restrict__ state) {
    const uint32_t j = lane_id();
    const unsigned long long all = (K == 64) ? ~0ull : ((1ull << K) - 1ull);
    const unsigned long long f = state[0];
    const bool stop = j < static_cast<uint32_t>(K) && !((f >> j) & 1ull) && deltas[j] < tol;
    const unsigned long long nf = f | ballot(stop);
    if (j == 0) {
        if (f != all) state[1] += 1ull;
        state[0] = nf;
    }
}

// Folds per-block partials in block order into out[0] (sum or max); one block.
template <int kMode>
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double* __restrict__ partials, uint32_t count,
                                                              double* __restrict__ out) {
    __shared__ double s[256];
    // blockIdx.y selects a column: partials[y * count ..], out[y]
    partials += static_cast<size_t>(blockIdx.y) * count;
    out += blockIdx.y;
    double acc = 0.0;
    // Each thread folds a contiguous run, then the runs are folded in thread order.
    const uint32_t per = (count + 255) / 256;
    const uint32_t b = threadIdx.x * per;
    for (uint32_t i = b; i < b + per && i < count; ++i) {
        if (kMode == 0) acc += partials[i]; else acc = partials[i] > acc ? partials[i] : acc;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < 256; ++i) {
            if (kMode == 0) t += s[i]; else t = s[i] > t ? s[i] : t;
        }
        out[0] = t;
    }
}

// Max |a - b| over float vectors (the reference's IsConvergent transform_reduce,
// apps/pagerank_blas.cu:63-78), written as per-block partials for reduce_partials_kernel<1>.
static __global__ void __launch_bounds__(256) max_abs_diff_kernel(uint32_t n, const float* __restrict__ a,
                                                           const float* __restrict__ b,
                                                           double* __restrict__ partials) {
    __shared__ double s_part[256 / kWave];
    double acc = 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const double dlt = static_cast<double>(fabsf(a[v] - b[v]));
        acc = dlt > acc ? dlt : acc;
    }
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) {
        const double o = __shfl_down(acc, off);
        acc = o > acc ? o : acc;
    }
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = s_part[0];
        for (int w = 1; w < 256 / kWave; ++w) t = s_part[w] > t ? s_part[w] : t;
        partials[blockIdx.x] = t;
    }
}

// Reached-vertex statistics: per block, the count of d[v] != INF and the sum of their out-degrees, added
// to slots_cnt[blockIdx.x] and slots_edges[blockIdx.x].
static __global__ void __launch_bounds__(256) reached_stats_kernel(uint32_t n, const uint32_t* __restrict__ d,
                                                                   const uint32_t* __restrict__ row_off,
                                                                   unsigned long long* __restrict__ slots_cnt,
                                                                   unsigned long long* __restrict__ slots_edges) {
    __shared__ unsigned long long s_red[2][256 / kWave];
    unsigned long long cnt = 0, edges = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        if (d[v] != kInf) {
            ++cnt;
            edges += row_off[v + 1] - row_off[v];
        }
    }
    cnt = wave_sum(cnt);
    edges = wave_sum(edges);
    if (lane_id() == 0) {
        s_red[0][threadIdx.x / kWave] = cnt;
        s_red[1][threadIdx.x / kWave] = edges;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        slots_cnt[blockIdx.x] += s_red[0][0] + s_red[0][1] + s_red[0][2] + s_red[0][3];
        slots_edges[blockIdx.x] += s_red[1][0] + s_red[1][1] + s_red[1][2] + s_red[1][3];
    }
}

// The end of a run on the hub-first copy of a graph: d[v] = the copy's result at inv[v] (kWord: the high
// half of its 64-bit distance word, else a u32 depth), with the reached-vertex stats of d counted as
// reached_stats_kernel counts them (row_off in d's ids), added atomically into slot blockIdx.x % slots.
// A workgroup takes 1024 consecutive v per step, each thread four of them 256 apart, so every load and
// store instruction covers 256 contiguous bytes and a thread's four gathers are in flight together; the
// grid covers n in one pass.
template <bool kWord>
static __global__ void __launch_bounds__(256) map_back_stats_kernel(uint32_t n, const uint32_t* __restrict__ inv,
                                                                    const void* __restrict__ src,
                                                                    uint32_t* __restrict__ d,
                                                                    const uint32_t* __restrict__ row_off,
                                                                    unsigned long long* __restrict__ slots_cnt,
                                                                    unsigned long long* __restrict__ slots_edges,
                                                                    uint32_t slots) {
    __shared__ unsigned long long s_red[2][256 / kWave];
    unsigned long long cnt = 0, edges = 0;
    const uint32_t tiles = (n + 1023) / 1024;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint32_t v[4], iv[4], x[4], r0[4], r1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = t * 1024 + k * 256 + threadIdx.x;
            const uint32_t c = v[k] < n ? v[k] : n - 1;  // clamped: every load is issued
            iv[k] = inv[c];
            r0[k] = row_off[c];
            r1[k] = row_off[c + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (kWord)
                x[k] = static_cast<uint32_t>(static_cast<const unsigned long long*>(src)[iv[k]] >> 32);
            else
                x[k] = static_cast<const uint32_t*>(src)[iv[k]];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (v[k] < n) {
                d[v[k]] = x[k];
                if (x[k] != kInf) {
                    ++cnt;
                    edges += r1[k] - r0[k];
                }
            }
        }
    }
    cnt = wave_sum(cnt);
    edges = wave_sum(edges);
    if (lane_id() == 0) {
        s_red[0][threadIdx.x / kWave] = cnt;
        s_red[1][threadIdx.x / kWave] = edges;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&slots_cnt[blockIdx.x % slots], s_red[0][0] + s_red[0][1] + s_red[0][2] + s_red[0][3]);
        atomicAdd(&slots_edges[blockIdx.x % slots], s_red[1][0] + s_red[1][1] + s_red[1][2] + s_red[1][3]);
    }
}

}  // namespace g1
