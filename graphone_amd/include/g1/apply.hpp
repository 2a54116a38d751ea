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
static __global__ void __launch_bounds__(256) pr_init_kernel(uint32_t n, const uint32_t* __restrict__ row_off,
                                                      const uint32_t* __restrict__ wts, float alpha,
                                                      float* __restrict__ iw, float* __restrict__ r, float r0,
                                                      float* __restrict__ x, uint32_t seed_vertex) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride) {
        float s = 0.0f;
        const uint32_t b = row_off[u], e = row_off[u + 1];
        if (wts) {
            for (uint32_t k = b; k < e; ++k) s += static_cast<float>(wts[k]);
        } else {
            s = static_cast<float>(e - b);
        }
        const float inv = (s != 0.0f) ? alpha / s : 0.0f;
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

// PageRank / personalized PageRank apply over K columns (K = 1 for a single run) with the canonical
// FP64 L1 reduction: block b covers vertices [256b, 256b + 256), thread t vertex 256b + t, and each
// column's 256 |r' - r| values are folded by the same fixed tree (wave shuffles, then the four waves in
// order) into partials[j * nblocks + b]. A batched column therefore gets bit for bit the norm its
// single-source run gets. Column j with bit j of `frozen` set keeps its rank and x (it has converged).
//   r'[v][j] = (v == seeds[j] ? base_seed : base_all) + y[v][j]
//   x[v][j]  = r'[v][j] * iw[v]
template <int K>
__global__ void __launch_bounds__(256) rank_apply_kernel(uint32_t n, const float* __restrict__ Y, float base_all,
                                                         const uint32_t* __restrict__ seeds, float base_seed,
                                                         const float* __restrict__ iw, const float* __restrict__ R_old,
                                                         float* __restrict__ R_new, float* __restrict__ X,
                                                         double* __restrict__ partials, uint32_t nblocks,
                                                         unsigned long long frozen) {
    __shared__ double s_part[256 / kWave];
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    const float iwv = v < n ? iw[v] : 0.0f;
    for (int j = 0; j < K; ++j) {
        double d = 0.0;
        if (v < n) {
            const size_t at = static_cast<size_t>(v) * K + j;
            if ((frozen >> j) & 1ull) {
                R_new[at] = R_old[at];
            } else {
                const float b = (v == seeds[j]) ? base_seed : base_all;
                const float rn = b + Y[at];
                R_new[at] = rn;
                X[at] = rn * iwv;
                d = fabs(static_cast<double>(rn) - static_cast<double>(R_old[at]));
            }
        }
#pragma unroll
        for (int off = kWave / 2; off >= 1; off >>= 1) d += __shfl_down(d, off);
        if (lane_id() == 0) s_part[threadIdx.x / kWave] = d;
        __syncthreads();
        if (threadIdx.x == 0)
            partials[static_cast<size_t>(j) * nblocks + blockIdx.x] = ((s_part[0] + s_part[1]) + s_part[2]) + s_part[3];
        __syncthreads();
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

}  // namespace g1
