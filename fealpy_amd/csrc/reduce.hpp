// Deterministic block / grid reductions: a grid that depends on n only,
// fixed tree order inside a block, per-block partials summed in index order.
#pragma once
#include "common.hpp"

namespace fa {

constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 2048;  // upper bound of a reduction grid (partials capacity): 8 waves per SIMD

// reduction grid for n items: depends on n only, so the summation order is fixed
inline unsigned red_grid(int64_t n) {
    int64_t g = (n + RED_BLOCK - 1) / RED_BLOCK;
    return static_cast<unsigned>(g < 1 ? 1 : (g > RED_GRID ? RED_GRID : g));
}

__device__ inline double block_sum(double v) {
    __shared__ double ws[RED_BLOCK / WAVE];
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < WAVE) {
        t = (threadIdx.x < blockDim.x / WAVE) ? ws[threadIdx.x] : 0.0;
        t = wave_sum(t);
    }
    return t;  // valid in thread 0
}

// Thread 0 stores the block's partial sums; partial array i (length
// gridDim.x) starts at partials + i * gridDim.x.  A one-block finishing kernel
// launched next on the same stream adds them in index order (finish_sum).
// Two kernels instead of a last-block-arrives counter: a device-scope atomic
// per workgroup on one address serialises across the XCDs and cost more than
// the reduction itself at thousands of workgroups.
template <int K>
__device__ inline void store_partials(const double (&v)[K], double* partials) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < K; ++i) partials[i * gridDim.x + blockIdx.x] = v[i];
    }
}

__device__ inline void store_partial(double v, double* partials) {
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// the same with the partial stored under the block index b whose rows the
// block summed: a kernel that deals the rows of block b = xcd_block() to
// physical block blockIdx.x stores each partial where the kernel with the
// plain order stores it, so the partials and their sum are bitwise the same
__device__ inline void store_partial_at(double v, double* partials, unsigned b) {
    if (threadIdx.x == 0) partials[b] = v;
}

// sum of partials[0..n) in a fixed order by one block; valid in thread 0
__device__ inline double finish_sum(const double* partials, int n) {
    constexpr int U = RED_GRID / RED_BLOCK;
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {  // all loads in flight before the adds
        const int i = threadIdx.x + u * RED_BLOCK;
        v[u] = i < n ? partials[i] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) s += v[u];
    return block_sum(s);
}

}  // namespace fa
