// Deterministic block / grid reductions: fixed grid, fixed tree order, the
// last block to arrive sums the per-block partials in index order.
#pragma once
#include "common.hpp"

namespace fa {

constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 1024;

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

// Thread 0 stores this block's partial(s); every thread of the last block to
// arrive gets true.  `k` partial arrays of length gridDim.x are laid out
// back to back.
template <int K>
__device__ inline bool publish_partials(const double (&v)[K], double* partials, unsigned* counter) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < K; ++i) partials[i * gridDim.x + blockIdx.x] = v[i];
        __threadfence();
        unsigned prev = atomicAdd(counter, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    return last;
}

__device__ inline bool publish_partial(double v, double* partials, unsigned* counter) {
    const double a[1] = {v};
    return publish_partials<1>(a, partials, counter);
}

__device__ inline double sum_partials(double* partials, int n) {
    __threadfence();
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        s += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return block_sum(s);
}

}  // namespace fa
