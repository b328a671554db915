// SELL-64: sliced ELLPACK whose slice is one wavefront of 64 consecutive rows.
// Row r = 64*s + i stores its k-th entry at slice_ptr[s] + 64*k + i, so the
// 64 lanes of a wave read 64 consecutive words per k (fully coalesced); the
// slice width is the longest row of the slice, shorter rows are padded with
// (col = own row or 0, value 0).
#pragma once
#include "common.hpp"

namespace fa {

template <class T, class X>
__device__ inline T sell_row_dot(const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ cols,
                                 const T* __restrict__ vals, int64_t row, const X* __restrict__ x) {
    const int64_t s = row >> 6;
    const int i = static_cast<int>(row & 63);
    const int32_t p0 = slice_ptr[s];
    const int w = (slice_ptr[s + 1] - p0) >> 6;
    const int32_t* c = cols + p0 + i;
    const T* v = vals + p0 + i;
    T acc = T(0);
    int k = 0;
    for (; k + 4 <= w; k += 4) {
        const int32_t c0 = c[64 * k], c1 = c[64 * (k + 1)], c2 = c[64 * (k + 2)], c3 = c[64 * (k + 3)];
        const T v0 = v[64 * k], v1 = v[64 * (k + 1)], v2 = v[64 * (k + 2)], v3 = v[64 * (k + 3)];
        const T x0 = static_cast<T>(x[c0]), x1 = static_cast<T>(x[c1]), x2 = static_cast<T>(x[c2]),
                x3 = static_cast<T>(x[c3]);
        acc += v0 * x0;
        acc += v1 * x1;
        acc += v2 * x2;
        acc += v3 * x3;
    }
    for (; k < w; ++k) acc += v[64 * k] * static_cast<T>(x[c[64 * k]]);
    return acc;
}

// row dot of the compressed format: sum_k vals[k] * x[row + offs[k]], FP32
// accumulation; rows of the last slice past n are not evaluated
__device__ inline float sell16_row_dot(const int32_t* __restrict__ slice_ptr, const int16_t* __restrict__ offs,
                                       const _Float16* __restrict__ vals, int64_t row, int64_t n,
                                       const float* __restrict__ x) {
    if (row >= n) return 0.f;
    const int64_t s = row >> 6;
    const int i = static_cast<int>(row & 63);
    const int32_t p0 = slice_ptr[s];
    const int w = (slice_ptr[s + 1] - p0) >> 6;
    const int16_t* o = offs + p0 + i;
    const _Float16* v = vals + p0 + i;
    float acc = 0.f;
    int k = 0;
    for (; k + 4 <= w; k += 4) {
        const int64_t c0 = row + o[64 * k], c1 = row + o[64 * (k + 1)], c2 = row + o[64 * (k + 2)],
                      c3 = row + o[64 * (k + 3)];
        const float v0 = static_cast<float>(v[64 * k]), v1 = static_cast<float>(v[64 * (k + 1)]),
                    v2 = static_cast<float>(v[64 * (k + 2)]), v3 = static_cast<float>(v[64 * (k + 3)]);
        const float x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc += v0 * x0;
        acc += v1 * x1;
        acc += v2 * x2;
        acc += v3 * x3;
    }
    for (; k < w; ++k) acc += static_cast<float>(v[64 * k]) * x[row + o[64 * k]];
    return acc;
}

}  // namespace fa
