// SELL-64: sliced ELLPACK whose slice is one wavefront of 64 consecutive rows.
// Row r = 64*s + i stores its k-th entry at slice_ptr[s] + 64*k + i, so the
// 64 lanes of a wave read 64 consecutive words per k (fully coalesced); the
// slice width is the longest row of the slice, shorter rows are padded with
// (col = own row or 0, value 0).
//
// Row dots load a batch of ROW_BATCH entries (columns and values), then the
// ROW_BATCH gathered x values, then accumulate in entry order: a row of up to
// ROW_BATCH entries costs two dependent memory latencies instead of one pair
// per few entries, which is what bounds the small multigrid levels (few
// waves per CU, nothing to hide latency behind).  Positions past the slice
// width re-read the row's last entry (no branch, so the loads issue back to
// back) and are dropped from the sum; the sum order is the entry order
// k = 0, 1, ...
#pragma once
#include "common.hpp"

namespace fa {

constexpr int ROW_BATCH = 16;

// batches of B entries over one row's w entries (c, v: the row's first
// column / value, stride 64)
template <class T, int B, class XF>
__device__ inline T sell_batches(const int32_t* __restrict__ c, const T* __restrict__ v, int w, XF xf) {
    T acc = T(0);
    for (int k0 = 0; k0 < w; k0 += B) {
        int32_t cc[B];
        T vv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int k = min(k0 + u, w - 1);
            cc[u] = c[64 * k];
            vv[u] = v[64 * k];
        }
        T xx[B];
#pragma unroll
        for (int u = 0; u < B; ++u) xx[u] = xf(cc[u]);
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k0 + u < w) acc += vv[u] * xx[u];
    }
    return acc;
}

// Slices at most SHORT_ROW entries wide (transfer operators: 1-4 entries a
// row) take one batch of SHORT_ROW instead of B, so a 2-entry row issues 4
// loads, not 16.  The width is that of the slice (one wavefront), so the
// choice never diverges within a wave, and the sum is the same.
constexpr int SHORT_ROW = 4;

// sum_k vals[k] * xf(cols[k]) over the row, xf a gather functor (the vector,
// or a value computed from several vectors at that column)
template <class T, int B, class XF>
__device__ inline T sell_row_dot_f(const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ cols,
                                   const T* __restrict__ vals, int64_t row, XF xf) {
    const int64_t s = row >> 6;
    const int i = static_cast<int>(row & 63);
    const int32_t p0 = slice_ptr[s];
    const int w = (slice_ptr[s + 1] - p0) >> 6;
    const int32_t* c = cols + p0 + i;
    const T* v = vals + p0 + i;
    if constexpr (B > SHORT_ROW) {
        if (w <= SHORT_ROW) return sell_batches<T, SHORT_ROW>(c, v, w, xf);
    }
    return sell_batches<T, B>(c, v, w, xf);
}

template <class T, class X, int B = ROW_BATCH>
__device__ inline T sell_row_dot(const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ cols,
                                 const T* __restrict__ vals, int64_t row, const X* __restrict__ x) {
    return sell_row_dot_f<T, B>(slice_ptr, cols, vals, row, [=](int32_t j) { return static_cast<T>(x[j]); });
}

// row dot on int16 column offsets (column = row + offs[k]) with full-precision
// values: the outer operator stored on the slices of the compressed smoother
// format; rows of the last slice past n are not evaluated
template <class T, int B = ROW_BATCH>
__device__ inline T sell_row_dot_off(const int32_t* __restrict__ slice_ptr, const int16_t* __restrict__ offs,
                                     const T* __restrict__ vals, int64_t row, int64_t n, const T* __restrict__ x) {
    if (row >= n) return T(0);
    const int64_t s = row >> 6;
    const int i = static_cast<int>(row & 63);
    const int32_t p0 = slice_ptr[s];
    const int w = (slice_ptr[s + 1] - p0) >> 6;
    const int16_t* o = offs + p0 + i;
    const T* v = vals + p0 + i;
    T acc = T(0);
    for (int k0 = 0; k0 < w; k0 += B) {
        int oo[B];
        T vv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int k = min(k0 + u, w - 1);
            oo[u] = o[64 * k];
            vv[u] = v[64 * k];
        }
        T xx[B];
#pragma unroll
        for (int u = 0; u < B; ++u) xx[u] = x[row + oo[u]];
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k0 + u < w) acc += vv[u] * xx[u];
    }
    return acc;
}

// Column sources of the FP16 smoother format: int16 offsets from the row
// (4 bytes an entry with the FP16 value) or the full SELL's int32 columns
// when a column lies too far from its row (6 bytes an entry).
// xcd: whether kernels over this format order workgroups by XCD (measured
// faster for the short-row operators that fit int16 offsets, slower for the
// long-row operators that need int32 columns)
struct OffsetCols {
    static constexpr bool xcd = true;
    const int16_t* __restrict__ o;
    __device__ int64_t operator()(int64_t p, int64_t row) const { return row + o[p]; }
};

struct IndexCols {
    static constexpr bool xcd = false;
    const int32_t* __restrict__ c;
    __device__ int64_t operator()(int64_t p, int64_t) const { return c[p]; }
};

// row dot of the FP16 format: sum_k vals[k] * xf(col_k), FP32 accumulation;
// rows of the last slice past n are not evaluated
template <class CS, class XF>
__device__ inline float sell16_row_dot_f(const int32_t* __restrict__ slice_ptr, CS cs,
                                         const _Float16* __restrict__ vals, int64_t row, int64_t n, XF xf) {
    if (row >= n) return 0.f;
    const int64_t s = row >> 6;
    const int i = static_cast<int>(row & 63);
    const int32_t p0 = slice_ptr[s] + i;
    const int w = (slice_ptr[s + 1] - slice_ptr[s]) >> 6;
    float acc = 0.f;
    for (int k0 = 0; k0 < w; k0 += ROW_BATCH) {
        int32_t cc[ROW_BATCH];  // column indices are int32 (SELL) or row + int16
        float vv[ROW_BATCH];
#pragma unroll
        for (int u = 0; u < ROW_BATCH; ++u) {
            const int64_t p = p0 + 64 * min(k0 + u, w - 1);
            cc[u] = static_cast<int32_t>(cs(p, row));
            vv[u] = static_cast<float>(vals[p]);
        }
        float xx[ROW_BATCH];
#pragma unroll
        for (int u = 0; u < ROW_BATCH; ++u) xx[u] = xf(cc[u]);
#pragma unroll
        for (int u = 0; u < ROW_BATCH; ++u)
            if (k0 + u < w) acc += vv[u] * xx[u];
    }
    return acc;
}

template <class CS>
__device__ inline float sell16_row_dot(const int32_t* __restrict__ slice_ptr, CS cs, const _Float16* __restrict__ vals,
                                       int64_t row, int64_t n, const float* __restrict__ x) {
    return sell16_row_dot_f(slice_ptr, cs, vals, row, n, [=](int64_t j) { return x[j]; });
}

}  // namespace fa
