// SELL-64 construction from CSR and SELL SpMV.
#include "common.hpp"
#include "sell.hpp"

namespace fa {

// entries per slice = 64 * (longest row of the slice); one wave per slice
__global__ void __launch_bounds__(256) k_sell_widths(int64_t n, const int32_t* __restrict__ indptr,
                                                    int32_t* __restrict__ entries) {
    const int64_t nslices = (n + 63) / 64;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < nslices * 64;
         row += (int64_t)gridDim.x * blockDim.x) {
        int len = row < n ? indptr[row + 1] - indptr[row] : 0;
        for (int o = WAVE / 2; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
        if ((row & 63) == 0) entries[row >> 6] = 64 * len;
    }
}

// Padding entries get value 0 and a column that is valid for x: the row itself
// when it is a column index (square matrices: same cache line as the
// diagonal), else column 0 (e.g. a prolongation with fewer columns than rows).
template <class A, class B>
__global__ void __launch_bounds__(256) k_sell_fill(int64_t n, int64_t ncols, const int32_t* __restrict__ indptr,
                                                  const int32_t* __restrict__ indices, const A* __restrict__ vin,
                                                  const int32_t* __restrict__ slice_ptr, int32_t* __restrict__ cols,
                                                  B* __restrict__ vals) {
    const int64_t nslices = (n + 63) / 64;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < nslices * 64;
         row += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = row >> 6;
        const int i = static_cast<int>(row & 63);
        const int32_t p0 = slice_ptr[s];
        const int w = (slice_ptr[s + 1] - p0) >> 6;
        const int32_t r0 = row < n ? indptr[row] : 0;
        const int len = row < n ? indptr[row + 1] - r0 : 0;
        const int32_t pad = row < ncols ? static_cast<int32_t>(row) : 0;
        for (int k = 0; k < w; ++k) {
            const int64_t o = p0 + 64 * k + i;
            if (cols) cols[o] = k < len ? indices[r0 + k] : pad;
            if (vals) vals[o] = k < len ? static_cast<B>(vin[r0 + k]) : B(0);
        }
    }
}

// Values-only refill with the slice staged in LDS: the CSR entries of a
// 64-row slice are one contiguous range, read with coalesced loads; each
// lane then writes its row's entries at the slice's column stride.
// One wavefront per slice; slices whose range exceeds SF_STAGE use the
// per-row path above.
constexpr int SF_STAGE = 2048;

template <class A, class B>
__global__ void __launch_bounds__(WAVE) k_sell_fill_staged(int64_t n, const int32_t* __restrict__ indptr,
                                                          const A* __restrict__ vin,
                                                          const int32_t* __restrict__ slice_ptr,
                                                          B* __restrict__ vals) {
    __shared__ A stage[SF_STAGE];
    const int lane = threadIdx.x;
    const int64_t nslices = (n + 63) / 64;
    for (int64_t s = blockIdx.x; s < nslices; s += gridDim.x) {
        const int64_t row = s * 64 + lane;
        const int64_t rlast = (s * 64 + 64 < n) ? s * 64 + 64 : n;
        const int32_t base = indptr[s * 64];
        const int32_t total = indptr[rlast] - base;
        const int32_t p0 = slice_ptr[s];
        const int w = (slice_ptr[s + 1] - p0) >> 6;
        const int32_t r0 = row < n ? indptr[row] - base : 0;
        const int len = row < n ? indptr[row + 1] - base - r0 : 0;
        if (total <= SF_STAGE) {
            constexpr int U = 8;  // loads of a lane in flight before its LDS stores
            for (int t0 = 0; t0 < total; t0 += U * WAVE) {
                A v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = t0 + u * WAVE + lane;
                    v[u] = t < total ? vin[base + t] : A(0);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = t0 + u * WAVE + lane;
                    if (t < total) stage[t] = v[u];
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            for (int k = 0; k < w; ++k) vals[p0 + 64 * k + lane] = k < len ? static_cast<B>(stage[r0 + k]) : B(0);
        } else {
            constexpr int U = 8;  // a lane's loads of U entries in flight before its stores
            for (int k0 = 0; k0 < w; k0 += U) {
                A v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = k0 + u < len ? vin[base + r0 + k0 + u] : A(0);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k0 + u < w) vals[p0 + 64 * (k0 + u) + lane] = static_cast<B>(v[u]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_sell_spmv(int64_t n, const int32_t* __restrict__ slice_ptr,
                                                  const int32_t* __restrict__ cols, const T* __restrict__ vals,
                                                  const T* __restrict__ x, T* __restrict__ y) {
    // FP32: XCD-ordered workgroups (measured faster); FP64 keeps dispatch order (measured slower with it)
    const int64_t b0 = sizeof(T) == 4 ? xcd_block() : blockIdx.x;
    for (int64_t row = b0 * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T v = sell_row_dot<T, T>(slice_ptr, cols, vals, row, x);
        if (row < n) y[row] = v;
    }
}

// compressed fill: offs = col - row, vals = a_ij / a_ii in FP16 (one thread per row;
// V: the CSR value type)
template <class V>
__global__ void __launch_bounds__(256) k_sell16_fill(int64_t n, const int32_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices,
                                                    const V* __restrict__ vin,
                                                    const int32_t* __restrict__ slice_ptr,
                                                    int16_t* __restrict__ offs, _Float16* __restrict__ vals,
                                                    int32_t* __restrict__ overflow) {
    const int64_t nslices = (n + 63) / 64;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < nslices * 64;
         row += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = row >> 6;
        const int i = static_cast<int>(row & 63);
        const int32_t p0 = slice_ptr[s];
        const int w = (slice_ptr[s + 1] - p0) >> 6;
        const int32_t r0 = row < n ? indptr[row] : 0;
        const int len = row < n ? indptr[row + 1] - r0 : 0;
        double d = 0.0;
        for (int k = 0; k < len; ++k)
            if (indices[r0 + k] == row) d = static_cast<double>(vin[r0 + k]);
        const double inv = d != 0.0 ? 1.0 / d : 1.0;
        bool bad = false;
        for (int k = 0; k < w; ++k) {
            const int64_t o = p0 + 64 * k + i;
            int64_t off = 0;
            double v = 0.0;
            if (k < len) {
                off = static_cast<int64_t>(indices[r0 + k]) - row;
                v = static_cast<double>(vin[r0 + k]) * inv;
            }
            if (off > 32767 || off < -32767) {
                bad = true;
                off = 0;
            }
            offs[o] = static_cast<int16_t>(off);
            vals[o] = static_cast<_Float16>(v);
        }
        if (bad && overflow) overflow[0] = 1;
    }
}

// compressed refill of an already checked pattern, slice staged in LDS
// (values and columns of the slice's contiguous CSR range)
template <class V>
__global__ void __launch_bounds__(WAVE) k_sell16_refill(int64_t n, const int32_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ indices,
                                                       const V* __restrict__ vin,
                                                       const int32_t* __restrict__ slice_ptr,
                                                       _Float16* __restrict__ vals) {
    __shared__ double sv[SF_STAGE];
    __shared__ int32_t si[SF_STAGE];
    const int lane = threadIdx.x;
    const int64_t nslices = (n + 63) / 64;
    for (int64_t s = blockIdx.x; s < nslices; s += gridDim.x) {
        const int64_t row = s * 64 + lane;
        const int64_t rlast = (s * 64 + 64 < n) ? s * 64 + 64 : n;
        const int32_t base = indptr[s * 64];
        const int32_t total = indptr[rlast] - base;
        const int32_t p0 = slice_ptr[s];
        const int w = (slice_ptr[s + 1] - p0) >> 6;
        const int32_t r0 = row < n ? indptr[row] - base : 0;
        const int len = row < n ? indptr[row + 1] - base - r0 : 0;
        const bool staged = total <= SF_STAGE;
        if (staged) {
            for (int t = lane; t < total; t += WAVE) {
                sv[t] = static_cast<double>(vin[base + t]);
                si[t] = indices[base + t];
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        double d = 0.0;
        for (int k = 0; k < len; ++k) {
            const int32_t c = staged ? si[r0 + k] : indices[base + r0 + k];
            if (c == row) d = staged ? sv[r0 + k] : static_cast<double>(vin[base + r0 + k]);
        }
        const double inv = d != 0.0 ? 1.0 / d : 1.0;
        for (int k = 0; k < w; ++k) {
            double v = 0.0;
            if (k < len) v = (staged ? sv[r0 + k] : static_cast<double>(vin[base + r0 + k])) * inv;
            vals[p0 + 64 * k + lane] = static_cast<_Float16>(v);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

}  // namespace fa

using namespace fa;

extern "C" int fa_sell16_fill(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices,
                              const void* values, const int32_t* slice_ptr, int16_t* offs, void* vals, void* work,
                              size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_sell16_fill: bad dtype %d", dtype);
    FA_REQUIRE(n > 0 && indptr && indices && values && slice_ptr && offs && vals, "fa_sell16_fill: bad arguments");
    FA_REQUIRE(!work || work_bytes >= sizeof(int32_t), "fa_sell16_fill: workspace too small");
    hipStream_t s = S(stream);
    int32_t* flag = static_cast<int32_t*>(work);
    if (!flag) {  // pattern already checked: values only, no sync
        const dim3 g(grid_for((n + 63) / 64, 1, 1 << 20));
        if (dtype == FA_F64)
            hipLaunchKernelGGL(k_sell16_refill<double>, g, dim3(WAVE), 0, s, n, indptr, indices,
                               static_cast<const double*>(values), slice_ptr, static_cast<_Float16*>(vals));
        else
            hipLaunchKernelGGL(k_sell16_refill<float>, g, dim3(WAVE), 0, s, n, indptr, indices,
                               static_cast<const float*>(values), slice_ptr, static_cast<_Float16*>(vals));
        FA_LAUNCH("k_sell16_refill");
        return FA_OK;
    }
    FA_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
    const dim3 g(grid_for(((n + 63) / 64) * 64, 256, 1 << 16));
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_sell16_fill<double>, g, dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), slice_ptr, offs, static_cast<_Float16*>(vals), flag);
    else
        hipLaunchKernelGGL(k_sell16_fill<float>, g, dim3(256), 0, s, n, indptr, indices,
                           static_cast<const float*>(values), slice_ptr, offs, static_cast<_Float16*>(vals), flag);
    FA_LAUNCH("k_sell16_fill");
    int32_t h = 0;
    FA_HIP(hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (h) {
        set_error("fa_sell16_fill: a column lies more than 32767 from its row");
        return FA_ECAPACITY;
    }
    return FA_OK;
}

extern "C" size_t fa_sell_work_bytes(int64_t n) {
    const int64_t ns = (n + 63) / 64;
    return align_up((ns + 1) * sizeof(int32_t)) + exclusive_scan_work_bytes(ns) + 1024;
}

extern "C" int fa_sell_count(int64_t n, const int32_t* indptr, int32_t* slice_ptr, int64_t* entries_host,
                             void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(n > 0, "fa_sell_count: empty matrix");
    hipStream_t s = S(stream);
    const int64_t ns = (n + 63) / 64;
    Arena ar(work, work_bytes);
    int32_t* ent = ar.take<int32_t>(ns + 1);
    const size_t sb = exclusive_scan_work_bytes(ns);
    void* sw = ar.take<char>(sb);
    if (!ent || !sw) {
        set_error("fa_sell_count: workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_sell_widths, dim3(grid_for(ns * 64, 256, 1 << 16)), dim3(256), 0, s, n, indptr, ent);
    FA_LAUNCH("k_sell_widths");
    int st = exclusive_scan_i32(ent, slice_ptr, ns, sw, sb, s);
    if (st != FA_OK) return st;
    int32_t tot = 0;
    FA_HIP(hipMemcpyAsync(&tot, slice_ptr + ns, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (entries_host) *entries_host = tot;
    return FA_OK;
}

extern "C" int fa_sell_fill(int dtype_in, int dtype_out, int64_t n, int64_t ncols, const int32_t* indptr,
                            const int32_t* indices, const void* values, const int32_t* slice_ptr, int32_t* cols,
                            void* vals, fa_stream_t stream) {
    FA_REQUIRE(n > 0 && ncols > 0, "fa_sell_fill: empty matrix");
    FA_REQUIRE(!vals || values, "fa_sell_fill: values missing");
    hipStream_t s = S(stream);
    if (!cols && vals) {  // values-only refresh: slices staged in LDS
        const unsigned gs = grid_for((n + 63) / 64, 1, 1 << 20);
#define FA_SS(A_, B_)                                                                                       \
    hipLaunchKernelGGL((k_sell_fill_staged<A_, B_>), dim3(gs), dim3(WAVE), 0, s, n, indptr,                 \
                       static_cast<const A_*>(values), slice_ptr, static_cast<B_*>(vals))
        if (dtype_in == FA_F64 && dtype_out == FA_F64) FA_SS(double, double);
        else if (dtype_in == FA_F64 && dtype_out == FA_F32) FA_SS(double, float);
        else if (dtype_in == FA_F32 && dtype_out == FA_F32) FA_SS(float, float);
        else if (dtype_in == FA_F32 && dtype_out == FA_F64) FA_SS(float, double);
        else {
            set_error("fa_sell_fill: bad dtypes");
            return FA_EINVAL;
        }
#undef FA_SS
        FA_LAUNCH("k_sell_fill_staged");
        return FA_OK;
    }
    const unsigned grid = grid_for(((n + 63) / 64) * 64, 256, 1 << 16);
#define FA_SF(A_, B_)                                                                                       \
    hipLaunchKernelGGL((k_sell_fill<A_, B_>), dim3(grid), dim3(256), 0, s, n, ncols, indptr, indices,      \
                       static_cast<const A_*>(values), slice_ptr, cols, static_cast<B_*>(vals))
    if (dtype_in == FA_F64 && dtype_out == FA_F64) FA_SF(double, double);
    else if (dtype_in == FA_F64 && dtype_out == FA_F32) FA_SF(double, float);
    else if (dtype_in == FA_F32 && dtype_out == FA_F32) FA_SF(float, float);
    else if (dtype_in == FA_F32 && dtype_out == FA_F64) FA_SF(float, double);
    else {
        set_error("fa_sell_fill: bad dtypes");
        return FA_EINVAL;
    }
#undef FA_SF
    FA_LAUNCH("k_sell_fill");
    return FA_OK;
}

extern "C" int fa_sell_spmv(int dtype, const fa_sell_t* A, const void* x, void* y, fa_stream_t stream) {
    FA_REQUIRE(A && A->n > 0, "fa_sell_spmv: empty matrix");
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(((A->n + 63) / 64) * 64, 256, 1 << 16);
    timer_begin(FA_TIMER_SELL_SPMV, s);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_sell_spmv<double>, dim3(grid), dim3(256), 0, s, A->n, A->slice_ptr, A->cols,
                           static_cast<const double*>(A->vals), static_cast<const double*>(x),
                           static_cast<double*>(y));
    else if (dtype == FA_F32)
        hipLaunchKernelGGL(k_sell_spmv<float>, dim3(xcd_grid(grid)), dim3(256), 0, s, A->n, A->slice_ptr, A->cols,
                           static_cast<const float*>(A->vals), static_cast<const float*>(x), static_cast<float*>(y));
    else {
        set_error("fa_sell_spmv: bad dtype %d", dtype);
        return FA_EINVAL;
    }
    timer_end(FA_TIMER_SELL_SPMV, s);
    FA_LAUNCH("k_sell_spmv");
    return FA_OK;
}

