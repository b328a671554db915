// Dirichlet BC, CSR SpMV and vector kernels (kernel "d" building blocks).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "reduce.hpp"

namespace fa {

// ---- SpMV -------------------------------------------------------------------
template <class T, int G>
__global__ void __launch_bounds__(256) k_spmv(int64_t n, const int32_t* __restrict__ indptr,
                                             const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                             const T* __restrict__ x, T* __restrict__ y) {
    const int lane = threadIdx.x % G;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; r < n;
         r += (int64_t)gridDim.x * blockDim.x / G) {
        const int32_t k1 = indptr[r + 1];
        T s = T(0);
        for (int32_t k = indptr[r] + lane; k < k1; k += G) s += vals[k] * x[indices[k]];
        s = group_sum<G>(s);
        if (lane == 0) y[r] = s;
    }
}

template <class T>
static int spmv_launch(int64_t n, const int32_t* indptr, const int32_t* indices, const T* vals, const T* x, T* y,
                       int G, hipStream_t s) {
    const unsigned grid = grid_for(n * G, 256, 1 << 20);
    switch (G) {
        case 4: hipLaunchKernelGGL((k_spmv<T, 4>), dim3(grid), dim3(256), 0, s, n, indptr, indices, vals, x, y); break;
        case 8: hipLaunchKernelGGL((k_spmv<T, 8>), dim3(grid), dim3(256), 0, s, n, indptr, indices, vals, x, y); break;
        case 16: hipLaunchKernelGGL((k_spmv<T, 16>), dim3(grid), dim3(256), 0, s, n, indptr, indices, vals, x, y); break;
        case 32: hipLaunchKernelGGL((k_spmv<T, 32>), dim3(grid), dim3(256), 0, s, n, indptr, indices, vals, x, y); break;
        default: hipLaunchKernelGGL((k_spmv<T, 64>), dim3(grid), dim3(256), 0, s, n, indptr, indices, vals, x, y); break;
    }
    FA_LAUNCH("k_spmv");
    return FA_OK;
}

// Y = A X for nv right-hand sides stored row-major (X[col*nv + v]), the
// layout of SciPy's csr_matvecs; G lanes per row, each entry's columns of X
// read by one lane, vectors taken VB at a time
template <class T, int G, int VB>
__global__ void __launch_bounds__(256) k_spmm(int64_t n, int64_t nv, const int32_t* __restrict__ indptr,
                                             const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                             const T* __restrict__ x, T* __restrict__ y) {
    const int lane = threadIdx.x % G;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; r < n;
         r += (int64_t)gridDim.x * blockDim.x / G) {
        const int32_t k0 = indptr[r], k1 = indptr[r + 1];
        for (int64_t v0 = 0; v0 < nv; v0 += VB) {
            T s[VB];
#pragma unroll
            for (int u = 0; u < VB; ++u) s[u] = T(0);
            for (int32_t k = k0 + lane; k < k1; k += G) {
                const T a = vals[k];
                const T* xr = x + static_cast<int64_t>(indices[k]) * nv + v0;
#pragma unroll
                for (int u = 0; u < VB; ++u)
                    if (v0 + u < nv) s[u] += a * xr[u];
            }
#pragma unroll
            for (int u = 0; u < VB; ++u) {
                const T t = group_sum<G>(s[u]);
                if (lane == 0 && v0 + u < nv) y[r * nv + v0 + u] = t;
            }
        }
    }
}

static int lanes_for(int64_t n, int64_t nnz) {
    const double avg = n > 0 ? double(nnz) / double(n) : 1.0;
    if (avg <= 6) return 4;
    if (avg <= 12) return 8;
    if (avg <= 24) return 16;
    if (avg <= 48) return 32;
    return 64;
}

// ---- Dirichlet BC (G lanes per row) ------------------------------------------------
constexpr int BCG = 16;

__global__ void __launch_bounds__(256) k_bc_count(int64_t n, const int32_t* __restrict__ indptr,
                                                 const int32_t* __restrict__ indices, const uint8_t* __restrict__ bd,
                                                 int32_t* __restrict__ cnt) {
    const int lane = threadIdx.x % BCG;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / BCG; r < n;
         r += (int64_t)gridDim.x * blockDim.x / BCG) {
        int c = 0;
        if (!bd[r])
            for (int32_t k = indptr[r] + lane; k < indptr[r + 1]; k += BCG) c += bd[indices[k]] ? 0 : 1;
        c = group_sum<BCG>(c);
        if (lane == 0) cnt[r] = bd[r] ? 1 : c;
    }
}

// kept entries keep their order; nidx == NULL refills values only; U: the
// type of the filtered values (an FP32 copy of an FP64 matrix is rounded once)
template <class T, class U>
__global__ void __launch_bounds__(256) k_bc_fill(int64_t n, const int32_t* __restrict__ indptr,
                                                const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                                const uint8_t* __restrict__ bd, const int32_t* __restrict__ nptr,
                                                int32_t* __restrict__ nidx, U* __restrict__ nval) {
    // two of a group's rows at a time, every load of a row issued before any
    // of its results is needed (row bounds with the boundary flag, values
    // with the column indices): three rounds of loads per pair of rows
    constexpr int R = 2;
    const int lane = threadIdx.x % BCG;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x / BCG;
    for (int64_t r0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / BCG; r0 < n; r0 += R * stride) {
        // row data of both rows (a row past n reads row n - 1 and is skipped
        // below): all loads issued, then used
        int32_t o0[R], k0[R], k1[R];
        bool rb[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t r = min(r0 + u * stride, n - 1);
            o0[u] = nptr[r];
            rb[u] = bd[r] != 0;
            k0[u] = indptr[r];
            k1[u] = indptr[r + 1];
        }
        // the first BCG entries of both rows: column indices and values, then
        // the columns' boundary flags, all loads issued before the results
        // are used (positions clamped into the row; none for an empty row)
        int32_t c[R];
        T v[R];
        bool all = true;
#pragma unroll
        for (int u = 0; u < R; ++u) all = all && k1[u] > k0[u];
        if (all) {  // the usual case: no empty row, the loads in one straight run
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int32_t k = k0[u] + min(lane, k1[u] - k0[u] - 1);
                c[u] = indices[k];
                v[u] = vals[k];
            }
        } else {
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int32_t k = k0[u] + min(lane, k1[u] - k0[u] - 1);
                c[u] = 0;
                v[u] = T(0);
                if (k1[u] > k0[u]) {
                    c[u] = indices[k];
                    v[u] = vals[k];
                }
            }
        }
        bool cb[R];
#pragma unroll
        for (int u = 0; u < R; ++u) cb[u] = bd[c[u]] != 0;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t r = r0 + u * stride;
            if (r >= n) break;  // the same for the group's lanes
            if (rb[u]) {
                if (lane == 0) {
                    if (nidx) nidx[o0[u]] = static_cast<int32_t>(r);
                    nval[o0[u]] = U(1);
                }
                continue;
            }
            int32_t base = o0[u];
            for (int32_t kb = k0[u]; kb < k1[u]; kb += BCG) {
                const int32_t k = kb + lane;
                int32_t cc = c[u];
                T vv = v[u];
                bool keep = false;
                if (k < k1[u]) {
                    if (kb != k0[u]) {  // entries past the first BCG: loaded here
                        cc = indices[k];
                        vv = vals[k];
                        keep = !bd[cc];
                    } else {
                        keep = !cb[u];
                    }
                }
                const int inc = group_inclusive_scan<BCG>(keep ? 1 : 0);
                if (keep) {
                    if (nidx) nidx[base + inc - 1] = cc;
                    nval[base + inc - 1] = static_cast<U>(vv);
                }
                base += __shfl(inc, BCG - 1, BCG);  // group total
            }
        }
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_bc_rhs(int64_t n, const int32_t* __restrict__ indptr,
                                               const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                               const uint8_t* __restrict__ bd, const T* __restrict__ uh,
                                               const T* __restrict__ f, T* __restrict__ out) {
    const int lane = threadIdx.x % BCG;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / BCG; r < n;
         r += (int64_t)gridDim.x * blockDim.x / BCG) {
        T s = T(0);
        if (!bd[r])
            for (int32_t k = indptr[r] + lane; k < indptr[r + 1]; k += BCG) s += vals[k] * uh[indices[k]];
        s = group_sum<BCG>(s);
        if (lane == 0) out[r] = bd[r] ? uh[r] : f[r] - s;
    }
}

// k_bc_fill (values only) and k_bc_rhs in one pass over A: the same lane
// order for both, so the filtered values and f - A g equal the two separate
// kernels' bit for bit
template <class T>
__global__ void __launch_bounds__(256) k_bc_apply(int64_t n, const int32_t* __restrict__ indptr,
                                                 const int32_t* __restrict__ indices, const T* __restrict__ vals,
                                                 const uint8_t* __restrict__ bd, const int32_t* __restrict__ nptr,
                                                 T* __restrict__ nval, const T* __restrict__ uh,
                                                 const T* __restrict__ f, T* __restrict__ out) {
    const int lane = threadIdx.x % BCG;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / BCG; r < n;
         r += (int64_t)gridDim.x * blockDim.x / BCG) {
        const int32_t o0 = nptr[r];
        if (bd[r]) {
            if (lane == 0) {
                nval[o0] = T(1);
                out[r] = uh[r];
            }
            continue;
        }
        int32_t base = o0;
        T acc = T(0);
        const int32_t k1 = indptr[r + 1];
        for (int32_t k0 = indptr[r]; k0 < k1; k0 += BCG) {
            const int32_t k = k0 + lane;
            bool keep = false;
            T v = T(0);
            if (k < k1) {
                const int32_t c = indices[k];
                v = vals[k];
                acc += v * uh[c];
                keep = !bd[c];
            }
            const int inc = group_inclusive_scan<BCG>(keep ? 1 : 0);
            if (keep) nval[base + inc - 1] = v;
            base += __shfl(inc, BCG - 1, BCG);
        }
        acc = group_sum<BCG>(acc);
        if (lane == 0) out[r] = f[r] - acc;
    }
}

// Pattern-preserving form of k_bc_apply: the values stay on A's pattern, a
// boundary row becomes the unit row (1 on the diagonal, 0 elsewhere) and a
// boundary column of an interior row becomes 0; f - A g is formed in the same
// lane order as k_bc_apply, so it equals that kernel's bit for bit.  Each
// entry is read and written by the same lane, so nval may be vals itself.
template <class T>
__global__ void __launch_bounds__(256) k_bc_keep(int64_t n, const int32_t* __restrict__ indptr,
                                                const int32_t* __restrict__ indices, const T* vals,
                                                const uint8_t* __restrict__ bd, T* nval, const T* __restrict__ uh,
                                                const T* __restrict__ f, T* __restrict__ out) {
    const int lane = threadIdx.x % BCG;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / BCG; r < n;
         r += (int64_t)gridDim.x * blockDim.x / BCG) {
        const int32_t k1 = indptr[r + 1];
        if (bd[r]) {
            for (int32_t k = indptr[r] + lane; k < k1; k += BCG) nval[k] = indices[k] == r ? T(1) : T(0);
            if (lane == 0) out[r] = uh[r];
            continue;
        }
        T acc = T(0);
        for (int32_t k0 = indptr[r]; k0 < k1; k0 += BCG) {
            const int32_t k = k0 + lane;
            if (k < k1) {
                const int32_t c = indices[k];
                const T v = vals[k];
                acc += v * uh[c];
                nval[k] = bd[c] ? T(0) : v;
            }
        }
        acc = group_sum<BCG>(acc);
        if (lane == 0) out[r] = f[r] - acc;
    }
}

// ---- vector kernels --------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(RED_BLOCK) k_dot(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                  double* __restrict__ partials) {
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += double(x[i]) * double(y[i]);
    store_partial(block_sum(s), partials);
}

__global__ void __launch_bounds__(RED_BLOCK) k_dot_finish(const double* __restrict__ partials, int nparts,
                                                         double* __restrict__ out) {
    const double t = finish_sum(partials, nparts);
    if (threadIdx.x == 0) *out = t;
}

template <class T>
__global__ void k_axpby(int64_t n, double a, const T* __restrict__ x, double b, T* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = static_cast<T>(a * double(x[i]) + b * double(y[i]));
}

// x += d off the Dirichlet rows (the correction step of iterative refinement:
// boundary values stay exactly those of the boundary condition)
template <class D>
__global__ void k_refine_update(int64_t n, double* __restrict__ x, const D* __restrict__ d,
                                const uint8_t* __restrict__ bd) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!bd[i]) x[i] += static_cast<double>(d[i]);
}

template <class A, class B>
__global__ void k_cast(int64_t n, const A* __restrict__ x, B* __restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = static_cast<B>(x[i]);
}

}  // namespace fa

using namespace fa;

extern "C" int fa_spmv(int dtype, int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                       const void* values, const void* x, void* y, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_spmv: bad dtype %d", dtype);
    if (n == 0) return FA_OK;
    const int G = tuned(FA_TUNE_SPMV_LANES, fa_tune_key(FA_TUNE_SPMV_LANES, n, nnz), lanes_for(n, nnz));
    if (dtype == FA_F64)
        return spmv_launch<double>(n, indptr, indices, static_cast<const double*>(values),
                                   static_cast<const double*>(x), static_cast<double*>(y), G, S(stream));
    return spmv_launch<float>(n, indptr, indices, static_cast<const float*>(values), static_cast<const float*>(x),
                              static_cast<float*>(y), G, S(stream));
}

extern "C" int fa_spmm(int dtype, int64_t n, int64_t nnz, int64_t nvec, const int32_t* indptr, const int32_t* indices,
                       const void* values, const void* x, void* y, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_spmm: bad dtype %d", dtype);
    FA_REQUIRE(nvec >= 1, "fa_spmm: nvec must be >= 1");
    if (n == 0) return FA_OK;
    const unsigned grid = grid_for(n * 8, 256, 1 << 20);
    if (dtype == FA_F64)
        hipLaunchKernelGGL((k_spmm<double, 8, 4>), dim3(grid), dim3(256), 0, S(stream), n, nvec, indptr, indices,
                           static_cast<const double*>(values), static_cast<const double*>(x), static_cast<double*>(y));
    else
        hipLaunchKernelGGL((k_spmm<float, 8, 4>), dim3(grid), dim3(256), 0, S(stream), n, nvec, indptr, indices,
                           static_cast<const float*>(values), static_cast<const float*>(x), static_cast<float*>(y));
    FA_LAUNCH("k_spmm");
    return FA_OK;
}

extern "C" size_t fa_dirichlet_work_bytes(int64_t n) {
    return align_up((n + 1) * sizeof(int32_t)) + exclusive_scan_work_bytes(n) + 1024;
}

extern "C" int fa_dirichlet_count(int64_t n, const int32_t* indptr, const int32_t* indices, const uint8_t* is_bd,
                                  int32_t* new_indptr, int64_t* new_nnz_host, void* work, size_t work_bytes,
                                  fa_stream_t stream) {
    FA_REQUIRE(n > 0, "fa_dirichlet_count: empty matrix");
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    int32_t* cnt = ar.take<int32_t>(n + 1);
    size_t sb = exclusive_scan_work_bytes(n);
    void* sw = ar.take<char>(sb);
    if (!cnt || !sw) {
        set_error("fa_dirichlet_count: workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_bc_count, dim3(grid_for(n * BCG, 256, 1 << 16)), dim3(256), 0, s, n, indptr, indices, is_bd,
                       cnt);
    FA_LAUNCH("k_bc_count");
    int st = exclusive_scan_i32(cnt, new_indptr, n, sw, sb, s);
    if (st != FA_OK) return st;
    int32_t nnz = 0;
    FA_HIP(hipMemcpyAsync(&nnz, new_indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (new_nnz_host) *new_nnz_host = nnz;
    return FA_OK;
}

extern "C" int fa_dirichlet_fill(int dtype, int dtype_out, int64_t n, const int32_t* indptr, const int32_t* indices,
                                 const void* values, const uint8_t* is_bd, const int32_t* new_indptr,
                                 int32_t* new_indices, void* new_values, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_dirichlet_fill: bad dtype %d", dtype);
    FA_REQUIRE(dtype_out == dtype || (dtype == FA_F64 && dtype_out == FA_F32),
               "fa_dirichlet_fill: values of dtype %d cannot be stored as dtype %d", dtype, dtype_out);
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n * BCG, 256, 1 << 16);
    if (dtype == FA_F64 && dtype_out == FA_F64)
        hipLaunchKernelGGL((k_bc_fill<double, double>), dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), is_bd, new_indptr, new_indices,
                           static_cast<double*>(new_values));
    else if (dtype == FA_F64)
        hipLaunchKernelGGL((k_bc_fill<double, float>), dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), is_bd, new_indptr, new_indices,
                           static_cast<float*>(new_values));
    else
        hipLaunchKernelGGL((k_bc_fill<float, float>), dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const float*>(values), is_bd, new_indptr, new_indices,
                           static_cast<float*>(new_values));
    FA_LAUNCH("k_bc_fill");
    return FA_OK;
}

extern "C" int fa_dirichlet_rhs(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices,
                                const void* values, const uint8_t* is_bd, const void* uh, const void* f, void* f_out,
                                fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_dirichlet_rhs: bad dtype %d", dtype);
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n * BCG, 256, 1 << 16);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_bc_rhs<double>, dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), is_bd, static_cast<const double*>(uh),
                           static_cast<const double*>(f), static_cast<double*>(f_out));
    else
        hipLaunchKernelGGL(k_bc_rhs<float>, dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const float*>(values), is_bd, static_cast<const float*>(uh),
                           static_cast<const float*>(f), static_cast<float*>(f_out));
    FA_LAUNCH("k_bc_rhs");
    return FA_OK;
}

extern "C" int fa_dirichlet_apply(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices,
                                  const void* values, const uint8_t* is_bd, const int32_t* new_indptr,
                                  void* new_values, const void* uh, const void* f, void* f_out, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_dirichlet_apply: bad dtype %d", dtype);
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n * BCG, 256, 1 << 16);
    timer_begin(FA_TIMER_BC_APPLY, s);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_bc_apply<double>, dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), is_bd, new_indptr, static_cast<double*>(new_values),
                           static_cast<const double*>(uh), static_cast<const double*>(f), static_cast<double*>(f_out));
    else
        hipLaunchKernelGGL(k_bc_apply<float>, dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const float*>(values), is_bd, new_indptr, static_cast<float*>(new_values),
                           static_cast<const float*>(uh), static_cast<const float*>(f), static_cast<float*>(f_out));
    timer_end(FA_TIMER_BC_APPLY, s);
    FA_LAUNCH("k_bc_apply");
    return FA_OK;
}

extern "C" size_t fa_reduce_work_bytes(void) { return 3 * align_up(RED_GRID * sizeof(double)) + 1024; }

extern "C" int fa_dot(int dtype, int64_t n, const void* x, const void* y, double* out, void* work,
                      size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_dot: bad dtype %d", dtype);
    Arena ar(work, work_bytes);
    double* partials = ar.take<double>(RED_GRID);
    if (!partials) {
        set_error("fa_dot: workspace too small");
        return FA_ENOMEM;
    }
    hipStream_t s = S(stream);
    const unsigned g = red_grid(n);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_dot<double>, dim3(g), dim3(RED_BLOCK), 0, s, n, static_cast<const double*>(x),
                           static_cast<const double*>(y), partials);
    else
        hipLaunchKernelGGL(k_dot<float>, dim3(g), dim3(RED_BLOCK), 0, s, n, static_cast<const float*>(x),
                           static_cast<const float*>(y), partials);
    hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(RED_BLOCK), 0, s, partials, (int)g, out);
    FA_LAUNCH("k_dot");
    return FA_OK;
}

extern "C" int fa_axpby(int dtype, int64_t n, double a, const void* x, double b, void* y, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_axpby: bad dtype %d", dtype);
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n, 256, 1 << 16);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_axpby<double>, dim3(grid), dim3(256), 0, s, n, a, static_cast<const double*>(x), b,
                           static_cast<double*>(y));
    else
        hipLaunchKernelGGL(k_axpby<float>, dim3(grid), dim3(256), 0, s, n, a, static_cast<const float*>(x), b,
                           static_cast<float*>(y));
    FA_LAUNCH("k_axpby");
    return FA_OK;
}

extern "C" int fa_cast(int dtype_in, int dtype_out, int64_t n, const void* x, void* y, fa_stream_t stream) {
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n, 256, 1 << 16);
    if (dtype_in == FA_F64 && dtype_out == FA_F32)
        hipLaunchKernelGGL((k_cast<double, float>), dim3(grid), dim3(256), 0, s, n, static_cast<const double*>(x),
                           static_cast<float*>(y));
    else if (dtype_in == FA_F32 && dtype_out == FA_F64)
        hipLaunchKernelGGL((k_cast<float, double>), dim3(grid), dim3(256), 0, s, n, static_cast<const float*>(x),
                           static_cast<double*>(y));
    else {
        set_error("fa_cast: unsupported dtypes %d -> %d", dtype_in, dtype_out);
        return FA_EINVAL;
    }
    FA_LAUNCH("k_cast");
    return FA_OK;
}


extern "C" int fa_refine_update(int dtype_d, int64_t n, double* x, const void* d, const uint8_t* is_bd,
                                fa_stream_t stream) {
    FA_REQUIRE(dtype_d == FA_F64 || dtype_d == FA_F32, "fa_refine_update: bad dtype %d", dtype_d);
    FA_REQUIRE(n >= 0 && (n == 0 || (x && d && is_bd)), "fa_refine_update: bad arguments");
    if (n == 0) return FA_OK;
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n, 256, 1 << 16);
    if (dtype_d == FA_F64)
        hipLaunchKernelGGL(k_refine_update<double>, dim3(grid), dim3(256), 0, s, n, x, static_cast<const double*>(d),
                           is_bd);
    else
        hipLaunchKernelGGL(k_refine_update<float>, dim3(grid), dim3(256), 0, s, n, x, static_cast<const float*>(d),
                           is_bd);
    FA_LAUNCH("k_refine_update");
    return FA_OK;
}

// r = b - A x off the Dirichlet rows, 0 on them, stored as Y, and the
// workgroup partial sums of r^2 (a grid fixed by n): the FP64 residual of
// iterative refinement on an assembled FP64 matrix (G lanes per row, lane
// partials by a fixed butterfly)
template <int G, class Y>
__global__ void __launch_bounds__(256) k_csr_residual(int64_t n, const int32_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ indices,
                                                     const double* __restrict__ vals, const double* __restrict__ x,
                                                     const double* __restrict__ b, const uint8_t* __restrict__ bd,
                                                     Y* __restrict__ y, double* __restrict__ part) {
    // a group's rows r, r + stride, ... taken U at a time with their loads in
    // flight together (the grid is capped, so a group walks many rows); the
    // rows, the lane order within a row and the order of the r^2 sum are
    // those of one row at a time
    constexpr int U = 4;
    const int lane = threadIdx.x % G;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x / G;
    double nrm = 0.0;
    for (int64_t r0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; r0 < n; r0 += U * stride) {
        // row data of the U rows (a row past n reads row n - 1 and counts
        // as empty): all loads issued, then used
        int32_t k0[U], k1[U];
        double bv[U];
        uint8_t dv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = min(r0 + u * stride, n - 1);
            k0[u] = indptr[r];
            k1[u] = indptr[r + 1];
            bv[u] = b[r];
            dv[u] = bd[r];
        }
        int len[U];
        int m = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            len[u] = r0 + u * stride < n ? k1[u] - k0[u] : 0;
            m = max(m, len[u]);
        }
        double s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = 0.0;
        for (int j = lane; j < m; j += G) {
            // loads for all U rows before any is used (positions clamped
            // into the row; none for an empty row), the sums masked
            double v[U], xv[U];
            int32_t c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int32_t k = k0[u] + min(j, len[u] - 1);
                v[u] = 0.0;
                c[u] = 0;
                if (len[u] > 0) {
                    v[u] = vals[k];
                    c[u] = indices[k];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j < len[u]) s[u] += v[u] * xv[u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t r = r0 + u * stride;
            if (r >= n) break;  // the same for the group's lanes
            double t = group_sum<G>(s[u]);
            t = dv[u] ? 0.0 : bv[u] - t;
            if (lane == 0) {
                y[r] = static_cast<Y>(t);
                nrm += t * t;
            }
        }
    }
    store_partial(block_sum(nrm), part);
}

__global__ void k_residual_finish(const double* __restrict__ part, int n, double* __restrict__ out) {
    const double s = finish_sum(part, n);
    if (threadIdx.x == 0) *out = s;
}

extern "C" size_t fa_csr_residual_work_bytes(void) { return align_up(RED_GRID * sizeof(double)) + 256; }

extern "C" int fa_csr_residual(int dtype_y, int64_t n, const int32_t* indptr, const int32_t* indices,
                               const double* values, const double* x, const double* b, const uint8_t* is_bd, void* y,
                               double* norm2, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(dtype_y == FA_F64 || dtype_y == FA_F32, "fa_csr_residual: bad dtype %d", dtype_y);
    FA_REQUIRE(n > 0 && indptr && indices && values && x && b && is_bd && y && norm2, "fa_csr_residual: bad arguments");
    Arena ar(work, work_bytes);
    double* part = ar.take<double>(RED_GRID);
    if (!part) {
        set_error("fa_csr_residual: workspace too small");
        return FA_ENOMEM;
    }
    hipStream_t s = S(stream);
    constexpr int G = 8;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((n * G + 255) / 256, RED_GRID));
    if (dtype_y == FA_F64)
        hipLaunchKernelGGL((k_csr_residual<G, double>), dim3(grid), dim3(256), 0, s, n, indptr, indices, values, x, b,
                           is_bd, static_cast<double*>(y), part);
    else
        hipLaunchKernelGGL((k_csr_residual<G, float>), dim3(grid), dim3(256), 0, s, n, indptr, indices, values, x, b,
                           is_bd, static_cast<float*>(y), part);
    FA_LAUNCH("k_csr_residual");
    hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(RED_BLOCK), 0, s, part, static_cast<int>(grid), norm2);
    FA_LAUNCH("k_residual_finish");
    return FA_OK;
}

extern "C" int fa_dirichlet_apply_keep(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices,
                                       const void* values, const uint8_t* is_bd, void* new_values, const void* uh,
                                       const void* f, void* f_out, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_dirichlet_apply_keep: bad dtype %d", dtype);
    FA_REQUIRE(n >= 0 && indptr && indices && values && is_bd && new_values && uh && f && f_out,
               "fa_dirichlet_apply_keep: bad arguments");
    if (n == 0) return FA_OK;
    hipStream_t s = S(stream);
    const unsigned grid = grid_for(n * BCG, 256, 1 << 16);
    timer_begin(FA_TIMER_BC_APPLY, s);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_bc_keep<double>, dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), is_bd, static_cast<double*>(new_values),
                           static_cast<const double*>(uh), static_cast<const double*>(f), static_cast<double*>(f_out));
    else
        hipLaunchKernelGGL(k_bc_keep<float>, dim3(grid), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const float*>(values), is_bd, static_cast<float*>(new_values),
                           static_cast<const float*>(uh), static_cast<const float*>(f), static_cast<float*>(f_out));
    timer_end(FA_TIMER_BC_APPLY, s);
    FA_LAUNCH("k_bc_keep");
    return FA_OK;
}
