// Kernel "d": Galerkin coarse operators, multigrid V-cycle and the PCG driver.
#include <algorithm>
#include <type_traits>
#include <climits>
#include <cmath>

#include <new>

#include <mutex>

#include "common.hpp"
#include "reduce.hpp"
#include "sell.hpp"

namespace fa {

// ---------------------------------------------------------------------------
// sort buf[0, m) ascending (padded to a power of two with INT_MAX) and
// compact its distinct values to the front; returns how many there are.
// One wavefront; m <= CAP (a power of two).  The compaction writes each value
// at a position no larger than the one it is read from, and a position is
// overwritten only after every read of it, so it runs in place.
template <int CAP>
__device__ inline int rap_sort_unique(int32_t* buf, int m) {
    const int lane = threadIdx.x;
    int P = 2;
    while (P < m) P <<= 1;
    for (int t = m + lane; t < P; t += WAVE) buf[t] = INT_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < P; t += WAVE) {
                int p = t ^ j;
                if (p > t) {
                    int32_t a = buf[t], b = buf[p];
                    if ((a > b) == ((t & k) == 0)) {
                        buf[t] = b;
                        buf[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    int count = 0;
    for (int t0 = 0; t0 < P; t0 += WAVE) {
        const int t = t0 + lane;
        const int32_t v = t < P ? buf[t] : INT_MAX;
        const int32_t prev = (t > 0 && t < P) ? buf[t - 1] : INT_MIN;
        const bool keep = v != INT_MAX && (t == 0 || prev != v);
        __syncthreads();
        const int inc = wave_inclusive_scan(keep ? 1 : 0);
        if (keep) buf[count + inc - 1] = v;
        count += __shfl(inc, WAVE - 1);
        __syncthreads();
    }
    return count;
}

// R A P: symbolic (count / fill) with one wavefront per coarse row.
// Candidates J are enumerated R-entry by R-entry (sequential), lanes over the
// entries of A's row, each lane appending the columns of its P row through a
// wave prefix sum; then bitonic sort + unique in LDS (as the pattern kernel).
// When the next batch would not fit, the list is first sorted and compacted
// to its distinct columns, so CAP bounds the distinct columns of the row plus
// one batch, not all candidates.
template <int CAP>
__global__ void __launch_bounds__(WAVE) k_rap_rows(int64_t nrows, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ri, const int32_t* __restrict__ ap,
                                                   const int32_t* __restrict__ ai, const int32_t* __restrict__ pp,
                                                   const int32_t* __restrict__ pi, int32_t* __restrict__ rowcount,
                                                   const int32_t* __restrict__ cptr, int32_t* __restrict__ cidx,
                                                   int32_t* __restrict__ max_row, int32_t* __restrict__ overflow) {
    __shared__ int32_t buf[CAP];
    const int lane = threadIdx.x;
    int32_t row_max = 0;
    for (int64_t I = blockIdx.x; I < nrows; I += gridDim.x) {
        int m = 0;
        bool ovf = false;
        for (int32_t u = rp[I]; u < rp[I + 1]; ++u) {
            const int32_t i = ri[u];
            const int32_t a0 = ap[i], a1 = ap[i + 1];
            for (int32_t t0 = a0; t0 < a1; t0 += WAVE) {
                const int32_t t = t0 + lane;
                int32_t j = -1, cnt = 0;
                if (t < a1) {
                    j = ai[t];
                    cnt = pp[j + 1] - pp[j];
                }
                const int inc = wave_inclusive_scan(cnt);
                const int tot = __shfl(inc, WAVE - 1);
                if (m + tot > CAP && !ovf) {      // make room: keep the distinct columns only
                    __syncthreads();
                    m = rap_sort_unique<CAP>(buf, m);
                }
                if (m + tot <= CAP) {
                    for (int c = 0; c < cnt; ++c) buf[m + inc - cnt + c] = pi[pp[j] + c];
                    m += tot;
                } else {
                    ovf = true;
                }
            }
        }
        __syncthreads();
        if (ovf) {
            if (lane == 0) atomicOr(overflow, 1);
            m = 0;
        }
        const int32_t count0 = rap_sort_unique<CAP>(buf, m);
        const int32_t base = cidx ? cptr[I] : 0;
        for (int t = lane; cidx && t < count0; t += WAVE) cidx[base + t] = buf[t];
        int32_t count = count0;
        if (count == 0) {  // empty coarse row (Dirichlet node): unit diagonal
            if (cidx && lane == 0) cidx[base] = static_cast<int32_t>(I);
            count = 1;
        }
        if (!cidx && lane == 0) rowcount[I] = count;
        row_max = max(row_max, count);
        __syncthreads();
    }
    if (max_row && lane == 0) atomicMax(max_row, row_max);
}

// R A P numeric: one wavefront per coarse row I.  The (R entry u, A entry t)
// pairs of the row are numbered flat (R order, then A order) and dealt to the
// lanes round-robin; lane l accumulates r_u a_t p_q into its private LDS row at
// the position of column J = col(p_q).  The row value is the sum over lanes by
// a fixed binary tree, so the result does not depend on timing.
// (T: the value type of A and C; products and sums are formed in FP64)
template <class T>
__global__ void __launch_bounds__(WAVE) k_rap_values(int64_t nrows, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ri, const double* __restrict__ rv,
                                                     const int32_t* __restrict__ ap, const int32_t* __restrict__ ai,
                                                     const T* __restrict__ av, const int32_t* __restrict__ pp,
                                                     const int32_t* __restrict__ pi, const double* __restrict__ pv,
                                                     const int32_t* __restrict__ cptr, const int32_t* __restrict__ cidx,
                                                     int lmax, T* __restrict__ cval) {
    extern __shared__ char smem[];
    int32_t* cols = reinterpret_cast<int32_t*>(smem);
    double* acc = reinterpret_cast<double*>(smem + align_up(sizeof(int32_t) * lmax, 16));  // [WAVE][lmax]
    __shared__ int32_t s_off[WAVE + 1], s_a0[WAVE];
    __shared__ double s_r[WAVE];
    const int lane = threadIdx.x;
    for (int64_t I = blockIdx.x; I < nrows; I += gridDim.x) {
        const int32_t c0 = cptr[I];
        const int L = cptr[I + 1] - c0;
        for (int t = lane; t < L; t += WAVE) cols[t] = cidx[c0 + t];
        for (int t = 0; t < L; ++t) acc[lane * lmax + t] = 0.0;
        __syncthreads();
        const int32_t r0 = rp[I], r1 = rp[I + 1];
        for (int32_t u0 = r0; u0 < r1; u0 += WAVE) {
            const int32_t u = u0 + lane;
            int len = 0;
            if (u < r1) {
                const int32_t i = ri[u];
                const int32_t a0 = ap[i];
                s_a0[lane] = a0;
                s_r[lane] = rv[u];
                len = ap[i + 1] - a0;
            }
            const int inc = wave_inclusive_scan(len);
            s_off[lane + 1] = inc;
            if (lane == 0) s_off[0] = 0;
            const int total = __shfl(inc, WAVE - 1);
            const int nu = min(WAVE, static_cast<int>(r1 - u0));
            __syncthreads();
            for (int f = lane; f < total; f += WAVE) {
                int lo = 0, hi = nu - 1;  // last R entry whose offset is <= f
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= f) lo = mid;
                    else hi = mid - 1;
                }
                const int32_t t = s_a0[lo] + (f - s_off[lo]);
                const int32_t j = ai[t];
                const double ra = s_r[lo] * static_cast<double>(av[t]);
                for (int32_t q = pp[j]; q < pp[j + 1]; ++q) {
                    const int32_t J = pi[q];
                    int a = 0, b = L;
                    while (a < b) {
                        const int mid = (a + b) >> 1;
                        if (cols[mid] < J) a = mid + 1;
                        else b = mid;
                    }
                    if (a < L && cols[a] == J) acc[lane * lmax + a] += ra * pv[q];
                }
            }
            __syncthreads();
        }
        for (int h = WAVE / 2; h > 0; h >>= 1) {  // lanes [0, h) += lanes [h, 2h)
            for (int w = lane; w < h * L; w += WAVE) {
                const int l = w / L, t = w - l * L;
                acc[l * lmax + t] += acc[(l + h) * lmax + t];
            }
            __syncthreads();
        }
        const bool empty = (r0 == r1);
        for (int t = lane; t < L; t += WAVE) cval[c0 + t] = static_cast<T>(empty ? 1.0 : acc[t]);
        __syncthreads();
    }
}

// R A P numeric for wide coarse rows (the per-lane tables above would not fit
// in LDS): one accumulator per coarse entry.  The (R entry u, A entry t)
// pairs are taken one after another in R order, then A order; the lanes
// share out the entries of P's row j(t), whose columns are distinct, so no
// two lanes add to one accumulator in a step, and every entry is summed in
// the fixed (u, t) order.
template <class T>
__global__ void __launch_bounds__(WAVE) k_rap_values_seq(int64_t nrows, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ri, const double* __restrict__ rv,
                                                         const int32_t* __restrict__ ap, const int32_t* __restrict__ ai,
                                                         const T* __restrict__ av, const int32_t* __restrict__ pp,
                                                         const int32_t* __restrict__ pi, const double* __restrict__ pv,
                                                         const int32_t* __restrict__ cptr,
                                                         const int32_t* __restrict__ cidx, int lmax,
                                                         T* __restrict__ cval) {
    extern __shared__ char smem[];
    int32_t* cols = reinterpret_cast<int32_t*>(smem);
    double* acc = reinterpret_cast<double*>(smem + align_up(sizeof(int32_t) * lmax, 16));
    const int lane = threadIdx.x;
    for (int64_t I = blockIdx.x; I < nrows; I += gridDim.x) {
        const int32_t c0 = cptr[I];
        const int L = cptr[I + 1] - c0;
        for (int t = lane; t < L; t += WAVE) {
            cols[t] = cidx[c0 + t];
            acc[t] = 0.0;
        }
        __syncthreads();
        const int32_t r0 = rp[I], r1 = rp[I + 1];
        for (int32_t u = r0; u < r1; ++u) {
            const int32_t i = ri[u];
            const double ru = rv[u];
            for (int32_t t = ap[i]; t < ap[i + 1]; ++t) {
                const int32_t j = ai[t];
                const double ra = ru * static_cast<double>(av[t]);
                for (int32_t q0 = pp[j]; q0 < pp[j + 1]; q0 += WAVE) {
                    const int32_t q = q0 + lane;
                    if (q < pp[j + 1]) {
                        const int32_t J = pi[q];
                        int a = 0, b = L;
                        while (a < b) {
                            const int mid = (a + b) >> 1;
                            if (cols[mid] < J) a = mid + 1;
                            else b = mid;
                        }
                        if (a < L && cols[a] == J) acc[a] += ra * pv[q];
                    }
                    __syncthreads();
                }
            }
        }
        const bool empty = (r0 == r1);
        for (int t = lane; t < L; t += WAVE) cval[c0 + t] = static_cast<T>(empty ? 1.0 : acc[t]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Chebyshev smoothing of D^-1 A (recurrence of Saad, Iterative Methods, 12.3):
//   FROM_ZERO: x = c2 * D^-1 b, d = x
//   otherwise: res = b - A x_in; d = c1 d + c2 D^-1 res; x_out = x_in + d
//   (FIRST: the c1 d term is absent, so d need not be initialised)
//
// The last step of the finest level's post-smoothing may also produce the
// partial sums of r . x_out for the outer CG (DOT): r is the outer residual
// (FP64 or FP32), one partial per workgroup; such a launch uses the
// reduction grid red_grid(n) of RED_BLOCK threads, so the sum order is the
// one of a separate dot kernel over the same vectors.
struct DotOut {
    const void* r;
    double* partials;
    int r_f64;
};

__device__ inline double dot_term(const DotOut& o, int64_t row, double v) {
    const double rv = o.r_f64 ? static_cast<const double*>(o.r)[row] : static_cast<const float*>(o.r)[row];
    return rv * v;
}

template <class T, bool FROM_ZERO, bool FIRST, bool DOT = false>
__global__ void __launch_bounds__(256) k_cheb(int64_t n, const int32_t* __restrict__ sp, const int32_t* __restrict__ sc,
                                              const T* __restrict__ sv, const T* __restrict__ dinv,
                                              const T* __restrict__ b, const T* __restrict__ xin, T* __restrict__ xout,
                                              T* __restrict__ d, T c1, T c2, const int* __restrict__ done,
                                              DotOut dot = DotOut{}) {
    if (done && *done) return;
    double acc = 0.0;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        if (FROM_ZERO) {
            if (row < n) {
                const T v = c2 * dinv[row] * b[row];
                d[row] = v;
                xout[row] = v;
            }
        } else {
            const T ax = sell_row_dot<T, T>(sp, sc, sv, row, xin);
            if (row < n) {
                const T res = b[row] - ax;
                const T dn = (FIRST ? T(0) : c1 * d[row]) + c2 * dinv[row] * res;
                d[row] = dn;
                const T xn = xin[row] + dn;
                xout[row] = xn;
                if (DOT) acc += dot_term(dot, row, double(xn));
            }
        }
    }
    if (DOT) store_partial(block_sum(acc), dot.partials);
}

// First two steps of a sweep that starts from zero, in one pass: the first
// step's x0 = c0 D^-1 b is formed at every gathered column instead of being
// stored, then res = b - A x0, d = c1 x0 + c2 D^-1 res, x_out = x0 + d.  Same
// operations and rounding as the two separate steps.
template <class T>
__global__ void __launch_bounds__(256) k_cheb_z(int64_t n, const int32_t* __restrict__ sp,
                                                const int32_t* __restrict__ sc, const T* __restrict__ sv,
                                                const T* __restrict__ dinv, const T* __restrict__ b,
                                                T* __restrict__ xout, T* __restrict__ d, T c0, T c1, T c2,
                                                const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T ax = sell_row_dot_f<T, ROW_BATCH>(sp, sc, sv, row, [=](int32_t j) { return c0 * dinv[j] * b[j]; });
        if (row < n) {
            const T x0 = c0 * dinv[row] * b[row];
            const T res = b[row] - ax;
            const T dn = c1 * x0 + c2 * dinv[row] * res;
            d[row] = dn;
            xout[row] = x0 + dn;
        }
    }
}

// The same steps on the compressed D^-1 A (FP32 cycles):
//   cheb:     d = c1 d + c2 (dinv b - (D^-1 A) x_in); x_out = x_in + d
//   cheb_z:   the first two steps from zero (x0 = c0 dinv b formed per column)
//   residual: r = b - diag * ((D^-1 A) x)
// CS: the column source (int16 offsets, or the full SELL's int32 columns)
template <bool FIRST, bool DOT, class CS>
__global__ void __launch_bounds__(256) k_cheb16(int64_t n, const int32_t* __restrict__ sp, CS so,
                                               const _Float16* __restrict__ sv,
                                               const float* __restrict__ dinv, const float* __restrict__ b,
                                               const float* __restrict__ xin, float* __restrict__ xout,
                                               float* __restrict__ d, float c1, float c2,
                                               const int* __restrict__ done, DotOut dot = DotOut{}) {
    if (done && *done) return;
    double acc = 0.0;
    const unsigned b0 = CS::xcd ? xcd_block() : blockIdx.x;
    for (int64_t row = b0 * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const float ax = sell16_row_dot(sp, so, sv, row, n, xin);
        if (row < n) {
            const float dn = (FIRST ? 0.f : c1 * d[row]) + c2 * (dinv[row] * b[row] - ax);
            d[row] = dn;
            const float xn = xin[row] + dn;
            xout[row] = xn;
            if (DOT) acc += dot_term(dot, row, double(xn));
        }
    }
    if (DOT) store_partial_at(block_sum(acc), dot.partials, b0);
}

template <class CS>
__global__ void __launch_bounds__(256) k_cheb16_z(int64_t n, const int32_t* __restrict__ sp, CS so,
                                                 const _Float16* __restrict__ sv,
                                                 const float* __restrict__ dinv, const float* __restrict__ b,
                                                 float* __restrict__ xout, float* __restrict__ d, float c0, float c1,
                                                 float c2, const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t row = (CS::xcd ? xcd_block() : blockIdx.x) * (int64_t)blockDim.x + threadIdx.x;
         row < ((n + 63) & ~63LL); row += (int64_t)gridDim.x * blockDim.x) {
        const float ax = sell16_row_dot_f(sp, so, sv, row, n, [=](int64_t j) { return c0 * dinv[j] * b[j]; });
        if (row < n) {
            const float x0 = c0 * dinv[row] * b[row];
            const float dn = c1 * x0 + c2 * (dinv[row] * b[row] - ax);
            d[row] = dn;
            xout[row] = x0 + dn;
        }
    }
}

template <class CS>
__global__ void __launch_bounds__(256) k_residual16(int64_t n, const int32_t* __restrict__ sp, CS so,
                                                   const _Float16* __restrict__ sv,
                                                   const float* __restrict__ diag, const float* __restrict__ b,
                                                   const float* __restrict__ x, float* __restrict__ r,
                                                   const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t row = (CS::xcd ? xcd_block() : blockIdx.x) * (int64_t)blockDim.x + threadIdx.x;
         row < ((n + 63) & ~63LL); row += (int64_t)gridDim.x * blockDim.x) {
        const float ax = sell16_row_dot(sp, so, sv, row, n, x);
        if (row < n) r[row] = b[row] - diag[row] * ax;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_residual(int64_t n, const int32_t* __restrict__ sp,
                                                 const int32_t* __restrict__ sc, const T* __restrict__ sv,
                                                 const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ r,
                                                 const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T ax = sell_row_dot<T, T>(sp, sc, sv, row, x);
        if (row < n) r[row] = b[row] - ax;
    }
}

// y = A x (ADD: y += A x)
template <class T, bool ADD>
__global__ void __launch_bounds__(256) k_sell_apply(int64_t n, const int32_t* __restrict__ sp,
                                                   const int32_t* __restrict__ sc, const T* __restrict__ sv,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T v = sell_row_dot<T, T>(sp, sc, sv, row, x);
        if (row < n) y[row] = ADD ? y[row] + v : v;
    }
}

// ---- transfers of nested from_box grids, from the grid -------------------------
// Fine grid (nx, ny, nz) (even counts), coarse grid (nx/2, ny/2, nz/2).  The
// P1 prolongation gives fine node f the coarse node f/2 (weight 1) when every
// coordinate of f is even, else 1/2 of each end of the Kuhn edge through f,
// floor(f/2) and floor(f/2) + (f mod 2); rows of Dirichlet fine nodes and
// columns of Dirichlet coarse nodes are absent.  R = P^T.  Each sum adds its
// terms in the SELL rows' column order (ascending node index), and the
// weights 1/2 and 1 make every product exact, so the values equal the SELL
// transfers'.
__device__ inline int64_t grid_node(int x, int y, int z, int ny, int nz) {
    return (static_cast<int64_t>(x) * (ny + 1) + y) * (nz + 1) + z;
}

// b_c = R r over the coarse nodes: the fine nodes 2I + o, o = 0 (weight 1)
// or a Kuhn edge direction (weight 1/2), in lexicographic order of o
template <class T>
__global__ void __launch_bounds__(256) k_box_restrict(int mx, int my, int mz, const uint8_t* __restrict__ bd_f,
                                                     const uint8_t* __restrict__ bd_c, const T* __restrict__ r,
                                                     T* __restrict__ bc, const int* __restrict__ done) {
    if (done && *done) return;
    const int nx = 2 * mx, ny = 2 * my, nz = 2 * mz;
    // node counts of the levels are below 2^31 (int32 SELL indices): 32-bit
    // index arithmetic
    const uint32_t nmz = mz + 1, nmyz = static_cast<uint32_t>(my + 1) * nmz, nc = static_cast<uint32_t>(mx + 1) * nmyz;
    for (uint32_t I = xcd_block() * blockDim.x + threadIdx.x; I < nc; I += gridDim.x * blockDim.x) {
        const int ix = static_cast<int>(I / nmyz);
        const uint32_t rem = I - ix * nmyz;
        const int iy = static_cast<int>(rem / nmz);
        const int iz = static_cast<int>(rem - iy * nmz);
        const int fx = 2 * ix, fy = 2 * iy, fz = 2 * iz;
        T v[15];
        uint8_t fb[15];
        bool in[15];
        int q = 0;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    const bool pos = dx >= 0 && dy >= 0 && dz >= 0, neg = dx <= 0 && dy <= 0 && dz <= 0;
                    if (!(pos || neg)) continue;
                    const int x = fx + dx, y = fy + dy, z = fz + dz;
                    in[q] = x >= 0 && x <= nx && y >= 0 && y <= ny && z >= 0 && z <= nz;
                    const int64_t f = in[q] ? grid_node(x, y, z, ny, nz) : grid_node(fx, fy, fz, ny, nz);
                    v[q] = r[f];
                    fb[q] = bd_f[f];
                    ++q;
                }
        const uint8_t cb = bd_c[I];
        T acc = T(0);
        q = 0;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    const bool pos = dx >= 0 && dy >= 0 && dz >= 0, neg = dx <= 0 && dy <= 0 && dz <= 0;
                    if (!(pos || neg)) continue;
                    const T w = (dx == 0 && dy == 0 && dz == 0) ? T(1) : T(0.5);
                    if (in[q] && !fb[q]) acc += w * v[q];
                    ++q;
                }
        bc[I] = cb ? T(0) : acc;
    }
}

// an empty asm that takes the values as register operands: the loads that
// produce them are all issued before this point, so none is sunk into a
// branch behind another's wait
template <class V>
__device__ inline void pin_load(V& v) {
    asm volatile("" : "+v"(v));
}
template <class... V>
__device__ inline void pin_loads(V&... v) {
    (pin_load(v), ...);
}

// x_f += P x_c over the fine nodes
template <class T>
__global__ void __launch_bounds__(256) k_box_prolong(int nx, int ny, int nz, const uint8_t* __restrict__ bd_f,
                                                    const uint8_t* __restrict__ bd_c, const T* __restrict__ xc,
                                                    T* __restrict__ xf, const int* __restrict__ done) {
    if (done && *done) return;
    const int my = ny / 2, mz = nz / 2;
    // node counts below 2^31: 32-bit index arithmetic
    const uint32_t nz1 = nz + 1, nyz = static_cast<uint32_t>(ny + 1) * nz1, nf = static_cast<uint32_t>(nx + 1) * nyz;
    for (uint32_t f = xcd_block() * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        const int x = static_cast<int>(f / nyz);
        const uint32_t rem = f - x * nyz;
        const int y = static_cast<int>(rem / nz1);
        const int z = static_cast<int>(rem - y * nz1);
        const int ox = x & 1, oy = y & 1, oz = z & 1;
        const bool two = ox | oy | oz;
        const int64_t c0 = grid_node(x >> 1, y >> 1, z >> 1, my, mz);
        const int64_t c1 = two ? grid_node((x >> 1) + ox, (y >> 1) + oy, (z >> 1) + oz, my, mz) : c0;
        int fb = bd_f[f], b0 = bd_c[c0], b1 = bd_c[c1];
        T x0 = xc[c0], x1 = xc[c1];
        const T y0 = xf[f];
        // every address is valid: the six loads go out together, before the
        // branches on the flags (which would otherwise wait on bd_f first)
        pin_loads(fb, b0, b1, x0, x1);
        const T w = two ? T(0.5) : T(1);
        T acc = T(0);
        if (!fb) {
            if (!b0) acc += w * x0;
            if (two && !b1) acc += w * x1;
        }
        xf[f] = y0 + acc;
    }
}

// ---------------------------------------------------------------------------
// coarsest level: dense Cholesky (one workgroup)
__global__ void __launch_bounds__(256) k_csr_to_dense(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                     double* __restrict__ dense) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x)
        dense[e] = 0.0;
}

__global__ void __launch_bounds__(256) k_csr_scatter_dense(int64_t n, const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ idx,
                                                          const double* __restrict__ val, double* __restrict__ dense) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) dense[r * n + idx[k]] += val[k];
}

// In-place Gauss-Jordan inversion of an SPD matrix (no pivoting needed), one
// workgroup.  Step k: scale row k by 1/a_kk, stage row k and column k, then the
// rank-1 update of every other row.
constexpr int COARSE_MAX = 2048;
constexpr int COARSE_LDS_MAX = 89;  // (n*n + 2n) doubles fit in 64 KiB of LDS
constexpr int COARSE_ONE_BLOCK = 256;  // dense coarse apply by one workgroup up to this size

__device__ inline void gj_step(int n, int k, double* m, double* rowk, double* colk) {
    const double piv = m[k * n + k];
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double v = ((j == k) ? 1.0 : m[k * n + j]) / piv;
        m[k * n + j] = v;
        rowk[j] = v;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) colk[i] = (i == k) ? 0.0 : m[i * n + k];
    __syncthreads();
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int i = t / n, j = t - i * n;
        if (i == k) continue;
        const double f = colk[i];
        m[t] = (j == k) ? -f * rowk[j] : m[t] - f * rowk[j];
    }
    __syncthreads();
}

// whole matrix in LDS (n <= COARSE_LDS_MAX)
__global__ void __launch_bounds__(256) k_spd_inverse_lds(int n, double* __restrict__ a) {
    extern __shared__ double lds[];
    double* m = lds;
    double* rowk = lds + n * n;
    double* colk = rowk + n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) m[t] = a[t];
    __syncthreads();
    for (int k = 0; k < n; ++k) gj_step(n, k, m, rowk, colk);
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) a[t] = m[t];
}

// Blocked Gauss-Jordan for COARSE_LDS_MAX < n <= COARSE_MAX, pivot blocks of
// GJB columns K = [k0, k0+b), four launches per block:
//   (1) A_KK <- P = A_KK^-1              (one workgroup, LDS)
//   (2) A_Kj <- P A_Kj,   j outside K    (row panel)
//   (3) A_ij <- A_ij - A_iK A_Kj, i, j outside K (old A_iK, new A_Kj)
//   (4) A_iK <- -A_iK P,  i outside K    (column panel)
// which is the scalar step of gj_step applied to blocks; after the last
// block the matrix holds its inverse.  SPD: every pivot block is invertible.
constexpr int GJB = 32;

__global__ void __launch_bounds__(256) k_gj_pivot(int n, int k0, int b, double* __restrict__ a) {
    __shared__ double m[GJB * GJB], rowk[GJB], colk[GJB];
    for (int t = threadIdx.x; t < b * b; t += blockDim.x) m[t] = a[(int64_t)(k0 + t / b) * n + k0 + t % b];
    __syncthreads();
    for (int k = 0; k < b; ++k) gj_step(b, k, m, rowk, colk);
    for (int t = threadIdx.x; t < b * b; t += blockDim.x) a[(int64_t)(k0 + t / b) * n + k0 + t % b] = m[t];
}

// row panel: one workgroup per block of GJB columns j outside K; the b x GJB
// sub-panel is read into LDS before any of it is overwritten
__global__ void __launch_bounds__(256) k_gj_rowpanel(int n, int k0, int b, double* __restrict__ a) {
    __shared__ double P[GJB * GJB], sub[GJB * GJB];
    for (int t = threadIdx.x; t < b * b; t += blockDim.x) P[t] = a[(int64_t)(k0 + t / b) * n + k0 + t % b];
    for (int j0 = blockIdx.x * GJB; j0 < n; j0 += gridDim.x * GJB) {
        __syncthreads();
        for (int t = threadIdx.x; t < b * GJB; t += blockDim.x) {
            const int r = t / GJB, j = j0 + t % GJB;
            sub[t] = j < n ? a[(int64_t)(k0 + r) * n + j] : 0.0;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < b * GJB; t += blockDim.x) {
            const int r = t / GJB, jj = t % GJB, j = j0 + jj;
            if (j >= n || (j >= k0 && j < k0 + b)) continue;
            double s = 0.0;
            for (int q = 0; q < b; ++q) s += P[r * b + q] * sub[q * GJB + jj];
            a[(int64_t)(k0 + r) * n + j] = s;
        }
    }
}

__global__ void __launch_bounds__(256) k_gj_trailing(int n, int k0, int b, double* __restrict__ a) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = static_cast<int>(e / n), j = static_cast<int>(e - (int64_t)i * n);
        if ((i >= k0 && i < k0 + b) || (j >= k0 && j < k0 + b)) continue;
        double s = 0.0;
        for (int t = 0; t < b; ++t) s += a[(int64_t)i * n + k0 + t] * a[(int64_t)(k0 + t) * n + j];
        a[e] -= s;
    }
}

// column panel: one workgroup per block of GJB rows i outside K
__global__ void __launch_bounds__(256) k_gj_colpanel(int n, int k0, int b, double* __restrict__ a) {
    __shared__ double P[GJB * GJB], sub[GJB * GJB];
    for (int t = threadIdx.x; t < b * b; t += blockDim.x) P[t] = a[(int64_t)(k0 + t / b) * n + k0 + t % b];
    for (int i0 = blockIdx.x * GJB; i0 < n; i0 += gridDim.x * GJB) {
        __syncthreads();
        for (int t = threadIdx.x; t < GJB * b; t += blockDim.x) {
            const int i = i0 + t / b, q = t % b;
            sub[t] = i < n ? a[(int64_t)i * n + k0 + q] : 0.0;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < GJB * b; t += blockDim.x) {
            const int ii = t / b, c = t % b, i = i0 + ii;
            if (i >= n || (i >= k0 && i < k0 + b)) continue;
            double s = 0.0;
            for (int q = 0; q < b; ++q) s += sub[ii * b + q] * P[q * b + c];
            a[(int64_t)i * n + k0 + c] = -s;
        }
    }
}

// x = Ainv b on a large coarsest level: one wavefront per row
template <class T>
__global__ void __launch_bounds__(256) k_dense_gemv(int n, const double* __restrict__ Ainv, const T* __restrict__ b,
                                                   T* __restrict__ x, const int* __restrict__ done) {
    if (done && *done) return;
    const int lane = threadIdx.x & (WAVE - 1);
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < n; i += gridDim.x * blockDim.x / WAVE) {
        const double* row = Ainv + (int64_t)i * n;
        double s = 0.0;
        for (int j = lane; j < n; j += WAVE) s += row[j] * static_cast<double>(b[j]);
        s = wave_sum(s);
        if (lane == 0) x[i] = static_cast<T>(s);
    }
}

// x = Ainv b on the coarsest level (dense matvec, one workgroup)
template <class T>
__global__ void __launch_bounds__(256) k_coarse_apply(int n, const double* __restrict__ Ainv, const T* __restrict__ b,
                                                     T* __restrict__ x, const int* __restrict__ done) {
    if (done && *done) return;
    __shared__ double bs[COARSE_MAX];
    for (int i = threadIdx.x; i < n; i += blockDim.x) bs[i] = static_cast<double>(b[i]);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double* row = Ainv + (int64_t)i * n;
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += row[j] * bs[j];
        x[i] = static_cast<T>(s);
    }
}

// ---------------------------------------------------------------------------
// power iteration on D^-1 A: w = D^-1 A v; lambda ~ (v, A v) / (v, D v); v <- w / ||w||
template <class T>
__global__ void __launch_bounds__(RED_BLOCK) k_power(int64_t n, const int32_t* __restrict__ sp,
                                                    const int32_t* __restrict__ sc, const T* __restrict__ sv,
                                                    const T* __restrict__ dinv, const T* __restrict__ v,
                                                    T* __restrict__ w, double* __restrict__ partials) {
    double vav = 0.0, vdv = 0.0, ww = 0.0;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T av = sell_row_dot<T, T>(sp, sc, sv, row, v);
        if (row < n) {
            const double vi = v[row], di = dinv[row];
            const double wi = di * static_cast<double>(av);
            w[row] = static_cast<T>(wi);
            vav += vi * static_cast<double>(av);
            vdv += vi * vi / di;
            ww += wi * wi;
        }
    }
    const double s[3] = {block_sum(vav), block_sum(vdv), block_sum(ww)};
    store_partials<3>(s, partials);
}

// out[0] = (v, A v) / (v, D v), out[1] = 1 / ||w||
__global__ void __launch_bounds__(RED_BLOCK) k_power_finish(const double* __restrict__ partials, int nparts,
                                                           double* __restrict__ out) {
    const double a = finish_sum(partials, nparts);
    const double b = finish_sum(partials + nparts, nparts);
    const double c = finish_sum(partials + 2 * nparts, nparts);
    if (threadIdx.x == 0) {
        out[0] = a / b;
        out[1] = 1.0 / sqrt(c);
    }
}

template <class T>
__global__ void k_scale(int64_t n, const T* __restrict__ w, T* __restrict__ v, const double* __restrict__ s) {
    const T f = static_cast<T>(s[1]);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = w[i] * f;
}

// pseudo-random start vector (integer hash -> [0.5, 1.5)): excites every mode,
// unlike a smooth start, which converges slowly to the top of the spectrum
template <class T>
__global__ void k_power_init(int64_t n, T* __restrict__ v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t h = static_cast<uint32_t>(i) * 2654435761u;
        h ^= h >> 16;
        h *= 2246822519u;
        h ^= h >> 13;
        v[i] = static_cast<T>(0.5 + (h & 0xffffff) / 16777216.0);
    }
}

// Jacobi data of a CSR FP64 operator, G lanes per row: dinv = 1/a_ii (1 where
// the diagonal is absent or zero) and the Gershgorin bound of D^-1 A,
// max_i sum_j |a_ij| / |a_ii|, which bounds its spectrum from above.
template <class T, int G>
__global__ void __launch_bounds__(RED_BLOCK) k_jacobi_setup(int64_t n, const int32_t* __restrict__ ptr,
                                                           const int32_t* __restrict__ idx,
                                                           const double* __restrict__ val, T* __restrict__ dinv,
                                                           double* __restrict__ partials) {
    __shared__ double ws[RED_BLOCK / WAVE];
    const int lane = threadIdx.x % G;
    double m = 0.0;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G; r < n;
         r += (int64_t)gridDim.x * blockDim.x / G) {
        double s = 0.0, d = 0.0;
        for (int32_t k = ptr[r] + lane; k < ptr[r + 1]; k += G) {
            const double v = val[k];
            s += fabs(v);
            if (idx[k] == r) d = v;
        }
        s = group_sum<G>(s);
        d = group_sum<G>(d);
        if (lane == 0) {
            dinv[r] = static_cast<T>(d != 0.0 ? 1.0 / d : 1.0);
            if (d != 0.0) m = fmax(m, s / fabs(d));
        }
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & (WAVE - 1)) == 0) ws[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < RED_BLOCK / WAVE; ++w) b = fmax(b, ws[w]);
        partials[blockIdx.x] = b;
    }
}

__global__ void __launch_bounds__(RED_BLOCK) k_max_finish(const double* __restrict__ partials, int nparts,
                                                         double* __restrict__ out) {
    __shared__ double ws[RED_BLOCK / WAVE];
    double m = 0.0;
    for (int i = threadIdx.x; i < nparts; i += RED_BLOCK) m = fmax(m, partials[i]);
    for (int o = WAVE / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & (WAVE - 1)) == 0) ws[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < RED_BLOCK / WAVE; ++w) m = fmax(m, ws[w]);
        *out = m;
    }
}

// Smoother data of a level in one pass over its FP64 SELL copy (one row per
// thread, coalesced across the slice): dinv = 1/a_ii (1 where the diagonal
// is absent or zero; the diagonal is the row's first entry in column `row`,
// real entries precede the padding), diag = 1/dinv for the compressed
// residual, the Gershgorin bound max_i sum_j |a_ij| / |a_ii| (one partial
// maximum per workgroup), and the cycle's operator values: FP16 D^-1 A on
// the compressed slices (h16) and/or the FP32 copy of the SELL values (v32).
template <class T, class V>
__global__ void __launch_bounds__(RED_BLOCK) k_smoother_setup(int64_t n, const int32_t* __restrict__ sp,
                                                             const int32_t* __restrict__ cols,
                                                             const V* __restrict__ v64, T* __restrict__ dinv,
                                                             T* __restrict__ diag, _Float16* __restrict__ h16,
                                                             float* __restrict__ v32, double* __restrict__ partials) {
    __shared__ double ws[RED_BLOCK / WAVE];
    double m = 0.0;
    const int64_t npad = (n + 63) & ~63LL;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < npad;
         row += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sl = row >> 6;
        const int32_t p0 = sp[sl] + static_cast<int32_t>(row & 63);
        const int w = (sp[sl + 1] - sp[sl]) >> 6;
        double d = 0.0, sum = 0.0;
        bool found = false;
        for (int k = 0; k < w; ++k) {
            const double v = static_cast<double>(v64[p0 + 64 * k]);
            sum += fabs(v);
            if (!found && cols[p0 + 64 * k] == row) {
                d = v;
                found = true;
            }
        }
        const double inv = d != 0.0 ? 1.0 / d : 1.0;
        if (row < n) {
            const T di = static_cast<T>(inv);
            dinv[row] = di;
            if (diag) diag[row] = T(1) / di;
            if (d != 0.0) m = fmax(m, sum / fabs(d));
        }
        if (h16 || v32) {
            for (int k = 0; k < w; ++k) {
                const double v = static_cast<double>(v64[p0 + 64 * k]);
                if (h16) h16[p0 + 64 * k] = static_cast<_Float16>(row < n ? v * inv : 0.0);
                if (v32) v32[p0 + 64 * k] = static_cast<float>(v);
            }
        }
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & (WAVE - 1)) == 0) ws[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < RED_BLOCK / WAVE; ++w) b = fmax(b, ws[w]);
        partials[blockIdx.x] = b;
    }
}

// ---------------------------------------------------------------------------
// Fused coarse cycle: the levels with at most FUSE_ROWS rows run as one
// workgroup, phases separated by block barriers, instead of ~7 launches per
// level whose cost on these sizes is the launch itself.  Same operations and
// order as the per-level kernels below.
constexpr int FUSE_MAX_LEVELS = 16;
constexpr int64_t FUSE_ROWS = 1024;  // one workgroup is latency-bound above ~1 row per thread
constexpr int FUSE_BLOCK = 1024;
constexpr size_t FUSE_LDS_MAX = 60 * 1024;  // level data kept in LDS up to this size

// entries a row dot loads in one batch in the fused coarse cycle: a whole
// coarse row (up to 16 entries) per round trip in FP32; 8 in FP64, and in the
// workgroup solves, where 16 would spill registers
template <class T>
constexpr int FUSE_BATCH = sizeof(T) == 4 ? 16 : 8;

struct FusedLevel {
    fa_sell_t A, R, P;
    const void* dinv;
    void *b, *x, *x2, *d, *r;
    double lmax;
};

struct FusedCycle {
    int nlev, degree;
    double lower, upper;
    const double* ainv;  // inverse of the coarsest operator, row-major
    int ainv_lds;        // the LDS variant copies ainv into LDS after the level data
    FusedLevel lv[FUSE_MAX_LEVELS];
};

// Smoothing of one level from the iterate in L.x (odd = false) or L.x2 (odd =
// true); returns where the result lies.  The level's two iterate buffers swap
// roles each step, so the caller keeps one bit a level instead of an array of
// pointers (a pointer array indexed by the level lives in scratch memory and
// puts one more memory round trip in front of every phase).
template <class T, int B>
__device__ inline bool blk_smooth(const FusedCycle& fc, const FusedLevel& L, const T* b, bool odd,
                                  bool from_zero) {
    T* cur = static_cast<T*>(odd ? L.x2 : L.x);
    T* nxt = static_cast<T*>(odd ? L.x : L.x2);
    const int64_t n = L.A.n, npad = (n + 63) & ~63LL;
    const T* sv = static_cast<const T*>(L.A.vals);
    const T* dinv = static_cast<const T*>(L.dinv);
    T* d = static_cast<T*>(L.d);
    const double lo = fc.lower * L.lmax, hi = fc.upper * L.lmax;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    for (int k = 0; k < fc.degree; ++k) {
        T c1, c2;
        double rho_new = rho;
        if (k == 0) {
            c1 = T(0);
            c2 = static_cast<T>(1.0 / theta);
        } else {
            rho_new = 1.0 / (2.0 * sigma - rho);
            c1 = static_cast<T>(rho_new * rho);
            c2 = static_cast<T>(2.0 * rho_new / delta);
        }
        if (k == 0 && from_zero) {
            for (int64_t row = threadIdx.x; row < n; row += blockDim.x) {
                const T v = c2 * dinv[row] * b[row];
                d[row] = v;
                cur[row] = v;
            }
        } else {
            for (int64_t row = threadIdx.x; row < npad; row += blockDim.x) {
                const T ax = sell_row_dot<T, T, B>(L.A.slice_ptr, L.A.cols, sv, row, cur);
                if (row < n) {
                    const T res = b[row] - ax;
                    const T dn = (k == 0 ? T(0) : c1 * d[row]) + c2 * dinv[row] * res;
                    d[row] = dn;
                    nxt[row] = cur[row] + dn;
                }
            }
            T* t = cur;
            cur = nxt;
            nxt = t;
            odd = !odd;
        }
        rho = rho_new;
        __syncthreads();
    }
    return odd;
}

template <class T, bool ADD, int B>
__device__ inline void blk_apply(const fa_sell_t& M, const T* x, T* y) {
    const int64_t npad = (M.n + 63) & ~63LL;
    const T* mv = static_cast<const T*>(M.vals);
    for (int64_t row = threadIdx.x; row < npad; row += blockDim.x) {
        const T v = sell_row_dot<T, T, B>(M.slice_ptr, M.cols, mv, row, x);
        if (row < M.n) y[row] = ADD ? y[row] + v : v;
    }
    __syncthreads();
}

// LDS bytes of the level data the LDS variant keeps on chip: slice offsets
// of A, R, P, dinv and the five level vectors
template <class T>
__host__ __device__ inline size_t fused_level_lds(const FusedLevel& L, bool last) {
    const int64_t n = L.A.n;
    size_t b = align_up(((n + 63) / 64 + 1) * sizeof(int32_t), 16) + 6 * align_up(n * sizeof(T), 16);
    if (!last)
        b += align_up(((L.R.n + 63) / 64 + 1) * sizeof(int32_t), 16) +
             align_up(((L.P.n + 63) / 64 + 1) * sizeof(int32_t), 16);
    return b;
}

template <class T>
__device__ inline const int32_t* lds_copy_i32(char*& p, const int32_t* src, int64_t m) {
    int32_t* dst = reinterpret_cast<int32_t*>(p);
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) dst[i] = src[i];
    p += align_up(m * sizeof(int32_t), 16);
    return dst;
}

template <class T>
__device__ inline T* lds_vec(char*& p, int64_t n, const void* src) {
    T* dst = reinterpret_cast<T*>(p);
    if (src)
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = static_cast<const T*>(src)[i];
    p += align_up(n * sizeof(T), 16);
    return dst;
}

// One V-cycle over the levels lv[0 .. fc.nlev) by the calling workgroup:
// pre-smoothing, residual and restriction down, the dense coarse solve,
// prolongation and post-smoothing up, phases separated by block barriers.
// Returns the level-0 vector that holds the result (L.x2, or L.x when level
// 0 is the coarsest).
template <class T, int B = FUSE_BATCH<T>>
__device__ inline T* fused_vcycle(const FusedCycle& fc, const FusedLevel* lv, const double* ainv) {
    const int nl = fc.nlev;
    // bit l: level l's current iterate is L.x2 (else L.x); the coarsest
    // level's solution is written to its x
    uint32_t odd = 0;
    auto cur = [&](int l) { return static_cast<T*>(((odd >> l) & 1u) ? lv[l].x2 : lv[l].x); };
    for (int l = 0; l < nl - 1; ++l) {
        const FusedLevel& L = lv[l];
        const T* b = static_cast<const T*>(L.b);
        if (blk_smooth<T, B>(fc, L, b, false, true)) odd |= 1u << l;
        const T* xl = cur(l);
        T* r = static_cast<T*>(L.r);
        const int64_t n = L.A.n, npad = (n + 63) & ~63LL;
        for (int64_t row = threadIdx.x; row < npad; row += blockDim.x) {
            const T ax = sell_row_dot<T, T, B>(L.A.slice_ptr, L.A.cols, static_cast<const T*>(L.A.vals), row, xl);
            if (row < n) r[row] = b[row] - ax;
        }
        __syncthreads();
        blk_apply<T, false, B>(L.R, r, static_cast<T*>(lv[l + 1].b));
    }
    {
        const FusedLevel& C = lv[nl - 1];
        const int n = static_cast<int>(C.A.n);
        const T* b = static_cast<const T*>(C.b);
        T* x = static_cast<T*>(C.x);
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double* row = ainv + (int64_t)i * n;
            double s = 0.0;
            int j = 0;
            // eight loads in flight, then the products added in column order
            for (; j + 8 <= n; j += 8) {
                double a[8], c[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    a[u] = row[j + u];
                    c[u] = static_cast<double>(b[j + u]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) s += a[u] * c[u];
            }
            for (; j < n; ++j) s += row[j] * static_cast<double>(b[j]);
            x[i] = static_cast<T>(s);
        }
        __syncthreads();
    }
    for (int l = nl - 2; l >= 0; --l) {
        const FusedLevel& L = lv[l];
        blk_apply<T, true, B>(L.P, cur(l + 1), cur(l));
        const bool o = blk_smooth<T, B>(fc, L, static_cast<const T*>(L.b), ((odd >> l) & 1u) != 0, false);
        odd = (odd & ~(1u << l)) | (static_cast<uint32_t>(o) << l);
    }
    return cur(0);
}

// LDS: every level vector, dinv and the slice offsets live in LDS (the
// right-hand side of the top level is copied in, its result copied out to
// `out`), so each phase waits on one global latency (the matrix entries)
// instead of three; the arithmetic is that of the global variant.
template <class T, bool LDS>
__global__ void __launch_bounds__(FUSE_BLOCK) k_fused_cycle(FusedCycle fc, void* out, const int* __restrict__ done) {
    if (done && *done) return;
    __shared__ FusedLevel slv[LDS ? FUSE_MAX_LEVELS : 1];
    const int nl = fc.nlev;
    const FusedLevel* lv = fc.lv;
    const double* ainv = fc.ainv;
    if (LDS) {
        extern __shared__ char smem[];
        char* p = smem;
        for (int l = 0; l < nl; ++l) {
            FusedLevel L = fc.lv[l];
            const int64_t n = L.A.n;
            L.A.slice_ptr = lds_copy_i32<T>(p, L.A.slice_ptr, (n + 63) / 64 + 1);
            if (l < nl - 1) {
                L.R.slice_ptr = lds_copy_i32<T>(p, L.R.slice_ptr, (L.R.n + 63) / 64 + 1);
                L.P.slice_ptr = lds_copy_i32<T>(p, L.P.slice_ptr, (L.P.n + 63) / 64 + 1);
            }
            L.dinv = lds_vec<T>(p, n, L.dinv);
            L.b = lds_vec<T>(p, n, l == 0 ? L.b : nullptr);
            L.x = lds_vec<T>(p, n, nullptr);
            L.x2 = lds_vec<T>(p, n, nullptr);
            L.d = lds_vec<T>(p, n, nullptr);
            L.r = lds_vec<T>(p, n, nullptr);
            if (threadIdx.x == 0) slv[l] = L;
        }
        if (fc.ainv_lds) {
            const int64_t nn = fc.lv[nl - 1].A.n * fc.lv[nl - 1].A.n;
            double* a = reinterpret_cast<double*>(p);
            for (int64_t i = threadIdx.x; i < nn; i += blockDim.x) a[i] = fc.ainv[i];
            ainv = a;
        }
        __syncthreads();
        lv = slv;
    }
    T* res = fused_vcycle<T>(fc, lv, ainv);
    if (LDS) {
        T* o = static_cast<T*>(out);
        for (int64_t i = threadIdx.x; i < lv[0].A.n; i += blockDim.x) o[i] = res[i];
    }
}

// ---- one-workgroup PCG solves (fa_pcg_wg_batched) ------------------------------
// sum over the workgroup (FUSE_BLOCK threads) in a fixed tree; every thread
// gets the result
__device__ inline double wg_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x < WAVE) {
        double t = threadIdx.x < FUSE_BLOCK / WAVE ? red[threadIdx.x] : 0.0;
        t = wave_sum(t);
        if (threadIdx.x == 0) red[FUSE_BLOCK / WAVE] = t;
    }
    __syncthreads();
    return red[FUSE_BLOCK / WAVE];
}

__global__ void __launch_bounds__(FUSE_BLOCK) k_pcg_wg(fa_wg_solve_t* __restrict__ reqs) {
    __shared__ FusedCycle sfc;
    __shared__ double red[FUSE_BLOCK / WAVE + 1];
    fa_wg_solve_t& Q = reqs[blockIdx.x];
    if (threadIdx.x == 0) {
        sfc.nlev = Q.nlevels;
        sfc.degree = Q.degree;
        sfc.lower = Q.lower;
        sfc.upper = Q.upper;
        sfc.ainv = Q.coarse_inv;
        for (int l = 0; l < Q.nlevels; ++l) {
            const fa_wg_level_t& W = Q.levels[l];
            sfc.lv[l] = FusedLevel{W.A, W.R, W.P, W.dinv, W.b, W.x, W.x2, W.d, W.r, W.lmax};
        }
    }
    __syncthreads();
    const fa_sell_t A = Q.A;
    const int64_t n = A.n, npad = (n + 63) & ~63LL;
    const double* sv = static_cast<const double*>(A.vals);
    const double* b = Q.b;
    double* x = Q.x;
    double* r = Q.r;
    double* p = Q.p;
    double* Ap = Q.Ap;
    float* ru = static_cast<float*>(sfc.lv[0].b);
    const double atol = Q.atol, rtol = Q.rtol;
    const long long maxit = Q.maxit;
    // r = b - A x, its FP32 copy, ||b||
    double bb = 0.0;
    for (int64_t row = threadIdx.x; row < npad; row += blockDim.x) {
        const double ax = sell_row_dot<double, double>(A.slice_ptr, A.cols, sv, row, x);
        if (row < n) {
            const double bv = b[row], rv = bv - ax;
            r[row] = rv;
            ru[row] = static_cast<float>(rv);
            bb += bv * bv;
        }
    }
    const double bnorm = sqrt(wg_sum(bb, red));
    long long niter = 0;
    double rn = 0.0;
    int status = 0;
    if (bnorm < 1e-15) {
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[i] = 0.0;
        status = 2;
    } else {
        const float* z = fused_vcycle<float, 8>(sfc, sfc.lv, sfc.ainv);
        double s = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const double zi = static_cast<double>(z[i]);
            p[i] = zi;
            s += r[i] * zi;
        }
        double rz = wg_sum(s, red);
        for (;;) {
            // Ap = A p and p . Ap
            s = 0.0;
            for (int64_t row = threadIdx.x; row < npad; row += blockDim.x) {
                const double v = sell_row_dot<double, double>(A.slice_ptr, A.cols, sv, row, p);
                if (row < n) {
                    Ap[row] = v;
                    s += p[row] * v;
                }
            }
            const double alpha = rz / wg_sum(s, red);
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
                x[i] += alpha * p[i];
                const double rv = r[i] - alpha * Ap[i];
                r[i] = rv;
                ru[i] = static_cast<float>(rv);
            }
            __syncthreads();
            z = fused_vcycle<float, 8>(sfc, sfc.lv, sfc.ainv);
            s = 0.0;
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += r[i] * static_cast<double>(z[i]);
            const double rzn = wg_sum(s, red);
            ++niter;
            rn = sqrt(rzn);
            if (rn < atol || rn < rtol * bnorm) { status = 1; break; }
            if (niter >= maxit) { status = 3; break; }
            if (!isfinite(rn) || !isfinite(alpha)) { status = 4; break; }
            const double beta = rzn / rz;
            rz = rzn;
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = static_cast<double>(z[i]) + beta * p[i];
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        Q.niter = niter;
        Q.residual = rn;
        Q.status = status;
    }
}

// longest row for which the first two pre-smoothing steps run fused
constexpr int Z01_MAX_WIDTH = 32;

// ---------------------------------------------------------------------------
// V-cycle driver
template <class T>
struct Cycle {
    const fa_mg_t* mg;
    hipStream_t s;
    const int* done;

    static unsigned grid(int64_t n) { return xcd_grid(grid_for(((n + 63) / 64) * 64, 256, 1 << 16)); }

    // compressed operator present (FP32 cycles only)
    static bool compressed(const fa_mg_level_t& L) {
        return std::is_same<T, float>::value && L.Ah.vals != nullptr && L.diag != nullptr;
    }

    // column source of the FP16 operator: its int16 offsets, or (offs NULL:
    // a column too far from its row) the level's int32 SELL columns
    template <class F>
    static void with_cols(const fa_mg_level_t& L, F f) {
        if (L.Ah.offs) f(OffsetCols{L.Ah.offs});
        else f(IndexCols{L.A.cols});
    }

    int coarse(const fa_mg_level_t& L, const T* b, T* x) {
        const int n = static_cast<int>(L.A.n);
        if (n <= COARSE_ONE_BLOCK)
            hipLaunchKernelGGL(k_coarse_apply<T>, dim3(1), dim3(256), 0, s, n, mg->coarse_chol, b, x, done);
        else
            hipLaunchKernelGGL(k_dense_gemv<T>, dim3(grid_for((int64_t)n * WAVE, 256, 4096)), dim3(256), 0, s, n,
                               mg->coarse_chol, b, x, done);
        FA_LAUNCH("coarse solve");
        return FA_OK;
    }

    // Chebyshev sweep of `deg` steps on level L: from zero (pre) or from *cur (post).
    int smooth(const fa_mg_level_t& L, const T* b, T*& cur, T*& nxt, bool from_zero, T* last_out = nullptr,
               const DotOut* dot = nullptr) {
        return smooth_deg(L, b, cur, nxt, from_zero, mg->degree, last_out, dot);
    }

    // last_out: the last step writes there (and cur then points to it); dot:
    // that step also stores the partials of r . x_out (reduction grid)
    int smooth_deg(const fa_mg_level_t& L, const T* b, T*& cur, T*& nxt, bool from_zero, int degree,
                   T* last_out = nullptr, const DotOut* dot = nullptr) {
        const double lo = mg->lower * L.lmax, hi = mg->upper * L.lmax;
        const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
        double rho = 1.0 / sigma;
        const T* dinv = static_cast<const T*>(L.dinv);
        const T* sv = static_cast<const T*>(L.A.vals);
        T* d = static_cast<T*>(L.d);
        const unsigned g = grid(L.A.n);
        const unsigned gd = red_grid(L.A.n);
        const bool cmp = compressed(L);
        // the finest level's compressed steps are timed as one kernel family
        const int tk = (cmp && &L == &mg->levels[0]) ? FA_TIMER_SMOOTH0 : 0;
        for (int k = 0; k < degree; ++k) {
            T c1, c2;
            double rho_new = rho;
            if (k == 0) {
                c1 = T(0);
                c2 = static_cast<T>(1.0 / theta);
            } else {
                rho_new = 1.0 / (2.0 * sigma - rho);
                c1 = static_cast<T>(rho_new * rho);
                c2 = static_cast<T>(2.0 * rho_new / delta);
            }
            if (k == 0 && from_zero) {
                // steps 0 and 1 in one pass, forming step 0 at every gathered
                // column: two gathers an entry instead of one, which pays for
                // the saved launch on short rows only
                if (degree >= 2 && L.width > 0 && L.width <= Z01_MAX_WIDTH) {
                    const T c0 = c2;
                    rho_new = 1.0 / (2.0 * sigma - rho);
                    const T d1 = static_cast<T>(rho_new * rho), d2 = static_cast<T>(2.0 * rho_new / delta);
                    T* out = (degree == 2 && last_out) ? last_out : nxt;
                    if (cmp)
                        with_cols(L, [&](auto cs) {
                            timer_begin(tk, s);
                            hipLaunchKernelGGL(k_cheb16_z<decltype(cs)>, dim3(g), dim3(256), 0, s, L.A.n,
                                               L.Ah.slice_ptr, cs, static_cast<const _Float16*>(L.Ah.vals),
                                               reinterpret_cast<const float*>(dinv), reinterpret_cast<const float*>(b),
                                               reinterpret_cast<float*>(out), reinterpret_cast<float*>(d), (float)c0,
                                               (float)d1, (float)d2, done);
                            timer_end(tk, s);
                        });
                    else
                        hipLaunchKernelGGL(k_cheb_z<T>, dim3(g), dim3(256), 0, s, L.A.n, L.A.slice_ptr, L.A.cols, sv,
                                           dinv, b, out, d, c0, d1, d2, done);
                    if (out == nxt) std::swap(cur, nxt);
                    else cur = out;
                    rho = rho_new;
                    ++k;
                    continue;
                }
                hipLaunchKernelGGL((k_cheb<T, true, true>), dim3(g), dim3(256), 0, s, L.A.n, L.A.slice_ptr, L.A.cols,
                                   sv, dinv, b, cur, cur, d, c1, c2, done);
                continue;  // result written in place into cur
            }
            const bool last = (k == degree - 1);
            T* out = (last && last_out) ? last_out : nxt;
            const bool with_dot = last && dot;
            if (cmp) {
                const _Float16* hv = static_cast<const _Float16*>(L.Ah.vals);
                const float* fb = reinterpret_cast<const float*>(b);
                const float* fdinv = reinterpret_cast<const float*>(dinv);
                const float* fcur = reinterpret_cast<const float*>(cur);
                float* fout = reinterpret_cast<float*>(out);
                float* fd = reinterpret_cast<float*>(d);
                with_cols(L, [&](auto cs) {
                    using CS = decltype(cs);
                    timer_begin(tk, s);
                    if (with_dot) {
                        if (k == 0)
                            hipLaunchKernelGGL((k_cheb16<true, true, CS>), dim3(gd), dim3(RED_BLOCK), 0, s, L.A.n,
                                               L.Ah.slice_ptr, cs, hv, fdinv, fb, fcur, fout, fd, (float)c1,
                                               (float)c2, done, *dot);
                        else
                            hipLaunchKernelGGL((k_cheb16<false, true, CS>), dim3(gd), dim3(RED_BLOCK), 0, s, L.A.n,
                                               L.Ah.slice_ptr, cs, hv, fdinv, fb, fcur, fout, fd, (float)c1,
                                               (float)c2, done, *dot);
                    } else if (k == 0) {
                        hipLaunchKernelGGL((k_cheb16<true, false, CS>), dim3(g), dim3(256), 0, s, L.A.n,
                                           L.Ah.slice_ptr, cs, hv, fdinv, fb, fcur, fout, fd, (float)c1, (float)c2,
                                           done, DotOut{});
                    } else {
                        hipLaunchKernelGGL((k_cheb16<false, false, CS>), dim3(g), dim3(256), 0, s, L.A.n,
                                           L.Ah.slice_ptr, cs, hv, fdinv, fb, fcur, fout, fd, (float)c1, (float)c2,
                                           done, DotOut{});
                    }
                    timer_end(tk, s);
                });
            } else if (with_dot) {
                if (k == 0)
                    hipLaunchKernelGGL((k_cheb<T, false, true, true>), dim3(gd), dim3(RED_BLOCK), 0, s, L.A.n,
                                       L.A.slice_ptr, L.A.cols, sv, dinv, b, cur, out, d, c1, c2, done, *dot);
                else
                    hipLaunchKernelGGL((k_cheb<T, false, false, true>), dim3(gd), dim3(RED_BLOCK), 0, s, L.A.n,
                                       L.A.slice_ptr, L.A.cols, sv, dinv, b, cur, out, d, c1, c2, done, *dot);
            } else if (k == 0) {
                hipLaunchKernelGGL((k_cheb<T, false, true>), dim3(g), dim3(256), 0, s, L.A.n, L.A.slice_ptr,
                                   L.A.cols, sv, dinv, b, cur, out, d, c1, c2, done, DotOut{});
            } else {
                hipLaunchKernelGGL((k_cheb<T, false, false>), dim3(g), dim3(256), 0, s, L.A.n, L.A.slice_ptr,
                                   L.A.cols, sv, dinv, b, cur, out, d, c1, c2, done, DotOut{});
            }
            if (out == nxt) std::swap(cur, nxt);
            else cur = out;
            rho = rho_new;
        }
        FA_LAUNCH("k_cheb");
        return FA_OK;
    }

    // levels l..nlevels-1 in one workgroup when all are small; the result of
    // level l lands in x2 (an odd number of smoother swaps: degree-1 before,
    // degree after), or in x when l is the coarsest level
    int fused(int l, T** out) {
        FusedCycle fc{};
        fc.nlev = mg->nlevels - l;
        fc.degree = mg->degree;
        fc.lower = mg->lower;
        fc.upper = mg->upper;
        fc.ainv = mg->coarse_chol;
        for (int k = 0; k < fc.nlev; ++k) {
            const fa_mg_level_t& L = mg->levels[l + k];
            fc.lv[k] = FusedLevel{L.A, L.R, L.P, L.dinv, L.b, L.x, L.x2, L.d, L.r, L.lmax};
        }
        const fa_mg_level_t& L = mg->levels[l];
        T* res = static_cast<T*>(fc.nlev == 1 ? L.x : L.x2);
        size_t lds = 0;
        for (int k = 0; k < fc.nlev; ++k) lds += fused_level_lds<T>(fc.lv[k], k == fc.nlev - 1);
        // the coarsest inverse joins the level data in LDS when it fits
        const int64_t nc = fc.lv[fc.nlev - 1].A.n;
        const size_t ainv_bytes = static_cast<size_t>(nc * nc) * sizeof(double);
        fc.ainv_lds = (lds + ainv_bytes <= FUSE_LDS_MAX) ? 1 : 0;
        if (fc.ainv_lds) lds += ainv_bytes;
        if (fc.nlev > 1 && lds <= FUSE_LDS_MAX)
            hipLaunchKernelGGL((k_fused_cycle<T, true>), dim3(1), dim3(FUSE_BLOCK), lds, s, fc, res, done);
        else
            hipLaunchKernelGGL((k_fused_cycle<T, false>), dim3(1), dim3(FUSE_BLOCK), 0, s, fc, res, done);
        FA_LAUNCH("k_fused_cycle");
        *out = res;
        return FA_OK;
    }

    bool fusable(int l) const {
        if (mg->nlevels - l > FUSE_MAX_LEVELS || mg->degree < 1) return false;
        if (!mg->coarse_chol || mg->levels[mg->nlevels - 1].A.n > COARSE_ONE_BLOCK) return false;
        for (int k = l; k < mg->nlevels; ++k)
            if (mg->levels[k].A.n > FUSE_ROWS) return false;
        return true;
    }

    // returns (through out) the level-l vector that holds the result; on the
    // finest level a non-null `z` receives the result from the last smoothing
    // step (and `dot` its partials), *fused tells whether that happened
    int run(int l, const T* b, T** out, T* z = nullptr, const DotOut* dot = nullptr, bool* fused_out = nullptr) {
        const fa_mg_level_t& L = mg->levels[l];
        T* cur = static_cast<T*>(L.x);
        T* nxt = static_cast<T*>(L.x2);
        if (fused_out) *fused_out = false;
        if (l > 0 && fusable(l)) return fused(l, out);
        if (l == mg->nlevels - 1) {
            int st = mg->coarse_chol ? coarse(L, b, cur) : smooth_deg(L, b, cur, nxt, true, mg->coarse_degree);
            *out = cur;
            return st;
        }
        int st = smooth(L, b, cur, nxt, true);
        if (st) return st;
        const fa_mg_level_t& C = mg->levels[l + 1];
        T* r = static_cast<T*>(L.r);
        const unsigned g = grid(L.A.n);
        if (compressed(L))
            with_cols(L, [&](auto cs) {
                const int tk = l == 0 ? FA_TIMER_SMOOTH0 : 0;
                timer_begin(tk, s);
                hipLaunchKernelGGL(k_residual16<decltype(cs)>, dim3(g), dim3(256), 0, s, L.A.n, L.Ah.slice_ptr, cs,
                                   static_cast<const _Float16*>(L.Ah.vals), static_cast<const float*>(L.diag),
                                   reinterpret_cast<const float*>(b), reinterpret_cast<const float*>(cur),
                                   reinterpret_cast<float*>(r), done);
                timer_end(tk, s);
            });
        else
            hipLaunchKernelGGL(k_residual<T>, dim3(g), dim3(256), 0, s, L.A.n, L.A.slice_ptr, L.A.cols,
                               static_cast<const T*>(L.A.vals), b, cur, r, done);
        // the stencils only where both levels are the grids' node sets
        const bool structured =
            L.grid[0] > 0 && L.grid[1] > 0 && L.grid[2] > 0 && L.bd_fine && L.bd_coarse &&
            (L.grid[0] | L.grid[1] | L.grid[2]) % 2 == 0 &&
            L.A.n < INT32_MAX && static_cast<int64_t>(L.grid[0] + 1) * (L.grid[1] + 1) * (L.grid[2] + 1) == L.A.n &&
            static_cast<int64_t>(L.grid[0] / 2 + 1) * (L.grid[1] / 2 + 1) * (L.grid[2] / 2 + 1) == C.A.n;
        if (structured)
            hipLaunchKernelGGL(k_box_restrict<T>, dim3(xcd_grid(grid(C.A.n))), dim3(256), 0, s, L.grid[0] / 2,
                               L.grid[1] / 2, L.grid[2] / 2, L.bd_fine, L.bd_coarse, r, static_cast<T*>(C.b), done);
        else
            hipLaunchKernelGGL((k_sell_apply<T, false>), dim3(grid(L.R.n)), dim3(256), 0, s, L.R.n, L.R.slice_ptr,
                               L.R.cols, static_cast<const T*>(L.R.vals), r, static_cast<T*>(C.b), done);
        FA_LAUNCH("restrict");
        T* xc = nullptr;
        st = run(l + 1, static_cast<const T*>(C.b), &xc);
        if (st) return st;
        if (structured)
            hipLaunchKernelGGL(k_box_prolong<T>, dim3(xcd_grid(grid(L.A.n))), dim3(256), 0, s, L.grid[0], L.grid[1],
                               L.grid[2], L.bd_fine, L.bd_coarse, xc, cur, done);
        else
            hipLaunchKernelGGL((k_sell_apply<T, true>), dim3(grid(L.P.n)), dim3(256), 0, s, L.P.n, L.P.slice_ptr,
                               L.P.cols, static_cast<const T*>(L.P.vals), xc, cur, done);
        FA_LAUNCH("prolongate");
        const bool direct = z && mg->degree >= 1 && z != cur && z != nxt;
        st = smooth(L, b, cur, nxt, false, direct ? z : nullptr, direct ? dot : nullptr);
        if (direct && fused_out) *fused_out = true;
        *out = cur;
        return st;
    }
};

// partials of r . z (one per workgroup of the reduction grid)
template <class R, class Z>
__global__ void __launch_bounds__(RED_BLOCK) k_dot_part(int64_t n, const R* __restrict__ r, const Z* __restrict__ z,
                                                       double* __restrict__ partials, const int* __restrict__ done) {
    if (done && *done) return;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += double(r[i]) * double(z[i]);
    store_partial(block_sum(acc), partials);
}

template <class T>
__global__ void k_copy(int64_t n, const T* __restrict__ x, T* __restrict__ y, const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x[i];
}

// z = V(r).  With `dot`, the partials of dot.r . z are stored too (by the
// last smoothing step, or by a separate pass when the cycle has none).
template <class T>
static int vcycle_t(const fa_mg_t* mg, const void* r, void* z, const int* done, hipStream_t s,
                    const DotOut* dot) {
    Cycle<T> c{mg, s, done};
    T* out = nullptr;
    bool direct = false;
    int st = c.run(0, static_cast<const T*>(r), &out, static_cast<T*>(z), dot, &direct);
    if (st) return st;
    const int64_t n = mg->levels[0].A.n;
    if (out != z)
        hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(n, 256, 1 << 16)), dim3(256), 0, s, n, out, static_cast<T*>(z),
                           done);
    FA_LAUNCH("k_copy");
    if (dot && !direct) {
        if (dot->r_f64)
            hipLaunchKernelGGL((k_dot_part<double, T>), dim3(red_grid(n)), dim3(RED_BLOCK), 0, s, n,
                               static_cast<const double*>(dot->r), static_cast<const T*>(z), dot->partials, done);
        else
            hipLaunchKernelGGL((k_dot_part<float, T>), dim3(red_grid(n)), dim3(RED_BLOCK), 0, s, n,
                               static_cast<const float*>(dot->r), static_cast<const T*>(z), dot->partials, done);
        FA_LAUNCH("k_dot_part");
    }
    return FA_OK;
}

static int vcycle_any(const fa_mg_t* mg, const void* r, void* z, const int* done, hipStream_t s,
                      const DotOut* dot = nullptr) {
    if (mg->dtype == FA_F64) return vcycle_t<double>(mg, r, z, done, s, dot);
    return vcycle_t<float>(mg, r, z, done, s, dot);
}

// one V-cycle for other translation units (the batched PCG)
int mg_vcycle(const fa_mg_t* mg, const void* r, void* z, const int* done, hipStream_t s) {
    return vcycle_any(mg, r, z, done, s);
}

// ---------------------------------------------------------------------------
// PCG (outer dtype T, cycle dtype U)
struct PcgState {
    double rz, pAp, alpha, beta, rnorm, bnorm, atol, rtol;
    double rnorm0;  // sqrt(r.z) of the first direction: the start of the host's rate estimate
    long long niter, maxit;
    int done;  // 0 running, 1 converged, 2 ||b|| < 1e-15, 3 maxit, 4 breakdown
};

// One PCG iteration = spmv(+p.Ap partials) | alpha + update | [cycle] |
// r.z partials | check | p update.  The one-block *_finish kernels write the
// state, and so does workgroup 0 of the update (alpha, p.Ap, which no other
// workgroup of that kernel reads); the other grid kernels only read it.
enum PcgFinish { PCG_INIT = 0, PCG_RZ0 = 1, PCG_CHECK = 3 };

template <int WHAT>
__global__ void __launch_bounds__(RED_BLOCK) k_pcg_finish(const double* __restrict__ partials, int nparts,
                                                         PcgState* st) {
    if (WHAT != PCG_INIT && st->done) return;
    const double t = finish_sum(partials, nparts);
    double t2 = 0.0;
    if (WHAT == PCG_INIT) t2 = finish_sum(partials + nparts, nparts);
    if (threadIdx.x != 0) return;
    if (WHAT == PCG_INIT) {
        st->rz = t;  // replaced by r.z when preconditioned
        st->rnorm0 = sqrt(t);
        st->bnorm = sqrt(t2);
        st->niter = 0;
        st->done = (st->bnorm < 1e-15) ? 2 : 0;
    } else if (WHAT == PCG_RZ0) {
        st->rz = t;
        st->rnorm0 = sqrt(t);
    } else {
        const double rn = sqrt(t);
        st->rnorm = rn;
        st->niter += 1;
        if (rn < st->atol || rn < st->rtol * st->bnorm) st->done = 1;
        else if (st->niter >= st->maxit) st->done = 3;
        else if (!isfinite(rn) || !isfinite(st->alpha)) st->done = 4;
        st->beta = t / st->rz;
        st->rz = t;
    }
}

template <class T, class U>
__global__ void __launch_bounds__(RED_BLOCK) k_pcg_init(int64_t n, const int32_t* __restrict__ sp,
                                                       const int32_t* __restrict__ sc, const T* __restrict__ sv,
                                                       const T* __restrict__ b, const T* __restrict__ x,
                                                       T* __restrict__ r, U* __restrict__ ru, T* __restrict__ p,
                                                       double* __restrict__ partials, bool prec) {
    double rr = 0.0, bb = 0.0;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T ax = sell_row_dot<T, T>(sp, sc, sv, row, x);
        if (row < n) {
            const T bv = b[row];
            const T rv = bv - ax;
            r[row] = rv;
            if (!prec) p[row] = rv;
            else if (ru) ru[row] = static_cast<U>(rv);
            rr += double(rv) * double(rv);
            bb += double(bv) * double(bv);
        }
    }
    const double v[2] = {block_sum(rr), block_sum(bb)};
    store_partials<2>(v, partials);
}

// the same for x0 = 0: x = 0, r = b (b - A 0 is b exactly), with no product
template <class T, class U>
__global__ void __launch_bounds__(RED_BLOCK) k_pcg_init_zero(int64_t n, const T* __restrict__ b, T* __restrict__ x,
                                                            T* __restrict__ r, U* __restrict__ ru, T* __restrict__ p,
                                                            double* __restrict__ partials, bool prec) {
    double rr = 0.0, bb = 0.0;
    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        if (row < n) {
            const T bv = b[row];
            x[row] = T(0);
            r[row] = bv;
            if (!prec) p[row] = bv;
            else if (ru) ru[row] = static_cast<U>(bv);
            rr += double(bv) * double(bv);
            bb += double(bv) * double(bv);
        }
    }
    const double v[2] = {block_sum(rr), block_sum(bb)};
    store_partials<2>(v, partials);
}

template <class T>
__global__ void k_pcg_trivial(int64_t n, T* __restrict__ x, const PcgState* st) {
    if (st->done != 2) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = T(0);
}

// partials of r.z (after the first preconditioner application), p = z
template <class T, class U>
__global__ void __launch_bounds__(RED_BLOCK) k_pcg_rz0(int64_t n, const T* __restrict__ r, const U* __restrict__ z,
                                                      T* __restrict__ p, double* __restrict__ partials,
                                                      const PcgState* st) {
    if (st->done) return;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T zi = static_cast<T>(z[i]);
        p[i] = zi;
        acc += double(r[i]) * double(zi);
    }
    store_partial(block_sum(acc), partials);
}

// Ap = A p and partials of p.Ap.  so != NULL: column indices are row + so[k]
// (int16 offsets on the same slices; 2 bytes an entry instead of 4)
template <class T>
__global__ void __launch_bounds__(RED_BLOCK) k_pcg_spmv(int64_t n, const int32_t* __restrict__ sp,
                                                       const int32_t* __restrict__ sc, const int16_t* __restrict__ so,
                                                       const T* __restrict__ sv, const T* __restrict__ p,
                                                       T* __restrict__ Ap, double* __restrict__ partials,
                                                       const PcgState* st) {
    if (st->done) return;
    double acc = 0.0;
    // short rows on int16 offsets: rows dealt to the XCDs in contiguous
    // ranges, partials stored under the block index of the rows
    const unsigned b0 = so ? xcd_block() : blockIdx.x;
    for (int64_t row = b0 * (int64_t)blockDim.x + threadIdx.x; row < ((n + 63) & ~63LL);
         row += (int64_t)gridDim.x * blockDim.x) {
        const T v = so ? sell_row_dot_off<T>(sp, so, sv, row, n, p) : sell_row_dot<T, T>(sp, sc, sv, row, p);
        if (row < n) {
            Ap[row] = v;
            acc += double(p[row]) * double(v);
        }
    }
    store_partial_at(block_sum(acc), partials, b0);
}

// alpha = r.z / p.Ap from the SpMV's partials (every workgroup sums them in
// the finish kernel's order, so all get that kernel's alpha; workgroup 0 keeps
// it in the state), then x += alpha p; r -= alpha Ap; then either partials of
// r.r (no preconditioner) or the cast of r for the cycle (none when ru is
// null: the cycle then reads r itself)
template <class T, class U, bool PREC>
__global__ void __launch_bounds__(RED_BLOCK) k_pcg_update(int64_t n, T* __restrict__ x, T* __restrict__ r,
                                                         U* __restrict__ ru, const T* __restrict__ p,
                                                         const T* __restrict__ Ap,
                                                         const double* __restrict__ pap_parts, int np_pap,
                                                         double* __restrict__ partials, PcgState* st) {
    if (st->done) return;
    __shared__ double s_alpha;
    const double pAp = finish_sum(pap_parts, np_pap);
    if (threadIdx.x == 0) {
        s_alpha = st->rz / pAp;
        if (blockIdx.x == 0) {
            st->pAp = pAp;
            st->alpha = s_alpha;
        }
    }
    __syncthreads();
    const T alpha = static_cast<T>(s_alpha);
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const T rv = r[i] - alpha * Ap[i];
        r[i] = rv;
        if (PREC) {
            if (ru) ru[i] = static_cast<U>(rv);
        } else {
            acc += double(rv) * double(rv);
        }
    }
    if (!PREC) store_partial(block_sum(acc), partials);
}

template <class T, class U>
__global__ void k_pcg_pupdate(int64_t n, T* __restrict__ p, const U* __restrict__ z, const PcgState* st) {
    if (st->done) return;
    const T beta = static_cast<T>(st->beta);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = static_cast<T>(z[i]) + beta * p[i];
}

__global__ void k_pcg_set(PcgState* st, double atol, double rtol, long long maxit) {
    st->atol = atol;
    st->rtol = rtol;
    st->maxit = maxit;
    st->done = 0;
    st->niter = 0;
    st->rnorm = 0.0;  // the residual reported when ||b|| < 1e-15 stops the solve before any iteration
}

// A captured PCG iteration: the kernels of one iteration recorded on a private
// capture stream (nothing runs there) into a graph, then replayed on the
// caller's stream.  Every kernel of an iteration reads the convergence state
// from device memory, so the replayed sequence is the eager one.
struct PcgGraph {
    hipStream_t cs = nullptr;
    hipGraphExec_t exec = nullptr;
};

// record body(cs) into a graph and make pg->exec run it: an update of the
// existing executable graph when the topology is unchanged, else a new one
template <class F>
static int pcg_capture(PcgGraph* pg, F body) {
    FA_HIP(hipStreamBeginCapture(pg->cs, hipStreamCaptureModeThreadLocal));
    const int e = body(pg->cs);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(pg->cs, &g);
    if (e != FA_OK || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (e != FA_OK) return e;
        set_error("fa_pcg_graph: stream capture failed: %s", hipGetErrorString(ec));
        return FA_EHIP;
    }
    if (pg->exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(pg->exec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(pg->exec);
            pg->exec = nullptr;
        }
    }
    hipError_t ei = hipSuccess;
    if (!pg->exec) ei = hipGraphInstantiate(&pg->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        pg->exec = nullptr;
        set_error("fa_pcg_graph: hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        return FA_EHIP;
    }
    return FA_OK;
}

// iterations until sqrt(r.z) reaches max(atol, rtol ||b||) at the mean rate
// since the first direction, clamped to [1, most]
static int64_t predicted_chunk(const PcgState& h, int64_t most) {
    const double tol = std::max(h.atol, h.rtol * h.bnorm);
    if (!(h.niter > 0 && h.rnorm > 0.0 && h.rnorm0 > h.rnorm && tol > 0.0)) return most;
    const double rate = std::log(h.rnorm / h.rnorm0) / static_cast<double>(h.niter);  // < 0
    const double need = std::log(tol / h.rnorm) / rate;
    if (!std::isfinite(need)) return most;
    return std::max<int64_t>(1, std::min<int64_t>(most, static_cast<int64_t>(std::ceil(need))));
}

template <class T, class U>
static int pcg_run(const fa_sell_t* A, const fa_mg_t* M, const T* b, T* x, double atol, double rtol, int64_t maxit,
                   int check_every, int64_t* niter_host, double* res_host, void* work, size_t work_bytes,
                   hipStream_t s, PcgGraph* pg = nullptr, bool x0_zero = false) {
    const int64_t n = A->n;
    Arena ar(work, work_bytes);
    T* r = ar.take<T>(n);
    T* p = ar.take<T>(n);
    T* Ap = ar.take<T>(n);
    // the cycle's right-hand side: r itself when the cycle runs in r's type
    // (the update then writes no copy), else r's cast
    constexpr bool same = std::is_same<T, U>::value;
    U* ru = M ? (same ? reinterpret_cast<U*>(r) : ar.take<U>(n)) : nullptr;
    U* z = M ? ar.take<U>(n) : nullptr;
    double* partials = ar.take<double>(2 * RED_GRID);
    PcgState* st = ar.take<PcgState>(1);
    if (!r || !p || !Ap || (M && (!ru || !z)) || !partials || !st) {
        set_error("fa_pcg: workspace too small (%zu bytes)", work_bytes);
        return FA_ENOMEM;
    }
    const int* done = &st->done;
    const unsigned g = red_grid(n);
    const int np = static_cast<int>(g);
    const dim3 grid(g), block(RED_BLOCK), one(1);
    const T* sv = static_cast<const T*>(A->vals);
    const int64_t cap = maxit > 0 ? maxit : (1LL << 62);
    hipLaunchKernelGGL(k_pcg_set, one, one, 0, s, st, atol, rtol, (long long)cap);
    if (x0_zero)
        hipLaunchKernelGGL((k_pcg_init_zero<T, U>), grid, block, 0, s, n, b, x, r, same ? nullptr : ru, p, partials,
                           M != nullptr);
    else
        hipLaunchKernelGGL((k_pcg_init<T, U>), grid, block, 0, s, n, A->slice_ptr, A->cols, sv, b, x, r,
                           same ? nullptr : ru, p, partials, M != nullptr);
    hipLaunchKernelGGL(k_pcg_finish<PCG_INIT>, one, block, 0, s, partials, np, st);
    hipLaunchKernelGGL(k_pcg_trivial<T>, grid, block, 0, s, n, x, st);
    FA_LAUNCH("k_pcg_init");
    if (M) {
        int e = vcycle_any(M, ru, z, done, s);
        if (e) return e;
        hipLaunchKernelGGL((k_pcg_rz0<T, U>), grid, block, 0, s, n, r, z, p, partials, st);
        hipLaunchKernelGGL(k_pcg_finish<PCG_RZ0>, one, block, 0, s, partials, np, st);
        FA_LAUNCH("k_pcg_rz0");
    }
    const DotOut rz_out{r, partials, std::is_same<T, double>::value ? 1 : 0};
    // the finest level's compressed smoother operator lies on A's own slices:
    // its int16 column offsets index A's values as well
    const int16_t* offs = (M && M->levels[0].Ah.vals && M->levels[0].Ah.offs &&
                           M->levels[0].Ah.slice_ptr == A->slice_ptr && M->levels[0].Ah.n == n)
                              ? M->levels[0].Ah.offs
                              : nullptr;
    // p.Ap partials in the second partial array: the update reads them while
    // writing r.r partials (no preconditioner) into the first
    double* pap = partials + RED_GRID;
    auto iteration = [&](hipStream_t q) -> int {
        timer_begin(FA_TIMER_PCG_SPMV, q);
        hipLaunchKernelGGL(k_pcg_spmv<T>, grid, block, 0, q, n, A->slice_ptr, A->cols, offs, sv, p, Ap, pap, st);
        timer_end(FA_TIMER_PCG_SPMV, q);
        if (M) {
            hipLaunchKernelGGL((k_pcg_update<T, U, true>), grid, block, 0, q, n, x, r, same ? nullptr : ru, p, Ap, pap,
                               np, partials, st);
            timer_begin(FA_TIMER_VCYCLE, q);
            int e = vcycle_any(M, ru, z, done, q, &rz_out);  // z = M r and the partials of r.z
            timer_end(FA_TIMER_VCYCLE, q);
            if (e) return e;
            hipLaunchKernelGGL(k_pcg_finish<PCG_CHECK>, one, block, 0, q, partials, np, st);
            hipLaunchKernelGGL((k_pcg_pupdate<T, U>), grid, block, 0, q, n, p, z, st);
        } else {
            hipLaunchKernelGGL((k_pcg_update<T, T, false>), grid, block, 0, q, n, x, r, (T*)nullptr, p, Ap, pap, np,
                               partials, st);
            hipLaunchKernelGGL(k_pcg_finish<PCG_CHECK>, one, block, 0, q, partials, np, st);
            hipLaunchKernelGGL((k_pcg_pupdate<T, T>), grid, block, 0, q, n, p, r, st);
        }
        FA_LAUNCH("pcg iteration");
        return FA_OK;
    };
    if (pg) {
        int e = pcg_capture(pg, iteration);
        if (e) return e;
    }
    // The host looks at the state after `next` more iterations: check_every
    // at first, then the count that the mean reduction rate so far predicts
    // to reach the tolerance (at least 1, at most check_every).  Iterations
    // launched past convergence return at once but still cost their launches
    // (about 30 a V-cycle), so a good guess saves them; the iterates and the
    // result do not depend on when the host looks.
    PcgState h{};
    int64_t it = 0, next = check_every;
    for (;;) {
        const int64_t chunk = std::min<int64_t>(next, cap - it);
        for (int64_t k = 0; k < chunk; ++k) {
            if (pg) {
                FA_HIP(hipGraphLaunch(pg->exec, s));
            } else {
                int e = iteration(s);
                if (e) return e;
            }
        }
        it += chunk;
        FA_HIP(hipMemcpyAsync(&h, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        if (h.done || it >= cap) break;
        next = predicted_chunk(h, check_every);
    }
    if (niter_host) *niter_host = h.niter;
    if (res_host) *res_host = h.rnorm;
    return (h.done == 3 || h.done == 4) ? FA_ENOCONV : FA_OK;
}

}  // namespace fa

using namespace fa;

extern "C" size_t fa_rap_work_bytes(int64_t nrows) {
    return align_up((nrows + 1) * sizeof(int32_t)) + exclusive_scan_work_bytes(nrows) + 2048;
}

template <int CAP>
static void launch_rap_rows(int64_t nrows, const int32_t* rp, const int32_t* ri, const int32_t* ap,
                            const int32_t* ai, const int32_t* pp, const int32_t* pi, int32_t* rowcount,
                            const int32_t* cptr, int32_t* cidx, int32_t* max_row, int32_t* overflow, hipStream_t s) {
    hipLaunchKernelGGL(k_rap_rows<CAP>, dim3(grid_for(nrows, 1, 1 << 22)), dim3(WAVE), 0, s, nrows, rp, ri, ap, ai, pp,
                       pi, rowcount, cptr, cidx, max_row, overflow);
}

static int rap_rows(int32_t max_cand, int64_t nrows, const int32_t* rp, const int32_t* ri, const int32_t* ap,
                    const int32_t* ai, const int32_t* pp, const int32_t* pi, int32_t* rowcount, const int32_t* cptr,
                    int32_t* cidx, int32_t* max_row, int32_t* overflow, hipStream_t s) {
    if (max_cand <= 512) launch_rap_rows<512>(nrows, rp, ri, ap, ai, pp, pi, rowcount, cptr, cidx, max_row, overflow, s);
    else if (max_cand <= 2048)
        launch_rap_rows<2048>(nrows, rp, ri, ap, ai, pp, pi, rowcount, cptr, cidx, max_row, overflow, s);
    else if (max_cand <= 8192)
        launch_rap_rows<8192>(nrows, rp, ri, ap, ai, pp, pi, rowcount, cptr, cidx, max_row, overflow, s);
    else {
        set_error("fa_rap: %d candidates per coarse row exceed the kernel capacity (8192)", max_cand);
        return FA_ECAPACITY;
    }
    FA_LAUNCH("k_rap_rows");
    return FA_OK;
}

__global__ void k_zero_i32(int32_t* p, int n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

extern "C" int fa_rap_count(int64_t nrows, int64_t ncols, const int32_t* r_ptr, const int32_t* r_idx,
                            const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx,
                            int32_t max_cand, int32_t* c_ptr, int64_t* nnz_host, int32_t* max_row_host, void* work,
                            size_t work_bytes, fa_stream_t stream) {
    (void)ncols;
    FA_REQUIRE(nrows > 0, "fa_rap_count: empty operator");
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    int32_t* cnt = ar.take<int32_t>(nrows + 1);
    int32_t* flags = ar.take<int32_t>(4);
    const size_t sb = exclusive_scan_work_bytes(nrows);
    void* sw = ar.take<char>(sb);
    if (!cnt || !flags || !sw) {
        set_error("fa_rap_count: workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, flags, 4);
    int st = rap_rows(max_cand, nrows, r_ptr, r_idx, a_ptr, a_idx, p_ptr, p_idx, cnt, nullptr, nullptr, flags,
                      flags + 1, s);
    if (st) return st;
    st = exclusive_scan_i32(cnt, c_ptr, nrows, sw, sb, s);
    if (st) return st;
    int32_t h[2] = {0, 0}, nnz = 0;
    FA_HIP(hipMemcpyAsync(h, flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&nnz, c_ptr + nrows, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (h[1]) {
        set_error("fa_rap_count: max_cand %d is smaller than a coarse row's candidate count", max_cand);
        return FA_ECAPACITY;
    }
    if (nnz_host) *nnz_host = nnz;
    if (max_row_host) *max_row_host = h[0];
    return FA_OK;
}

extern "C" int fa_rap_fill(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const int32_t* a_ptr,
                           const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx, int32_t max_cand,
                           const int32_t* c_ptr, int32_t* c_idx, fa_stream_t stream) {
    return rap_rows(max_cand, nrows, r_ptr, r_idx, a_ptr, a_idx, p_ptr, p_idx, nullptr, c_ptr, c_idx, nullptr,
                    nullptr, S(stream));
}

extern "C" int fa_rap_values(int dtype, int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx,
                             const double* r_val, const int32_t* a_ptr, const int32_t* a_idx, const void* a_val,
                             const int32_t* p_ptr, const int32_t* p_idx, const double* p_val, const int32_t* c_ptr,
                             const int32_t* c_idx, int32_t max_row, void* c_val, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_rap_values: bad dtype %d", dtype);
    FA_REQUIRE(max_row > 0, "fa_rap_values: bad max_row");
    const size_t lds = align_up(sizeof(int32_t) * max_row, 16) + sizeof(double) * WAVE * max_row;
    const dim3 grid(grid_for(nrows, 1, 1 << 22));
    if (lds > 64 * 1024) {  // wide coarse rows: one accumulator per entry, pairs in sequence
        const size_t lds1 = align_up(sizeof(int32_t) * max_row, 16) + sizeof(double) * max_row;
        if (lds1 > 64 * 1024) {
            set_error("fa_rap_values: coarse rows of %d entries exceed the LDS capacity", max_row);
            return FA_ECAPACITY;
        }
        if (dtype == FA_F64)
            hipLaunchKernelGGL(k_rap_values_seq<double>, grid, dim3(WAVE), lds1, S(stream), nrows, r_ptr, r_idx,
                               r_val, a_ptr, a_idx, static_cast<const double*>(a_val), p_ptr, p_idx, p_val, c_ptr,
                               c_idx, max_row, static_cast<double*>(c_val));
        else
            hipLaunchKernelGGL(k_rap_values_seq<float>, grid, dim3(WAVE), lds1, S(stream), nrows, r_ptr, r_idx,
                               r_val, a_ptr, a_idx, static_cast<const float*>(a_val), p_ptr, p_idx, p_val, c_ptr,
                               c_idx, max_row, static_cast<float*>(c_val));
        FA_LAUNCH("k_rap_values_seq");
        return FA_OK;
    }
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_rap_values<double>, grid, dim3(WAVE), lds, S(stream), nrows, r_ptr, r_idx, r_val, a_ptr,
                           a_idx, static_cast<const double*>(a_val), p_ptr, p_idx, p_val, c_ptr, c_idx, max_row,
                           static_cast<double*>(c_val));
    else
        hipLaunchKernelGGL(k_rap_values<float>, grid, dim3(WAVE), lds, S(stream), nrows, r_ptr, r_idx, r_val, a_ptr,
                           a_idx, static_cast<const float*>(a_val), p_ptr, p_idx, p_val, c_ptr, c_idx, max_row,
                           static_cast<float*>(c_val));
    FA_LAUNCH("k_rap_values");
    return FA_OK;
}

extern "C" int fa_csr_to_dense(int64_t n, const int32_t* indptr, const int32_t* indices, const double* values,
                               double* dense, fa_stream_t stream) {
    hipStream_t s = S(stream);
    hipLaunchKernelGGL(k_csr_to_dense, dim3(grid_for(n * n, 256, 4096)), dim3(256), 0, s, n, indptr, indices, values,
                       dense);
    hipLaunchKernelGGL(k_csr_scatter_dense, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, indptr, indices,
                       values, dense);
    FA_LAUNCH("k_csr_to_dense");
    return FA_OK;
}

extern "C" int fa_dense_spd_inverse(int64_t n, double* a, fa_stream_t stream) {
    FA_REQUIRE(n > 0 && n <= COARSE_MAX, "fa_dense_spd_inverse: n = %lld outside 1..%d", (long long)n, COARSE_MAX);
    hipStream_t s = S(stream);
    if (n <= COARSE_LDS_MAX) {
        hipLaunchKernelGGL(k_spd_inverse_lds, dim3(1), dim3(256), (n * n + 2 * n) * sizeof(double), s, (int)n, a);
    } else {
        const int nn = static_cast<int>(n);
        const unsigned gp = grid_for(nn, GJB, 1024), gt = grid_for((int64_t)nn * nn, 256, 1 << 14);
        for (int k0 = 0; k0 < nn; k0 += GJB) {
            const int b = std::min(GJB, nn - k0);
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, s, nn, k0, b, a);
            hipLaunchKernelGGL(k_gj_rowpanel, dim3(gp), dim3(256), 0, s, nn, k0, b, a);
            hipLaunchKernelGGL(k_gj_trailing, dim3(gt), dim3(256), 0, s, nn, k0, b, a);
            hipLaunchKernelGGL(k_gj_colpanel, dim3(gp), dim3(256), 0, s, nn, k0, b, a);
        }
    }
    FA_LAUNCH("k_spd_inverse");
    return FA_OK;
}

extern "C" int fa_mg_lmax(const fa_mg_t* mg, int level, int iters, double* lmax_host, void* work, size_t work_bytes,
                          fa_stream_t stream) {
    FA_REQUIRE(mg && level >= 0 && level < mg->nlevels && iters > 0, "fa_mg_lmax: bad arguments");
    hipStream_t s = S(stream);
    const fa_mg_level_t& L = mg->levels[level];
    Arena ar(work, work_bytes);
    double* partials = ar.take<double>(3 * RED_GRID);
    double* out = ar.take<double>(2);
    if (!partials || !out) {
        set_error("fa_mg_lmax: workspace too small");
        return FA_ENOMEM;
    }
    const unsigned g = grid_for(L.A.n, 256, 1 << 16);
    const unsigned gr = red_grid(L.A.n);
    double lam = 0.0;
#define FA_PW(T)                                                                                                   \
    {                                                                                                              \
        T* v = static_cast<T*>(L.x);                                                                               \
        T* w = static_cast<T*>(L.x2);                                                                              \
        hipLaunchKernelGGL(k_power_init<T>, dim3(g), dim3(256), 0, s, L.A.n, v);                                  \
        for (int k = 0; k < iters; ++k) {                                                                          \
            hipLaunchKernelGGL(k_power<T>, dim3(gr), dim3(RED_BLOCK), 0, s, L.A.n, L.A.slice_ptr, L.A.cols,        \
                               static_cast<const T*>(L.A.vals), static_cast<const T*>(L.dinv), v, w, partials);   \
            hipLaunchKernelGGL(k_power_finish, dim3(1), dim3(RED_BLOCK), 0, s, partials, (int)gr, out);           \
            hipLaunchKernelGGL(k_scale<T>, dim3(g), dim3(256), 0, s, L.A.n, w, v, out);                           \
        }                                                                                                          \
    }
    if (mg->dtype == FA_F64) FA_PW(double) else FA_PW(float)
#undef FA_PW
    FA_LAUNCH("k_power");
    FA_HIP(hipMemcpyAsync(&lam, out, sizeof(double), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (lmax_host) *lmax_host = lam;
    return FA_OK;
}

extern "C" int fa_jacobi_setup(int dtype_out, int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                               const double* values, void* dinv, double* bound_host, void* work, size_t work_bytes,
                               fa_stream_t stream) {
    FA_REQUIRE(n > 0, "fa_jacobi_setup: empty matrix");
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    double* partials = ar.take<double>(RED_GRID);
    double* out = ar.take<double>(1);
    if (!partials || !out) {
        set_error("fa_jacobi_setup: workspace too small");
        return FA_ENOMEM;
    }
    const double avg = double(nnz) / double(n);
    const unsigned grid = grid_for(n * 16, RED_BLOCK, RED_GRID);
#define FA_JS(T, G)                                                                                               \
    hipLaunchKernelGGL((k_jacobi_setup<T, G>), dim3(grid), dim3(RED_BLOCK), 0, s, n, indptr, indices, values,   \
                       static_cast<T*>(dinv), partials)
    if (dtype_out == FA_F64) {
        if (avg <= 12) FA_JS(double, 8); else if (avg <= 24) FA_JS(double, 16); else FA_JS(double, 32);
    } else {
        if (avg <= 12) FA_JS(float, 8); else if (avg <= 24) FA_JS(float, 16); else FA_JS(float, 32);
    }
#undef FA_JS
    hipLaunchKernelGGL(k_max_finish, dim3(1), dim3(RED_BLOCK), 0, s, partials, (int)grid, out);
    FA_LAUNCH("k_jacobi_setup");
    double h = 0.0;
    FA_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (bound_host) *bound_host = h;
    return FA_OK;
}

extern "C" int fa_smoother_setup(int dtype, int dtype_in, const fa_sell_t* A, void* dinv, void* diag,
                                 void* vals_f32, const fa_sell16_t* Ah, double* bound_dev, void* work,
                                 size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(A && A->n > 0 && A->slice_ptr && A->cols && A->vals && dinv && bound_dev,
               "fa_smoother_setup: bad arguments");
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_smoother_setup: dtype must be FP64 or FP32");
    FA_REQUIRE(dtype_in == FA_F64 || dtype_in == FA_F32, "fa_smoother_setup: dtype_in must be FP64 or FP32");
    FA_REQUIRE(!Ah || !Ah->vals || (Ah->n == A->n && Ah->slice_ptr == A->slice_ptr),
               "fa_smoother_setup: the compressed operator must share the slices of A");
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    double* partials = ar.take<double>(RED_GRID);
    if (!partials) {
        set_error("fa_smoother_setup: workspace too small");
        return FA_ENOMEM;
    }
    const int64_t npad = (A->n + 63) & ~63LL;
    const unsigned grid = red_grid(npad);
    _Float16* h16 = (Ah && Ah->vals) ? static_cast<_Float16*>(const_cast<void*>(Ah->vals)) : nullptr;
#define FA_SS(T_, V_)                                                                                         \
    hipLaunchKernelGGL((k_smoother_setup<T_, V_>), dim3(grid), dim3(RED_BLOCK), 0, s, A->n, A->slice_ptr, A->cols, \
                       static_cast<const V_*>(A->vals), static_cast<T_*>(dinv), static_cast<T_*>(diag), h16,     \
                       static_cast<float*>(vals_f32), partials)
    if (dtype == FA_F64) {
        if (dtype_in == FA_F64) FA_SS(double, double); else FA_SS(double, float);
    } else {
        if (dtype_in == FA_F64) FA_SS(float, double); else FA_SS(float, float);
    }
#undef FA_SS
    hipLaunchKernelGGL(k_max_finish, dim3(1), dim3(RED_BLOCK), 0, s, partials, (int)grid, bound_dev);
    FA_LAUNCH("k_smoother_setup");
    return FA_OK;
}

extern "C" int fa_mg_vcycle(const fa_mg_t* mg, const void* r, void* z, fa_stream_t stream) {
    FA_REQUIRE(mg && mg->nlevels >= 1 && mg->levels, "fa_mg_vcycle: empty hierarchy");
    FA_REQUIRE(mg->coarse_chol || mg->coarse_degree >= 1,
               "fa_mg_vcycle: the coarsest level needs its inverse or a smoothing degree");
    FA_REQUIRE(!mg->coarse_chol || mg->levels[mg->nlevels - 1].A.n <= COARSE_MAX,
               "fa_mg_vcycle: coarsest level too large for the dense inverse");
    return vcycle_any(mg, r, z, nullptr, S(stream));
}

extern "C" size_t fa_pcg_work_bytes(int dtype, int64_t n) {
    const size_t e = dtype == FA_F32 ? 4 : 8;
    return 3 * align_up(n * e) + 2 * align_up(n * 8) + align_up(2 * RED_GRID * sizeof(double)) +
           align_up(sizeof(PcgState)) + 2048;
}

static int pcg_dispatch(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol,
                        double rtol, int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
                        void* work, size_t work_bytes, hipStream_t s, PcgGraph* pg, bool x0_zero = false) {
    FA_REQUIRE(A && A->n > 0 && check_every > 0, "fa_pcg: bad arguments");
    if (M) {
        FA_REQUIRE(M->nlevels >= 1 && M->levels && (M->coarse_chol || M->coarse_degree >= 1),
                   "fa_pcg: incomplete multigrid hierarchy");
        FA_REQUIRE(M->levels[0].A.n == A->n, "fa_pcg: preconditioner size mismatch");
        FA_REQUIRE(!M->coarse_chol || M->levels[M->nlevels - 1].A.n <= COARSE_MAX,
                   "fa_pcg: coarsest level too large for the dense inverse");
    }
    const int ud = M ? M->dtype : dtype;
#define FA_PCG(T, U)                                                                                         \
    return pcg_run<T, U>(A, M, static_cast<const T*>(b), static_cast<T*>(x), atol, rtol, maxit, check_every, \
                         niter_host, residual_host, work, work_bytes, s, pg, x0_zero)
    if (dtype == FA_F64 && ud == FA_F64) FA_PCG(double, double);
    if (dtype == FA_F64 && ud == FA_F32) FA_PCG(double, float);
    if (dtype == FA_F32 && ud == FA_F32) FA_PCG(float, float);
#undef FA_PCG
    set_error("fa_pcg: unsupported dtype pair (%d, %d)", dtype, ud);
    return FA_EINVAL;
}

extern "C" int fa_pcg(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol,
                      double rtol, int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
                      void* work, size_t work_bytes, fa_stream_t stream) {
    return pcg_dispatch(dtype, A, M, b, x, atol, rtol, maxit, check_every, niter_host, residual_host, work,
                        work_bytes, S(stream), nullptr);
}

extern "C" int fa_pcg_from_zero(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x,
                                double atol, double rtol, int64_t maxit, int32_t check_every, int64_t* niter_host,
                                double* residual_host, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(x, "fa_pcg_from_zero: bad arguments");
    return pcg_dispatch(dtype, A, M, b, x, atol, rtol, maxit, check_every, niter_host, residual_host, work,
                        work_bytes, S(stream), nullptr, true);
}

// ---- fa_pcg_static: the PCG workspace in a file-scope device array --------
// For callers that cannot allocate.  The array holds the workspace of
// fa_pcg_work_bytes(FA_F64, FA_PCG_STATIC_MAX_ROWS); every GPU the library
// runs on keeps its own copy until the process ends.  In host code the
// array's name is only a stand-in, so its device address is looked up once
// per device with hipGetSymbolAddress and only that address is used.
constexpr size_t PCG_STATIC_BYTES = FA_PCG_STATIC_BYTES;
static __device__ __align__(256) unsigned char scratch[PCG_STATIC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_on_current_device(void** out) {
    int dev = -1;
    FA_HIP(hipGetDevice(&dev));
    FA_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "fa_pcg_static: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FA_OK;
}
}  // namespace

extern "C" int fa_pcg_static(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol,
                             double rtol, int64_t maxit, int32_t check_every, int64_t* niter_host,
                             double* residual_host, fa_stream_t stream) {
    FA_REQUIRE(A && A->n > 0, "fa_pcg_static: bad arguments");
    FA_REQUIRE(A->n <= FA_PCG_STATIC_MAX_ROWS, "fa_pcg_static: %lld rows exceed the static workspace (at most %d)",
               (long long)A->n, FA_PCG_STATIC_MAX_ROWS);
    const size_t need = fa_pcg_work_bytes(dtype, A->n);
    FA_REQUIRE(need <= PCG_STATIC_BYTES, "fa_pcg_static: workspace of %zu bytes exceeds the static %zu", need,
               PCG_STATIC_BYTES);
    void* work = nullptr;
    const int e = scratch_on_current_device(&work);
    if (e != FA_OK) return e;
    return pcg_dispatch(dtype, A, M, b, x, atol, rtol, maxit, check_every, niter_host, residual_host, work, need,
                        S(stream), nullptr);
}

struct fa_pcg_graph_s {
    PcgGraph g;
};

extern "C" int fa_pcg_graph_create(fa_pcg_graph_t** out) {
    FA_REQUIRE(out, "fa_pcg_graph_create: out is NULL");
    *out = nullptr;
    fa_pcg_graph_t* h = new (std::nothrow) fa_pcg_graph_t();
    if (!h) {
        set_error("fa_pcg_graph_create: out of host memory");
        return FA_ENOMEM;
    }
    const hipError_t e = hipStreamCreateWithFlags(&h->g.cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        set_error("fa_pcg_graph_create: hipStreamCreateWithFlags failed: %s", hipGetErrorString(e));
        return FA_EHIP;
    }
    *out = h;
    return FA_OK;
}

extern "C" void fa_pcg_graph_destroy(fa_pcg_graph_t* graph) {
    if (!graph) return;
    if (graph->g.exec) (void)hipGraphExecDestroy(graph->g.exec);
    if (graph->g.cs) (void)hipStreamDestroy(graph->g.cs);
    delete graph;
}

extern "C" int fa_pcg_graph(int dtype, const fa_sell_t* A, const fa_mg_t* M, const void* b, void* x, double atol,
                            double rtol, int64_t maxit, int32_t check_every, int64_t* niter_host,
                            double* residual_host, void* work, size_t work_bytes, fa_pcg_graph_t* graph,
                            fa_stream_t stream) {
    FA_REQUIRE(graph && graph->g.cs, "fa_pcg_graph: no graph handle (fa_pcg_graph_create)");
    return pcg_dispatch(dtype, A, M, b, x, atol, rtol, maxit, check_every, niter_host, residual_host, work,
                        work_bytes, S(stream), &graph->g);
}

extern "C" int fa_pcg_wg_batched(int nreq, const fa_wg_solve_t* reqs_host, fa_wg_solve_t* reqs_dev,
                                 fa_stream_t stream) {
    FA_REQUIRE(nreq > 0 && reqs_host && reqs_dev, "fa_pcg_wg_batched: bad arguments");
    static_assert(FA_WG_MAX_LEVELS == FUSE_MAX_LEVELS, "level capacity");
    for (int q = 0; q < nreq; ++q) {
        const fa_wg_solve_t& Q = reqs_host[q];
        FA_REQUIRE(Q.A.n > 0 && Q.A.slice_ptr && Q.A.cols && Q.A.vals && Q.b && Q.x && Q.r && Q.p && Q.Ap &&
                       Q.maxit > 0 && Q.nlevels >= 1 && Q.nlevels <= FA_WG_MAX_LEVELS && Q.degree >= 1 &&
                       Q.coarse_inv,
                   "fa_pcg_wg_batched: request %d: bad descriptor", q);
        FA_REQUIRE(Q.levels[0].A.n == Q.A.n, "fa_pcg_wg_batched: request %d: the finest level has %lld rows, A %lld",
                   q, (long long)Q.levels[0].A.n, (long long)Q.A.n);
        for (int l = 0; l < Q.nlevels; ++l) {
            const fa_wg_level_t& W = Q.levels[l];
            FA_REQUIRE(W.A.n > 0 && W.A.slice_ptr && W.A.vals && W.dinv && W.b && W.x && W.x2 && W.d && W.r,
                       "fa_pcg_wg_batched: request %d: level %d incomplete", q, l);
            if (l + 1 < Q.nlevels)
                FA_REQUIRE(W.R.n == Q.levels[l + 1].A.n && W.R.ncols == W.A.n && W.P.n == W.A.n &&
                               W.P.ncols == Q.levels[l + 1].A.n && W.R.vals && W.P.vals,
                           "fa_pcg_wg_batched: request %d: transfers of level %d do not match", q, l);
        }
        FA_REQUIRE(Q.levels[Q.nlevels - 1].A.n <= COARSE_MAX,
                   "fa_pcg_wg_batched: request %d: coarsest level above %d rows", q, COARSE_MAX);
    }
    hipLaunchKernelGGL(k_pcg_wg, dim3(nreq), dim3(FUSE_BLOCK), 0, S(stream), reqs_dev);
    FA_LAUNCH("k_pcg_wg");
    return FA_OK;
}
