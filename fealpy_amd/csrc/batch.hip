// Batched solves for serving: many independent Dirichlet problems packed at
// offsets into one block-diagonal system, so every kernel of a PCG iteration
// and of its V-cycle runs once for the whole batch instead of once per
// request (the reference's counterpart is the batch axis of Form and of its
// cg, fealpy/fem/form.py:39-52 and fealpy/solver/cg.py:69-123, whose
// columns iterate in lockstep with their own alpha and beta).
//
// Packing: request q occupies the rows [off_q, off_q + n_q) of every packed
// operator, off_q a multiple of the segment (64 rows on coarse levels, 256 on
// the finest, so a workgroup of the PCG kernels holds rows of one request);
// rows in between are empty.  fa_pack_sell concatenates SELL-64 matrices
// (slice pointers shifted by the entries before, column indices by the
// request's column offset); fa_copy_segments moves vectors in and out.
//
// fa_pcg_batched: preconditioned CG on the packed system with one state per
// request: every reduction (b.b, r.z, p.Ap) is summed per request from the
// partials of its own workgroups in a fixed order, alpha, beta and the stop
// test are per request, and the kernels skip the workgroups of a request
// that has stopped, so its x stays the iterate at which it converged.  The
// V-cycle is the ordinary one on the packed hierarchy (block diagonal: no
// level couples two requests).
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "reduce.hpp"
#include "sell.hpp"

namespace fa {

constexpr int BSEG = RED_BLOCK;  // rows per PCG workgroup on the finest packed level

// ---- packing -------------------------------------------------------------------
// blockIdx.y = request; the slice pointers (and trailing empty slices), then
// the entries, grid-strided over blockIdx.x
__global__ void __launch_bounds__(256) k_pack_sell(const fa_pack_sell_t* __restrict__ items, int elem_bytes,
                                                  int32_t* __restrict__ usp, int32_t* __restrict__ ucols,
                                                  char* __restrict__ uvals) {
    const fa_pack_sell_t it = items[blockIdx.y];
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
    const int64_t nsp = it.nslices + it.extra_slices;
    for (int64_t s = t0; s < nsp; s += st) {
        const int64_t k = s < it.nslices ? s : it.nslices;
        usp[it.slice_base + s] = static_cast<int32_t>(it.entry_base + it.sp[k] - it.sp[0]);
    }
    const int32_t e0 = it.sp[0];
    for (int64_t e = t0; e < it.entries; e += st) {
        ucols[it.entry_base + e] = it.cols[e0 + e] + it.col_off;
        const char* src = static_cast<const char*>(it.vals) + (e0 + e) * elem_bytes;
        char* dst = uvals + (it.entry_base + e) * elem_bytes;
        if (elem_bytes == 8) *reinterpret_cast<double*>(dst) = *reinterpret_cast<const double*>(src);
        else *reinterpret_cast<float*>(dst) = *reinterpret_cast<const float*>(src);
    }
}

__global__ void k_pack_sell_end(int32_t* __restrict__ usp, int64_t nslices_total, int64_t entries_total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) usp[nslices_total] = static_cast<int32_t>(entries_total);
}

// blockIdx.y = segment: dst[0, len) = src[0, len) (src NULL: zeros), then
// zeros up to len + zero_tail
__global__ void __launch_bounds__(256) k_copy_segments(const fa_copy_seg_t* __restrict__ segs, int elem_bytes) {
    const fa_copy_seg_t sg = segs[blockIdx.y];
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
    const int64_t tot = sg.len + sg.zero_tail;
    for (int64_t i = t0; i < tot; i += st) {
        const bool has = sg.src && i < sg.len;
        if (elem_bytes == 8) {
            static_cast<double*>(sg.dst)[i] = has ? static_cast<const double*>(sg.src)[i] : 0.0;
        } else if (elem_bytes == 4) {
            static_cast<float*>(sg.dst)[i] = has ? static_cast<const float*>(sg.src)[i] : 0.0f;
        } else {
            static_cast<int32_t*>(sg.dst)[i] = has ? static_cast<const int32_t*>(sg.src)[i] : 0;
        }
    }
}

// ---- batched PCG ---------------------------------------------------------------------
struct BState {
    double rz, pAp, alpha, beta, rnorm, bnorm;
    long long niter;
    int done;  // 0 running, 1 converged, 2 ||b|| < 1e-15, 3 maxit, 4 breakdown
};

// per-request sums of the partials of its workgroups [boff[q], boff[q+1]), in
// index order per thread, then the block's fixed tree
__device__ inline double seg_sum(const double* __restrict__ part, int32_t b0, int32_t b1) {
    double s = 0.0;
    for (int32_t i = b0 + static_cast<int32_t>(threadIdx.x); i < b1; i += blockDim.x) s += part[i];
    return block_sum(s);
}

enum BFinish { B_INIT = 0, B_RZ0 = 1, B_PAP = 2, B_CHECK = 3 };

// one workgroup for all requests, one after another; then the all-done flag
template <int WHAT>
__global__ void __launch_bounds__(RED_BLOCK) k_bpcg_finish(int nreq, const int32_t* __restrict__ boff,
                                                          const double* __restrict__ pa,
                                                          const double* __restrict__ pb, BState* __restrict__ st,
                                                          double atol, double rtol, long long maxit,
                                                          int* __restrict__ all_done) {
    __shared__ int s_run;
    if (threadIdx.x == 0) s_run = 0;
    for (int q = 0; q < nreq; ++q) {
        BState& S = st[q];
        const bool skip = (WHAT != B_INIT) && S.done;   // uniform: every thread reads the same state
        double t = 0.0, t2 = 0.0;
        if (!skip) {
            t = seg_sum(pa, boff[q], boff[q + 1]);
            if (WHAT == B_INIT) t2 = seg_sum(pb, boff[q], boff[q + 1]);
        }
        if (threadIdx.x == 0 && !skip) {
            if (WHAT == B_INIT) {
                S.rz = t;
                S.bnorm = sqrt(t2);
                S.niter = 0;
                S.rnorm = 0.0;
                S.done = S.bnorm < 1e-15 ? 2 : 0;
            } else if (WHAT == B_RZ0) {
                S.rz = t;
            } else if (WHAT == B_PAP) {
                S.pAp = t;
                S.alpha = S.rz / t;
            } else {
                const double rn = sqrt(t);
                S.rnorm = rn;
                S.niter += 1;
                if (rn < atol || rn < rtol * S.bnorm) S.done = 1;
                else if (S.niter >= maxit) S.done = 3;
                else if (!isfinite(rn) || !isfinite(S.alpha)) S.done = 4;
                S.beta = t / S.rz;
                S.rz = t;
            }
        }
        if (threadIdx.x == 0 && !S.done) s_run = 1;
        __syncthreads();
    }
    if (threadIdx.x == 0) *all_done = s_run ? 0 : 1;
}

// r = b - A x, ru = r in the cycle type, p = r; partials of r.r and b.b
template <class U>
__global__ void __launch_bounds__(BSEG) k_bpcg_init(fa_sell_t A, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   U* __restrict__ ru, double* __restrict__ p,
                                                   double* __restrict__ pa, double* __restrict__ pb) {
    const int64_t row = blockIdx.x * (int64_t)BSEG + threadIdx.x;
    const double ax = sell_row_dot<double, double>(A.slice_ptr, A.cols, static_cast<const double*>(A.vals), row, x);
    const double bv = b[row];
    const double rv = bv - ax;
    r[row] = rv;
    ru[row] = static_cast<U>(rv);
    p[row] = rv;
    const double v[2] = {block_sum(rv * rv), block_sum(bv * bv)};
    if (threadIdx.x == 0) {
        pa[blockIdx.x] = v[0];
        pb[blockIdx.x] = v[1];
    }
}

// partials of r.z (and p = z for the first direction when FIRST)
template <class U, bool FIRST>
__global__ void __launch_bounds__(BSEG) k_bpcg_rz(const double* __restrict__ r, const U* __restrict__ z,
                                                 double* __restrict__ p, double* __restrict__ pa,
                                                 const int32_t* __restrict__ breq, const BState* __restrict__ st) {
    if (st[breq[blockIdx.x]].done) return;
    const int64_t row = blockIdx.x * (int64_t)BSEG + threadIdx.x;
    const double zv = static_cast<double>(z[row]);
    if (FIRST) p[row] = zv;
    const double s = block_sum(r[row] * zv);
    if (threadIdx.x == 0) pa[blockIdx.x] = s;
}

// Ap = A p and partials of p.Ap
__global__ void __launch_bounds__(BSEG) k_bpcg_spmv(fa_sell_t A, const double* __restrict__ p, double* __restrict__ Ap,
                                                   double* __restrict__ pa, const int32_t* __restrict__ breq,
                                                   const BState* __restrict__ st) {
    if (st[breq[blockIdx.x]].done) return;
    const int64_t row = blockIdx.x * (int64_t)BSEG + threadIdx.x;
    const double v = sell_row_dot<double, double>(A.slice_ptr, A.cols, static_cast<const double*>(A.vals), row, p);
    Ap[row] = v;
    const double s = block_sum(p[row] * v);
    if (threadIdx.x == 0) pa[blockIdx.x] = s;
}

template <class U>
__global__ void __launch_bounds__(BSEG) k_bpcg_update(double* __restrict__ x, double* __restrict__ r,
                                                     U* __restrict__ ru, const double* __restrict__ p,
                                                     const double* __restrict__ Ap, const int32_t* __restrict__ breq,
                                                     const BState* __restrict__ st) {
    const BState& S = st[breq[blockIdx.x]];
    if (S.done) return;
    const int64_t row = blockIdx.x * (int64_t)BSEG + threadIdx.x;
    const double a = S.alpha;
    x[row] += a * p[row];
    const double rv = r[row] - a * Ap[row];
    r[row] = rv;
    ru[row] = static_cast<U>(rv);
}

template <class U>
__global__ void __launch_bounds__(BSEG) k_bpcg_pupdate(double* __restrict__ p, const U* __restrict__ z,
                                                      const int32_t* __restrict__ breq,
                                                      const BState* __restrict__ st) {
    const BState& S = st[breq[blockIdx.x]];
    if (S.done) return;
    const int64_t row = blockIdx.x * (int64_t)BSEG + threadIdx.x;
    p[row] = static_cast<double>(z[row]) + S.beta * p[row];
}

__global__ void k_bpcg_zero_flag(int* f) {
    if (threadIdx.x == 0) *f = 0;
}

}  // namespace fa

using namespace fa;

extern "C" int fa_pack_sell(int elem_bytes, int nreq, const fa_pack_sell_t* items_host, const fa_pack_sell_t* items_dev,
                            int64_t nslices_total, int64_t entries_total, int32_t* slice_ptr, int32_t* cols,
                            void* vals, fa_stream_t stream) {
    FA_REQUIRE((elem_bytes == 4 || elem_bytes == 8) && nreq > 0 && items_host && items_dev && slice_ptr && cols &&
                   vals && nreq <= 65535,
               "fa_pack_sell: bad arguments");
    int64_t most = 0;
    for (int q = 0; q < nreq; ++q) {
        const fa_pack_sell_t& it = items_host[q];
        FA_REQUIRE(it.entry_base + it.entries <= entries_total && it.slice_base + it.nslices + it.extra_slices <=
                       nslices_total,
                   "fa_pack_sell: request %d does not fit the packed arrays", q);
        most = std::max<int64_t>(most, std::max<int64_t>(it.entries, it.nslices + it.extra_slices + 1));
    }
    hipStream_t s = S(stream);
    const dim3 grid(grid_for(most, 256, 4096), nreq);
    hipLaunchKernelGGL(k_pack_sell, grid, dim3(256), 0, s, items_dev, elem_bytes, slice_ptr, cols,
                       static_cast<char*>(vals));
    hipLaunchKernelGGL(k_pack_sell_end, dim3(1), dim3(64), 0, s, slice_ptr, nslices_total, entries_total);
    FA_LAUNCH("k_pack_sell");
    return FA_OK;
}

extern "C" int fa_copy_segments(int elem_bytes, int nseg, const fa_copy_seg_t* segs_host, const fa_copy_seg_t* segs_dev,
                                fa_stream_t stream) {
    FA_REQUIRE((elem_bytes == 4 || elem_bytes == 8) && nseg > 0 && nseg <= 65535 && segs_host && segs_dev,
               "fa_copy_segments: bad arguments");
    int64_t most = 1;
    for (int q = 0; q < nseg; ++q) most = std::max<int64_t>(most, segs_host[q].len + segs_host[q].zero_tail);
    hipLaunchKernelGGL(k_copy_segments, dim3(grid_for(most, 256, 4096), nseg), dim3(256), 0, S(stream), segs_dev,
                       elem_bytes);
    FA_LAUNCH("k_copy_segments");
    return FA_OK;
}

extern "C" size_t fa_pcg_batched_work_bytes(int64_t n, int nreq) {
    const int64_t nb = (n + BSEG - 1) / BSEG;
    return 3 * align_up(n * 8) + 2 * align_up(n * 4) + 2 * align_up(nb * 8) + align_up(nreq * sizeof(BState)) +
           align_up(sizeof(int)) + 4096;
}

extern "C" int fa_pcg_batched(const fa_sell_t* A, const fa_mg_t* M, int nreq, const int32_t* block_req,
                              const int32_t* block_off, const double* b, double* x, double atol, double rtol,
                              int64_t maxit, int32_t check_every, int64_t* niter_host, double* residual_host,
                              int32_t* status_host, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(A && M && nreq > 0 && block_req && block_off && b && x && check_every > 0,
               "fa_pcg_batched: bad arguments");
    FA_REQUIRE(A->n > 0 && A->n % BSEG == 0, "fa_pcg_batched: the packed system must hold whole segments of %d rows",
               BSEG);
    FA_REQUIRE(M->dtype == FA_F32 && M->nlevels >= 1 && M->levels[0].A.n == A->n,
               "fa_pcg_batched: needs an FP32 cycle over the packed system");
    const int64_t n = A->n;
    const int64_t nb = n / BSEG;
    Arena ar(work, work_bytes);
    double* r = ar.take<double>(n);
    double* p = ar.take<double>(n);
    double* Ap = ar.take<double>(n);
    float* ru = ar.take<float>(n);
    float* z = ar.take<float>(n);
    double* pa = ar.take<double>(nb);
    double* pb = ar.take<double>(nb);
    BState* st = ar.take<BState>(nreq);
    int* all_done = ar.take<int>(1);
    if (!r || !p || !Ap || !ru || !z || !pa || !pb || !st || !all_done) {
        set_error("fa_pcg_batched: workspace too small (%zu bytes)", work_bytes);
        return FA_ENOMEM;
    }
    hipStream_t s = S(stream);
    const long long cap = maxit > 0 ? maxit : (1LL << 62);
    const dim3 g(static_cast<unsigned>(nb)), blk(BSEG), one(1), red(RED_BLOCK);
    hipLaunchKernelGGL(k_bpcg_zero_flag, one, dim3(64), 0, s, all_done);
    hipLaunchKernelGGL((k_bpcg_init<float>), g, blk, 0, s, *A, b, x, r, ru, p, pa, pb);
    hipLaunchKernelGGL(k_bpcg_finish<B_INIT>, one, red, 0, s, nreq, block_off, pa, pb, st, atol, rtol, cap, all_done);
    FA_LAUNCH("k_bpcg_init");
    int e = mg_vcycle(M, ru, z, all_done, s);
    if (e) return e;
    hipLaunchKernelGGL((k_bpcg_rz<float, true>), g, blk, 0, s, r, z, p, pa, block_req, st);
    hipLaunchKernelGGL(k_bpcg_finish<B_RZ0>, one, red, 0, s, nreq, block_off, pa, pb, st, atol, rtol, cap, all_done);
    FA_LAUNCH("k_bpcg_rz0");
    int h = 0;
    int64_t it = 0;
    for (;;) {
        const int64_t chunk = std::min<int64_t>(check_every, cap - it);
        for (int64_t k = 0; k < chunk; ++k) {
            hipLaunchKernelGGL(k_bpcg_spmv, g, blk, 0, s, *A, p, Ap, pa, block_req, st);
            hipLaunchKernelGGL(k_bpcg_finish<B_PAP>, one, red, 0, s, nreq, block_off, pa, pb, st, atol, rtol, cap,
                               all_done);
            hipLaunchKernelGGL((k_bpcg_update<float>), g, blk, 0, s, x, r, ru, p, Ap, block_req, st);
            e = mg_vcycle(M, ru, z, all_done, s);
            if (e) return e;
            hipLaunchKernelGGL((k_bpcg_rz<float, false>), g, blk, 0, s, r, z, p, pa, block_req, st);
            hipLaunchKernelGGL(k_bpcg_finish<B_CHECK>, one, red, 0, s, nreq, block_off, pa, pb, st, atol, rtol, cap,
                               all_done);
            hipLaunchKernelGGL((k_bpcg_pupdate<float>), g, blk, 0, s, p, z, block_req, st);
            FA_LAUNCH("batched pcg iteration");
        }
        it += chunk;
        FA_HIP(hipMemcpyAsync(&h, all_done, sizeof(int), hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        if (h || it >= cap) break;
    }
    std::vector<BState> hs(nreq);
    FA_HIP(hipMemcpyAsync(hs.data(), st, nreq * sizeof(BState), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < nreq; ++q) {
        if (niter_host) niter_host[q] = hs[q].niter;
        if (residual_host) residual_host[q] = hs[q].rnorm;
        if (status_host) status_host[q] = hs[q].done;
    }
    return FA_OK;
}
