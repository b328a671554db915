// Galerkin product C = R A P as a precomputed term list.
//
// The multigrid transfers take few distinct values (1 and 1/2 for nested P1
// grids and the P2 -> P1 transfer, 1, 2/3, 1/3 for P3 -> P1), so every coarse
// entry is a fixed weighted sum of fine entries:
//     c_IJ = sum over (i, j) with r_Ii a_ij p_jJ != 0 of (r_Ii p_jJ) a_ij.
// The symbolic phase lists, per coarse entry, the fine entry index t of each
// a_ij and a 4-bit code of the weight r_Ii p_jJ (one 32-bit word per term),
// in the enumeration order R entry -> A entry -> P entry.  The numeric phase
// (per assembly) is then one gather-sum per coarse entry in that order: no
// search, no atomics, a sum order fixed by the pattern.
//
// Terms of 64 consecutive coarse entries form a slice stored like SELL-64:
// term k of entry e at tslice[e / 64] + 64 k + e % 64, the slice as wide as
// its longest list, shorter lists padded with the zero-weight code.
#include <climits>
#include <cstdint>

#include "common.hpp"

namespace fa {

constexpr int TERM_IDX_BITS = 28;
constexpr uint32_t TERM_IDX_MASK = (1u << TERM_IDX_BITS) - 1;
constexpr uint32_t TERM_UNIT = 14;  // the unit diagonal of an empty coarse row
constexpr uint32_t TERM_PAD = 15;   // padding, weight 0
constexpr int TERM_MAX_VALUES = 3;  // distinct transfer values (codes ir * nw + ip < 9)

struct WeightSet {
    double v[TERM_MAX_VALUES];
    int n;
};

struct WeightTable {
    double w[16];
};

__device__ inline uint32_t weight_index(const WeightSet& ws, double x, int32_t* flags) {
    for (int k = 0; k < ws.n; ++k)
        if (x == ws.v[k]) return static_cast<uint32_t>(k);
    atomicOr(flags + 1, 1);  // a transfer value outside the set
    return 0;
}

__device__ inline int wave_inclusive_max(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v = max(v, u);
    }
    return v;
}

// One wavefront per coarse row I.  Enumerate the (J, term word) pairs of the
// row as in the symbolic product (R entries in order, lanes over the A row,
// each lane its P row), sort by (J, enumeration index), then walk the groups:
// group g is the row's g-th coarse entry (the pattern of fa_rap_fill).
// COUNT: counts[e] = terms of entry e.  FILL: write the words at their slots.
// flags: [0] capacity overflow, [1] transfer value outside the set, [2] A
// entry index past TERM_IDX_MASK.
template <int CAP, bool FILL>
__global__ void __launch_bounds__(WAVE) k_rap_terms(int64_t nrows, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ri, const double* __restrict__ rv,
                                                    const int32_t* __restrict__ ap, const int32_t* __restrict__ ai,
                                                    const int32_t* __restrict__ pp, const int32_t* __restrict__ pi,
                                                    const double* __restrict__ pv, WeightSet ws,
                                                    const int32_t* __restrict__ cptr, int32_t* __restrict__ counts,
                                                    const int32_t* __restrict__ tslice, uint32_t* __restrict__ terms,
                                                    int32_t* __restrict__ flags) {
    __shared__ uint64_t key[CAP];
    __shared__ uint32_t word[CAP];
    const int lane = threadIdx.x;
    for (int64_t I = blockIdx.x; I < nrows; I += gridDim.x) {
        int m = 0;
        bool ovf = false;
        for (int32_t u = rp[I]; u < rp[I + 1]; ++u) {
            const int32_t i = ri[u];
            const uint32_t ir = weight_index(ws, rv[u], flags);
            const int32_t a0 = ap[i], a1 = ap[i + 1];
            for (int32_t t0 = a0; t0 < a1; t0 += WAVE) {
                const int32_t t = t0 + lane;
                int32_t j = 0, cnt = 0;
                if (t < a1) {
                    j = ai[t];
                    cnt = pp[j + 1] - pp[j];
                    if (static_cast<uint32_t>(t) > TERM_IDX_MASK) atomicOr(flags + 2, 1);
                }
                const int inc = wave_inclusive_scan(cnt);
                const int tot = __shfl(inc, WAVE - 1);
                if (m + tot <= CAP) {
                    for (int c = 0; c < cnt; ++c) {
                        const int32_t q = pp[j] + c;
                        const int seq = m + inc - cnt + c;
                        const uint32_t code = ir * static_cast<uint32_t>(ws.n) + weight_index(ws, pv[q], flags);
                        key[seq] = (static_cast<uint64_t>(static_cast<uint32_t>(pi[q])) << 32) |
                                   static_cast<uint32_t>(seq);
                        word[seq] = (code << TERM_IDX_BITS) | (static_cast<uint32_t>(t) & TERM_IDX_MASK);
                    }
                } else {
                    ovf = true;
                }
                m += tot;
            }
        }
        __syncthreads();
        if (ovf) {
            if (lane == 0) atomicOr(flags, 1);
            m = 0;
        }
        const int32_t cbase = cptr[I];
        if (m == 0) {  // no candidates: the unit diagonal of a Dirichlet row
            if (lane == 0) {
                if (FILL) terms[tslice[cbase >> 6] + (cbase & 63)] = TERM_UNIT << TERM_IDX_BITS;
                else counts[cbase] = 1;
            }
            __syncthreads();
            continue;
        }
        int P = 2;
        while (P < m) P <<= 1;
        for (int t = m + lane; t < P; t += WAVE) key[t] = ~0ull;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < P; t += WAVE) {
                    const int p = t ^ j;
                    if (p > t) {
                        const uint64_t a = key[t], b = key[p];
                        if ((a > b) == ((t & k) == 0)) {
                            key[t] = b;
                            key[p] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        int rank_carry = 0, gs_carry = 0;
        for (int k0 = 0; k0 < m; k0 += WAVE) {
            const int k = k0 + lane;
            const bool valid = k < m;
            const uint32_t J = valid ? static_cast<uint32_t>(key[k] >> 32) : 0u;
            const bool start = valid && (k == 0 || static_cast<uint32_t>(key[k - 1] >> 32) != J);
            const int rank = rank_carry + wave_inclusive_scan(start ? 1 : 0) - 1;
            const int gs = max(gs_carry, wave_inclusive_max(start ? k : -1));
            if (valid) {
                const int32_t e = cbase + rank;
                const int pos = k - gs;
                if (FILL) {
                    const uint32_t seq = static_cast<uint32_t>(key[k] & 0xffffffffull);
                    terms[tslice[e >> 6] + 64 * pos + (e & 63)] = word[seq];
                } else {
                    const bool end = (k == m - 1) || static_cast<uint32_t>(key[k + 1] >> 32) != J;
                    if (end) counts[e] = pos + 1;
                }
            }
            rank_carry = __shfl(rank, WAVE - 1) + 1;
            gs_carry = __shfl(gs, WAVE - 1);
        }
        __syncthreads();
    }
}

// slice widths of the term layout: width[s] = 64 * max(counts of the slice)
__global__ void __launch_bounds__(256) k_term_widths(int64_t ne, const int32_t* __restrict__ counts,
                                                    int32_t* __restrict__ width) {
    const int64_t ns = (ne + 63) / 64;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ns * 64;
         e += (int64_t)gridDim.x * blockDim.x) {
        int c = e < ne ? counts[e] : 0;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o));
        if ((e & 63) == 0) width[e >> 6] = 64 * c;
    }
}

__global__ void __launch_bounds__(256) k_term_pad(int64_t n, uint32_t* __restrict__ terms) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        terms[i] = TERM_PAD << TERM_IDX_BITS;
}

// c[e] = sum_k w(code_k) a[t_k] in term order (code TERM_UNIT: the constant 1)
// (T: the value type of a and c; the sum is formed in FP64 either way)
template <class T>
__global__ void __launch_bounds__(256) k_rap_apply_terms(int64_t ne, const int32_t* __restrict__ tslice,
                                                        const uint32_t* __restrict__ terms, WeightTable wt,
                                                        const T* __restrict__ a, T* __restrict__ c) {
    __shared__ double tab[16];
    if (threadIdx.x < 16) tab[threadIdx.x] = wt.w[threadIdx.x];
    __syncthreads();
    constexpr int B = 16;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ((ne + 63) & ~63LL);
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = e >> 6;
        const int32_t base = tslice[s];
        const int w = (tslice[s + 1] - base) >> 6;
        const uint32_t* tp = terms + base + (e & 63);
        double acc = 0.0;
        for (int k0 = 0; k0 < w; k0 += B) {
            uint32_t ww[B];
#pragma unroll
            for (int u = 0; u < B; ++u) ww[u] = tp[64 * min(k0 + u, w - 1)];
            double av[B];
#pragma unroll
            for (int u = 0; u < B; ++u) av[u] = static_cast<double>(a[ww[u] & TERM_IDX_MASK]);
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const uint32_t code = ww[u] >> TERM_IDX_BITS;
                const double v = (code == TERM_UNIT) ? 1.0 : tab[code] * av[u];
                if (k0 + u < w) acc += v;
            }
        }
        if (e < ne) c[e] = static_cast<T>(acc);
    }
}

template <int CAP, bool FILL>
static void launch_terms(int64_t nrows, const int32_t* rp, const int32_t* ri, const double* rv, const int32_t* ap,
                         const int32_t* ai, const int32_t* pp, const int32_t* pi, const double* pv, const WeightSet& ws,
                         const int32_t* cptr, int32_t* counts, const int32_t* tslice, uint32_t* terms, int32_t* flags,
                         hipStream_t s) {
    hipLaunchKernelGGL((k_rap_terms<CAP, FILL>), dim3(grid_for(nrows, 1, 1 << 22)), dim3(WAVE), 0, s, nrows, rp, ri, rv,
                       ap, ai, pp, pi, pv, ws, cptr, counts, tslice, terms, flags);
}

template <bool FILL>
static int rap_terms(int32_t max_cand, int64_t nrows, const int32_t* rp, const int32_t* ri, const double* rv,
                     const int32_t* ap, const int32_t* ai, const int32_t* pp, const int32_t* pi, const double* pv,
                     const WeightSet& ws, const int32_t* cptr, int32_t* counts, const int32_t* tslice, uint32_t* terms,
                     int32_t* flags, hipStream_t s) {
    if (max_cand <= 512)
        launch_terms<512, FILL>(nrows, rp, ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tslice, terms, flags, s);
    else if (max_cand <= 2048)
        launch_terms<2048, FILL>(nrows, rp, ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tslice, terms, flags, s);
    else if (max_cand <= 8192)
        launch_terms<8192, FILL>(nrows, rp, ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tslice, terms, flags, s);
    else {
        set_error("fa_rap_terms: %d candidates per coarse row exceed the kernel capacity (8192)", max_cand);
        return FA_ECAPACITY;
    }
    FA_LAUNCH("k_rap_terms");
    return FA_OK;
}

static int weight_set(const double* values, int nvalues, WeightSet* ws) {
    if (nvalues < 1 || nvalues > TERM_MAX_VALUES) {
        set_error("fa_rap_terms: %d distinct transfer values (1 to %d supported)", nvalues, TERM_MAX_VALUES);
        return FA_ECAPACITY;
    }
    *ws = WeightSet{};
    for (int k = 0; k < nvalues; ++k) ws->v[k] = values[k];
    ws->n = nvalues;
    return FA_OK;
}

__global__ void __launch_bounds__(256) k_zero_counts(int64_t n, int32_t* __restrict__ c) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c[i] = 0;
}

__global__ void k_zero_flags(int32_t* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

static int check_flags(const int32_t* flags, hipStream_t s, const char* who) {
    int32_t h[4] = {0, 0, 0, 0};
    FA_HIP(hipMemcpyAsync(h, flags, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (h[0]) {
        set_error("%s: max_cand is smaller than a coarse row's candidate count", who);
        return FA_ECAPACITY;
    }
    if (h[1]) {
        set_error("%s: a transfer value is not in the given value set", who);
        return FA_EINVAL;
    }
    if (h[2]) {
        set_error("%s: fine operator has more than 2^28 entries", who);
        return FA_ECAPACITY;
    }
    return FA_OK;
}

}  // namespace fa

using namespace fa;

extern "C" size_t fa_rap_terms_work_bytes(int64_t nnz_c) {
    const int64_t ns = (nnz_c + 63) / 64;
    return align_up(nnz_c * sizeof(int32_t)) + align_up((ns + 1) * sizeof(int32_t)) + exclusive_scan_work_bytes(ns) +
           1024;
}

extern "C" int fa_rap_terms_count(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                                  const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr,
                                  const int32_t* p_idx, const double* p_val, int32_t max_cand, const int32_t* c_ptr,
                                  int64_t nnz_c, const double* values, int nvalues, int32_t* tslice,
                                  int64_t* nterms_host, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nrows > 0 && nnz_c > 0 && tslice, "fa_rap_terms_count: bad arguments");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    hipStream_t s = S(stream);
    const int64_t ns = (nnz_c + 63) / 64;
    Arena ar(work, work_bytes);
    int32_t* counts = ar.take<int32_t>(nnz_c);
    int32_t* width = ar.take<int32_t>(ns + 1);
    int32_t* flags = ar.take<int32_t>(4);
    const size_t sb = exclusive_scan_work_bytes(ns);
    void* sw = ar.take<char>(sb);
    if (!counts || !width || !flags || !sw) {
        set_error("fa_rap_terms_count: workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, flags);
    hipLaunchKernelGGL(k_zero_counts, dim3(grid_for(nnz_c, 256, 1 << 16)), dim3(256), 0, s, nnz_c, counts);
    st = rap_terms<false>(max_cand, nrows, r_ptr, r_idx, r_val, a_ptr, a_idx, p_ptr, p_idx, p_val, ws, c_ptr, counts,
                          nullptr, nullptr, flags, s);
    if (st) return st;
    hipLaunchKernelGGL(k_term_widths, dim3(grid_for(ns * 64, 256, 1 << 16)), dim3(256), 0, s, nnz_c, counts, width);
    FA_LAUNCH("k_term_widths");
    st = exclusive_scan_i32(width, tslice, ns, sw, sb, s);
    if (st) return st;
    st = check_flags(flags, s, "fa_rap_terms_count");
    if (st) return st;
    int32_t total = 0;
    FA_HIP(hipMemcpyAsync(&total, tslice + ns, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (total < 0) {
        set_error("fa_rap_terms_count: more than 2^31 terms");
        return FA_ECAPACITY;
    }
    if (nterms_host) *nterms_host = total;
    return FA_OK;
}

extern "C" int fa_rap_terms_fill(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                                 const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr, const int32_t* p_idx,
                                 const double* p_val, int32_t max_cand, const int32_t* c_ptr, int64_t nnz_c,
                                 const double* values, int nvalues, const int32_t* tslice, int64_t nterms,
                                 uint32_t* terms, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nrows > 0 && nnz_c > 0 && tslice && terms && nterms > 0, "fa_rap_terms_fill: bad arguments");
    FA_REQUIRE(work && work_bytes >= 4 * sizeof(int32_t), "fa_rap_terms_fill: workspace too small");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    hipStream_t s = S(stream);
    int32_t* flags = static_cast<int32_t*>(work);
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, flags);
    hipLaunchKernelGGL(k_term_pad, dim3(grid_for(nterms, 256, 1 << 16)), dim3(256), 0, s, nterms, terms);
    FA_LAUNCH("k_term_pad");
    st = rap_terms<true>(max_cand, nrows, r_ptr, r_idx, r_val, a_ptr, a_idx, p_ptr, p_idx, p_val, ws, c_ptr, nullptr,
                         tslice, terms, flags, s);
    if (st) return st;
    return check_flags(flags, s, "fa_rap_terms_fill");
}

extern "C" int fa_rap_values_terms(int dtype, int64_t nnz_c, const int32_t* tslice, const uint32_t* terms,
                                   const double* values, int nvalues, const void* a_val, void* c_val,
                                   fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_rap_values_terms: bad dtype %d", dtype);
    FA_REQUIRE(nnz_c > 0 && tslice && terms && a_val && c_val, "fa_rap_values_terms: bad arguments");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    WeightTable wt{};
    for (int ir = 0; ir < ws.n; ++ir)
        for (int ip = 0; ip < ws.n; ++ip) wt.w[ir * ws.n + ip] = ws.v[ir] * ws.v[ip];
    wt.w[TERM_UNIT] = 1.0;
    wt.w[TERM_PAD] = 0.0;
    const dim3 grid(grid_for(((nnz_c + 63) / 64) * 64, 256, 1 << 16));
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_rap_apply_terms<double>, grid, dim3(256), 0, S(stream), nnz_c, tslice, terms, wt,
                           static_cast<const double*>(a_val), static_cast<double*>(c_val));
    else
        hipLaunchKernelGGL(k_rap_apply_terms<float>, grid, dim3(256), 0, S(stream), nnz_c, tslice, terms, wt,
                           static_cast<const float*>(a_val), static_cast<float*>(c_val));
    FA_LAUNCH("k_rap_apply_terms");
    return FA_OK;
}
