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
// The terms of coarse entry e are stored contiguously, tptr[e] .. tptr[e+1]
// (CSR over the coarse entries, no padding): the numeric kernel gives each
// entry 16 lanes that read its list coalesced, so a diagonal entry with five
// times its neighbours' terms costs no padding in the others.
#include <climits>
#include <cstdint>

#include "common.hpp"

namespace fa {

constexpr int TERM_IDX_BITS = 28;
constexpr uint32_t TERM_IDX_MASK = (1u << TERM_IDX_BITS) - 1;
constexpr uint32_t TERM_UNIT = 14;  // the unit diagonal of an empty coarse row
constexpr int TERM_MAX_VALUES = 3;  // distinct transfer values (codes ir * nw + ip < 9)
// Staged form: the fine rows of R row I (in R's order) are staged one after
// another, and a term names its A entry by its slot in that block: a 16-bit
// word, slot in bits 0-8, weight code in bits 9-12.
constexpr int RAP_STAGE_MAX = 512;  // staged fine entries per coarse row
constexpr int RAP_ROW_WORDS = 1024;  // term words per coarse row (staged form)
constexpr int STAGE_SLOT_BITS = 9;
constexpr uint32_t STAGE_SLOT_MASK = (1u << STAGE_SLOT_BITS) - 1;

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
// entry index past TERM_IDX_MASK, [3] (STAGED) a coarse row past the staged
// form's bounds (fine rows holding more than RAP_STAGE_MAX entries, more than
// 64 R entries, 64 or more coarse entries, more than RAP_ROW_WORDS terms).
// STAGED: 16-bit words with staging slots.
template <int CAP, bool FILL, bool STAGED>
__global__ void __launch_bounds__(WAVE) k_rap_terms(int64_t nrows, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ri, const double* __restrict__ rv,
                                                    const int32_t* __restrict__ ap, const int32_t* __restrict__ ai,
                                                    const int32_t* __restrict__ pp, const int32_t* __restrict__ pi,
                                                    const double* __restrict__ pv, WeightSet ws,
                                                    const int32_t* __restrict__ cptr, int32_t* __restrict__ counts,
                                                    const int32_t* __restrict__ tptr, void* __restrict__ terms,
                                                    int32_t* __restrict__ flags) {
    __shared__ uint64_t key[CAP];
    __shared__ uint32_t word[CAP];
    const int lane = threadIdx.x;
    for (int64_t I = blockIdx.x; I < nrows; I += gridDim.x) {
        int m = 0;
        int off = 0;  // staging slots of the fine rows of R entries before u
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
                        word[seq] = STAGED ? (code << STAGE_SLOT_BITS) |
                                                 (static_cast<uint32_t>(off + t - a0) & STAGE_SLOT_MASK)
                                           : (code << TERM_IDX_BITS) | (static_cast<uint32_t>(t) & TERM_IDX_MASK);
                    }
                } else {
                    ovf = true;
                }
                m += tot;
            }
            off += a1 - a0;
        }
        if (STAGED && FILL && lane == 0) {
            const bool over = off > RAP_STAGE_MAX || rp[I + 1] - rp[I] > WAVE || cptr[I + 1] - cptr[I] >= WAVE ||
                              tptr[cptr[I + 1]] - tptr[cptr[I]] > RAP_ROW_WORDS;
            if (over) atomicOr(flags + 3, 1);
        }
        __syncthreads();
        if (ovf) {
            if (lane == 0) atomicOr(flags, 1);
            m = 0;
        }
        const int32_t cbase = cptr[I];
        if (m == 0) {  // no candidates: the unit diagonal of a Dirichlet row
            if (lane == 0) {
                if (FILL) {
                    if (STAGED) static_cast<uint16_t*>(terms)[tptr[cbase]] = TERM_UNIT << STAGE_SLOT_BITS;
                    else static_cast<uint32_t*>(terms)[tptr[cbase]] = TERM_UNIT << TERM_IDX_BITS;
                }
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
                    if (STAGED) static_cast<uint16_t*>(terms)[tptr[e] + pos] = static_cast<uint16_t>(word[seq]);
                    else static_cast<uint32_t*>(terms)[tptr[e] + pos] = word[seq];
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

// c[e] = sum_k w(code_k) a[t_k] over the terms of entry e (code TERM_UNIT:
// the constant 1).  RAP_G lanes per entry: lane h adds the terms h, h + G,
// h + 2G, ... in order, then the lane sums are combined by a fixed butterfly,
// so the sum order depends on the term count only and c is the same run to
// run.  T: the value type of a and c; the sum is formed in FP64 either way.
// Consecutive entries go to one XCD (their terms gather nearby fine rows).
constexpr int RAP_G = 16;
template <class T>
__global__ void __launch_bounds__(256) k_rap_apply_terms(int64_t ne, const int32_t* __restrict__ tptr,
                                                        const uint32_t* __restrict__ terms, WeightTable wt,
                                                        const T* __restrict__ a, T* __restrict__ c) {
    __shared__ double tab[16];
    if (threadIdx.x < 16) tab[threadIdx.x] = wt.w[threadIdx.x];
    __syncthreads();
    constexpr int G = RAP_G, U = 2;
    const int h = threadIdx.x & (G - 1);
    const int64_t per_block = blockDim.x / G;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * per_block;
    // the trip count is the same for all lanes of a group, so the butterfly
    // below runs with the whole group active
    for (int64_t e = xcd_block() * per_block + threadIdx.x / G; e < ne; e += stride) {
        const int32_t b = tptr[e];
        const int n = tptr[e + 1] - b;
        double acc = 0.0;
        for (int k0 = 0; k0 < n; k0 += G * U) {
            uint32_t ww[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * G + h;
                ww[u] = k < n ? terms[b + k] : 0u;
            }
            double av[U];
#pragma unroll
            for (int u = 0; u < U; ++u) av[u] = static_cast<double>(a[ww[u] & TERM_IDX_MASK]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t code = ww[u] >> TERM_IDX_BITS;
                const double v = (code == TERM_UNIT) ? 1.0 : tab[code] * av[u];
                if (k0 + u * G + h < n) acc += v;
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
        if (h == 0) c[e] = static_cast<T>(acc);
    }
}

// Staged form, one wave per coarse row I, for rows of at most 64 R entries,
// 64 coarse entries, RAP_ROW_WORDS term words and RAP_STAGE_MAX staged fine
// entries (the host checks these bounds; the kernel skips a row past them):
//   1. lane j loads R entry j of the row and the offset tptr of entry j;
//   2. the A row bounds of each R entry, and all the row's term words into
//      LDS (issued together);
//   3. the A rows themselves into LDS one after another (16 lanes a row, each
//      row read contiguously; offsets from a wave scan of the row lengths).
// Each entry is then summed from LDS as in k_rap_apply_terms (16 lanes an
// entry, lane h adding terms h, h + 16, ... in order, the same butterfly):
// the values of fa_rap_values_terms bit for bit, with three dependent rounds
// of loads per coarse row and no gathers from A.
template <class T>
__global__ void __launch_bounds__(WAVE) k_rap_apply_staged(int64_t nrows, const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ri,
                                                          const int32_t* __restrict__ ap,
                                                          const int32_t* __restrict__ cp,
                                                          const int32_t* __restrict__ tptr,
                                                          const uint16_t* __restrict__ terms, WeightTable wt,
                                                          const T* __restrict__ a, T* __restrict__ c) {
    __shared__ double tab[16];
    __shared__ T stage[RAP_STAGE_MAX];
    __shared__ uint16_t wbuf[RAP_ROW_WORDS];
    __shared__ int32_t ebuf[WAVE + 1];
    const int lane = threadIdx.x;
    if (lane < 16) tab[lane] = wt.w[lane];
    const int g = lane >> 4, h = lane & (RAP_G - 1);
    for (int64_t I = xcd_block(); I < nrows; I += gridDim.x) {
        const int32_t u0 = rp[I], u1 = rp[I + 1], e0 = cp[I], e1 = cp[I + 1];
        const int nr = u1 - u0, ne = e1 - e0;
        int32_t toff = 0;
        if (lane <= ne && lane <= WAVE) toff = tptr[e0 + min(lane, ne)];
        int32_t i = -1;
        if (lane < nr) i = ri[u0 + lane];
        const int32_t wbase = __shfl(toff, 0);
        const int32_t wend = __shfl(toff, min(ne, WAVE - 1));
        // a row the staged form does not take (the host selects rows within
        // the bounds, so this is a guard, not a path)
        const bool fits = nr <= WAVE && ne < WAVE && wend - wbase <= RAP_ROW_WORDS;
        if (!fits) continue;
        // round 2: the A row bounds and the first 8 x 64 words of the row,
        // all in flight together (values to LDS once both have arrived)
        const int nw = wend - wbase;
        uint16_t wv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = lane + q * WAVE;
            wv[q] = k < nw ? terms[wbase + k] : uint16_t(0);
        }
        int32_t a0 = 0, len = 0;
        if (i >= 0) {
            a0 = ap[i];
            len = ap[i + 1] - a0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = lane + q * WAVE;
            if (k < nw) wbuf[k] = wv[q];
        }
        for (int k = lane + 8 * WAVE; k < nw; k += WAVE) wbuf[k] = terms[wbase + k];
        // round 3: the A rows, 16 lanes a row; the first 16 entries of the
        // first 16 R entries at once, then whatever is left
        const int incl = wave_inclusive_scan(len);
        const int base = incl - len;
        T av[4];
        int dst[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = q * 4 + g;
            const int32_t aj = __shfl(a0, j);
            const int lj = j < nr ? __shfl(len, j) : 0;
            const int bj = __shfl(base, j);
            dst[q] = h < lj ? bj + h : -1;
            av[q] = dst[q] >= 0 ? a[aj + h] : T(0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (dst[q] >= 0 && dst[q] < RAP_STAGE_MAX) stage[dst[q]] = av[q];
        for (int jb = 0; jb < nr; jb += 4) {
            const int j = jb + g;
            const int32_t aj = __shfl(a0, j & (WAVE - 1));
            const int lj = j < nr ? __shfl(len, j & (WAVE - 1)) : 0;
            const int bj = __shfl(base, j & (WAVE - 1));
            for (int k = (jb < 16 ? RAP_G : 0) + h; k < lj; k += RAP_G)
                if (bj + k < RAP_STAGE_MAX) stage[bj + k] = a[aj + k];
        }
        if (lane <= ne) ebuf[lane] = toff - wbase;
        __syncthreads();
        for (int eb = 0; eb < ne; eb += 4) {
            const int el = eb + g;
            if (el < ne) {  // uniform within a group of 16 lanes
                const int b = ebuf[el];
                const int n = ebuf[el + 1] - b;
                double acc = 0.0;
                for (int k = h; k < n; k += RAP_G) {
                    const uint32_t w = wbuf[(b + k) & (RAP_ROW_WORDS - 1)];
                    const uint32_t code = w >> STAGE_SLOT_BITS;
                    acc += (code == TERM_UNIT) ? 1.0 : tab[code] * static_cast<double>(stage[w & STAGE_SLOT_MASK]);
                }
#pragma unroll
                for (int o = RAP_G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, RAP_G);
                if (h == 0) c[e0 + el] = static_cast<T>(acc);
            }
        }
        __syncthreads();  // the next row overwrites the staged values
    }
}

template <int CAP, bool FILL, bool STAGED>
static void launch_terms(int64_t nrows, const int32_t* rp, const int32_t* ri, const double* rv, const int32_t* ap,
                         const int32_t* ai, const int32_t* pp, const int32_t* pi, const double* pv, const WeightSet& ws,
                         const int32_t* cptr, int32_t* counts, const int32_t* tptr, void* terms, int32_t* flags,
                         hipStream_t s) {
    hipLaunchKernelGGL((k_rap_terms<CAP, FILL, STAGED>), dim3(grid_for(nrows, 1, 1 << 22)), dim3(WAVE), 0, s, nrows, rp,
                       ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tptr, terms, flags);
}

template <bool FILL, bool STAGED = false>
static int rap_terms(int32_t max_cand, int64_t nrows, const int32_t* rp, const int32_t* ri, const double* rv,
                     const int32_t* ap, const int32_t* ai, const int32_t* pp, const int32_t* pi, const double* pv,
                     const WeightSet& ws, const int32_t* cptr, int32_t* counts, const int32_t* tptr, void* terms,
                     int32_t* flags, hipStream_t s) {
    if (max_cand <= 512)
        launch_terms<512, FILL, STAGED>(nrows, rp, ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tptr, terms, flags, s);
    else if (max_cand <= 2048)
        launch_terms<2048, FILL, STAGED>(nrows, rp, ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tptr, terms, flags,
                                         s);
    else if (max_cand <= 8192)
        launch_terms<8192, FILL, STAGED>(nrows, rp, ri, rv, ap, ai, pp, pi, pv, ws, cptr, counts, tptr, terms, flags,
                                         s);
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
    FA_HIP(hipMemcpyAsync(h, flags, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
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
    if (h[3]) {
        set_error("%s: a coarse row exceeds the staged form (fine rows of more than %d entries, more than 64 R "
                  "entries, 64 or more coarse entries or more than %d terms)", who, RAP_STAGE_MAX, RAP_ROW_WORDS);
        return FA_ECAPACITY;
    }
    return FA_OK;
}

}  // namespace fa

using namespace fa;

extern "C" size_t fa_rap_terms_work_bytes(int64_t nnz_c) {
    return align_up(nnz_c * sizeof(int32_t)) + exclusive_scan_work_bytes(nnz_c) + 1024;
}

extern "C" int fa_rap_terms_count(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                                  const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr,
                                  const int32_t* p_idx, const double* p_val, int32_t max_cand, const int32_t* c_ptr,
                                  int64_t nnz_c, const double* values, int nvalues, int32_t* tptr,
                                  int64_t* nterms_host, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nrows > 0 && nnz_c > 0 && tptr, "fa_rap_terms_count: bad arguments");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    int32_t* counts = ar.take<int32_t>(nnz_c);
    int32_t* flags = ar.take<int32_t>(4);
    const size_t sb = exclusive_scan_work_bytes(nnz_c);
    void* sw = ar.take<char>(sb);
    if (!counts || !flags || !sw) {
        set_error("fa_rap_terms_count: workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, flags);
    hipLaunchKernelGGL(k_zero_counts, dim3(grid_for(nnz_c, 256, 1 << 16)), dim3(256), 0, s, nnz_c, counts);
    st = rap_terms<false>(max_cand, nrows, r_ptr, r_idx, r_val, a_ptr, a_idx, p_ptr, p_idx, p_val, ws, c_ptr, counts,
                          nullptr, nullptr, flags, s);
    if (st) return st;
    st = exclusive_scan_i32(counts, tptr, nnz_c, sw, sb, s);
    if (st) return st;
    st = check_flags(flags, s, "fa_rap_terms_count");
    if (st) return st;
    int32_t total = 0;
    FA_HIP(hipMemcpyAsync(&total, tptr + nnz_c, sizeof(int32_t), hipMemcpyDeviceToHost, s));
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
                                 const double* values, int nvalues, const int32_t* tptr, int64_t nterms,
                                 uint32_t* terms, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nrows > 0 && nnz_c > 0 && tptr && terms && nterms > 0, "fa_rap_terms_fill: bad arguments");
    FA_REQUIRE(work && work_bytes >= 4 * sizeof(int32_t), "fa_rap_terms_fill: workspace too small");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    hipStream_t s = S(stream);
    int32_t* flags = static_cast<int32_t*>(work);
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, flags);
    st = rap_terms<true>(max_cand, nrows, r_ptr, r_idx, r_val, a_ptr, a_idx, p_ptr, p_idx, p_val, ws, c_ptr, nullptr,
                         tptr, terms, flags, s);
    if (st) return st;
    return check_flags(flags, s, "fa_rap_terms_fill");
}

extern "C" int fa_rap_values_terms(int dtype, int64_t nnz_c, const int32_t* tptr, const uint32_t* terms,
                                   const double* values, int nvalues, const void* a_val, void* c_val,
                                   fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_rap_values_terms: bad dtype %d", dtype);
    FA_REQUIRE(nnz_c > 0 && tptr && terms && a_val && c_val, "fa_rap_values_terms: bad arguments");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    WeightTable wt{};
    for (int ir = 0; ir < ws.n; ++ir)
        for (int ip = 0; ip < ws.n; ++ip) wt.w[ir * ws.n + ip] = ws.v[ir] * ws.v[ip];
    wt.w[TERM_UNIT] = 1.0;
    const dim3 grid(xcd_grid(grid_for(nnz_c * RAP_G, 256, 1 << 16)));
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_rap_apply_terms<double>, grid, dim3(256), 0, S(stream), nnz_c, tptr, terms, wt,
                           static_cast<const double*>(a_val), static_cast<double*>(c_val));
    else
        hipLaunchKernelGGL(k_rap_apply_terms<float>, grid, dim3(256), 0, S(stream), nnz_c, tptr, terms, wt,
                           static_cast<const float*>(a_val), static_cast<float*>(c_val));
    FA_LAUNCH("k_rap_apply_terms");
    return FA_OK;
}

extern "C" int fa_rap_terms_fill_staged(int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx, const double* r_val,
                                        const int32_t* a_ptr, const int32_t* a_idx, const int32_t* p_ptr,
                                        const int32_t* p_idx, const double* p_val, int32_t max_cand,
                                        const int32_t* c_ptr, int64_t nnz_c, const double* values, int nvalues,
                                        const int32_t* tptr, int64_t nterms, uint16_t* terms, void* work,
                                        size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nrows > 0 && nnz_c > 0 && tptr && terms && nterms > 0, "fa_rap_terms_fill_staged: bad arguments");
    FA_REQUIRE(work && work_bytes >= 4 * sizeof(int32_t), "fa_rap_terms_fill_staged: workspace too small");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    hipStream_t s = S(stream);
    int32_t* flags = static_cast<int32_t*>(work);
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, flags);
    st = rap_terms<true, true>(max_cand, nrows, r_ptr, r_idx, r_val, a_ptr, a_idx, p_ptr, p_idx, p_val, ws, c_ptr,
                               nullptr, tptr, terms, flags, s);
    if (st) return st;
    return check_flags(flags, s, "fa_rap_terms_fill_staged");
}

extern "C" int fa_rap_values_staged(int dtype, int64_t nrows, const int32_t* r_ptr, const int32_t* r_idx,
                                    const int32_t* a_ptr, const int32_t* c_ptr, const int32_t* tptr,
                                    const uint16_t* terms, const double* values, int nvalues, const void* a_val,
                                    void* c_val, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_rap_values_staged: bad dtype %d", dtype);
    FA_REQUIRE(nrows > 0 && r_ptr && r_idx && a_ptr && c_ptr && tptr && terms && a_val && c_val,
               "fa_rap_values_staged: bad arguments");
    WeightSet ws;
    int st = weight_set(values, nvalues, &ws);
    if (st) return st;
    WeightTable wt{};
    for (int ir = 0; ir < ws.n; ++ir)
        for (int ip = 0; ip < ws.n; ++ip) wt.w[ir * ws.n + ip] = ws.v[ir] * ws.v[ip];
    wt.w[TERM_UNIT] = 1.0;
    const dim3 grid(xcd_grid(grid_for(nrows, 1, 1 << 20)));
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_rap_apply_staged<double>, grid, dim3(WAVE), 0, S(stream), nrows, r_ptr, r_idx, a_ptr,
                           c_ptr, tptr, terms, wt, static_cast<const double*>(a_val), static_cast<double*>(c_val));
    else
        hipLaunchKernelGGL(k_rap_apply_staged<float>, grid, dim3(WAVE), 0, S(stream), nrows, r_ptr, r_idx, a_ptr,
                           c_ptr, tptr, terms, wt, static_cast<const float*>(a_val), static_cast<float*>(c_val));
    FA_LAUNCH("k_rap_apply_staged");
    return FA_OK;
}
