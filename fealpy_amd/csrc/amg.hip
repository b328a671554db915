// Aggregation for algebraic multigrid (the role of the reference's AMG setup,
// fealpy/solver/amg.py and the algebraic branch of GAMGSolver,
// fealpy/solver/gamg_solver.py:104-121).
//
// Strength of connection: j is a strong neighbour of i when
// |a_ij| >= theta sqrt(|a_ii a_jj|) (symmetric for symmetric A).  Rows with
// no strong off-diagonal entry (Dirichlet rows among them) stay out of every
// aggregate.  Aggregates are grown from a distance-2 maximal independent set
// (MIS-2) of the strength graph, computed in parallel rounds as in Bell,
// Dalton and Olson, "Exposing fine-grained parallelism in algebraic multigrid
// methods" (SIAM J. Sci. Comput. 34, 2012): every undecided node carries a
// key (a hash of its index as priority, the index as tie-break); a node whose
// key is the largest within two strong hops becomes a root, nodes within two
// hops of a root leave the candidate set.  Then every node joins a root
// neighbour (the one with the largest key), and the remaining nodes join a
// neighbour that joined in that step.  Every round depends only on the
// previous round's state, so the aggregates are the same run to run.
#include "common.hpp"

namespace fa {

constexpr uint64_t KEY_ROOT = ~0ULL;
constexpr uint64_t KEY_OUT = 0ULL;  // decided non-root, or not aggregated
enum : int8_t { ST_UNDECIDED = 0, ST_ROOT = 1, ST_OUT = 2, ST_ISOLATED = 3 };

__device__ inline uint32_t hash32(uint32_t x) {  // a fixed integer mix (murmur3 finaliser)
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

__device__ inline uint64_t undecided_key(int64_t i) {
    return (1ULL << 62) | (static_cast<uint64_t>(hash32(static_cast<uint32_t>(i)) >> 2) << 31) |
           static_cast<uint64_t>(i & 0x7fffffff);
}

// strong-entry mask (1 byte per entry) and the isolated flag of each row
template <class T>
__global__ void __launch_bounds__(256) k_amg_strength(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ idx, const T* __restrict__ val,
                                                     double theta, uint8_t* __restrict__ strong,
                                                     int8_t* __restrict__ state) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = 0.0;
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
            if (idx[k] == i) d = fabs(static_cast<double>(val[k]));
        int cnt = 0;
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            const int32_t j = idx[k];
            bool s = false;
            if (j != i) {
                double dj = 0.0;
                for (int32_t m = ptr[j]; m < ptr[j + 1]; ++m)
                    if (idx[m] == j) dj = fabs(static_cast<double>(val[m]));
                const double a = fabs(static_cast<double>(val[k]));
                s = a != 0.0 && a >= theta * sqrt(d * dj);
            }
            strong[k] = s ? 1 : 0;
            cnt += s ? 1 : 0;
        }
        state[i] = cnt ? ST_UNDECIDED : ST_ISOLATED;
    }
}

__device__ inline uint64_t state_key(int8_t st, int64_t i) {
    return st == ST_UNDECIDED ? undecided_key(i) : (st == ST_ROOT ? KEY_ROOT : KEY_OUT);
}

// m[i] = max of the keys of i and its strong neighbours; `in` = NULL takes
// the keys from the state (first hop), else from the previous hop's maxima
__global__ void __launch_bounds__(256) k_amg_max(int64_t n, const int32_t* __restrict__ ptr,
                                                const int32_t* __restrict__ idx, const uint8_t* __restrict__ strong,
                                                const int8_t* __restrict__ state, const uint64_t* __restrict__ in,
                                                uint64_t* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t m = in ? in[i] : state_key(state[i], i);
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            if (!strong[k]) continue;
            const int32_t j = idx[k];
            const uint64_t v = in ? in[j] : state_key(state[j], j);
            m = v > m ? v : m;
        }
        out[i] = m;
    }
}

// undecided nodes: the maximum of their two-hop neighbourhood becomes a root,
// a root within two hops takes them out; counts the still undecided
__global__ void __launch_bounds__(256) k_amg_decide(int64_t n, const uint64_t* __restrict__ m2,
                                                   int8_t* __restrict__ state, int32_t* __restrict__ undecided) {
    int cnt = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (state[i] != ST_UNDECIDED) continue;
        const uint64_t m = m2[i];
        if (m == undecided_key(i)) state[i] = ST_ROOT;
        else if (m == KEY_ROOT) state[i] = ST_OUT;
        else ++cnt;
    }
    cnt = wave_sum(cnt);
    if ((threadIdx.x & (WAVE - 1)) == 0 && cnt) atomicAdd(undecided, cnt);
}

__global__ void __launch_bounds__(256) k_amg_root_flags(int64_t n, const int8_t* __restrict__ state,
                                                       int32_t* __restrict__ flag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = state[i] == ST_ROOT ? 1 : 0;
}

// pass 0: roots take their number (exclusive scan of the root flags);
// pass 1: other nodes join the strong root neighbour with the largest key;
// pass 2: nodes still out join the strong neighbour (assigned in pass 1) with
// the largest key.  `src` is the previous pass's assignment, `dst` this one's.
__global__ void __launch_bounds__(256) k_amg_assign(int64_t n, int pass, const int32_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ idx, const uint8_t* __restrict__ strong,
                                                   const int8_t* __restrict__ state, const int32_t* __restrict__ rid,
                                                   const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int8_t st = state[i];
        if (pass == 0) {
            dst[i] = st == ST_ROOT ? rid[i] : -1;
            continue;
        }
        int32_t a = src[i];
        if (a < 0 && st != ST_ISOLATED) {
            uint64_t best = 0;
            for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
                if (!strong[k]) continue;
                const int32_t j = idx[k];
                const bool ok = pass == 1 ? state[j] == ST_ROOT : src[j] >= 0;
                if (!ok) continue;
                const uint64_t key = undecided_key(j);
                if (key > best) {
                    best = key;
                    a = src[j];
                }
            }
        }
        dst[i] = a;
    }
}

__global__ void __launch_bounds__(256) k_amg_count_unassigned(int64_t n, const int8_t* __restrict__ state,
                                                             const int32_t* __restrict__ agg,
                                                             int32_t* __restrict__ bad) {
    int cnt = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        cnt += (state[i] != ST_ISOLATED && agg[i] < 0) ? 1 : 0;
    cnt = wave_sum(cnt);
    if ((threadIdx.x & (WAVE - 1)) == 0 && cnt) atomicAdd(bad, cnt);
}

__global__ void k_amg_zero(int32_t* p) {
    if (threadIdx.x == 0) *p = 0;
}

// smoothed prolongation P = (I - omega D^-1 A) P_tent on the pattern of
// A P_tent (which holds P_tent's entries, as a_ii != 0 on aggregated rows):
// p_ij = [agg(i) == j] - omega / a_ii * (A P_tent)_ij
template <class T>
__global__ void __launch_bounds__(256) k_amg_smooth_p(int64_t n, const int32_t* __restrict__ a_ptr,
                                                     const int32_t* __restrict__ a_idx, const T* __restrict__ a_val,
                                                     const int32_t* __restrict__ agg, const int32_t* __restrict__ s_ptr,
                                                     const int32_t* __restrict__ s_idx, const double* __restrict__ s_val,
                                                     double omega, double* __restrict__ p_val) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = 0.0;
        for (int32_t k = a_ptr[i]; k < a_ptr[i + 1]; ++k)
            if (a_idx[k] == i) d = static_cast<double>(a_val[k]);
        const double w = d != 0.0 ? omega / d : 0.0;
        for (int32_t k = s_ptr[i]; k < s_ptr[i + 1]; ++k)
            p_val[k] = (s_idx[k] == agg[i] ? 1.0 : 0.0) - w * s_val[k];
    }
}

}  // namespace fa

using namespace fa;

extern "C" size_t fa_amg_aggregate_work_bytes(int64_t n, int64_t nnz) {
    return align_up(nnz) + align_up(n) + 2 * align_up(n * 8) + 3 * align_up((n + 1) * 4) +
           align_up(exclusive_scan_work_bytes(n)) + 1024;
}

extern "C" int fa_amg_aggregate(int dtype, int64_t n, const int32_t* indptr, const int32_t* indices,
                                const void* values, double theta, int32_t* agg, int64_t* naggregates_host,
                                int32_t* rounds_host, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_amg_aggregate: bad dtype %d", dtype);
    FA_REQUIRE(n > 0 && n < INT32_MAX && indptr && indices && values && agg && naggregates_host,
               "fa_amg_aggregate: bad arguments");
    hipStream_t s = S(stream);
    int32_t nnz = 0;
    FA_HIP(hipMemcpyAsync(&nnz, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    Arena ar(work, work_bytes);
    uint8_t* strong = ar.take<uint8_t>(nnz > 0 ? nnz : 1);
    int8_t* state = ar.take<int8_t>(n);
    uint64_t* m1 = ar.take<uint64_t>(n);
    uint64_t* m2 = ar.take<uint64_t>(n);
    int32_t* flag = ar.take<int32_t>(n + 1);
    int32_t* rid = ar.take<int32_t>(n + 1);
    int32_t* tmp = ar.take<int32_t>(n + 1);
    const size_t sb = exclusive_scan_work_bytes(n);
    void* sw = ar.take<char>(sb);
    int32_t* counter = ar.take<int32_t>(4);
    if (!strong || !state || !m1 || !m2 || !flag || !rid || !tmp || !sw || !counter) {
        set_error("fa_amg_aggregate: workspace too small");
        return FA_ENOMEM;
    }
    const unsigned g = grid_for(n, 256, 1 << 16);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_amg_strength<double>, dim3(g), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const double*>(values), theta, strong, state);
    else
        hipLaunchKernelGGL(k_amg_strength<float>, dim3(g), dim3(256), 0, s, n, indptr, indices,
                           static_cast<const float*>(values), theta, strong, state);
    FA_LAUNCH("k_amg_strength");
    int rounds = 0;
    for (;;) {  // MIS-2 rounds; each one decides at least the global maximum of the undecided
        hipLaunchKernelGGL(k_amg_max, dim3(g), dim3(256), 0, s, n, indptr, indices, strong, state,
                           (const uint64_t*)nullptr, m1);
        hipLaunchKernelGGL(k_amg_max, dim3(g), dim3(256), 0, s, n, indptr, indices, strong, state, m1, m2);
        hipLaunchKernelGGL(k_amg_zero, dim3(1), dim3(64), 0, s, counter);
        hipLaunchKernelGGL(k_amg_decide, dim3(g), dim3(256), 0, s, n, m2, state, counter);
        FA_LAUNCH("k_amg_decide");
        int32_t left = 0;
        FA_HIP(hipMemcpyAsync(&left, counter, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        FA_HIP(hipStreamSynchronize(s));
        ++rounds;
        if (left == 0) break;
        if (rounds > 10000) {
            set_error("fa_amg_aggregate: the independent-set rounds did not finish");
            return FA_ECAPACITY;
        }
    }
    hipLaunchKernelGGL(k_amg_root_flags, dim3(g), dim3(256), 0, s, n, state, flag);
    int st = exclusive_scan_i32(flag, rid, n, sw, sb, s);
    if (st) return st;
    hipLaunchKernelGGL(k_amg_assign, dim3(g), dim3(256), 0, s, n, 0, indptr, indices, strong, state, rid,
                       (const int32_t*)nullptr, tmp);
    hipLaunchKernelGGL(k_amg_assign, dim3(g), dim3(256), 0, s, n, 1, indptr, indices, strong, state, rid, tmp, flag);
    hipLaunchKernelGGL(k_amg_assign, dim3(g), dim3(256), 0, s, n, 2, indptr, indices, strong, state, rid, flag, agg);
    hipLaunchKernelGGL(k_amg_zero, dim3(1), dim3(64), 0, s, counter + 1);
    hipLaunchKernelGGL(k_amg_count_unassigned, dim3(g), dim3(256), 0, s, n, state, agg, counter + 1);
    FA_LAUNCH("k_amg_assign");
    int32_t h[2] = {0, 0};
    FA_HIP(hipMemcpyAsync(&h[0], rid + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&h[1], counter + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (h[1]) {
        set_error("fa_amg_aggregate: %d nodes were left without an aggregate", h[1]);
        return FA_ECAPACITY;
    }
    *naggregates_host = h[0];
    if (rounds_host) *rounds_host = rounds;
    return FA_OK;
}

extern "C" int fa_amg_smooth_prolongation(int dtype, int64_t n, const int32_t* a_indptr, const int32_t* a_indices,
                                          const void* a_values, const int32_t* agg, const int32_t* s_indptr,
                                          const int32_t* s_indices, const double* s_values, double omega,
                                          double* p_values, fa_stream_t stream) {
    FA_REQUIRE(dtype == FA_F64 || dtype == FA_F32, "fa_amg_smooth_prolongation: bad dtype %d", dtype);
    FA_REQUIRE(n > 0 && a_indptr && agg && s_indptr && p_values, "fa_amg_smooth_prolongation: bad arguments");
    const unsigned g = grid_for(n, 256, 1 << 16);
    if (dtype == FA_F64)
        hipLaunchKernelGGL(k_amg_smooth_p<double>, dim3(g), dim3(256), 0, S(stream), n, a_indptr, a_indices,
                           static_cast<const double*>(a_values), agg, s_indptr, s_indices, s_values, omega, p_values);
    else
        hipLaunchKernelGGL(k_amg_smooth_p<float>, dim3(g), dim3(256), 0, S(stream), n, a_indptr, a_indices,
                           static_cast<const float*>(a_values), agg, s_indptr, s_indices, s_values, omega, p_values);
    FA_LAUNCH("k_amg_smooth_p");
    return FA_OK;
}
