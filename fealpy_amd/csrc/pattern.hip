// Kernel "b": integer sparsity build.
//
// 1. DOF -> (cell, local) incidence by counting sort: per-DOF counts (integer
//    atomics), exclusive scan, scatter through per-DOF cursors, then a per-DOF
//    sort so the incidence of every DOF is in ascending (cell, local) order --
//    deterministic regardless of the scatter order.
// 2. One wavefront per global row: gather the columns of every incident cell
//    into LDS, bitonic-sort them, drop duplicates.  Count pass -> scan ->
//    fill pass.  The result equals the reference's lexsort/unique pattern
//    (fealpy/sparse/coo_tensor.py:184-213) with columns sorted inside rows.
#include <climits>

#include "common.hpp"

namespace fa {

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void k_incidence_count(const int32_t* __restrict__ c2d, int64_t ne, int64_t gdof,
                                  int32_t* __restrict__ count, int32_t* __restrict__ bad) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t d = c2d[e];
        if (d < 0 || d >= gdof) {
            atomicOr(bad, 1);
            continue;
        }
        atomicAdd(&count[d], 1);
    }
}

__global__ void k_incidence_scatter(const int32_t* __restrict__ c2d, int64_t ne, const int32_t* __restrict__ ptr,
                                    int32_t* __restrict__ cursor, int32_t* __restrict__ adj, int64_t gdof) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t d = c2d[e];
        if (d < 0 || d >= gdof) continue;
        int32_t slot = atomicAdd(&cursor[d], 1);
        adj[ptr[d] + slot] = static_cast<int32_t>(e);
    }
}

// insertion sort of each DOF's incidence list (short lists) + max degree
__global__ void k_incidence_sort(const int32_t* __restrict__ ptr, int32_t* __restrict__ adj, int64_t gdof,
                                 int32_t* __restrict__ max_deg) {
    int32_t local_max = 0;
    for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < gdof; d += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a0 = ptr[d], a1 = ptr[d + 1];
        local_max = max(local_max, a1 - a0);
        for (int32_t i = a0 + 1; i < a1; ++i) {
            int32_t v = adj[i];
            int32_t j = i - 1;
            while (j >= a0 && adj[j] > v) {
                adj[j + 1] = adj[j];
                --j;
            }
            adj[j + 1] = v;
        }
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) local_max = max(local_max, __shfl_xor(local_max, o));
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicMax(max_deg, local_max);
}

// One wavefront (block of 64) per row.  CAP = LDS capacity in candidates.
template <int CAP>
__global__ void __launch_bounds__(WAVE) k_pattern_rows(const int32_t* __restrict__ c2d, int32_t ldof, int64_t gdof,
                                                       const int32_t* __restrict__ ptr, const int32_t* __restrict__ adj,
                                                       int32_t* __restrict__ rowcount, const int32_t* __restrict__ indptr,
                                                       int32_t* __restrict__ indices, int32_t* __restrict__ max_row) {
    __shared__ int32_t buf[CAP];
    const int lane = threadIdx.x;
    int32_t row_max = 0;
    for (int64_t r = blockIdx.x; r < gdof; r += gridDim.x) {
        const int32_t a0 = ptr[r];
        const int32_t m = (ptr[r + 1] - a0) * ldof;  // m <= CAP guaranteed by the host
        int P = 2;
        while (P < m) P <<= 1;
        for (int t = lane; t < P; t += WAVE) {
            int32_t v = INT_MAX;
            if (t < m) {
                int32_t e = adj[a0 + t / ldof];
                v = c2d[(e / ldof) * ldof + t % ldof];
            }
            buf[t] = v;
        }
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < P; t += WAVE) {
                    int p = t ^ j;
                    if (p > t) {
                        int32_t a = buf[t], b = buf[p];
                        bool up = (t & k) == 0;
                        if ((a > b) == up) {
                            buf[t] = b;
                            buf[p] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // unique: keep first of each run, skip padding
        int32_t base = indices ? indptr[r] : 0;
        int32_t count = 0;
        for (int t0 = 0; t0 < P; t0 += WAVE) {
            int t = t0 + lane;
            int32_t v = (t < P) ? buf[t] : INT_MAX;
            bool keep = (v != INT_MAX) && (t == 0 || buf[t - 1] != v);
            int inc = wave_inclusive_scan(keep ? 1 : 0);
            if (indices && keep) indices[base + count + inc - 1] = v;
            count += __shfl(inc, WAVE - 1);
        }
        if (!indices && lane == 0) rowcount[r] = count;
        row_max = max(row_max, count);
        __syncthreads();
    }
    if (max_row && lane == 0) atomicMax(max_row, row_max);
}

template <int CAP>
static void launch_rows(const int32_t* c2d, int32_t ldof, int64_t gdof, const int32_t* ptr, const int32_t* adj,
                        int32_t* rowcount, const int32_t* indptr, int32_t* indices, int32_t* max_row, hipStream_t s) {
    unsigned grid = grid_for(gdof, 1, 1 << 22);
    hipLaunchKernelGGL(k_pattern_rows<CAP>, dim3(grid), dim3(WAVE), 0, s, c2d, ldof, gdof, ptr, adj, rowcount,
                       indptr, indices, max_row);
}

static int dispatch_rows(int64_t m, const int32_t* c2d, int32_t ldof, int64_t gdof, const int32_t* ptr,
                         const int32_t* adj, int32_t* rowcount, const int32_t* indptr, int32_t* indices,
                         int32_t* max_row, hipStream_t s) {
    if (m <= 128) launch_rows<128>(c2d, ldof, gdof, ptr, adj, rowcount, indptr, indices, max_row, s);
    else if (m <= 512) launch_rows<512>(c2d, ldof, gdof, ptr, adj, rowcount, indptr, indices, max_row, s);
    else if (m <= 2048) launch_rows<2048>(c2d, ldof, gdof, ptr, adj, rowcount, indptr, indices, max_row, s);
    else if (m <= 8192) launch_rows<8192>(c2d, ldof, gdof, ptr, adj, rowcount, indptr, indices, max_row, s);
    else {
        set_error("a row has %lld column candidates; the pattern kernel holds at most 8192", (long long)m);
        return FA_ECAPACITY;
    }
    FA_LAUNCH("k_pattern_rows");
    return FA_OK;
}

}  // namespace fa

using namespace fa;

extern "C" size_t fa_adjacency_work_bytes(int64_t nc, int32_t ldof, int64_t gdof) {
    (void)nc;
    (void)ldof;
    return align_up((gdof + 1) * sizeof(int32_t)) * 2 + exclusive_scan_work_bytes(gdof) + 1024;
}

extern "C" int fa_adjacency_build(const int32_t* cell2dof, int64_t nc, int32_t ldof, int64_t gdof, int32_t* adj_ptr,
                                  int32_t* adj, int32_t* max_degree_host, void* work, size_t work_bytes,
                                  fa_stream_t stream) {
    FA_REQUIRE(nc >= 0 && ldof > 0 && gdof > 0, "fa_adjacency_build: bad sizes nc=%lld ldof=%d gdof=%lld",
               (long long)nc, ldof, (long long)gdof);
    FA_REQUIRE(nc * (int64_t)ldof < INT_MAX, "fa_adjacency_build: nc*ldof must be < 2^31");
    FA_REQUIRE(gdof < INT_MAX, "fa_adjacency_build: gdof must be < 2^31");
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    int32_t* count = ar.take<int32_t>(gdof + 1);
    int32_t* flags = ar.take<int32_t>(4);
    size_t scan_bytes = exclusive_scan_work_bytes(gdof);
    void* scan_ws = ar.take<char>(scan_bytes);
    if (!count || !flags || !scan_ws) {
        set_error("fa_adjacency_build: workspace too small (%zu bytes)", work_bytes);
        return FA_ENOMEM;
    }
    const int64_t ne = nc * ldof;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(gdof + 1, 256, 4096)), dim3(256), 0, s, count, gdof + 1, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flags, (int64_t)4, 0);
    hipLaunchKernelGGL(k_incidence_count, dim3(grid_for(ne, 256, 8192)), dim3(256), 0, s, cell2dof, ne, gdof, count,
                       flags);
    FA_LAUNCH("k_incidence_count");
    int st = exclusive_scan_i32(count, adj_ptr, gdof, scan_ws, scan_bytes, s);
    if (st != FA_OK) return st;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(gdof, 256, 4096)), dim3(256), 0, s, count, gdof, 0);
    hipLaunchKernelGGL(k_incidence_scatter, dim3(grid_for(ne, 256, 8192)), dim3(256), 0, s, cell2dof, ne, adj_ptr,
                       count, adj, gdof);
    FA_LAUNCH("k_incidence_scatter");
    hipLaunchKernelGGL(k_incidence_sort, dim3(grid_for(gdof, 256, 8192)), dim3(256), 0, s, adj_ptr, adj, gdof,
                       flags + 1);
    FA_LAUNCH("k_incidence_sort");
    int32_t h[2];
    FA_HIP(hipMemcpyAsync(h, flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    FA_REQUIRE(h[0] == 0, "fa_adjacency_build: cell2dof has entries outside [0, gdof)");
    if (max_degree_host) *max_degree_host = h[1];
    return FA_OK;
}

extern "C" size_t fa_pattern_work_bytes(int64_t gdof) {
    return align_up((gdof + 1) * sizeof(int32_t)) + exclusive_scan_work_bytes(gdof) + 1024;
}

extern "C" int fa_pattern_count(const int32_t* cell2dof, int64_t nc, int32_t ldof, int64_t gdof,
                                const int32_t* adj_ptr, const int32_t* adj, int32_t max_degree, int32_t* indptr,
                                int64_t* nnz_host, int32_t* max_row_host, void* work, size_t work_bytes,
                                fa_stream_t stream) {
    FA_REQUIRE(gdof > 0 && ldof > 0, "fa_pattern_count: bad sizes");
    FA_REQUIRE((int64_t)nc * ldof * ldof < INT_MAX, "fa_pattern_count: nc*ldof^2 must be < 2^31");
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    int32_t* rowcount = ar.take<int32_t>(gdof + 1);
    int32_t* flags = ar.take<int32_t>(4);
    size_t scan_bytes = exclusive_scan_work_bytes(gdof);
    void* scan_ws = ar.take<char>(scan_bytes);
    if (!rowcount || !flags || !scan_ws) {
        set_error("fa_pattern_count: workspace too small (%zu bytes)", work_bytes);
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flags, (int64_t)4, 0);
    int st = dispatch_rows((int64_t)max_degree * ldof, cell2dof, ldof, gdof, adj_ptr, adj, rowcount, nullptr,
                           nullptr, flags, s);
    if (st != FA_OK) return st;
    st = exclusive_scan_i32(rowcount, indptr, gdof, scan_ws, scan_bytes, s);
    if (st != FA_OK) return st;
    int32_t nnz32 = 0, mrow = 0;
    FA_HIP(hipMemcpyAsync(&nnz32, indptr + gdof, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&mrow, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (nnz_host) *nnz_host = nnz32;
    if (max_row_host) *max_row_host = mrow;
    return FA_OK;
}

extern "C" int fa_pattern_fill(const int32_t* cell2dof, int64_t nc, int32_t ldof, int64_t gdof,
                               const int32_t* adj_ptr, const int32_t* adj, int32_t max_degree, const int32_t* indptr,
                               int32_t* indices, fa_stream_t stream) {
    (void)nc;
    FA_REQUIRE(gdof > 0 && ldof > 0, "fa_pattern_fill: bad sizes");
    return dispatch_rows((int64_t)max_degree * ldof, cell2dof, ldof, gdof, adj_ptr, adj, nullptr, indptr, indices,
                         nullptr, S(stream));
}

// ---- P1 pattern of a from_box tetrahedral grid ------------------------------
// TetrahedronMesh.from_box numbers node (ix, iy, iz) as (ix (ny+1) + iy)(nz+1)
// + iz; two nodes share a Kuhn cell exactly when their offset is 0 or one of
// the 14 directions +-(1,0,0), +-(0,1,0), +-(0,0,1), +-(1,1,0), +-(1,0,1),
// +-(0,1,1), +-(1,1,1) and both lie in the grid, and lexicographic offset
// order is ascending column order.  So the CSR pattern (the one the generic
// kernels build from the incidence lists) follows from the grid alone.
namespace fa {
__device__ inline bool box_edge(int dx, int dy, int dz) {
    // 0 or a Kuhn edge direction: all nonzero components of one sign
    const bool pos = dx >= 0 && dy >= 0 && dz >= 0, neg = dx <= 0 && dy <= 0 && dz <= 0;
    return pos || neg;
}

template <bool FILL>
__global__ void __launch_bounds__(256) k_box_pattern_rows(int nx, int ny, int nz, int32_t* __restrict__ rowcount,
                                                         const int32_t* __restrict__ indptr,
                                                         int32_t* __restrict__ indices, int64_t nnz) {
    const int64_t nz1 = nz + 1, nyz = static_cast<int64_t>(ny + 1) * nz1, nn = static_cast<int64_t>(nx + 1) * nyz;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nn; r += (int64_t)gridDim.x * blockDim.x) {
        const int ix = static_cast<int>(r / nyz);
        const int64_t rem = r - ix * nyz;
        const int iy = static_cast<int>(rem / nz1);
        const int iz = static_cast<int>(rem - iy * nz1);
        int32_t k = FILL ? indptr[r] : 0;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    if (!box_edge(dx, dy, dz)) continue;
                    if (ix + dx < 0 || ix + dx > nx || iy + dy < 0 || iy + dy > ny || iz + dz < 0 || iz + dz > nz)
                        continue;
                    if (FILL && k < nnz) indices[k] = static_cast<int32_t>(r + dx * nyz + dy * nz1 + dz);
                    ++k;
                }
        if (!FILL) rowcount[r] = k;
    }
}
}  // namespace fa

extern "C" size_t fa_p1_box_pattern_work_bytes(int nx, int ny, int nz) {
    const int64_t nn = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
    return align_up((nn + 1) * sizeof(int32_t)) + exclusive_scan_work_bytes(nn) + 1024;
}

extern "C" int fa_p1_box_pattern(int nx, int ny, int nz, int32_t* indptr, int32_t* indices, int64_t nnz, void* work,
                                 size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && indptr && indices, "fa_p1_box_pattern: bad arguments");
    const int64_t nn = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
    FA_REQUIRE(nn < INT32_MAX, "fa_p1_box_pattern: grid too large for int32 indices");
    int64_t want = nn;
    {
        const int o[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        for (const auto& e : o) want += 2 * static_cast<int64_t>(nx + 1 - e[0]) * (ny + 1 - e[1]) * (nz + 1 - e[2]);
    }
    FA_REQUIRE(nnz == want && want < INT32_MAX, "fa_p1_box_pattern: %lld nonzeros given, the grid has %lld",
               (long long)nnz, (long long)want);
    hipStream_t s = S(stream);
    Arena ar(work, work_bytes);
    int32_t* rowcount = ar.take<int32_t>(nn + 1);
    const size_t scan_bytes = exclusive_scan_work_bytes(nn);
    void* scan_ws = ar.take<char>(scan_bytes);
    if (!rowcount || !scan_ws) {
        set_error("fa_p1_box_pattern: workspace too small (%zu bytes)", work_bytes);
        return FA_ENOMEM;
    }
    const unsigned grid = grid_for(nn, 256, 1 << 20);
    hipLaunchKernelGGL(k_box_pattern_rows<false>, dim3(grid), dim3(256), 0, s, nx, ny, nz, rowcount, nullptr, nullptr,
                       nnz);
    FA_LAUNCH("k_box_pattern_rows");
    int st = exclusive_scan_i32(rowcount, indptr, nn, scan_ws, scan_bytes, s);
    if (st != FA_OK) return st;
    hipLaunchKernelGGL(k_box_pattern_rows<true>, dim3(grid), dim3(256), 0, s, nx, ny, nz, nullptr, indptr, indices,
                       nnz);
    FA_LAUNCH("k_box_pattern_rows");
    return FA_OK;
}
