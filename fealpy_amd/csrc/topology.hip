// Mesh topology on the device: the unique faces / edges of a simplex mesh,
// numbered in the lexicographic order of their sorted vertex tuples, stored
// with the orientation of their first occurrence, with entity -> cell
// ([c0, c1, l0, l1] from the first and last occurrence) and cell -> entity
// maps.  Same numbering as MeshDS.construct + flocc
// (fealpy/mesh/mesh_data_structure.py:428-464, fealpy/mesh/utils.py:81-110),
// which sort every local entity with a stable lexsort.
//
// Here the local entities are bucketed by their smallest vertex (a counting
// sort: count, exclusive scan, scatter), and each bucket is sorted by (second
// vertex, third vertex, occurrence index) in LDS by one wavefront.  Buckets
// are ascending in the smallest vertex, so the concatenated buckets are in
// lexicographic order, and the occurrence index as the last key makes the
// order equal to the stable sort's: the scatter's arrival order does not
// matter.  A flag scan then numbers the unique tuples.
#include "common.hpp"

namespace fa {

constexpr int TOPO_BUCKET_MAX = 1024;  // occurrences sharing one smallest vertex (LDS capacity)

// the m (<= 3) vertices of local entity l of cell c, ascending
__device__ inline void entity_sorted(const int32_t* __restrict__ cell, int nv, const int32_t* __restrict__ local,
                                     int m, int64_t c, int l, int32_t t[3]) {
    t[0] = cell[c * nv + local[l * m + 0]];
    t[1] = cell[c * nv + local[l * m + 1]];
    t[2] = m == 3 ? cell[c * nv + local[l * m + 2]] : 0;
    if (m == 2) {
        if (t[1] < t[0]) {
            const int32_t s = t[0];
            t[0] = t[1];
            t[1] = s;
        }
        return;
    }
    if (t[1] < t[0]) { const int32_t s = t[0]; t[0] = t[1]; t[1] = s; }
    if (t[2] < t[1]) { const int32_t s = t[1]; t[1] = t[2]; t[2] = s; }
    if (t[1] < t[0]) { const int32_t s = t[0]; t[0] = t[1]; t[1] = s; }
}

__device__ inline uint64_t entity_key(const int32_t t[3]) {
    return (static_cast<uint64_t>(static_cast<uint32_t>(t[1])) << 32) | static_cast<uint32_t>(t[2]);
}

__global__ void __launch_bounds__(256) k_topo_zero(int64_t n, int32_t* __restrict__ a) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

// count[v] = occurrences whose smallest vertex is v
__global__ void __launch_bounds__(256) k_topo_count(int64_t nocc, const int32_t* __restrict__ cell, int nv,
                                                   const int32_t* __restrict__ local, int m, int nloc,
                                                   int32_t* __restrict__ count) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nocc; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t t[3];
        entity_sorted(cell, nv, local, m, e / nloc, static_cast<int>(e % nloc), t);
        atomicAdd(&count[t[0]], 1);
    }
}

// place every occurrence in its bucket (arrival order within a bucket is
// arbitrary: the bucket sort below orders it completely)
__global__ void __launch_bounds__(256) k_topo_scatter(int64_t nocc, const int32_t* __restrict__ cell, int nv,
                                                     const int32_t* __restrict__ local, int m, int nloc,
                                                     const int32_t* __restrict__ off, int32_t* __restrict__ cursor,
                                                     uint64_t* __restrict__ key, int32_t* __restrict__ occ) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nocc; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t t[3];
        entity_sorted(cell, nv, local, m, e / nloc, static_cast<int>(e % nloc), t);
        const int32_t p = off[t[0]] + atomicAdd(&cursor[t[0]], 1);
        key[p] = entity_key(t);
        occ[p] = static_cast<int32_t>(e);
    }
}

__device__ inline bool pair_less(uint64_t ka, int32_t ia, uint64_t kb, int32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// one wavefront per bucket: bitonic sort of (key, occurrence) in LDS, then the
// group-start flags (a tuple differing from its predecessor in the bucket;
// the first of a bucket always starts a group)
__global__ void __launch_bounds__(WAVE) k_topo_bucket_sort(int64_t nn, const int32_t* __restrict__ off,
                                                          uint64_t* __restrict__ key, int32_t* __restrict__ occ,
                                                          int32_t* __restrict__ flag, int32_t* __restrict__ bad) {
    __shared__ uint64_t sk[TOPO_BUCKET_MAX];
    __shared__ int32_t si[TOPO_BUCKET_MAX];
    const int lane = threadIdx.x;
    for (int64_t v = blockIdx.x; v < nn; v += gridDim.x) {
        const int32_t b0 = off[v];
        const int len = off[v + 1] - b0;
        if (len == 0) continue;
        if (len > TOPO_BUCKET_MAX) {
            if (lane == 0) atomicOr(bad, 1);
            continue;
        }
        if (len == 1) {
            if (lane == 0) flag[b0] = 1;
            continue;
        }
        int P = 2;
        while (P < len) P <<= 1;
        for (int t = lane; t < P; t += WAVE) {
            sk[t] = t < len ? key[b0 + t] : ~0ULL;
            si[t] = t < len ? occ[b0 + t] : 0x7fffffff;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < P; t += WAVE) {
                    const int u = t ^ j;
                    if (u > t) {
                        const bool up = (t & k) == 0;
                        const uint64_t ka = sk[t], kb = sk[u];
                        const int32_t ia = si[t], ib = si[u];
                        const bool swap = up ? pair_less(kb, ib, ka, ia) : pair_less(ka, ia, kb, ib);
                        if (swap) {
                            sk[t] = kb;
                            sk[u] = ka;
                            si[t] = ib;
                            si[u] = ia;
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
        for (int t = lane; t < len; t += WAVE) {
            key[b0 + t] = sk[t];
            occ[b0 + t] = si[t];
            flag[b0 + t] = (t == 0 || sk[t] != sk[t - 1]) ? 1 : 0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// gid[p] (exclusive scan of the flags) numbers the groups: the group of
// position p is gid[p] + flag[p] - 1.  Writes the entity (first occurrence's
// orientation), entity -> cell and cell -> entity.
__global__ void __launch_bounds__(256) k_topo_fill(int64_t nocc, const int32_t* __restrict__ cell, int nv,
                                                  const int32_t* __restrict__ local, int m, int nloc,
                                                  const int32_t* __restrict__ occ, const int32_t* __restrict__ flag,
                                                  const int32_t* __restrict__ gid, int32_t* __restrict__ entity,
                                                  int32_t* __restrict__ cell2entity, int32_t* __restrict__ entity2cell) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nocc; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t f = flag[p];
        const int32_t g = gid[p] + f - 1;
        const int32_t e = occ[p];
        const int64_t c = e / nloc;
        const int l = e % nloc;
        cell2entity[e] = g;
        if (f) {  // first occurrence of the group: its orientation and (c0, l0)
            for (int a = 0; a < m; ++a) entity[static_cast<int64_t>(g) * m + a] = cell[c * nv + local[l * m + a]];
            entity2cell[static_cast<int64_t>(g) * 4 + 0] = static_cast<int32_t>(c);
            entity2cell[static_cast<int64_t>(g) * 4 + 2] = l;
        }
        if (p + 1 == nocc || flag[p + 1]) {  // last occurrence: (c1, l1)
            entity2cell[static_cast<int64_t>(g) * 4 + 1] = static_cast<int32_t>(c);
            entity2cell[static_cast<int64_t>(g) * 4 + 3] = l;
        }
    }
}

struct TopoWork {
    int32_t* off;     // nn + 1
    int32_t* cursor;  // nn (counts, then scatter cursors)
    uint64_t* key;    // nocc
    int32_t* occ;     // nocc
    int32_t* flag;    // nocc
    int32_t* gid;     // nocc + 1
    int32_t* bad;     // 1
    void* scan;
    size_t scan_bytes;
};

static void topo_layout(Arena& ar, int64_t nocc, int64_t nn, TopoWork* w) {
    w->off = ar.take<int32_t>(nn + 1);
    w->cursor = ar.take<int32_t>(nn);
    w->key = ar.take<uint64_t>(nocc);
    w->occ = ar.take<int32_t>(nocc);
    w->flag = ar.take<int32_t>(nocc);
    w->gid = ar.take<int32_t>(nocc + 1);
    w->bad = ar.take<int32_t>(1);
    w->scan_bytes = exclusive_scan_work_bytes(nocc > nn ? nocc : nn);
    w->scan = ar.take<char>(w->scan_bytes);
}

static bool topo_carve(void* work, size_t work_bytes, int64_t nocc, int64_t nn, TopoWork* w) {
    Arena ar(work, work_bytes);
    topo_layout(ar, nocc, nn, w);
    return w->off && w->cursor && w->key && w->occ && w->flag && w->gid && w->bad && w->scan;
}

}  // namespace fa

using namespace fa;

extern "C" size_t fa_topology_work_bytes(int64_t nocc, int64_t nn) {
    char probe = 0;  // layout only: the sizes the carve takes from a workspace
    Arena ar(&probe, SIZE_MAX / 2);
    TopoWork w;
    topo_layout(ar, nocc, nn, &w);
    return ar.used + 256;
}

extern "C" int fa_topology_sort(int m, int nloc, const int32_t* local, int nv, const int32_t* cell, int64_t nc,
                                int64_t nn, int64_t* nuniq_host, void* work, size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE((m == 2 || m == 3) && nloc > 0 && nv > 0 && local && cell && nc > 0 && nn > 0 && nuniq_host,
               "fa_topology_sort: bad arguments");
    const int64_t nocc = nc * nloc;
    FA_REQUIRE(nocc < INT32_MAX && nn < INT32_MAX, "fa_topology_sort: %lld occurrences exceed int32 indices",
               (long long)nocc);
    TopoWork w;
    if (!topo_carve(work, work_bytes, nocc, nn, &w)) {
        set_error("fa_topology_sort: workspace too small");
        return FA_ENOMEM;
    }
    hipStream_t s = S(stream);
    const unsigned g = grid_for(nocc, 256, 1 << 16);
    const unsigned gn = grid_for(nn, 256, 1 << 16);
    hipLaunchKernelGGL(k_topo_zero, dim3(gn), dim3(256), 0, s, nn, w.cursor);
    hipLaunchKernelGGL(k_topo_zero, dim3(1), dim3(64), 0, s, 1, w.bad);
    hipLaunchKernelGGL(k_topo_count, dim3(g), dim3(256), 0, s, nocc, cell, nv, local, m, nloc, w.cursor);
    FA_LAUNCH("k_topo_count");
    int st = exclusive_scan_i32(w.cursor, w.off, nn, w.scan, w.scan_bytes, s);
    if (st) return st;
    hipLaunchKernelGGL(k_topo_zero, dim3(gn), dim3(256), 0, s, nn, w.cursor);
    hipLaunchKernelGGL(k_topo_scatter, dim3(g), dim3(256), 0, s, nocc, cell, nv, local, m, nloc, w.off, w.cursor, w.key,
                       w.occ);
    FA_LAUNCH("k_topo_scatter");
    hipLaunchKernelGGL(k_topo_bucket_sort, dim3(grid_for(nn, 1, 1 << 20)), dim3(WAVE), 0, s, nn, w.off, w.key, w.occ,
                       w.flag, w.bad);
    FA_LAUNCH("k_topo_bucket_sort");
    st = exclusive_scan_i32(w.flag, w.gid, nocc, w.scan, w.scan_bytes, s);
    if (st) return st;
    int32_t h[2] = {0, 0};
    FA_HIP(hipMemcpyAsync(&h[0], w.gid + nocc, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipMemcpyAsync(&h[1], w.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FA_HIP(hipStreamSynchronize(s));
    if (h[1]) {
        set_error("fa_topology_sort: more than %d occurrences share a smallest vertex", TOPO_BUCKET_MAX);
        return FA_ECAPACITY;
    }
    *nuniq_host = h[0];
    return FA_OK;
}

extern "C" int fa_topology_fill(int m, int nloc, const int32_t* local, int nv, const int32_t* cell, int64_t nc,
                                int64_t nn, int32_t* entity, int32_t* cell2entity, int32_t* entity2cell, void* work,
                                size_t work_bytes, fa_stream_t stream) {
    FA_REQUIRE((m == 2 || m == 3) && nloc > 0 && nv > 0 && local && cell && nc > 0 && nn > 0 && entity &&
                   cell2entity && entity2cell,
               "fa_topology_fill: bad arguments");
    const int64_t nocc = nc * nloc;
    TopoWork w;
    if (!topo_carve(work, work_bytes, nocc, nn, &w)) {
        set_error("fa_topology_fill: workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_topo_fill, dim3(grid_for(nocc, 256, 1 << 16)), dim3(256), 0, S(stream), nocc, cell, nv, local,
                       m, nloc, w.occ, w.flag, w.gid, entity, cell2entity, entity2cell);
    FA_LAUNCH("k_topo_fill");
    return FA_OK;
}
