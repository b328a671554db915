// Exclusive scan of int32 counts (tile scan + recursive scan of tile sums),
// and the thread-local error string of the C-ABI.
#include <cstdarg>
#include <cstdio>

#include "common.hpp"

namespace fa {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// block-wide exclusive scan of one value per thread; returns the block total in *total
__device__ inline int block_exclusive_scan(int v, int* total) {
    __shared__ int wsum[SCAN_BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    int inc = wave_inclusive_scan(v);
    if (lane == WAVE - 1) wsum[wid] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / WAVE; ++w) {
        int s = wsum[w];
        off += (w < wid) ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                           int64_t n, int32_t* __restrict__ tile_sums) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        s += v[k];
    }
    int total;
    int ex = block_exclusive_scan(s, &total);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) {
        if (tile_sums) tile_sums[blockIdx.x] = total;
        else out[n] = total;  // single-tile case: write the grand total
    }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_add_offsets(int32_t* __restrict__ out, int64_t n,
                                                            const int32_t* __restrict__ offs, int64_t ntiles) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * SCAN_TILE;
    const int add = offs[blockIdx.x];
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_BLOCK) {
        int64_t i = base + k;
        if (i < n) out[i] += add;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = offs[ntiles];
}

size_t exclusive_scan_work_bytes(int64_t n) {
    size_t bytes = 0;
    while (n > SCAN_TILE) {
        int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
        bytes += align_up(nt * sizeof(int32_t)) + align_up((nt + 1) * sizeof(int32_t));
        n = nt;
    }
    return bytes + 256;
}

static int scan_rec(const int32_t* in, int32_t* out, int64_t n, Arena& ar, hipStream_t s) {
    if (n <= SCAN_TILE) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_BLOCK), 0, s, in, out, n, (int32_t*)nullptr);
        FA_LAUNCH("k_scan_tiles");
        return FA_OK;
    }
    int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    int32_t* sums = ar.take<int32_t>(nt);
    int32_t* offs = ar.take<int32_t>(nt + 1);
    if (!sums || !offs) {
        set_error("scan workspace too small");
        return FA_ENOMEM;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, s, in, out, n, sums);
    FA_LAUNCH("k_scan_tiles");
    int st = scan_rec(sums, offs, nt, ar, s);
    if (st != FA_OK) return st;
    hipLaunchKernelGGL(k_add_offsets, dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, s, out, n, offs, nt);
    FA_LAUNCH("k_add_offsets");
    return FA_OK;
}

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, void* work, size_t work_bytes, hipStream_t s) {
    Arena ar(work, work_bytes);
    return scan_rec(in, out, n, ar, s);
}

}  // namespace fa

extern "C" const char* fa_last_error(void) { return fa::g_err; }
extern "C" int fa_version(void) { return 1; }
