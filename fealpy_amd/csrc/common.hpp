// Shared helpers of libfealpy_amd (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "fealpy_amd.h"

namespace fa {

constexpr int WAVE = 64;

void set_error(const char* fmt, ...);

// autotuned tile of a kernel family for a shape class (fa_tune_set), or fallback
int tuned(int family, int key, int fallback);

// one V-cycle z = M r on stream s; kernels return at once while *done != 0
int mg_vcycle(const fa_mg_t* mg, const void* r, void* z, const int* done, hipStream_t s);

// In-step kernel timer (fa_timer_arm / fa_timer_read): when the kernel family
// `kind` (FA_TIMER_*) is armed, timer_begin / timer_end record HIP events on
// `s` around its launch; otherwise they do nothing.
void timer_begin(int kind, hipStream_t s);
void timer_end(int kind, hipStream_t s);

inline hipStream_t S(fa_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int launch_status(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return FA_EHIP;
    }
    return FA_OK;
}

#define FA_REQUIRE(cond, ...)              \
    do {                                   \
        if (!(cond)) {                     \
            ::fa::set_error(__VA_ARGS__);  \
            return FA_EINVAL;              \
        }                                  \
    } while (0)

#define FA_HIP(expr)                                                            \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::fa::set_error("%s failed: %s", #expr, hipGetErrorString(e_));     \
            return FA_EHIP;                                                     \
        }                                                                       \
    } while (0)

#define FA_LAUNCH(name) \
    do { int st_ = ::fa::launch_status(name); if (st_ != FA_OK) return st_; } while (0)

__host__ __device__ inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Carve typed sub-buffers out of a caller workspace.
struct Arena {
    char* base;
    size_t size;
    size_t used = 0;
    Arena(void* b, size_t s) : base(static_cast<char*>(b)), size(s) {}
    template <class T>
    T* take(size_t count) {
        size_t off = align_up(used);
        used = off + count * sizeof(T);
        return used <= size ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

// ---- wave-level primitives ------------------------------------------------
template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// inclusive scan over a whole wavefront
// XCD-aware workgroup order.  The dispatcher deals consecutive workgroups to
// the 8 XCDs in turn, each with its own L2; renumbering workgroup b of a grid
// of G (G a multiple of 8) as (b % 8) * (G / 8) + b / 8 gives every XCD one
// contiguous range of rows, so the neighbouring rows whose x entries a
// sparse row gathers sit in the same L2.  Any other grid keeps its order.
__device__ inline unsigned xcd_block() {
    const unsigned G = gridDim.x, b = blockIdx.x;
    return (G & 7u) ? b : (b & 7u) * (G >> 3) + (b >> 3);
}

inline unsigned xcd_grid(unsigned g) { return (g + 7u) & ~7u; }

__device__ inline int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// reduction inside an aligned group of G lanes (G power of two <= 64)
template <int G, class T>
__device__ inline T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// inclusive scan inside an aligned group of G lanes
template <int G>
__device__ inline int group_inclusive_scan(int v) {
    const int lane = threadIdx.x % G;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        int t = __shfl_up(v, o, G);
        if (lane >= o) v += t;
    }
    return v;
}

// scan of int32 counts: out[i] = sum_{j<i} in[j], out[n] = total (out has n+1 slots)
int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, void* work, size_t work_bytes,
                       hipStream_t s);
size_t exclusive_scan_work_bytes(int64_t n);

}  // namespace fa
