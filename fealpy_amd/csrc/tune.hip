// Tile-shape table of the autotuner (fealpy_amd/tune.py): per kernel family
// and shape class, the tile the last tuning run measured fastest.  Kernels
// read it at launch and fall back to their built-in choice where it is unset.
#include <atomic>

#include "common.hpp"

namespace fa {

namespace {
std::atomic<int> g_tune[FA_TUNE_FAMILIES][FA_TUNE_KEYS];
}

int tuned(int family, int key, int fallback) {
    if (family < 0 || family >= FA_TUNE_FAMILIES || key < 0 || key >= FA_TUNE_KEYS) return fallback;
    const int v = g_tune[family][key].load(std::memory_order_relaxed);
    return v > 0 ? v : fallback;
}

}  // namespace fa

using namespace fa;

extern "C" int fa_tune_set(int family, int key, int value) {
    FA_REQUIRE(family >= 0 && family < FA_TUNE_FAMILIES && key >= 0 && key < FA_TUNE_KEYS && value >= 0,
               "fa_tune_set: bad family %d / key %d / value %d", family, key, value);
    g_tune[family][key].store(value, std::memory_order_relaxed);
    return FA_OK;
}

extern "C" int fa_tune_get(int family, int key) {
    if (family < 0 || family >= FA_TUNE_FAMILIES || key < 0 || key >= FA_TUNE_KEYS) return 0;
    return g_tune[family][key].load(std::memory_order_relaxed);
}

extern "C" int fa_tune_key(int family, int64_t a, int64_t b) {
    switch (family) {
        case FA_TUNE_SPMV_LANES: {  // a = rows, b = entries: average row length class
            const double avg = a > 0 ? double(b) / double(a) : 1.0;
            return avg <= 6 ? 0 : avg <= 12 ? 1 : avg <= 24 ? 2 : avg <= 48 ? 3 : 4;
        }
        case FA_TUNE_SUM_ROWS_LANES:  // a = topological dimension, b = longest CSR row
            return static_cast<int>((a == 3 ? 4 : 0) + (b <= 16 ? 0 : b <= 32 ? 1 : b <= 64 ? 2 : 3));
        default: return -1;
    }
}
