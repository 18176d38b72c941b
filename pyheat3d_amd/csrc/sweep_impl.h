// Internal launch interface shared by the sweep translation units.
#pragma once
#include "ph3_common.h"
#include "../../include/pyheat3d_hip.h"

#include <atomic>

namespace ph3 {

struct SweepArgs {
    int storage, periodic, dims, nz, ny, nx;
    float c0, c1;
    const void* u;
    void* o;
    const float* b;
    ph3_tile tile;
    hipStream_t stream;
    ph3_plan* plan = nullptr;   // set: describe the launch in *plan instead of making it
};

// Fills *a.plan (when set) and returns true: the caller then returns without launching.
inline bool planned(const SweepArgs& a, int kind, int k, int vec, int rows, int waves, int zch,
                    int tile_x, int tile_y, int out_x, int out_y, int threads, long long groups) {
    if (!a.plan) return false;
    *a.plan = ph3_plan{kind, k, vec, rows, waves, zch, tile_x, tile_y, out_x, out_y, threads,
                       groups};
    return true;
}

int launch_naive(const SweepArgs& a);
int launch_zmarch(const SweepArgs& a);
int launch_tb(const SweepArgs& a, int k);
bool tb_supported(int dims, int k);

// Workgroups of `kernel` (`block` threads) resident on the current device at once.
inline long resident_groups(const void* kernel, int block) {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1)
        nb = 1;
    return (long)cus * nb;
}

// resident_groups of one kernel, queried once per device (slot = device id mod 16; a stale
// slot after a device change only mis-sizes chunks, never results). Keyed by the kernel
// itself, so kernels sharing a signature do not share a cache.
template <auto Kernel>
inline long cached_resident(int block) {
    static std::atomic<long> cache[16];
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::atomic<long>& slot = cache[dev & 15];
    long r = slot.load(std::memory_order_relaxed);
    if (!r) {
        r = resident_groups((const void*)Kernel, block);
        slot.store(r, std::memory_order_relaxed);
    }
    return r;
}

// March chunk for `tiles` tiles over n_march planes. Workgroups run in waves of `resident`;
// a wave of chunk ch costs about ch + fill planes (fill: the halo planes a chunk loads or the
// pipeline it fills), so among the chunks >= min_chunk giving at least 4 waves (fewer measured
// slower: the last workgroups of a single wave finish unevenly) the one minimising
// waves * (ch + fill) is taken. A fixed chunk leaves the last wave part-empty for most sizes
// that are not powers of two (at 768^3 the 64-plane chunk filled 84% of the last wave).
// A grid too small for 4 waves even at min_chunk is latency-bound instead: its time is about
// one workgroup's march, so it gets the shortest chunks (no shorter than small_chunk) whose
// workgroups still fit in one wave, at most four per CU of the 256 (more only add chunk
// overhead). 128^3 z-march: 16 workgroups of 32 planes before, 50 GLUPS; now 331.
inline int fill_chunk(int n_march, long tiles, long resident, int min_chunk, int fill,
                      int small_chunk) {
    constexpr long kMinWaves = 4;
    int best = min_chunk < n_march ? min_chunk : n_march;
    long best_cost = -1;
    for (int ch = best; ch <= n_march; ++ch) {
        const long chunks = (n_march + ch - 1) / ch;
        const long waves = (tiles * chunks + resident - 1) / resident;
        if (waves < kMinWaves) break;   // longer chunks only give fewer waves
        const long cost = waves * (ch + fill);
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = ch; }
    }
    if (best_cost >= 0) return best;
    const long target = resident < 1024 ? resident : 1024;
    long chunks = target / tiles;                             // whole chunks that fit one wave
    if (chunks < 1) chunks = 1;
    long ch = (n_march + chunks - 1) / chunks;
    if (ch < small_chunk) ch = small_chunk;
    if (ch > best) ch = best;
    return ch < 1 ? 1 : (int)ch;
}

// Widest lane vector <= v whose 64-lane wave still fits the row: a wave of 64*v sites over a
// narrower row leaves lanes idle (nx = 64 with 4 sites per lane: 48 of 64).
inline int fit_vec(int v, int nx) {
    while (v > 1 && 64 * v > nx) v >>= 1;
    return v;
}

}  // namespace ph3

// Expands BODY once per (storage type T, periodic flag P, source flag S) combination.
#define PH3_DISPATCH_TPS(a, ...)                                                              \
    do {                                                                                       \
        const bool src_ = (a).b != nullptr;                                                    \
        switch ((a).storage) {                                                                 \
            case ph3::kF32: {                                                                  \
                using T = float;                                                               \
                PH3_DISPATCH_PS_(a, src_, __VA_ARGS__);                                               \
            } break;                                                                           \
            case ph3::kF16: {                                                                  \
                using T = _Float16;                                                            \
                PH3_DISPATCH_PS_(a, src_, __VA_ARGS__);                                               \
            } break;                                                                           \
            case ph3::kBF16: {                                                                 \
                using T = ph3::bf16_t;                                                         \
                PH3_DISPATCH_PS_(a, src_, __VA_ARGS__);                                               \
            } break;                                                                           \
            default:                                                                           \
                ph3host::set_error("storage %d is not 0 (f32), 1 (f16) or 2 (bf16)",          \
                                   (a).storage);                                               \
                return PH3_EINVAL;                                                             \
        }                                                                                      \
    } while (0)

#define PH3_DISPATCH_PS_(a, src_, ...)                                                        \
    if ((a).periodic) {                                                                        \
        constexpr bool P = true;                                                               \
        if (src_) { constexpr bool S = true; __VA_ARGS__; } else { constexpr bool S = false; __VA_ARGS__; }  \
    } else {                                                                                   \
        constexpr bool P = false;                                                              \
        if (src_) { constexpr bool S = true; __VA_ARGS__; } else { constexpr bool S = false; __VA_ARGS__; }  \
    }
