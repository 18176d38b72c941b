// Internal launch interface shared by the sweep translation units.
#pragma once
#include "ph3_common.h"
#include "../../include/pyheat3d_hip.h"

namespace ph3 {

struct SweepArgs {
    int storage, periodic, dims, nz, ny, nx;
    float c0, c1;
    const void* u;
    void* o;
    const float* b;
    ph3_tile tile;
    hipStream_t stream;
};

int launch_naive(const SweepArgs& a);
int launch_zmarch(const SweepArgs& a);
int launch_tb(const SweepArgs& a, int k);
bool tb_supported(int dims, int k);

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
