// Shared device helpers for the pyheat3d HIP kernels (gfx950 / CDNA4, wave64).
//
// Every kernel computes the update in the canonical order of the reference
// (src/pyheat3d/problem.py:3-10):
//     s = (xm + xp); s = s + (ym + yp); [3D] s = s + (zm + zp);  v = (c0*u) + (c1*s); [v = v + b]
// in float32, each operation separately rounded. Multiply-add contraction is disabled for the
// whole library (-ffp-contract=off and the pragma below) so results equal the NumPy oracle bit
// for bit. Storage may be float32, float16 or bfloat16; values are widened on load and rounded
// to nearest even on store.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#pragma clang fp contract(off)

namespace ph3 {

enum Storage : int { kF32 = 0, kF16 = 1, kBF16 = 2 };

// bfloat16 stored as its 16-bit pattern.
struct bf16_t { uint16_t bits; };

template <typename T> struct Codec;

template <> struct Codec<float> {
    static constexpr int code = kF32;
    __device__ __forceinline__ static float load(float v) { return v; }
    __device__ __forceinline__ static float store(float v) { return v; }
};

template <> struct Codec<_Float16> {
    static constexpr int code = kF16;
    __device__ __forceinline__ static float load(_Float16 v) { return (float)v; }
    // v_cvt_f16_f32 rounds to nearest even in the default mode.
    __device__ __forceinline__ static _Float16 store(float v) { return (_Float16)v; }
};

template <> struct Codec<bf16_t> {
    static constexpr int code = kBF16;
    __device__ __forceinline__ static float load(bf16_t v) {
        return __uint_as_float((uint32_t)v.bits << 16);
    }
    // Round to nearest even on the float32 bit pattern (the bfloat16 definition used by the
    // reference, backends/cupy_backend.py:103-107).
    __device__ __forceinline__ static bf16_t store(float v) {
        const uint32_t b = __float_as_uint(v);
        bf16_t r;
        r.bits = (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
        return r;
    }
};

// A pointer that a kernel reads from device memory (a solve descriptor) is a generic pointer to
// the compiler, and accesses through it become flat instructions. Flat loads count in both the
// vector-memory and the LDS/scalar wait counters, so every wait for an LDS access, and the wait
// for each prefetched row, then waits for all loads in flight (the batched tile passes ran at
// half the speed of the same tile bodies launched with kernel-argument pointers). Re-deriving
// the pointer from a global-address-space one lets the compiler emit global instructions; the
// value is also made wave-uniform (scalar registers), which buffer resources need. The pointer
// must be uniform over the wave.
template <typename T>
__device__ __forceinline__ T* as_global(T* p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    using GT = __attribute__((address_space(1))) T;
    return (T*)(GT*)(((unsigned long long)hi << 32) | lo);
}

template <typename T>
__device__ __forceinline__ float ld(const T* __restrict__ p, size_t i) { return Codec<T>::load(p[i]); }

template <typename T>
__device__ __forceinline__ void st(T* __restrict__ p, size_t i, float v) { p[i] = Codec<T>::store(v); }

// The update with every operation rounded separately (contraction is off).
__device__ __forceinline__ float update(float c0, float c1, float u, float s) {
    const float a = c0 * u;
    const float m = c1 * s;
    return a + m;
}

// Index of the calling wave in its workgroup, in a scalar register: conditions on it are then
// uniform branches or scalar selects instead of exec-masked blocks (the compiler cannot tell
// that threadIdx.x >> 6 is the same in every lane of a wave).
__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Whole-wave lane shifts by one lane with DPP (wave_shr:1 / wave_shl:1, GFX9-family controls
// available on gfx950): a VALU move instead of an LDS-crossbar permute. Lane 0 (resp. 63) has
// no source lane; with bound_ctrl it receives 0 and no register has to be zeroed first.
// Callers use their own halo value there or never store what depends on it.
__device__ __forceinline__ float from_prev_lane(float v) {  // lane i gets lane i-1
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(
        __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_next_lane(float v) {  // lane i gets lane i+1
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(
        __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, true));
}

}  // namespace ph3

// Host-side error plumbing shared by every translation unit of the library.
namespace ph3host {
void set_error(const char* fmt, ...);
int check(hipError_t e, const char* what);
}  // namespace ph3host

#define PH3_CHECK(call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return ph3host::check(e_, #call);                                \
    } while (0)

#define PH3_LAUNCH_CHECK(what)                                                                 \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess) return ph3host::check(e_, what);                                 \
    } while (0)
