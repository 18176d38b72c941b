// Vector loads and stores of V contiguous stored values, widened to / rounded from float32.
// V * sizeof(T) must be 2, 4, 8 or 16 bytes, and the address must be aligned to that size.
#pragma once
#include "ph3_common.h"

namespace ph3 {

template <int BYTES> struct RawVec;
template <> struct RawVec<2> { using type = uint16_t; };
template <> struct RawVec<4> { using type = uint32_t; };
template <> struct RawVec<8> { using type = uint2; };
template <> struct RawVec<16> { using type = uint4; };
struct Raw32 { uint4 lo, hi; };  // two 16-byte accesses (float32 source at V = 8)
template <> struct RawVec<32> { using type = Raw32; };

template <typename T, int V>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, size_t i, float (&out)[V]) {
    if constexpr (V == 1) {
        out[0] = Codec<T>::load(p[i]);
    } else {
        using R = typename RawVec<V * sizeof(T)>::type;
        const R raw = *reinterpret_cast<const R*>(p + i);
        T tmp[V];
        __builtin_memcpy(tmp, &raw, sizeof(R));
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] = Codec<T>::load(tmp[e]);
    }
}

template <typename T, int V>
__device__ __forceinline__ void store_vec(T* __restrict__ p, size_t i, const float (&v)[V]) {
    if constexpr (V == 1) {
        p[i] = Codec<T>::store(v[0]);
    } else {
        using R = typename RawVec<V * sizeof(T)>::type;
        T tmp[V];
#pragma unroll
        for (int e = 0; e < V; ++e) tmp[e] = Codec<T>::store(v[e]);
        R raw;
        __builtin_memcpy(&raw, tmp, sizeof(R));
        *reinterpret_cast<R*>(p + i) = raw;
    }
}

// Masked stores without a branch. A raw buffer store whose offset lies at or past the buffer's
// num_records bytes is dropped by the hardware, so lanes (or whole steps) that must not write
// get kDropOffset instead of being masked off. Every step of a loop then issues the same number
// of stores, and the compiler's waits for loads issued before those stores stay exact (behind
// an exec-masked branch it cannot count them and waits for the stores as well).
constexpr uint32_t kDropOffset = 0x80000000u;
constexpr size_t kMaxBufferBytes = 0x7FFFFFFFu;   // num_records must stay below kDropOffset

typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// Buffer resource over `bytes` bytes from p (0 bytes: every store through it is dropped).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const void* p, uint32_t bytes) {
    // word 3: 32-bit data format, the gfx9-family value for raw (untyped) access
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

template <typename T, int V>
__device__ __forceinline__ void store_vec_buf(__amdgpu_buffer_rsrc_t r, uint32_t off_bytes,
                                              const float (&v)[V], uint32_t soff_bytes = 0) {
    const int so = (int)soff_bytes;
    constexpr int B = V * (int)sizeof(T);
    static_assert(B == 2 || B == 4 || B == 8 || B == 16, "buffer stores of 2-16 bytes");
    T tmp[V];
#pragma unroll
    for (int e = 0; e < V; ++e) tmp[e] = Codec<T>::store(v[e]);
    if constexpr (B == 2) {
        unsigned short x;
        __builtin_memcpy(&x, tmp, 2);
        __builtin_amdgcn_raw_buffer_store_b16(x, r, off_bytes, so, 0);
    } else if constexpr (B == 4) {
        unsigned int x;
        __builtin_memcpy(&x, tmp, 4);
        __builtin_amdgcn_raw_buffer_store_b32(x, r, off_bytes, so, 0);
    } else if constexpr (B == 8) {
        u32x2_t x;
        __builtin_memcpy(&x, tmp, 8);
        __builtin_amdgcn_raw_buffer_store_b64(x, r, off_bytes, so, 0);
    } else {
        u32x4_t x;
        __builtin_memcpy(&x, tmp, 16);
        __builtin_amdgcn_raw_buffer_store_b128(x, r, off_bytes, so, 0);
    }
}

// A raw buffer load of V float32 values at byte offset off_bytes (4 * V bytes, aligned to that):
// the base address stays in scalar registers and each lane keeps a 32-bit offset, instead of a
// 64-bit address per load. The loaded vector is copied out whole: extracting its elements one
// by one and bit-casting each made this compiler shrink the load to one dword and copy that
// dword into every element (checked in the gfx950 assembly).
// soff_bytes: a wave-uniform part of the offset (scalar register), added by the hardware.
template <int V>
__device__ __forceinline__ void load_vec_buf(__amdgpu_buffer_rsrc_t r, uint32_t off_bytes,
                                             float (&out)[V], uint32_t soff_bytes = 0) {
    static_assert(V == 1 || V == 2 || V == 4 || V == 8, "buffer loads of 1, 2, 4 or 8 floats");
    const int so = (int)soff_bytes;
    if constexpr (V == 1) {
        out[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off_bytes, so, 0));
    } else if constexpr (V == 2) {
        const u32x2_t x = __builtin_amdgcn_raw_buffer_load_b64(r, off_bytes, so, 0);
        __builtin_memcpy(out, &x, 8);
    } else {
#pragma unroll
        for (int h = 0; h < V / 4; ++h) {
            const u32x4_t x = __builtin_amdgcn_raw_buffer_load_b128(r, off_bytes + 16 * h, so, 0);
            __builtin_memcpy(out + 4 * h, &x, 16);
        }
    }
}

// Make the wave wait here until a loaded value has arrived: an empty asm statement that reads
// the register counts as a use for the compiler's wait insertion. Applied to a loop's first
// prefetch, it makes the loop entry look like the back edge (loads complete, stores possibly in
// flight); otherwise the merge of the two at the loop header takes the entry's pending loads as
// the worst case and the header waits for every store the previous iteration left in flight.
__device__ __forceinline__ void arrived(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void arrived(uint32_t v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void arrived(uint16_t v) { arrived((uint32_t)v); }
__device__ __forceinline__ void arrived(const uint2& v) { arrived(v.x); arrived(v.y); }
__device__ __forceinline__ void arrived(const uint4& v) {
    arrived(v.x); arrived(v.y); arrived(v.z); arrived(v.w);
}
template <int V> __device__ __forceinline__ void arrived(const float (&v)[V]) {
#pragma unroll
    for (int e = 0; e < V; ++e) arrived(v[e]);
}

// A lane vector kept in its stored format between the load and the step that uses it: a
// conversion placed right after a load would make the wave wait for that load there.
template <typename T, int V> using Packed = typename RawVec<V * sizeof(T)>::type;

template <typename T, int V>
__device__ __forceinline__ Packed<T, V> load_packed(const T* __restrict__ p, size_t i) {
    return *reinterpret_cast<const Packed<T, V>*>(p + i);
}

template <typename T, int V>
__device__ __forceinline__ void unpack(const Packed<T, V>& raw, float (&out)[V]) {
    T tmp[V];
    __builtin_memcpy(tmp, &raw, sizeof(raw));
#pragma unroll
    for (int e = 0; e < V; ++e) out[e] = Codec<T>::load(tmp[e]);
}

}  // namespace ph3
