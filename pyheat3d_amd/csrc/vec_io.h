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
