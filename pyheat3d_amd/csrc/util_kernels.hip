// Device-side storage codec and the copy-bandwidth kernel.
#include "sweep_impl.h"

namespace ph3 {

template <typename T>
__global__ void encode_kernel(const float* __restrict__ src, T* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = Codec<T>::store(src[i]);
}

template <typename T>
__global__ void decode_kernel(const T* __restrict__ src, float* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = Codec<T>::load(src[i]);
}

// Streaming copy with 16-byte accesses: the ceiling the sweeps are compared against. Each
// lane moves 4 x 16 B per iteration of a grid-stride loop (4 loads issued before any store).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy16_kernel(const u32x4* __restrict__ src,
                                                     u32x4* __restrict__ dst, size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i);
        const u32x4 b = __builtin_nontemporal_load(src + i + stride);
        const u32x4 c = __builtin_nontemporal_load(src + i + 2 * stride);
        const u32x4 d = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void copy_tail_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                 size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

static unsigned grid_for(size_t n, unsigned block, unsigned cap) {
    size_t g = (n + block - 1) / block;
    if (g > cap) g = cap;
    return (unsigned)(g ? g : 1);
}

}  // namespace ph3

using namespace ph3;

extern "C" int ph3_encode(int storage, const float* src, void* dst, size_t n, void* stream) {
    if (!n) return PH3_OK;
    const unsigned g = grid_for(n, 256, 65536);
    hipStream_t s = (hipStream_t)stream;
    switch (storage) {
        case kF32: PH3_CHECK(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToDevice, s)); return PH3_OK;
        case kF16: encode_kernel<_Float16><<<g, 256, 0, s>>>(src, (_Float16*)dst, n); break;
        case kBF16: encode_kernel<bf16_t><<<g, 256, 0, s>>>(src, (bf16_t*)dst, n); break;
        default: ph3host::set_error("encode: storage %d", storage); return PH3_EINVAL;
    }
    PH3_LAUNCH_CHECK("encode launch");
    return PH3_OK;
}

extern "C" int ph3_decode(int storage, const void* src, float* dst, size_t n, void* stream) {
    if (!n) return PH3_OK;
    const unsigned g = grid_for(n, 256, 65536);
    hipStream_t s = (hipStream_t)stream;
    switch (storage) {
        case kF32: PH3_CHECK(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToDevice, s)); return PH3_OK;
        case kF16: decode_kernel<_Float16><<<g, 256, 0, s>>>((const _Float16*)src, dst, n); break;
        case kBF16: decode_kernel<bf16_t><<<g, 256, 0, s>>>((const bf16_t*)src, dst, n); break;
        default: ph3host::set_error("decode: storage %d", storage); return PH3_EINVAL;
    }
    PH3_LAUNCH_CHECK("decode launch");
    return PH3_OK;
}

extern "C" int ph3_copy_kernel(const void* src, void* dst, size_t bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (((uintptr_t)src | (uintptr_t)dst) & 15u) {
        ph3host::set_error("copy kernel needs 16-byte aligned buffers");
        return PH3_EINVAL;
    }
    const size_t n16 = bytes / 16, tail = bytes - n16 * 16;
    if (n16) {
        // 8 workgroups of 256 lanes per CU-slot on a 256-CU part
        copy16_kernel<<<grid_for(n16, 256, 256 * 8), 256, 0, s>>>((const u32x4*)src, (u32x4*)dst, n16);
        PH3_LAUNCH_CHECK("copy launch");
    }
    if (tail) {
        copy_tail_kernel<<<1, 64, 0, s>>>((const uint8_t*)src + n16 * 16, (uint8_t*)dst + n16 * 16, tail);
        PH3_LAUNCH_CHECK("copy tail launch");
    }
    return PH3_OK;
}
