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

// Streaming copy with 16-byte accesses: the ceiling the sweeps are compared against. One
// 16-byte element per lane over a full grid; on MI355X this measured 6.1 TB/s (read + write),
// ahead of grid-stride loops with 4 elements per lane (4.5-5.1 TB/s, with or without
// non-temporal hints) and of hipMemcpy device-to-device (4.9 TB/s).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy16_kernel(const u32x4* __restrict__ src,
                                                     u32x4* __restrict__ dst, size_t n16) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

__global__ void copy_tail_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                 size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Counter-based uniform generator: site i of stream `seed` is
//   U = (splitmix64(seed * 0xD1B54A32D192ED03 + i) >> 40) * 2^-24, exactly representable.
__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// kind 0: U[0,1) inside, 0 on the outer layer; kind 1: U[0,1) on the outer layer, 0 inside;
// kind 2: U[0,1) everywhere.
template <typename T>
__global__ void fill_random_kernel(T* __restrict__ dst, int dims, int nz, int ny, int nx,
                                   uint64_t seed, int kind) {
    const size_t n = (size_t)nz * ny * nx;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint64_t base = seed * 0xD1B54A32D192ED03ull;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int x = (int)(i % nx);
        const size_t r = i / nx;
        const int y = (int)(r % ny), z = (int)(r / ny);
        const bool edge = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 ||
                          (dims == 3 && (z == 0 || z == nz - 1));
        float v = (float)(splitmix64(base + i) >> 40) * 5.9604644775390625e-08f;
        if ((kind == 0 && edge) || (kind == 1 && !edge)) v = 0.f;
        dst[i] = Codec<T>::store(v);
    }
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

extern "C" int ph3_fill_random(int storage, void* dst, int dims, int nz, int ny, int nx,
                               unsigned long long seed, int kind, void* stream) {
    if (dims == 2) nz = 1;
    if (!dst || nx < 1 || ny < 1 || nz < 1 || kind < 0 || kind > 2) {
        ph3host::set_error("fill_random: bad arguments");
        return PH3_EINVAL;
    }
    const size_t n = (size_t)nz * ny * nx;
    const unsigned g = grid_for(n, 256, 65536);
    hipStream_t s = (hipStream_t)stream;
    switch (storage) {
        case kF32: fill_random_kernel<float><<<g, 256, 0, s>>>((float*)dst, dims, nz, ny, nx, seed, kind); break;
        case kF16: fill_random_kernel<_Float16><<<g, 256, 0, s>>>((_Float16*)dst, dims, nz, ny, nx, seed, kind); break;
        case kBF16: fill_random_kernel<bf16_t><<<g, 256, 0, s>>>((bf16_t*)dst, dims, nz, ny, nx, seed, kind); break;
        default: ph3host::set_error("fill_random: storage %d", storage); return PH3_EINVAL;
    }
    PH3_LAUNCH_CHECK("fill_random launch");
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
        if ((n16 + 255) / 256 > 0x7fffffffull) { ph3host::set_error("copy kernel: too large"); return PH3_EINVAL; }
        copy16_kernel<<<(unsigned)((n16 + 255) / 256), 256, 0, s>>>((const u32x4*)src, (u32x4*)dst, n16);
        PH3_LAUNCH_CHECK("copy launch");
    }
    if (tail) {
        copy_tail_kernel<<<1, 64, 0, s>>>((const uint8_t*)src + n16 * 16, (uint8_t*)dst + n16 * 16, tail);
        PH3_LAUNCH_CHECK("copy tail launch");
    }
    return PH3_OK;
}
