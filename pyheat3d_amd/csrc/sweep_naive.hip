// Rung "naive": one lattice site per thread and VEC sites per thread along x, every operand
// read through the cache hierarchy (counterpart of reference backends/stencil.cu:43-93).
//
// Workgroup = 64 x 4 lanes: one wave covers 64 consecutive x sites of one row, so every load
// instruction of a wave is one contiguous 256-byte (fp32) segment. Site v of a thread sits at
// x = (blockIdx.x * VEC + v) * 64 + lane, which keeps each of the VEC loads coalesced.
#include "sweep_impl.h"

namespace ph3 {

constexpr int kNaiveBX = 64, kNaiveBY = 4;

template <typename T, bool P, bool S, int VEC>
__global__ void __launch_bounds__(kNaiveBX * kNaiveBY)
naive2d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
        int ny, int nx, float c0, float c1)
{
    const int j = blockIdx.y * kNaiveBY + threadIdx.y;
    if (j >= ny) return;
    int jm, jp;
    if (P) { jm = j == 0 ? ny - 1 : j - 1; jp = j == ny - 1 ? 0 : j + 1; }
    else {
        if (j == 0 || j == ny - 1) return;
        jm = j - 1; jp = j + 1;
    }
    const size_t r = (size_t)j * nx, rm = (size_t)jm * nx, rp = (size_t)jp * nx;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        const int i = (blockIdx.x * VEC + v) * kNaiveBX + threadIdx.x;
        if (i >= nx) continue;
        int im, ip;
        if (P) { im = i == 0 ? nx - 1 : i - 1; ip = i == nx - 1 ? 0 : i + 1; }
        else {
            if (i == 0 || i == nx - 1) continue;
            im = i - 1; ip = i + 1;
        }
        float s = ld(u, r + im) + ld(u, r + ip);
        s = s + (ld(u, rm + i) + ld(u, rp + i));
        float w = update(c0, c1, ld(u, r + i), s);
        if (S) w = w + b[r + i];
        st(o, r + i, w);
    }
}

template <typename T, bool P, bool S, int VEC>
__global__ void __launch_bounds__(kNaiveBX * kNaiveBY)
naive3d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
        int nz, int ny, int nx, float c0, float c1)
{
    const int j = blockIdx.y * kNaiveBY + threadIdx.y;
    const int k = blockIdx.z;
    if (j >= ny) return;
    int jm, jp, km, kp;
    if (P) {
        jm = j == 0 ? ny - 1 : j - 1; jp = j == ny - 1 ? 0 : j + 1;
        km = k == 0 ? nz - 1 : k - 1; kp = k == nz - 1 ? 0 : k + 1;
    } else {
        if (j == 0 || j == ny - 1 || k == 0 || k == nz - 1) return;
        jm = j - 1; jp = j + 1; km = k - 1; kp = k + 1;
    }
    const size_t plane = (size_t)nx * ny;
    const size_t r = (size_t)k * plane + (size_t)j * nx;
    const size_t ry_m = (size_t)k * plane + (size_t)jm * nx, ry_p = (size_t)k * plane + (size_t)jp * nx;
    const size_t rz_m = (size_t)km * plane + (size_t)j * nx, rz_p = (size_t)kp * plane + (size_t)j * nx;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        const int i = (blockIdx.x * VEC + v) * kNaiveBX + threadIdx.x;
        if (i >= nx) continue;
        int im, ip;
        if (P) { im = i == 0 ? nx - 1 : i - 1; ip = i == nx - 1 ? 0 : i + 1; }
        else {
            if (i == 0 || i == nx - 1) continue;
            im = i - 1; ip = i + 1;
        }
        float s = ld(u, r + im) + ld(u, r + ip);
        s = s + (ld(u, ry_m + i) + ld(u, ry_p + i));
        s = s + (ld(u, rz_m + i) + ld(u, rz_p + i));
        float w = update(c0, c1, ld(u, r + i), s);
        if (S) w = w + b[r + i];
        st(o, r + i, w);
    }
}

template <typename T, bool P, bool S, int VEC>
static int naive_launch(const SweepArgs& a) {
    const dim3 block(kNaiveBX, kNaiveBY);
    const unsigned gx = (unsigned)((a.nx + kNaiveBX * VEC - 1) / (kNaiveBX * VEC));
    const unsigned gy = (unsigned)((a.ny + kNaiveBY - 1) / kNaiveBY);
    if (gy > 65535u || (a.dims == 3 && a.nz > 65535)) {
        ph3host::set_error("naive: grid too large (ny=%d nz=%d)", a.ny, a.nz);
        return PH3_EINVAL;
    }
    if (a.dims == 2) {
        naive2d<T, P, S, VEC><<<dim3(gx, gy), block, 0, a.stream>>>(
            (const T*)a.u, (T*)a.o, a.b, a.ny, a.nx, a.c0, a.c1);
    } else {
        naive3d<T, P, S, VEC><<<dim3(gx, gy, (unsigned)a.nz), block, 0, a.stream>>>(
            (const T*)a.u, (T*)a.o, a.b, a.nz, a.ny, a.nx, a.c0, a.c1);
    }
    PH3_LAUNCH_CHECK("naive sweep launch");
    return PH3_OK;
}

int launch_naive(const SweepArgs& a) {
    // Sites per thread (the reference's NAIVE_VEC, backends/cupy_backend.py:53, which is 1 for
    // 4-byte and 2 for 2-byte storage). Measured on MI355X (GLUPS, vec 1 / 2 / 4): float32 3D
    // 1024^3 326 / 420 / 461, 2D 16384^2 397 / 565 / 536; bfloat16 3D 396 / 503 / 593, 2D
    // 450 / 598 / 635. More independent loads in flight per lane is what pays.
    int vec = a.tile.vec ? a.tile.vec : fit_vec(a.storage == kF32 && a.dims == 2 ? 2 : 4, a.nx);
    PH3_DISPATCH_TPS(a, {
        switch (vec) {
            case 1: return naive_launch<T, P, S, 1>(a);
            case 2: return naive_launch<T, P, S, 2>(a);
            case 4: return naive_launch<T, P, S, 4>(a);
            default:
                ph3host::set_error("naive: vec %d not in {1, 2, 4}", vec);
                return PH3_EINVAL;
        }
    });
    return PH3_OK;
}

}  // namespace ph3
