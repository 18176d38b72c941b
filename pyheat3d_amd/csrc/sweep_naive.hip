// Rung "naive": one lattice site per thread and VEC sites per thread along x, every operand
// read through the cache hierarchy (counterpart of reference backends/stencil.cu:43-93).
//
// Workgroup = 64 x 4 lanes: one wave covers 64 consecutive x sites of one row, so every load
// instruction of a wave is one contiguous 256-byte (fp32) segment. Site v of a thread sits at
// x = (blockIdx.x * VEC + v) * 64 + lane, which keeps each of the VEC loads coalesced.
#include "sweep_impl.h"

namespace ph3 {

constexpr int kNaiveBX = 64, kNaiveBY = 4;

// Column of site v of this thread and its x neighbours. Every index is clamped into the row
// (columns past the end read the last site, Dirichlet walls read themselves), so all loads of a
// thread are unconditional and the compiler issues the 7*VEC loads before the first wait; only
// the store is guarded. A per-site guard around the loads would split the thread into VEC basic
// blocks that each wait for their 7 loads before the next ones issue.
template <bool P, int VEC>
struct NaiveCols {
    int i[VEC], ic[VEC], im[VEC], ip[VEC];
    __device__ __forceinline__ NaiveCols(int nx) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            i[v] = (blockIdx.x * VEC + v) * kNaiveBX + threadIdx.x;
            ic[v] = min(i[v], nx - 1);
            if (P) {
                im[v] = ic[v] == 0 ? nx - 1 : ic[v] - 1;
                ip[v] = ic[v] == nx - 1 ? 0 : ic[v] + 1;
            } else {
                im[v] = max(ic[v] - 1, 0);
                ip[v] = min(ic[v] + 1, nx - 1);
            }
        }
    }
    __device__ __forceinline__ bool writes(int v, int nx) const {
        return P ? i[v] < nx : (i[v] > 0 && i[v] < nx - 1);
    }
};

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
    const NaiveCols<P, VEC> x(nx);
    float xm[VEC], xp[VEC], ym[VEC], yp[VEC], uc[VEC], bs[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        xm[v] = ld(u, r + x.im[v]);
        xp[v] = ld(u, r + x.ip[v]);
        ym[v] = ld(u, rm + x.ic[v]);
        yp[v] = ld(u, rp + x.ic[v]);
        uc[v] = ld(u, r + x.ic[v]);
        if (S) bs[v] = b[r + x.ic[v]];
    }
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        float s = xm[v] + xp[v];
        s = s + (ym[v] + yp[v]);
        float w = update(c0, c1, uc[v], s);
        if (S) w = w + bs[v];
        if (x.writes(v, nx)) st(o, r + x.i[v], w);
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
    const NaiveCols<P, VEC> x(nx);
    float xm[VEC], xp[VEC], ym[VEC], yp[VEC], zm[VEC], zp[VEC], uc[VEC], bs[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        xm[v] = ld(u, r + x.im[v]);
        xp[v] = ld(u, r + x.ip[v]);
        ym[v] = ld(u, ry_m + x.ic[v]);
        yp[v] = ld(u, ry_p + x.ic[v]);
        zm[v] = ld(u, rz_m + x.ic[v]);
        zp[v] = ld(u, rz_p + x.ic[v]);
        uc[v] = ld(u, r + x.ic[v]);
        if (S) bs[v] = b[r + x.ic[v]];
    }
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
        float s = xm[v] + xp[v];
        s = s + (ym[v] + yp[v]);
        s = s + (zm[v] + zp[v]);
        float w = update(c0, c1, uc[v], s);
        if (S) w = w + bs[v];
        if (x.writes(v, nx)) st(o, r + x.i[v], w);
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
    if (planned(a, PH3_NAIVE, 1, VEC, kNaiveBY, 1, 0, kNaiveBX * VEC, kNaiveBY, kNaiveBX * VEC,
                kNaiveBY, kNaiveBX * kNaiveBY,
                (long long)gx * gy * (a.dims == 3 ? a.nz : 1)))
        return PH3_OK;
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
    // 4-byte and 2 for 2-byte storage). With every load of a thread issued before the first wait,
    // more sites per lane is more loads in flight, and 4 wins for every storage type. Measured on
    // MI355X (GLUPS, vec 1 / 2 / 4): float32 3D 1024^3 329 / 401 / 489, 2D 16384^2
    // 403 / 550 / 711; bfloat16 3D 396 / 521 / 692, 2D 456 / 631 / 819. Rows narrower than 256
    // sites take a narrower vector (fit_vec).
    int vec = a.tile.vec ? a.tile.vec : fit_vec(4, a.nx);
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
