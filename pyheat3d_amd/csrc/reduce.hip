// FP64 Jacobi residual: sum over updated sites of (b + c0*u + c1*S(u) - u)^2.
//
// Two deterministic passes. Pass 1 runs a fixed grid of kResBlocks workgroups; workgroup g owns
// rows g, g + G, ... (row = one x-line), lanes stride along x, and each workgroup folds its
// lanes in a fixed tree into partial[g]. Pass 2 is one workgroup that folds the partials in a
// fixed tree. The result depends only on the shape, never on timing, so a solve converges
// after the same number of sweeps alone or inside a batch.
#include "sweep_impl.h"

#include <mutex>

namespace ph3 {

constexpr int kResBlocks = 1024;
constexpr int kResThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kResThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    return sm[0];
}

template <typename T, bool P, bool S>
__global__ void __launch_bounds__(kResThreads)
residual_pass1(const T* __restrict__ u, const float* __restrict__ b, int dims, int nz, int ny,
               int nx, float c0, float c1, double* __restrict__ partial)
{
    __shared__ double sm[kResThreads];
    const double c0d = c0, c1d = c1;
    const size_t plane = (size_t)nx * ny;
    const long rows = (long)nz * ny;
    double acc = 0.0;
    for (long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int k = (int)(r / ny), j = (int)(r - (long)k * ny);
        if (!P && (j == 0 || j == ny - 1 || (dims == 3 && (k == 0 || k == nz - 1)))) continue;
        const int jm = P ? (j == 0 ? ny - 1 : j - 1) : j - 1;
        const int jp = P ? (j == ny - 1 ? 0 : j + 1) : j + 1;
        const size_t base = (size_t)k * plane;
        const size_t row = base + (size_t)j * nx;
        for (int i = threadIdx.x; i < nx; i += kResThreads) {
            if (!P && (i == 0 || i == nx - 1)) continue;
            const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
            double s = (double)ld(u, row + im) + (double)ld(u, row + ip);
            s += (double)ld(u, base + (size_t)jm * nx + i) + (double)ld(u, base + (size_t)jp * nx + i);
            if (dims == 3) {
                const int km = P ? (k == 0 ? nz - 1 : k - 1) : k - 1;
                const int kp = P ? (k == nz - 1 ? 0 : k + 1) : k + 1;
                s += (double)ld(u, (size_t)km * plane + (size_t)j * nx + i) +
                     (double)ld(u, (size_t)kp * plane + (size_t)j * nx + i);
            }
            const double uc = (double)ld(u, row + i);
            double res = c0d * uc + c1d * s - uc;
            if (S) res += (double)b[row + i];
            acc += res * res;
        }
    }
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kResThreads)
residual_pass2(const double* __restrict__ partial, double* __restrict__ out)
{
    __shared__ double sm[kResThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < kResBlocks; i += kResThreads) acc += partial[i];
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) *out = t;
}

}  // namespace ph3

extern "C" size_t ph3_residual_work_bytes(void) { return sizeof(double) * ph3::kResBlocks; }

extern "C" int ph3_residual(int storage, int periodic, int dims, int nz, int ny, int nx,
                            float c0, float c1, const void* u, const float* b, double* out,
                            void* work, void* stream)
{
    if (dims == 2) nz = 1;
    if ((dims != 2 && dims != 3) || nx < 3 || ny < 3 || (dims == 3 && nz < 3)) {
        ph3host::set_error("residual: bad shape dims=%d nz=%d ny=%d nx=%d", dims, nz, ny, nx);
        return PH3_EINVAL;
    }
    if (!u || !out || !work) { ph3host::set_error("residual: NULL pointer"); return PH3_EINVAL; }
    ph3::SweepArgs a{storage, periodic ? 1 : 0, dims, nz, ny, nx, c0, c1, u, nullptr, b,
                     ph3_tile{0, 0, 0, 0, 0}, (hipStream_t)stream};
    double* partial = (double*)work;
    PH3_DISPATCH_TPS(a, {
        ph3::residual_pass1<T, P, S><<<ph3::kResBlocks, ph3::kResThreads, 0, a.stream>>>(
            (const T*)u, b, dims, nz, ny, nx, c0, c1, partial);
    });
    PH3_LAUNCH_CHECK("residual pass 1");
    ph3::residual_pass2<<<1, ph3::kResThreads, 0, a.stream>>>(partial, out);
    PH3_LAUNCH_CHECK("residual pass 2");
    return PH3_OK;
}

// ---- the residual without any allocation ---------------------------------------------------
// Pass 1's workgroup partials live in this file-scope device array instead of a caller
// workspace. Every device holds its own copy at its own address, kept per device after the
// first call on it; host code passes that address to the kernels, never the array's name
// (which is a host-side stand-in). Pass 1 overwrites every partial, so nothing needs clearing.
static __device__ __align__(256) double scratch[ph3::kResBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_of_current_device(double** out) {
    int dev = 0;
    PH3_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        ph3host::set_error("residual_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return PH3_EUNSUP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return ph3host::check(e, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return PH3_OK;
}
}  // namespace

extern "C" size_t ph3_residual_noalloc_scratch_bytes(void) { return sizeof(scratch); }

extern "C" int ph3_residual_noalloc(int storage, int periodic, int dims, int nz, int ny, int nx,
                                    float c0, float c1, const void* u, const float* b,
                                    double* out, void* stream)
{
    // The kernels need one partial per pass-1 workgroup, a fixed count whatever the shape.
    static_assert(sizeof(scratch) == sizeof(double) * ph3::kResBlocks, "one partial per workgroup");
    if (ph3_residual_work_bytes() > sizeof(scratch)) {
        ph3host::set_error("residual_noalloc: %zu bytes of partials exceed the %zu-byte scratch",
                           ph3_residual_work_bytes(), sizeof(scratch));
        return PH3_EUNSUP;
    }
    double* partial = nullptr;
    if (int r = scratch_of_current_device(&partial)) return r;
    return ph3_residual(storage, periodic, dims, nz, ny, nx, c0, c1, u, b, out, partial, stream);
}
