// ph3_sweep: argument checks and dispatch to the rung kernels.
#include "sweep_impl.h"

using ph3host::set_error;

static int sweep_or_plan(int kind, int k, int storage, int periodic, int dims, int nz, int ny,
                         int nx, float c0, float c1, const void* u, void* o, const float* b,
                         const ph3_tile* tile, void* stream, ph3_plan* plan)
{
    if (dims != 2 && dims != 3) { set_error("dims must be 2 or 3, got %d", dims); return PH3_EINVAL; }
    if (dims == 2) nz = 1;
    // every axis needs 3 points (reference backends/base.py:36-39)
    if (nx < 3 || ny < 3 || (dims == 3 && nz < 3)) {
        set_error("every axis needs at least 3 points (nz=%d ny=%d nx=%d)", nz, ny, nx);
        return PH3_EINVAL;
    }
    if (!u || !o) { set_error("u and o must be device pointers"); return PH3_EINVAL; }
    if (u == o) { set_error("u and o must be distinct buffers"); return PH3_EINVAL; }
    if (storage < 0 || storage > 2) { set_error("storage %d not in {0,1,2}", storage); return PH3_EINVAL; }
    ph3::SweepArgs a{storage, periodic ? 1 : 0, dims, nz, ny, nx, c0, c1, u, o, b,
                     tile ? *tile : ph3_tile{0, 0, 0, 0, 0}, (hipStream_t)stream, plan};
    switch (kind) {
        case PH3_NAIVE: return ph3::launch_naive(a);
        case PH3_ZMARCH: return ph3::launch_zmarch(a);
        case PH3_TB:
            if (k == 1) return ph3::launch_zmarch(a);
            if (!ph3::tb_supported(dims, k)) {
                set_error("temporal blocking with k=%d is not built for %dD", k, dims);
                return PH3_EUNSUP;
            }
            return ph3::launch_tb(a, k);
    }
    set_error("unknown sweep kind %d", kind);
    return PH3_EINVAL;
}

extern "C" int ph3_tb_supported(int dims, int k) { return ph3::tb_supported(dims, k) ? 1 : 0; }

extern "C" int ph3_sweep(int kind, int k, int storage, int periodic, int dims, int nz, int ny,
                         int nx, float c0, float c1, const void* u, void* o, const float* b,
                         const ph3_tile* tile, void* stream)
{
    return sweep_or_plan(kind, k, storage, periodic, dims, nz, ny, nx, c0, c1, u, o, b, tile,
                         stream, nullptr);
}

extern "C" int ph3_sweep_plan(int kind, int k, int storage, int periodic, int dims, int nz,
                              int ny, int nx, const void* u, void* o, const float* b,
                              const ph3_tile* tile, ph3_plan* plan)
{
    if (!plan) { set_error("ph3_sweep_plan: plan is NULL"); return PH3_EINVAL; }
    return sweep_or_plan(kind, k, storage, periodic, dims, nz, ny, nx, 0.0f, 0.0f, u, o, b, tile,
                         nullptr, plan);
}
