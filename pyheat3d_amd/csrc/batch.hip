// Batched solves: many independent uneven 2D/3D grids advanced together (see the ph3_batch_*
// section of include/pyheat3d_hip.h for the contract).
//
// Tier 1, grids of at most PH3_BATCH_SMALL_SITES sites: one workgroup owns a grid and runs its
// sweeps back to back, separated only by workgroup barriers (the grid stays in L2); the FP64
// residual is a workgroup reduction in a fixed order. A persistent kernel hands grids to
// workgroups through an atomic counter, so a single launch advances every small grid.
// Tier 2, larger grids: one varlen launch per pass covering the tiles of every running grid of
// one (storage, dims, bc, source, steps per pass, tile shape) class: the fused temporal-blocking
// tiles (float32) or the z-march / y-march tiles (single steps), one workgroup per tile (one wave
// per strip for 2D fused passes), each finding its grid by binary search in a per-class tile
// prefix table. After every PH3_BATCH_CHECK sweeps of a class, two small kernels on the
// class's stream compute the residual of each of its Jacobi grids in fixed slices, then advance
// the step counters and apply the convergence / stagnation / cap test.
// All per-site arithmetic is the canonical update of ph3_common.h, so a grid gets the same bits
// whichever tier, tile or batch it runs in; tier and check points depend only on the grid.
#include "march_tiles.h"
#include "tb_tiles.h"

#include <math.h>
#include <type_traits>
#include <vector>

using ph3host::set_error;

namespace ph3 {

constexpr int kWG = 256;          // lanes per workgroup in every batch kernel
constexpr int kPassGroups = 2048; // workgroups of the persistent tier-1 / residual launches
constexpr int kMinSlices = 16;    // residual slices of a tier-2 Jacobi grid: at least ...
constexpr int kMaxSlices = 256;   // ... and at most these, one per kSliceSites sites
constexpr int kSliceSites = 32768;
constexpr int kTier2Tiles = 512; // tier-2 tiles a large grid is spread over
constexpr int kMaxChunk = 64;     // longest tier-2 march chunk of passes with k < 8
constexpr int kClassStreams = 7;  // streams for tier-2 classes beside the caller's
constexpr int kR3 = 4, kW3 = 4;   // tier-2 3D tile: 4 waves x 4 rows, 64 sites wide
constexpr int kR2 = 4, kW2 = 4;   // tier-2 2D tile: 4 waves x 64 columns, 4 rows per step
constexpr int kTbR3 = 8, kTbW3 = 4; // tier-2 3D temporal-blocking tile: 4 waves x 8 rows

__device__ __forceinline__ long long sites_of(const ph3_solve_desc& d) {
    return (long long)d.nz * d.ny * d.nx;
}

// Steps fused per tier-2 pass of a solve (0 or 1 = single steps).
__host__ __device__ inline int kpass_of(const ph3_solve_desc& d) { return d.kpass > 1 ? d.kpass : 1; }

// Index of the buffer holding the current state.
__host__ __device__ inline int cur_buf(const ph3_solve_desc& d) {
    return (int)((d.done / kpass_of(d)) & 1);
}

// The problem fields of a descriptor that the workgroup-level kernels read, as scalars. (A copy
// of the whole descriptor indexed by a run-time buffer number lives in scratch memory: 400 bytes
// per lane.)
struct Prob {
    int dims, nz, ny, nx, mode;
    bool P;
    float c0, c1;
    const float* src;
    void* b0;
    void* b1;
    long long steps;
    __device__ explicit Prob(const ph3_solve_desc& d)
        : dims(d.dims), nz(d.nz), ny(d.ny), nx(d.nx), mode(d.mode), P(d.periodic != 0),
          c0(d.c0), c1(d.c1), src(d.src ? as_global(d.src) : nullptr), b0(as_global(d.buf[0])),
          b1(as_global(d.buf[1])), steps(d.steps) {}
    __device__ void* buf(long long i) const { return (i & 1) ? b1 : b0; }
};

// ---- one workgroup: sweep and residual of a whole grid ---------------------------------------

// One sweep of a whole grid (at most PH3_BATCH_SMALL_SITES sites, so offsets fit in 32 bits)
// by one workgroup. Each wave takes kRowsAtOnce rows at a time and issues the loads of all of
// them before using any, instead of waiting on each row's loads in turn (a chain of dependent
// L2 round trips per sweep: 181^2 grids, when tier 1 still took them, ran 140 -> 188 solves/s
// on the bench mix). Rows that are out of range or fixed (Dirichlet) load from row 0 and
// store nothing.
constexpr int kRowsAtOnce = 8;

template <typename T>
__device__ void wg_sweep(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
                         int dims, int nz, int ny, int nx, bool P, float c0, float c1) {
    constexpr int U = kRowsAtOnce, NW = kWG / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int plane = nx * ny;
    const int rows = nz * ny;
    for (int row0 = wave; row0 < rows; row0 += NW * U) {
        int r0[U], rm[U], rp[U], zm[U], zp[U];
        bool ok[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int row = row0 + q * NW;
            const int k = row / ny, j = row - k * ny;
            ok[q] = row < rows &&
                    (P || !(j == 0 || j == ny - 1 || (dims == 3 && (k == 0 || k == nz - 1))));
            if (!ok[q]) { r0[q] = rm[q] = rp[q] = zm[q] = zp[q] = 0; continue; }
            const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
            const int base = k * plane;
            r0[q] = base + j * nx;
            rm[q] = base + jm * nx;
            rp[q] = base + jp * nx;
            zm[q] = zp[q] = 0;
            if (dims == 3) {
                const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
                zm[q] = km * plane + j * nx;
                zp[q] = kp * plane + j * nx;
            }
        }
        for (int i = lane; i < nx; i += 64) {
            const bool xin = P || (i > 0 && i < nx - 1);
            const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
            float cv[U], xl[U], xr[U], yl[U], yr[U], zl[U], zr[U], bv[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                cv[q] = ld(u, r0[q] + i);
                xl[q] = ld(u, r0[q] + im);
                xr[q] = ld(u, r0[q] + ip);
                yl[q] = ld(u, rm[q] + i);
                yr[q] = ld(u, rp[q] + i);
                if (dims == 3) { zl[q] = ld(u, zm[q] + i); zr[q] = ld(u, zp[q] + i); }
                if (b) bv[q] = b[r0[q] + i];
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                float s = xl[q] + xr[q];
                s = s + (yl[q] + yr[q]);
                if (dims == 3) s = s + (zl[q] + zr[q]);
                float w = update(c0, c1, cv[q], s);
                if (b) w = w + bv[q];
                if (ok[q] && xin) st(o, r0[q] + i, w);
            }
        }
    }
}

// Sum of the squared defect over the sites lane, lane + 64, ... of rows row0, row0 + rstride,
// ..., per lane, in FP64, in that order. The (row, 64-site block) items are taken U at a time
// and all their loads issued before the first is used (one item at a time, each wave waited for
// its 5-7 loads per site block). Sites off the row or on a fixed Dirichlet boundary add an exact
// 0, which leaves the sum as if they were skipped (it is never -0). Site indices of type I:
// 32-bit for grids below 2^31 sites (the base pointer in scalar registers), 64-bit otherwise.
// (U = 2 with 32-bit indices measured 2% faster on the Jacobi tail mixes than U = 4 with 64-bit
// or 32-bit ones.) row0 and rstride must be uniform over the wave.
template <typename T, typename I, int U>
__device__ double residual_items(const T* __restrict__ u, const float* __restrict__ b, int dims,
                                 int nz, int ny, int nx, bool P, float c0, float c1, int row0,
                                 int rstride, int lane) {
    const double c0d = c0, c1d = c1;
    const I plane = (I)nx * (I)ny;
    const int rows = nz * ny;
    const int nxb = (nx + 63) / 64;   // 64-site blocks per row
    int row = __builtin_amdgcn_readfirstlane(row0), xb = 0;   // the next item
    double acc = 0.0;
    while (row < rows) {
        I ic[U], xm[U], xp[U], ym[U], yp[U], zm[U], zp[U];
        bool ok[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const bool live = row < rows;
            const int r = live ? row : 0;
            const int k = r / ny, j = r - k * ny, i = xb * 64 + lane;
            ok[q] = live && i < nx &&
                    (P || (j > 0 && j < ny - 1 && i > 0 && i < nx - 1 &&
                           (dims == 2 || (k > 0 && k < nz - 1))));
            const int ii = ok[q] ? i : 0;
            const int im = ii == 0 ? nx - 1 : ii - 1, ip = ii == nx - 1 ? 0 : ii + 1;
            const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
            const I base = (I)k * plane, r0 = base + (I)j * (I)nx;
            ic[q] = r0 + (I)ii;
            xm[q] = r0 + (I)im;
            xp[q] = r0 + (I)ip;
            ym[q] = base + (I)jm * (I)nx + (I)ii;
            yp[q] = base + (I)jp * (I)nx + (I)ii;
            if (dims == 3) {
                const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
                zm[q] = (I)km * plane + (I)j * (I)nx + (I)ii;
                zp[q] = (I)kp * plane + (I)j * (I)nx + (I)ii;
            } else {
                zm[q] = zp[q] = ic[q];
            }
            if (live && ++xb == nxb) { xb = 0; row += rstride; }
        }
        float vc[U], vxm[U], vxp[U], vym[U], vyp[U], vzm[U], vzp[U], vb[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            vc[q] = ld(u, ic[q]);
            vxm[q] = ld(u, xm[q]);
            vxp[q] = ld(u, xp[q]);
            vym[q] = ld(u, ym[q]);
            vyp[q] = ld(u, yp[q]);
            if (dims == 3) { vzm[q] = ld(u, zm[q]); vzp[q] = ld(u, zp[q]); }
            vb[q] = b ? b[ic[q]] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            double s = (double)vxm[q] + (double)vxp[q];
            s += (double)vym[q] + (double)vyp[q];
            if (dims == 3) s += (double)vzm[q] + (double)vzp[q];
            const double uc = (double)vc[q];
            double r = c0d * uc + c1d * s - uc;
            if (b) r += (double)vb[q];
            acc += ok[q] ? r * r : 0.0;
        }
    }
    return acc;
}

template <typename T>
__device__ double residual_rows(const T* __restrict__ u, const float* __restrict__ b, int dims,
                                int nz, int ny, int nx, bool P, float c0, float c1, int row0,
                                int rstride, int lane) {
    if ((long long)nz * ny * nx < (1LL << 31))
        return residual_items<T, uint32_t, 2>(u, b, dims, nz, ny, nx, P, c0, c1, row0, rstride, lane);
    return residual_items<T, size_t, 2>(u, b, dims, nz, ny, nx, P, c0, c1, row0, rstride, lane);
}

// Fixed-order workgroup sum (every lane gets the result).
__device__ double wg_sum(double v, double* sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    const double t = sm[0];
    __syncthreads();
    return t;
}

template <typename T>
__device__ double wg_residual(const Prob& d, const void* u, double* sm) {
    const double part = residual_rows<T>((const T*)u, d.src, d.dims, d.nz, d.ny, d.nx, d.P,
                                         d.c0, d.c1, threadIdx.x >> 6, kWG / 64,
                                         threadIdx.x & 63);
    return wg_sum(part, sm);
}

// Convergence / stagnation / cap test after check number `checks` (1-based).
__device__ void apply_check(ph3_solve_desc& d, double r) {
    const int c = d.checks + 1;
    d.checks = c;
    d.res_sq = r;
    const double rel = d.res0_sq > 0.0 ? sqrt(r / d.res0_sq) : 0.0;
    int status = 0;
    if (rel <= d.tol) {
        status = 1;
    } else {
        if (d.stall_window > 0 && c >= d.stall_window && d.log_rate < 0.0) {
            const double old = d.ring[(c - d.stall_window) % PH3_STALL_RING];
            const double drop = 0.5 * log(fmax(r, 1e-300) / fmax(old, 1e-300));
            if (drop > d.stall_fraction * d.stall_window * d.log_rate) status = 2;
        }
        if (status == 0 && d.done >= d.steps) status = 3;
    }
    d.ring[c % PH3_STALL_RING] = r;
    d.status = status;
}

// ---- tier 1 ----------------------------------------------------------------------------------

template <typename T>
__device__ void small_solve(ph3_solve_desc* dp, int sweeps, double* sm) {
    __shared__ long long s_done;
    __shared__ int s_status;
    if (threadIdx.x == 0) { s_done = dp->done; s_status = dp->status; }
    __syncthreads();
    long long done = s_done;
    int status = s_status;
    const Prob d(*dp);  // the problem fields
    if (status != 0) return;
    const long long target = done + sweeps;
    while (status == 0 && done < target) {
        long long seg = d.mode == 1 ? PH3_BATCH_CHECK : target - done;
        if (seg > d.steps - done) seg = d.steps - done;
        if (seg <= 0) {  // nothing left to run: finished (heat) or at the sweep cap (Jacobi)
            status = d.mode == 0 ? 1 : 3;
            if (threadIdx.x == 0) dp->status = status;
            break;
        }
        for (long long i = 0; i < seg; ++i) {
            const long long cur = (done + i) & 1;
            wg_sweep<T>((const T*)d.buf(cur), (T*)d.buf(cur ^ 1), d.src, d.dims, d.nz, d.ny, d.nx,
                        d.P, d.c0, d.c1);
            __syncthreads();
        }
        done += seg;
        if (d.mode == 0) {
            if (done >= d.steps) status = 1;
        } else {
            const double r = wg_residual<T>(d, d.buf(done), sm);
            if (threadIdx.x == 0) {
                dp->done = done;
                apply_check(*dp, r);
                s_status = dp->status;
            }
            __syncthreads();
            status = s_status;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        dp->done = done;
        if (d.mode == 0) dp->status = status;
    }
}

template <typename T>
__global__ void __launch_bounds__(kWG)
batch_small(ph3_solve_desc* descs, const int* __restrict__ list, int n, int* counter, int sweeps) {
    __shared__ double sm[kWG];
    __shared__ int s_idx;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(counter, 1);
        __syncthreads();
        const int t = s_idx;
        __syncthreads();
        if (t >= n) return;
        small_solve<T>(descs + list[t], sweeps, sm);
        __syncthreads();
    }
}

__global__ void batch_reset(int* counter) { *counter = 0; }

// Initial residual of every listed Jacobi solve that has not been checked yet (one workgroup per
// solve, the same reduction as tier 1).
template <typename T>
__global__ void __launch_bounds__(kWG)
batch_init_residual(ph3_solve_desc* descs, const int* __restrict__ list, int n) {
    __shared__ double sm[kWG];
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        ph3_solve_desc* dp = descs + list[t];
        if (dp->mode != 1 || dp->checks != 0) continue;   // uniform over the workgroup
        const Prob d(*dp);
        const double r = wg_residual<T>(d, d.buf(dp->done / kpass_of(*dp)), sm);
        if (threadIdx.x == 0) {
            dp->res0_sq = r;
            dp->res_sq = r;
            dp->ring[0] = r;
            if (r == 0.0) dp->status = 1;
        }
        __syncthreads();
    }
}

// ---- tier 2 ----------------------------------------------------------------------------------

// Tier-2 tile shapes. Single steps: the z-march tile (3D, 4 waves x 4 rows, 64 sites wide) or
// the y-march tile (2D, 4 strips of 64 columns). Fused steps (float32, K > 1), by shape id
// (ph3_solve_desc.tile; 0 = the library's rule, see tier2_shape_for):
//   3D  1: 1-site lanes, 4 waves x 8 rows (64 x 32 extended tile, 256 lanes)
//       2: 2-site lanes, 4 waves x 8 rows (128 x 32)
//       3: 1-site lanes, 8 waves x 8 rows (64 x 64, 512 lanes: the single-grid tb4 shape)
//       4: 2-site lanes, 8 waves x 4 rows (128 x 32, 512 lanes)
//       5, 6: shapes 1 and 3 storing through buffers on the stored planes only (StoreMode)
//   2D  1: 1-site lanes, masked stores        2: 2-site lanes, masked stores
//       3: 1-site lanes, a store every step   4: 2-site lanes, a store every step
//       5: 2-site lanes, a store every step, each site loaded and stored on its own (any
//          width; tb_tiles.h tb2d_tile EW)
//       (one strip per wave, 4 waves per workgroup; store modes as in tb_tiles.h StoreMode)
// A wider shape recomputes a smaller share of halo sites; the narrow ones split small grids
// into more tiles. Results never depend on the shape.
struct T2Shape {
    int v, r, w, sm;   // sites per lane, rows per wave (3D), waves per workgroup, store mode
    bool ew = false;   // 2D: sites loaded and stored one by one
};

__host__ __device__ inline T2Shape tier2_shape(int dims, int K, int id) {
    if (K <= 1) return T2Shape{1, dims == 3 ? kR3 : kR2, dims == 3 ? kW3 : kW2, kStoreMasked};
    if (dims == 3) {
        switch (id) {
            case 2: return T2Shape{2, 8, 4, kStoreMasked};
            case 3: return T2Shape{1, 8, 8, kStoreMasked};
            case 4: return T2Shape{2, 4, 8, kStoreMasked};
            case 5: return T2Shape{1, 8, 4, kStorePlanes};
            case 6: return T2Shape{1, 8, 8, kStorePlanes};
            default: return T2Shape{1, 8, 4, kStoreMasked};
        }
    }
    switch (id) {
        case 2: return T2Shape{2, 0, kW2, kStoreMasked};
        case 3: return T2Shape{1, 0, kW2, kStoreAll};
        case 4: return T2Shape{2, 0, kW2, kStoreAll};
        case 5: return T2Shape{2, 0, kW2, kStoreAll, true};
        default: return T2Shape{1, 0, kW2, kStoreMasked};
    }
}

// Tier-2 tiling of one grid. The march chunk is chosen so that the grid spreads over about
// `tpg` tiles, a per-class figure the host derives from how many grids share the class (a lone
// large grid needs many short marches to fill the device; many grids together fill it with
// long ones, which recompute fewer of the 2K pipeline-fill planes of a fused pass), within
// [lo, hi] planes or rows. Results do not depend on the tiling.
struct T2Geom {
    int tx, ty, chunk;  // tiles along x, along y (1 in 2D), march chunk
    long long tiles;    // workgroup tiles of one pass over the grid (they set the chunk)
    int strips;         // 2D fused passes: strips of one row chunk (one wave each)
    long long items;    // work items of the pass kernel: 2D fused, strips x chunks; else tiles
};

__host__ __device__ inline T2Geom tier2_geom(int dims, int nz, int ny, int nx, int K, int few,
                                             const T2Shape& sh, int tpg, int max_chunk) {
    T2Geom g;
    g.strips = 0;
    if (dims == 3) {
        if (K <= 1) {
            g.tx = (nx + 63) / 64;
            g.ty = (ny + kR3 * kW3 - 1) / (kR3 * kW3);
        } else {
            // a tile of 64*V columns writes 64*V - 2*KV of them (TbGeom)
            const int kv = ((K + sh.v - 1) / sh.v) * sh.v, ox = 64 * sh.v - 2 * kv;
            const int oy = sh.w * sh.r - 2 * K;
            g.tx = (nx + ox - 1) / ox;
            g.ty = (ny + oy - 1) / oy;
        }
    } else {
        const int kv = ((K + sh.v - 1) / sh.v) * sh.v;
        const int ox = K <= 1 ? 64 : 64 * sh.v - 2 * kv;
        g.strips = (nx + ox - 1) / ox;
        g.tx = (g.strips + sh.w - 1) / sh.w;
        g.ty = 1;
    }
    const int march = dims == 3 ? nz : ny;
    const long long txy = (long long)g.tx * g.ty;
    const long long want = (tpg + txy - 1) / txy;
    int ch = (int)((march + want - 1) / want);
    // A k-step chunk recomputes 2k extra rows or planes. Deep 2D passes (k = 8) take chunks of
    // 64-128 rows; `few` (the class has fewer tiles than a pass has workgroups: a pass lasts as
    // long as one tile's march) lets them go down to 2k rows instead.
    const int lo = K >= 8 ? (few ? 2 * K : 8 * K) : (K <= 1 ? 8 : (2 * K > 8 ? 2 * K : 8));
    const int hi = max_chunk > 0 ? (max_chunk > lo ? max_chunk : lo) : (K >= 8 ? 16 * K : kMaxChunk);
    if (ch < lo) ch = lo;
    if (ch > hi) ch = hi;
    g.chunk = ch;
    const long long chunks = (march + ch - 1) / ch;
    g.tiles = txy * chunks;
    g.items = dims == 2 && K > 1 ? (long long)g.strips * chunks : g.tiles;
    return g;
}

// Work items of a tier-2 class per workgroup: a 2D fused pass gives each wave its own strip
// (its waves are independent, so no wave of a workgroup idles beside a partial tile); the other
// passes give each workgroup one tile.
__host__ __device__ constexpr int items_per_group(int dims, int K, int W) {
    return dims == 2 && K > 1 ? W : 1;
}

// One pass of a tier-2 class: V sites per lane, R rows per wave (3D fused), W waves, store mode
// SM, as tier2_shape gives them for the class's shape id. The launch covers every work item
// (prefix[n] of them) once: item i goes to workgroup i / items_per_group (a grid-stride loop
// past 2^20 workgroups), so the hardware dispatcher balances the tiles.
template <typename T, bool P, bool S, int DIMS, int K, int V, int R, int W, int SM, bool EW = false>
__global__ void __launch_bounds__(64 * W)
batch_pass(const ph3_solve_desc* __restrict__ descs, const int* __restrict__ list,
           const long long* __restrict__ prefix, int n, int p, int few, int tpg,
           int max_chunk) {
    const T2Shape sh{V, R, W, SM, EW};
    constexpr int IPG = items_per_group(DIMS, K, W);
    const long long total = prefix[n];
    // wave-uniform item offset inside the workgroup (scalar registers)
    const int sub = IPG > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    for (long long t = (long long)blockIdx.x * IPG + sub; t < total; t += (long long)gridDim.x * IPG) {
        int lo = 0, hi = n - 1;  // last i with prefix[i] <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const ph3_solve_desc& d = descs[list[lo]];
        if (d.status == 0 && d.done + (long long)p * K < d.steps) {
            const long long local = t - prefix[lo];
            const int cur = (int)((d.done / K + p) & 1);
            const T* u = as_global((const T*)d.buf[cur]);
            T* o = as_global((T*)d.buf[cur ^ 1]);
            const float* src = S ? as_global(d.src) : nullptr;
            const T2Geom g = tier2_geom(DIMS, d.nz, d.ny, d.nx, K, few, sh, tpg, max_chunk);
            if constexpr (DIMS == 3) {
                const int ix = (int)(local % g.tx);
                const long long rest = local / g.tx;
                const int iy = (int)(rest % g.ty), iz = (int)(rest / g.ty);
                if constexpr (K == 1)
                    zmarch3d_tile<T, P, S, 1, kR3, kW3>(u, o, src, d.nz, d.ny, d.nx, d.c0,
                                                        d.c1, g.chunk, ix, iy, iz, 0);
                else
                    tb3d_tile<T, P, S, V, R, W, K, SM>(u, o, src, d.nz, d.ny, d.nx, d.c0,
                                                      d.c1, g.chunk, ix, iy, iz);
            } else {
                if constexpr (K == 1)
                    ymarch2d_tile<T, P, S, 1, kR2, kW2>(u, o, src, d.ny, d.nx, d.c0, d.c1,
                                                        g.chunk, (int)(local % g.tx),
                                                        (int)(local / g.tx));
                else
                    tb2d_tile<T, P, S, V, K, SM, EW>(u, o, src, d.ny, d.nx, d.c0, d.c1, g.chunk,
                                                 (int)(local % g.strips),
                                                 (int)(local / g.strips));
            }
        }
        // the 3D tiles' and the y-march's LDS slots are reused by the next tile (2D fused
        // strips use no LDS, and their waves go on independently)
        if constexpr (IPG == 1) __syncthreads();
    }
}

// Step count of a running solve once the current segment of PH3_BATCH_CHECK sweeps is done.
__device__ inline long long done_after_segment(const ph3_solve_desc& d) {
    long long adv = d.steps - d.done;
    if (adv > PH3_BATCH_CHECK) adv = PH3_BATCH_CHECK;
    return d.done + adv;
}

// Residual slices of a tier-2 solve: 0 for heat; for Jacobi one per kSliceSites sites, within
// [kMinSlices, kMaxSlices]. A function of the solve alone, so its FP64 summation order (slices
// in order, each a fixed workgroup tree over interleaved rows) is the same in any batch.
__host__ __device__ inline int res_slices(const ph3_solve_desc& d) {
    if (d.mode != 1) return 0;
    const long long n = ((long long)d.nz * d.ny * d.nx + kSliceSites - 1) / kSliceSites;
    return (int)(n < kMinSlices ? kMinSlices : (n > kMaxSlices ? kMaxSlices : n));
}

// Slice `item` of the listed solves: `prefix` holds each solve's first slice (res_slices).
template <typename T>
__global__ void __launch_bounds__(kWG)
batch_residual(const ph3_solve_desc* __restrict__ descs, const int* __restrict__ list,
               const int* __restrict__ prefix, int n, double* __restrict__ partial) {
    __shared__ double sm[kWG];
    const int total = prefix[n];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = n - 1;  // last t with prefix[t] <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= item) lo = mid; else hi = mid - 1;
        }
        const int t = lo, sl = item - prefix[t], nsl = prefix[t + 1] - prefix[t];
        const ph3_solve_desc& d = descs[list[t]];
        if (d.mode != 1 || d.status != 0 || d.storage != Codec<T>::code) continue;
        // the state the segment's passes have written (the counters advance in batch_check)
        const int cur = (int)((done_after_segment(d) / kpass_of(d)) & 1);
        const double part = residual_rows<T>(as_global((const T*)d.buf[cur]),
                                             d.src ? as_global(d.src) : nullptr, d.dims, d.nz,
                                             d.ny, d.nx, d.periodic != 0, d.c0, d.c1,
                                             sl * (kWG / 64) + (threadIdx.x >> 6),
                                             nsl * (kWG / 64), threadIdx.x & 63);
        const double s = wg_sum(part, sm);
        if (threadIdx.x == 0) partial[item] = s;
    }
}

// After a segment: advance the step counter of every running solve, finish heat solves that
// reached their step count, and apply the convergence / stagnation / cap test to Jacobi ones.
__global__ void batch_check(ph3_solve_desc* descs, const int* __restrict__ list,
                            const int* __restrict__ prefix, int n,
                            const double* __restrict__ partial) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    ph3_solve_desc& d = descs[list[t]];
    if (d.status != 0) return;
    d.done = done_after_segment(d);
    if (d.mode != 1) {
        if (d.done >= d.steps) d.status = 1;
        return;
    }
    double r = 0.0;
    for (int s = prefix[t]; s < prefix[t + 1]; ++s) r += partial[s];
    apply_check(d, r);
}

}  // namespace ph3

// ---- host side -----------------------------------------------------------------------------

struct Tier2Class {
    int storage, dims, periodic, src, kpass;
    int shape = 1; // tier2_shape id of the class's fused passes (1 for single steps)
    int few = 0;   // tier2_geom's short-chunk mode (set from the class's tile count)
    int tpg = ph3::kTier2Tiles;  // tiles per grid the chunk aims at
    int max_chunk = 0;           // longest march chunk (0: the geometry's default)
    long long items = 0;         // work items of one pass over every grid of the class
    std::vector<int> idx;
    const int* d_list = nullptr;          // into the batch's int arena
    const long long* d_prefix = nullptr;  // into the batch's prefix arena
    const int* d_res_prefix = nullptr;    // first residual slice of each solve (+ total), int arena
    int res_total = 0;                    // residual slices of the class's Jacobi solves
    int partial_base = 0;                 // the class's first residual partial
};

// Upper bound on tier-2 classes: storage (3) x dims (2) x bc (2) x source (2) x kpass (4) x
// lane width (2).
constexpr int kMaxClasses = 192;

struct ph3_batch {
    int cap = 0, n = 0;
    ph3_solve_desc* d_desc = nullptr;
    int* d_counter = nullptr;
    double* d_partial = nullptr;
    // Per-set index tables live in two arenas sized for `cap` solves at creation: every list
    // of a set (tier-1 lists, initial-residual lists, the per-class lists and residual slice
    // prefixes) together holds at most 5 * cap + 1 ints, and the per-class tile prefixes at
    // most cap + kMaxClasses long longs; the residual partials kMaxSlices * cap doubles. A set rewrites them in place and never frees device
    // memory, so a graph recorded from an earlier round never points at freed memory.
    int* d_iarena = nullptr;
    long long* d_llarena = nullptr;
    const int* d_small[3] = {nullptr, nullptr, nullptr};  // tier-1 lists per storage
    int n_small[3] = {0, 0, 0};
    std::vector<Tier2Class> classes;
    hipStream_t side = nullptr;                     // tier 1 runs here, beside tier 2
    hipEvent_t fork = nullptr, join = nullptr;
    // tier-2 classes after the first run their passes on these, beside each other
    hipStream_t cls[ph3::kClassStreams] = {};
    hipEvent_t seg = nullptr, cls_join[ph3::kClassStreams] = {};
    // chunk policy per (dims - 2, log2 kpass): tiles a pass of the class aims at, and the
    // longest chunk (0 = defaults); set by ph3_batch_params
    int target_tiles[2][4] = {};
    int max_chunk[2][4] = {};
};

// Forget the current set (the batch then has no solves; rounds enqueue nothing).
static void clear_set(ph3_batch* b) {
    b->n = 0;
    b->classes.clear();
    for (int s = 0; s < 3; ++s) { b->d_small[s] = nullptr; b->n_small[s] = 0; }
}

extern "C" size_t ph3_solve_desc_bytes(void) { return sizeof(ph3_solve_desc); }

static void destroy_parts(ph3_batch* b) {
    (void)hipFree(b->d_desc); (void)hipFree(b->d_counter); (void)hipFree(b->d_partial);
    (void)hipFree(b->d_iarena); (void)hipFree(b->d_llarena);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->seg) (void)hipEventDestroy(b->seg);
    for (int i = 0; i < ph3::kClassStreams; ++i) {
        if (b->cls[i]) (void)hipStreamDestroy(b->cls[i]);
        if (b->cls_join[i]) (void)hipEventDestroy(b->cls_join[i]);
    }
}

extern "C" int ph3_batch_create(ph3_batch** out, int capacity) {
    if (!out || capacity < 1 || capacity > (1 << 28)) {
        set_error("ph3_batch_create: bad arguments (capacity %d)", capacity);
        return PH3_EINVAL;
    }
    ph3_batch* b = new ph3_batch();
    b->cap = capacity;
    const size_t cap = (size_t)capacity;
    hipError_t e = hipMalloc((void**)&b->d_desc, sizeof(ph3_solve_desc) * cap);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_counter, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_partial, sizeof(double) * ph3::kMaxSlices * cap);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_iarena, sizeof(int) * (5 * cap + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_llarena, sizeof(long long) * (cap + kMaxClasses));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->seg, hipEventDisableTiming);
    for (int i = 0; i < ph3::kClassStreams && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&b->cls[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->cls_join[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        destroy_parts(b);
        delete b;
        return ph3host::check(e, "ph3_batch_create");
    }
    *out = b;
    return PH3_OK;
}

extern "C" int ph3_batch_destroy(ph3_batch* b) {
    if (!b) return PH3_OK;
    clear_set(b);
    destroy_parts(b);
    delete b;
    return PH3_OK;
}

static int check_desc(const ph3_solve_desc& d, int i) {
    const bool shape_ok = (d.dims == 2 && d.nz == 1) || (d.dims == 3 && d.nz >= 3);
    if (!shape_ok || d.ny < 3 || d.nx < 3) { set_error("solve %d: bad shape", i); return PH3_EINVAL; }
    if (d.storage < 0 || d.storage > 2) { set_error("solve %d: bad storage", i); return PH3_EINVAL; }
    if (!d.buf[0] || !d.buf[1] || d.buf[0] == d.buf[1]) { set_error("solve %d: buffers", i); return PH3_EINVAL; }
    if (d.mode != 0 && d.mode != 1) { set_error("solve %d: mode %d", i, d.mode); return PH3_EINVAL; }
    if (d.tile < 0 || d.tile > 6 || (d.dims == 2 && d.tile > 5)) {
        set_error("solve %d: tile %d not in 0..%d", i, d.tile, d.dims == 2 ? 5 : 6);
        return PH3_EINVAL;
    }
    if (d.kpass < 0 || d.kpass > 8 || (d.kpass & (d.kpass - 1))) {
        set_error("solve %d: kpass %d not in {0, 1, 2, 4, 8}", i, d.kpass);
        return PH3_EINVAL;
    }
    if (d.steps < 0 || d.stall_window < 0 || d.stall_window >= PH3_STALL_RING) {
        set_error("solve %d: steps / stall_window out of range", i);
        return PH3_EINVAL;
    }
    // the fused tier-2 tiles read and write through buffer resources over one plane (3D) or row
    // (2D) of float32 values (the source is float32 whatever the storage)
    const size_t span = (d.dims == 3 ? (size_t)d.ny : 1) * (size_t)d.nx * sizeof(float);
    if (span > ph3::kMaxBufferBytes) {
        set_error("solve %d: a %s of %zu bytes exceeds the %zu-byte buffer window", i,
                  d.dims == 3 ? "plane" : "row", span, (size_t)ph3::kMaxBufferBytes);
        return PH3_EUNSUP;
    }
    return PH3_OK;
}

// Steps fused per tier-2 pass, a function of the solve alone: 1 for tier-1 grids and for 16-bit
// storage (whose rounding cadence is part of the result); for float32 the largest of 8 (2D) or
// 2 (3D) halvings that divides PH3_BATCH_CHECK and the step count. A caller's explicit kpass > 1
// is kept when it is valid.
static int choose_kpass(const ph3_solve_desc& d) {
    const long long sites = (long long)d.nz * d.ny * d.nx;
    if (sites <= PH3_BATCH_SMALL_SITES || d.storage != ph3::kF32) return 1;
    // 3D: 4-step passes from 128^3 sites (full-size heat mix 3012 -> 3714 solves/s), 2-step
    // below (smaller grids pay more for the deeper halo: Jacobi mixes 5% slower with 4)
    const int k3 = sites >= (1LL << 21) ? 4 : 2;
    int k = d.kpass > 1 ? d.kpass : (d.dims == 3 ? k3 : 8);
    if (d.dims == 3 && k > 4) k = 4;
    if (k > 8) k = 8;
    while (k > 1 && (PH3_BATCH_CHECK % k || d.steps % k)) k >>= 1;
    return k;
}

// Tile shape of a tier-2 solve (tier2_shape ids). The caller's `tile` is taken when it is one of
// the fused shapes; 2-site vector lanes need an even width above 64 and 8-byte aligned buffers,
// and fall back otherwise: 2D shape 4 to shape 5 (site-by-site access), the others to the
// 1-site shape of the same width class. Default (tile 0): in 3D,
// 2-site lanes for 2-step grids where allowed (a 128-wide tile then idles no more lanes than
// two 64-wide ones), and for 4-step grids the narrow 1-site shape storing whole planes (3D
// 128..256 Jacobi mix: 1132 GLUPS with masked stores, 1212); in 2D, 2-site lanes storing on
// every step (2D 600..2048 Jacobi mix: 2070 GLUPS with 1-site masked strips, 2512), as one
// vector access per lane where the width allows, else site by site (shape 5).
static int tier2_shape_for(const ph3_solve_desc& d) {
    if (d.kpass <= 1) return 1;
    auto al8 = [](const void* q) { return ((uintptr_t)q & 7u) == 0; };
    const bool two = d.nx > 64 && d.nx % 2 == 0 && al8(d.buf[0]) && al8(d.buf[1]) &&
                     (!d.src || al8(d.src));
    int id = d.tile;
    if (id == 0) id = d.dims == 3 ? (d.kpass == 2 ? 2 : 5) : (two ? 4 : 5);
    if (!two) {
        if (d.dims == 2 && id == 4) id = 5;               // the same lanes, site by site
        else if (id == 2 || id == 4) id -= 1;
    }
    return id;
}

extern "C" int ph3_batch_params(ph3_batch* b, int dims, int kpass, int target_tiles, int max_chunk) {
    if (!b || (dims != 2 && dims != 3) || kpass < 1 || kpass > 8 || (kpass & (kpass - 1)) ||
        target_tiles < 0 || max_chunk < 0 || max_chunk > 4096) {
        set_error("ph3_batch_params: bad arguments");
        return PH3_EINVAL;
    }
    int lk = 0;
    while ((1 << lk) < kpass) ++lk;
    b->target_tiles[dims - 2][lk] = target_tiles;
    b->max_chunk[dims - 2][lk] = max_chunk;
    return PH3_OK;
}

extern "C" int ph3_batch_set(ph3_batch* b, const ph3_solve_desc* descs, int n, void* stream) {
    if (!b || n < 0 || n > b->cap || (n && !descs)) { set_error("ph3_batch_set: bad arguments"); return PH3_EINVAL; }
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < n; ++i)
        if (int r = check_desc(descs[i], i)) return r;
    PH3_CHECK(hipStreamSynchronize(s));
    // The new set is built in locals and committed only once every upload and the initial
    // residuals have succeeded; from here until then the batch holds no solves.
    clear_set(b);
    std::vector<ph3_solve_desc> table(descs, descs + n);
    for (auto& d : table) d.kpass = choose_kpass(d);
    std::vector<int> small[3], init[3];
    std::vector<Tier2Class> classes;
    for (int i = 0; i < n; ++i) {
        const ph3_solve_desc& d = table[i];
        init[d.storage].push_back(i);
        const long long sites = (long long)d.nz * d.ny * d.nx;
        if (sites <= PH3_BATCH_SMALL_SITES) { small[d.storage].push_back(i); continue; }
        const int shape = tier2_shape_for(d);
        Tier2Class* c = nullptr;
        for (auto& k : classes)
            if (k.storage == d.storage && k.dims == d.dims && k.periodic == (d.periodic != 0) &&
                k.src == (d.src != nullptr) && k.kpass == d.kpass && k.shape == shape) c = &k;
        if (!c) {
            Tier2Class k;
            k.storage = d.storage; k.dims = d.dims; k.periodic = d.periodic != 0;
            k.src = d.src != nullptr; k.kpass = d.kpass; k.shape = shape;
            classes.push_back(k);
            c = &classes.back();
        }
        c->idx.push_back(i);
    }
    // host images of the two arenas; offsets are in elements
    std::vector<int> ih;
    std::vector<long long> lh;
    auto put = [&ih](const std::vector<int>& v) { const size_t o = ih.size(); ih.insert(ih.end(), v.begin(), v.end()); return o; };
    size_t o_small[3], o_init[3];
    for (int st = 0; st < 3; ++st) { o_small[st] = put(small[st]); o_init[st] = put(init[st]); }
    std::vector<size_t> o_list(classes.size()), o_prefix(classes.size()), o_res(classes.size());
    int partials = 0;
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        Tier2Class& c = classes[ci];
        const ph3::T2Shape sh = ph3::tier2_shape(c.dims, c.kpass, c.shape);
        int lk = 0;
        while ((1 << lk) < c.kpass) ++lk;
        const int target = b->target_tiles[c.dims - 2][lk] > 0 ? b->target_tiles[c.dims - 2][lk]
                                                                : ph3::kTier2Tiles;
        // tiles per grid: the class's target split over its grids, at most the lone-grid figure
        const long long per = (target + (long long)c.idx.size() - 1) / (long long)c.idx.size();
        c.tpg = (int)(per > ph3::kTier2Tiles ? ph3::kTier2Tiles : (per < 1 ? 1 : per));
        if (b->target_tiles[c.dims - 2][lk] <= 0) c.tpg = ph3::kTier2Tiles;   // default policy
        c.max_chunk = b->max_chunk[c.dims - 2][lk];
        long long total = 0;
        for (int i : c.idx) {
            const ph3_solve_desc& d = table[i];
            total += ph3::tier2_geom(d.dims, d.nz, d.ny, d.nx, d.kpass, 0, sh, c.tpg, c.max_chunk).tiles;
        }
        c.few = total < ph3::kPassGroups ? 1 : 0;
        o_list[ci] = put(c.idx);
        o_prefix[ci] = lh.size();
        lh.push_back(0);
        for (int i : c.idx) {
            const ph3_solve_desc& d = table[i];
            lh.push_back(lh.back() + ph3::tier2_geom(d.dims, d.nz, d.ny, d.nx, d.kpass, c.few, sh,
                                                     c.tpg, c.max_chunk).items);
        }
        c.items = lh.back();
        // residual slices of the class's solves, and where its partial sums start
        std::vector<int> rp(c.idx.size() + 1, 0);
        for (size_t t = 0; t < c.idx.size(); ++t) rp[t + 1] = rp[t] + ph3::res_slices(table[c.idx[t]]);
        o_res[ci] = put(rp);
        c.res_total = rp.back();
        c.partial_base = partials;
        partials += c.res_total;
    }
    if (ih.size() > 5 * (size_t)b->cap + 1 || lh.size() > (size_t)b->cap + kMaxClasses) {
        set_error("ph3_batch_set: index tables exceed the batch's arenas");
        return PH3_EINVAL;
    }
    // Uploads from these host vectors: the stream is synchronised before they go out of scope,
    // on success and on every failure after the first enqueued copy.
    auto fail = [s](hipError_t e, const char* what) {
        (void)hipStreamSynchronize(s);
        return ph3host::check(e, what);
    };
    hipError_t e = hipSuccess;
    if (n) e = hipMemcpyAsync(b->d_desc, table.data(), sizeof(ph3_solve_desc) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !ih.empty())
        e = hipMemcpyAsync(b->d_iarena, ih.data(), sizeof(int) * ih.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !lh.empty())
        e = hipMemcpyAsync(b->d_llarena, lh.data(), sizeof(long long) * lh.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(e, "ph3_batch_set: upload");
    // initial residuals of unchecked Jacobi solves
    for (int st = 0; st < 3; ++st) {
        const int ni = (int)init[st].size();
        if (!ni) continue;
        const int* list = b->d_iarena + o_init[st];
        const unsigned g = (unsigned)(ni < ph3::kPassGroups ? ni : ph3::kPassGroups);
        switch (st) {
            case ph3::kF32: ph3::batch_init_residual<float><<<g, ph3::kWG, 0, s>>>(b->d_desc, list, ni); break;
            case ph3::kF16: ph3::batch_init_residual<_Float16><<<g, ph3::kWG, 0, s>>>(b->d_desc, list, ni); break;
            default: ph3::batch_init_residual<ph3::bf16_t><<<g, ph3::kWG, 0, s>>>(b->d_desc, list, ni); break;
        }
        e = hipGetLastError();
        if (e != hipSuccess) return fail(e, "batch initial residual");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ph3host::check(e, "ph3_batch_set: synchronize");
    // commit
    for (int st = 0; st < 3; ++st) {
        b->n_small[st] = (int)small[st].size();
        b->d_small[st] = b->n_small[st] ? b->d_iarena + o_small[st] : nullptr;
    }
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        classes[ci].d_list = b->d_iarena + o_list[ci];
        classes[ci].d_prefix = b->d_llarena + o_prefix[ci];
        classes[ci].d_res_prefix = b->d_iarena + o_res[ci];
    }
    b->classes = std::move(classes);
    b->n = n;
    return PH3_OK;
}

template <typename T, bool P, bool S, int DIMS, int K, int V, int R, int W, int SM, bool EW = false>
static void launch_pass(const Tier2Class& c, const ph3_batch* b, int p, hipStream_t s) {
    // one workgroup per items_per_group items (a multiple of 8 workgroups, one per XCD in
    // turn), at most 2^20 (the kernel loops past that)
    constexpr int ipg = ph3::items_per_group(DIMS, K, W);
    long long g = (c.items + ipg - 1) / ipg;
    g = (g + 7) / 8 * 8;
    if (g > (1 << 20)) g = 1 << 20;
    if (g < 8) g = 8;
    ph3::batch_pass<T, P, S, DIMS, K, V, R, W, SM, EW><<<(unsigned)g, 64 * W, 0, s>>>(
        b->d_desc, c.d_list, c.d_prefix, (int)c.idx.size(), p, c.few, c.tpg, c.max_chunk);
}

// Fused passes of one K by shape id (tier2_shape).
template <typename T, bool P, bool S, int DIMS, int K>
static void pass_shape(const Tier2Class& c, const ph3_batch* b, int p, hipStream_t s) {
    using namespace ph3;
    if constexpr (DIMS == 3) {
        switch (c.shape) {
            case 2: return launch_pass<T, P, S, 3, K, 2, 8, 4, kStoreMasked>(c, b, p, s);
            case 3: return launch_pass<T, P, S, 3, K, 1, 8, 8, kStoreMasked>(c, b, p, s);
            case 4: return launch_pass<T, P, S, 3, K, 2, 4, 8, kStoreMasked>(c, b, p, s);
            case 5: return launch_pass<T, P, S, 3, K, 1, 8, 4, kStorePlanes>(c, b, p, s);
            case 6: return launch_pass<T, P, S, 3, K, 1, 8, 8, kStorePlanes>(c, b, p, s);
            default: return launch_pass<T, P, S, 3, K, 1, 8, 4, kStoreMasked>(c, b, p, s);
        }
    } else {
        switch (c.shape) {
            case 2: return launch_pass<T, P, S, 2, K, 2, 0, kW2, kStoreMasked>(c, b, p, s);
            case 3: return launch_pass<T, P, S, 2, K, 1, 0, kW2, kStoreAll>(c, b, p, s);
            case 4: return launch_pass<T, P, S, 2, K, 2, 0, kW2, kStoreAll>(c, b, p, s);
            case 5: return launch_pass<T, P, S, 2, K, 2, 0, kW2, kStoreAll, true>(c, b, p, s);
            default: return launch_pass<T, P, S, 2, K, 1, 0, kW2, kStoreMasked>(c, b, p, s);
        }
    }
}

template <typename T, bool P, bool S, int DIMS>
static void pass_k(const Tier2Class& c, const ph3_batch* b, int p, hipStream_t s) {
    using namespace ph3;
    if constexpr (std::is_same<T, float>::value) {
        if (c.kpass == 2) return pass_shape<T, P, S, DIMS, 2>(c, b, p, s);
        if (c.kpass == 4) return pass_shape<T, P, S, DIMS, 4>(c, b, p, s);
        if constexpr (DIMS == 2) {
            if (c.kpass == 8) return pass_shape<T, P, S, DIMS, 8>(c, b, p, s);
        }
    }
    launch_pass<T, P, S, DIMS, 1, 1, DIMS == 3 ? kR3 : kR2, DIMS == 3 ? kW3 : kW2, kStoreMasked>(c, b, p, s);
}

template <typename T>
static void pass_class(const Tier2Class& c, const ph3_batch* b, int p, hipStream_t s) {
    const bool P = c.periodic, S = c.src;
    if (c.dims == 3) {
        if (P) { if (S) pass_k<T, true, true, 3>(c, b, p, s); else pass_k<T, true, false, 3>(c, b, p, s); }
        else { if (S) pass_k<T, false, true, 3>(c, b, p, s); else pass_k<T, false, false, 3>(c, b, p, s); }
    } else {
        if (P) { if (S) pass_k<T, true, true, 2>(c, b, p, s); else pass_k<T, true, false, 2>(c, b, p, s); }
        else { if (S) pass_k<T, false, true, 2>(c, b, p, s); else pass_k<T, false, false, 2>(c, b, p, s); }
    }
}

extern "C" int ph3_batch_round(ph3_batch* b, int sweeps, void* stream) {
    if (!b || sweeps < 1) { set_error("ph3_batch_round: bad arguments"); return PH3_EINVAL; }
    hipStream_t s = (hipStream_t)stream;
    const int segs = (sweeps + PH3_BATCH_CHECK - 1) / PH3_BATCH_CHECK;
    const int rs = segs * PH3_BATCH_CHECK;
    // Tier 1 and tier 2 touch disjoint solves, so when both are present tier 1 runs on the
    // batch's side stream, forked from and joined back to `stream` with events (in a captured
    // graph: two parallel branches). Its few long-running workgroups then no longer keep the
    // tier-2 passes waiting.
    const bool any_small = b->n_small[0] + b->n_small[1] + b->n_small[2] > 0;
    const bool fork = any_small && !b->classes.empty();
    hipStream_t s1 = fork ? b->side : s;
    if (fork) {
        PH3_CHECK(hipEventRecord(b->fork, s));
        PH3_CHECK(hipStreamWaitEvent(s1, b->fork, 0));
    }
    // tier 1: one persistent launch per storage type present
    for (int st = 0; st < 3; ++st) {
        if (!b->n_small[st]) continue;
        ph3::batch_reset<<<1, 1, 0, s1>>>(b->d_counter);
        PH3_LAUNCH_CHECK("batch counter reset");
        const unsigned g = (unsigned)(b->n_small[st] < ph3::kPassGroups ? b->n_small[st] : ph3::kPassGroups);
        switch (st) {
            case ph3::kF32: ph3::batch_small<float><<<g, ph3::kWG, 0, s1>>>(b->d_desc, b->d_small[st], b->n_small[st], b->d_counter, rs); break;
            case ph3::kF16: ph3::batch_small<_Float16><<<g, ph3::kWG, 0, s1>>>(b->d_desc, b->d_small[st], b->n_small[st], b->d_counter, rs); break;
            default: ph3::batch_small<ph3::bf16_t><<<g, ph3::kWG, 0, s1>>>(b->d_desc, b->d_small[st], b->n_small[st], b->d_counter, rs); break;
        }
        PH3_LAUNCH_CHECK("batch tier-1 launch");
    }
    if (fork) PH3_CHECK(hipEventRecord(b->join, s1));
    // tier 2: per class, segments of PH3_BATCH_CHECK sweeps in k-step varlen passes, each
    // followed by the class's residual and check (advance / convergence test). Classes touch
    // disjoint solves, so each runs its whole round on its own stream (the first on `stream`),
    // forked and joined once per round with events: the classes' chains, residuals and checks
    // overlap, and a round takes about the longest chain.
    const int nc = (int)b->classes.size();
    if (nc > 0) {
        const int nside = nc - 1 < ph3::kClassStreams ? nc - 1 : ph3::kClassStreams;
        if (nside > 0) {
            PH3_CHECK(hipEventRecord(b->seg, s));
            for (int i = 0; i < nside; ++i) PH3_CHECK(hipStreamWaitEvent(b->cls[i], b->seg, 0));
        }
        for (int ci = 0; ci < nc; ++ci) {
            const auto& c = b->classes[ci];
            hipStream_t sc = ci == 0 ? s : b->cls[(ci - 1) % nside];
            const int n = (int)c.idx.size();
            const unsigned gl = (unsigned)((n + 255) / 256);
            const unsigned gr = (unsigned)(c.res_total < 2 * ph3::kPassGroups ? c.res_total : 2 * ph3::kPassGroups);
            double* part = b->d_partial + c.partial_base;
            for (int seg = 0; seg < segs; ++seg) {
                for (int p = 0; p < PH3_BATCH_CHECK / c.kpass; ++p) {  // k-step passes: fewer
                    switch (c.storage) {
                        case ph3::kF32: pass_class<float>(c, b, p, sc); break;
                        case ph3::kF16: pass_class<_Float16>(c, b, p, sc); break;
                        default: pass_class<ph3::bf16_t>(c, b, p, sc); break;
                    }
                    PH3_LAUNCH_CHECK("batch tier-2 pass");
                }
                if (c.res_total > 0) {   // the class has Jacobi solves
                    switch (c.storage) {
                        case ph3::kF32: ph3::batch_residual<float><<<gr, ph3::kWG, 0, sc>>>(b->d_desc, c.d_list, c.d_res_prefix, n, part); break;
                        case ph3::kF16: ph3::batch_residual<_Float16><<<gr, ph3::kWG, 0, sc>>>(b->d_desc, c.d_list, c.d_res_prefix, n, part); break;
                        default: ph3::batch_residual<ph3::bf16_t><<<gr, ph3::kWG, 0, sc>>>(b->d_desc, c.d_list, c.d_res_prefix, n, part); break;
                    }
                    PH3_LAUNCH_CHECK("batch residual");
                }
                ph3::batch_check<<<gl, 256, 0, sc>>>(b->d_desc, c.d_list, c.d_res_prefix, n, part);
                PH3_LAUNCH_CHECK("batch check");
            }
        }
        for (int i = 0; i < nside; ++i) {
            PH3_CHECK(hipEventRecord(b->cls_join[i], b->cls[i]));
            PH3_CHECK(hipStreamWaitEvent(s, b->cls_join[i], 0));
        }
    }
    if (fork) PH3_CHECK(hipStreamWaitEvent(s, b->join, 0));
    return PH3_OK;
}

extern "C" int ph3_batch_get(ph3_batch* b, ph3_solve_desc* descs, int n, void* stream) {
    if (!b || n < 0 || n > b->n || (n && !descs)) { set_error("ph3_batch_get: bad arguments"); return PH3_EINVAL; }
    if (n) PH3_CHECK(hipMemcpyAsync(descs, b->d_desc, sizeof(ph3_solve_desc) * n, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PH3_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return PH3_OK;
}
