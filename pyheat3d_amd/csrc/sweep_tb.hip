// Rung "tb<k>": k time steps fused per HBM pass (counterpart of reference
// backends/stencil.cu:163-385, tb3d and tb2d), redesigned for wave64.
//
// Overlapped tiling: every workgroup reads an extended tile, recomputes the k-deep halo
// redundantly and writes only the centre, so workgroups never exchange data. The valid region
// shrinks by one site per fused step, hence values computed near the tile edge from missing
// neighbours (the edge lanes / rows use their own value) never reach the written centre.
//
// 3D. W waves are stacked in y (wave w owns R rows); lane l owns V contiguous x sites. The
// workgroup marches z through a (z, t) wavefront: at iteration iz, plane iz enters level 0 and
// stage s (1..k) computes level s of plane iz - s from the 3-plane register queue of level
// s-1. Rows across waves go through a double-buffered LDS slot (one barrier per stage); x
// neighbours come from the lane itself or from lanes l-1 / l+1 by cross-lane shuffle. Level 0
// is read from HBM once and level k written once per pass: 2 * sizeof(T) / k bytes per update
// plus the halo over-fetch of the extended tile.
//
// 2D. Each wave owns an extended strip of 64*V columns and marches y with the same wavefront
// (rows play the role of planes); waves are independent, so there is no LDS and no barrier.
//
// Dirichlet: sites on or outside the grid boundary are fixed (a stage keeps their value);
// intermediate levels stay in float32 registers and only level k is rounded to storage.
#include "sweep_impl.h"
#include "vec_io.h"

namespace ph3 {

__device__ __forceinline__ int tb_logical_tile(int remap) {
    const int L = blockIdx.x;
    if (!remap) return L;
    const int per = gridDim.x >> 3;
    return (L & 7) * per + (L >> 3);
}

__device__ __forceinline__ int wrap_or_clamp(int i, int n, bool periodic) {
    if (periodic) { i %= n; return i < 0 ? i + n : i; }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// First site of a V-wide lane vector at x = i (i % V == 0, n % V == 0): wrapped, or clamped
// so that all V sites stay inside the row.
template <int V>
__device__ __forceinline__ int vec_col(int i, int n, bool periodic) {
    if (periodic) { i %= n; return i < 0 ? i + n : i; }
    return i < 0 ? 0 : (i > n - V ? n - V : i);
}

// halo of the extended tile in x, rounded up to whole lane vectors (keeps vector alignment)
template <int V, int K> struct TbGeom {
    static constexpr int KV = ((K + V - 1) / V) * V;
    static constexpr int OX = 64 * V - 2 * KV;  // output sites per tile along x
};

template <typename T, bool P, bool S, int V, int R, int W, int K>
__global__ void __launch_bounds__(64 * W)
tb3d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
     int nz, int ny, int nx, float c0, float c1, int zch, int tiles_x, int tiles_y, int remap)
{
    using G = TbGeom<V, K>;
    constexpr int OY = W * R - 2 * K;  // output rows per tile
    __shared__ float edge[2][W][2][64 * V];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int L = tb_logical_tile(remap);
    const int tx = L % tiles_x;
    const int rest = L / tiles_x;
    const int ty = rest % tiles_y;
    const int cz = rest / tiles_y;
    const int z0 = cz * zch, z1 = min(z0 + zch, nz);
    const int gx = tx * G::OX - G::KV + lane * V;   // first x site of this lane (may be outside)
    const int gy0 = ty * OY - K + wave * R;         // first row of this wave (may be outside)
    const int lx = vec_col<V>(gx, nx, P);           // vector-aligned: nx % V == 0, gx % V == 0
    const size_t plane = (size_t)nx * ny;

    size_t roff[R];
    uint32_t fixmask = 0;      // bit r*V+e: site (row r, element e) is a fixed Dirichlet site
    uint32_t wmask = 0;        // bit r*V+e: site is written (inside output region and grid)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gy = gy0 + r;
        roff[r] = (size_t)wrap_or_clamp(gy, ny, P) * nx;
        const int row = wave * R + r;
        const bool row_out = row >= K && row < W * R - K && gy >= 0 && gy < ny;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int x = gx + e;
            const int col = lane * V + e;
            const bool fixed = !P && (x <= 0 || x >= nx - 1 || gy <= 0 || gy >= ny - 1);
            if (fixed) fixmask |= 1u << (r * V + e);
            if (row_out && col >= G::KV && col < 64 * V - G::KV && x >= 0 && x < nx && !fixed)
                wmask |= 1u << (r * V + e);
        }
    }

    float q[K][3][R][V];   // level s: planes c-1, c, c+1 (c = plane computed by stage s+1)
    float pref[R][V];
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < V; ++e) q[s][p][r][e] = 0.f;
    {
        const size_t zo = (size_t)wrap_or_clamp(z0 - K, nz, P) * plane;
#pragma unroll
        for (int r = 0; r < R; ++r) load_vec<T, V>(u, zo + roff[r] + lx, pref[r]);
    }
    int par = 0;
    for (int iz = z0 - K; iz < z1 + K; ++iz) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int e = 0; e < V; ++e) {
                q[0][0][r][e] = q[0][1][r][e];
                q[0][1][r][e] = q[0][2][r][e];
                q[0][2][r][e] = pref[r][e];
            }
        if (iz + 1 < z1 + K) {
            const size_t zo = (size_t)wrap_or_clamp(iz + 1, nz, P) * plane;
#pragma unroll
            for (int r = 0; r < R; ++r) load_vec<T, V>(u, zo + roff[r] + lx, pref[r]);
        }
        float res[R][V];
#pragma unroll
        for (int s = 1; s <= K; ++s) {
            const int c = iz - s;                       // plane computed by this stage
            const bool zfixed = !P && (c <= 0 || c >= nz - 1);
            const float (&ctr)[R][V] = q[s - 1][1];
            // rows of the neighbouring waves (level s-1, plane c)
            float up[V], dn[V];
            if (W > 1) {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    edge[par][wave][0][lane * V + e] = ctr[0][e];
                    edge[par][wave][1][lane * V + e] = ctr[R - 1][e];
                }
                __syncthreads();
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    up[e] = wave > 0 ? edge[par][wave - 1][1][lane * V + e] : ctr[0][e];
                    dn[e] = wave < W - 1 ? edge[par][wave + 1][0][lane * V + e] : ctr[R - 1][e];
                }
                par ^= 1;
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) { up[e] = ctr[0][e]; dn[e] = ctr[R - 1][e]; }
            }
            float bsrc[R][V];
            if (S) {
                const size_t zo = (size_t)wrap_or_clamp(c, nz, P) * plane;
#pragma unroll
                for (int r = 0; r < R; ++r) load_vec<float, V>(b, zo + roff[r] + lx, bsrc[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float fromL = __shfl_up(ctr[r][V - 1], 1);
                const float fromR = __shfl_down(ctr[r][0], 1);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const float xm = e > 0 ? ctr[r][e - 1] : fromL;
                    const float xp = e < V - 1 ? ctr[r][e + 1] : fromR;
                    const float ym = r > 0 ? ctr[r - 1][e] : up[e];
                    const float yp = r < R - 1 ? ctr[r + 1][e] : dn[e];
                    float sum = xm + xp;
                    sum = sum + (ym + yp);
                    sum = sum + (q[s - 1][0][r][e] + q[s - 1][2][r][e]);
                    float v = update(c0, c1, ctr[r][e], sum);
                    if (S) v = v + bsrc[r][e];
                    if (!P && (zfixed || ((fixmask >> (r * V + e)) & 1u))) v = ctr[r][e];
                    res[r][e] = v;
                }
            }
            if (s < K) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        q[s][0][r][e] = q[s][1][r][e];
                        q[s][1][r][e] = q[s][2][r][e];
                        q[s][2][r][e] = res[r][e];
                    }
            }
        }
        const int c = iz - K;
        const bool zok = c >= z0 && c < z1 && (P || (c > 0 && c < nz - 1));
        if (zok && wmask) {
            const size_t zo = (size_t)c * plane;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                constexpr uint32_t full = (1u << V) - 1u;
                const uint32_t m = (wmask >> (r * V)) & full;
                if (m == full) {
                    store_vec<T, V>(o, zo + roff[r] + gx, res[r]);
                } else if (m) {
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        if ((m >> e) & 1u) st(o, zo + roff[r] + gx + e, res[r][e]);
                }
            }
        }
    }
}

template <typename T, bool P, bool S, int V, int K, int W>
__global__ void __launch_bounds__(64 * W)
tb2d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
     int ny, int nx, float c0, float c1, int ych, int tiles_x, int remap)
{
    using G = TbGeom<V, K>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int L = tb_logical_tile(remap);
    const int strip = (L % tiles_x) * W + wave;
    const int cy = L / tiles_x;
    const int y0 = cy * ych, y1 = min(y0 + ych, ny);
    const int gx = strip * G::OX - G::KV + lane * V;
    if (strip * G::OX - G::KV >= nx) return;          // whole wave past the grid (no barriers)
    const int lx = vec_col<V>(gx, nx, P);
    uint32_t fixmask = 0, wmask = 0;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int x = gx + e;
        const int col = lane * V + e;
        const bool fixed = !P && (x <= 0 || x >= nx - 1);
        if (fixed) fixmask |= 1u << e;
        if (col >= G::KV && col < 64 * V - G::KV && x >= 0 && x < nx && !fixed) wmask |= 1u << e;
    }
    float q[K][3][V];
    float pref[V];
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int e = 0; e < V; ++e) q[s][p][e] = 0.f;
    load_vec<T, V>(u, (size_t)wrap_or_clamp(y0 - K, ny, P) * nx + lx, pref);
    for (int iy = y0 - K; iy < y1 + K; ++iy) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
            q[0][0][e] = q[0][1][e];
            q[0][1][e] = q[0][2][e];
            q[0][2][e] = pref[e];
        }
        if (iy + 1 < y1 + K)
            load_vec<T, V>(u, (size_t)wrap_or_clamp(iy + 1, ny, P) * nx + lx, pref);
        float res[V];
#pragma unroll
        for (int s = 1; s <= K; ++s) {
            const int c = iy - s;
            const bool yfixed = !P && (c <= 0 || c >= ny - 1);
            const float (&ctr)[V] = q[s - 1][1];
            float bsrc[V];
            if (S) load_vec<float, V>(b, (size_t)wrap_or_clamp(c, ny, P) * nx + lx, bsrc);
            const float fromL = __shfl_up(ctr[V - 1], 1);
            const float fromR = __shfl_down(ctr[0], 1);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const float xm = e > 0 ? ctr[e - 1] : fromL;
                const float xp = e < V - 1 ? ctr[e + 1] : fromR;
                float sum = xm + xp;
                sum = sum + (q[s - 1][0][e] + q[s - 1][2][e]);
                float v = update(c0, c1, ctr[e], sum);
                if (S) v = v + bsrc[e];
                if (!P && (yfixed || ((fixmask >> e) & 1u))) v = ctr[e];
                res[e] = v;
            }
            if (s < K) {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    q[s][0][e] = q[s][1][e];
                    q[s][1][e] = q[s][2][e];
                    q[s][2][e] = res[e];
                }
            }
        }
        const int c = iy - K;
        if (wmask && c >= y0 && c < y1 && (P || (c > 0 && c < ny - 1))) {
            const size_t at = (size_t)c * nx + gx;
            constexpr uint32_t full = (1u << V) - 1u;
            if (wmask == full) {
                store_vec<T, V>(o, at, res);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if ((wmask >> e) & 1u) st(o, at + e, res[e]);
            }
        }
    }
}

// ---- launch ---------------------------------------------------------------------------------

static int tb_pick_vec(const SweepArgs& a, int elem_bytes, int maxv) {
    const uintptr_t addr = (uintptr_t)a.u | (uintptr_t)a.o;
    for (int v = maxv; v > 1; v >>= 1) {
        if (v * elem_bytes > 16) continue;
        const bool aligned = addr % (size_t)(v * elem_bytes) == 0 &&
                             (!a.b || (uintptr_t)a.b % (size_t)(v * 4) == 0);
        if (a.nx % v == 0 && aligned) return v;
    }
    return 1;
}

static int tb_chunk(int n, long tiles, int k, int forced) {
    if (forced > 0) return forced;
    // ~1024 workgroups, chunks of at least 16k planes so the 2k-plane pipeline fill is small
    long want = (1024 + tiles - 1) / tiles;
    if (want < 1) want = 1;
    int ch = (int)((n + want - 1) / want);
    const int lo = 16 * k;
    return ch < lo ? lo : ch;
}

template <typename T, bool P, bool S, int V, int R, int W, int K>
static int tb3_launch(const SweepArgs& a) {
    using G = TbGeom<V, K>;
    constexpr int OY = W * R - 2 * K;
    static_assert(OY > 0 && G::OX > 0, "tile too small for k");
    static_assert(R * V <= 32, "site masks are 32-bit");
    const int tiles_x = (a.nx + G::OX - 1) / G::OX;
    const int tiles_y = (a.ny + OY - 1) / OY;
    const long tiles = (long)tiles_x * tiles_y;
    const int zch = tb_chunk(a.nz, tiles, K, a.tile.zch);
    const long groups = tiles * ((a.nz + zch - 1) / zch);
    if (groups > 0x7fffffffL) { ph3host::set_error("tb3d: grid too large"); return PH3_EINVAL; }
    tb3d<T, P, S, V, R, W, K><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.nz, a.ny, a.nx, a.c0, a.c1, zch, tiles_x, tiles_y,
        groups % 8 == 0 ? 1 : 0);
    PH3_LAUNCH_CHECK("tb3d launch");
    return PH3_OK;
}

template <typename T, bool P, bool S, int V, int K, int W>
static int tb2_launch(const SweepArgs& a) {
    using G = TbGeom<V, K>;
    const int strips = (a.nx + G::OX - 1) / G::OX;
    const int tiles_x = (strips + W - 1) / W;
    const int ych = tb_chunk(a.ny, tiles_x, K, a.tile.zch);
    const long groups = (long)tiles_x * ((a.ny + ych - 1) / ych);
    if (groups > 0x7fffffffL) { ph3host::set_error("tb2d: grid too large"); return PH3_EINVAL; }
    tb2d<T, P, S, V, K, W><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.ny, a.nx, a.c0, a.c1, ych, tiles_x,
        groups % 8 == 0 ? 1 : 0);
    PH3_LAUNCH_CHECK("tb2d launch");
    return PH3_OK;
}

// 3D shapes (vec, rows per wave, waves): default (2, 4, 16); (1, 8, 8) for odd widths.
template <typename T, bool P, bool S, int K>
static int tb3_shape(const SweepArgs& a) {
    // Defaults measured on MI355X at 1024^3 float32 (vec, rows, waves): k=2 (4, 2, 16)
    // 981 GLUPS; k=3 (2, 4, 16) 1291 GLUPS; k=4 (1, 4, 16) 977 GLUPS (register-bound).
    const int vmax = K == 2 ? 4 : (K == 3 ? 2 : 1);
    int v = a.tile.vec ? a.tile.vec : tb_pick_vec(a, (int)sizeof(T), vmax);
    int r = a.tile.ty ? a.tile.ty : (v == 4 ? 2 : 4);
    int w = a.tile.waves ? a.tile.waves : 16;
    if (a.nx % v) { ph3host::set_error("tb3d: vec %d does not divide nx %d", v, a.nx); return PH3_EINVAL; }
#define PH3_TB3(V_, R_, W_) \
    if (v == V_ && r == R_ && w == W_) return tb3_launch<T, P, S, V_, R_, W_, K>(a)
    PH3_TB3(2, 4, 16); PH3_TB3(1, 8, 8); PH3_TB3(2, 4, 8); PH3_TB3(4, 2, 16); PH3_TB3(1, 4, 16);
#undef PH3_TB3
    ph3host::set_error("tb3d: (vec %d, rows %d, waves %d) not built; built: (2,4,16) (1,8,8) "
                       "(2,4,8) (4,2,16) (1,4,16)", v, r, w);
    return PH3_EINVAL;
}

template <typename T, bool P, bool S, int K>
static int tb2_shape(const SweepArgs& a) {
    int v = a.tile.vec ? a.tile.vec : tb_pick_vec(a, (int)sizeof(T), 16 / (int)sizeof(T));
    if (a.nx % v) { ph3host::set_error("tb2d: vec %d does not divide nx %d", v, a.nx); return PH3_EINVAL; }
    const int w = a.tile.waves ? a.tile.waves : 4;
    if (w != 4) { ph3host::set_error("tb2d: waves %d not built (4)", w); return PH3_EINVAL; }
    switch (v) {
        case 1: return tb2_launch<T, P, S, 1, K, 4>(a);
        case 2: return tb2_launch<T, P, S, 2, K, 4>(a);
        case 4: return tb2_launch<T, P, S, 4, K, 4>(a);
        case 8:
            if constexpr (sizeof(T) == 2) return tb2_launch<T, P, S, 8, K, 4>(a);
            break;
    }
    ph3host::set_error("tb2d: vec %d unsupported", v);
    return PH3_EINVAL;
}

bool tb_supported(int dims, int k) {
    if (dims == 3) return k == 2 || k == 3 || k == 4;
    if (dims == 2) return k == 2 || k == 3 || k == 4 || k == 8;
    return false;
}

template <int K>
static int tb_k(const SweepArgs& a) {
    PH3_DISPATCH_TPS(a, {
        if (a.dims == 3) {
            if constexpr (K <= 4) return tb3_shape<T, P, S, K>(a);
        } else {
            return tb2_shape<T, P, S, K>(a);
        }
    });
    ph3host::set_error("tb: k=%d not built for %dD", K, a.dims);
    return PH3_EUNSUP;
}

int launch_tb(const SweepArgs& a, int k) {
    if (a.tile.flags) { ph3host::set_error("tb: tile flags must be 0"); return PH3_EINVAL; }
    switch (k) {
        case 2: return tb_k<2>(a);
        case 3: return tb_k<3>(a);
        case 4: return tb_k<4>(a);
        case 8: return tb_k<8>(a);
    }
    ph3host::set_error("tb: k=%d not built", k);
    return PH3_EUNSUP;
}

}  // namespace ph3
