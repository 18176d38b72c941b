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
// neighbours come from the lane itself or from lanes l-1 / l+1 by a DPP wave shift. Level 0
// is read from HBM once and level k written once per pass: 2 * sizeof(T) / k bytes per update
// plus the halo over-fetch of the extended tile.
//
// 2D. Each wave owns an extended strip of 64*V columns and marches y with the same wavefront
// (rows play the role of planes); waves are independent, so there is no LDS and no barrier.
//
// Dirichlet: sites on or outside the grid boundary are fixed (a stage keeps their value);
// intermediate levels stay in float32 registers and only level k is rounded to storage.
#include "tb_tiles.h"

namespace ph3 {

// Tiles that exchange the edge rows of all k stages of a plane step at once, behind one
// barrier (X1 in tb3d_tile), chosen from a sweep of every built shape at 256^3-1024^3 float32
// (and bfloat16 at 768^3 / 1024^3), both boundaries, with and without a source: the 16-wave
// 2-site tiles of k >= 3 gain 4-27% (tb4 1024^3 periodic 1635 -> 2078 GLUPS, Dirichlet
// 1673 -> 2017), every tile with a source 12-34%, 16-wave 1-site Dirichlet tiles of k >= 3
// 7-12%; 1-site periodic tiles lose 10-20% and 8-wave tiles are within 2%. The 2 x k slots of
// every wave's two edge rows must fit the LDS (k * waves * vec <= 160: 160 KiB).
template <bool P, bool S, int V, int W, int K>
constexpr bool tb3_one_exchange() {
    if (K * W * V > 160) return false;
    if (S) return true;
    if (K >= 3 && W == 16 && V == 2) return true;
    return !P && K >= 3 && W == 16 && V == 1;
}

template <typename T, bool P, bool S, int V, int R, int W, int K>
__global__ void __launch_bounds__(64 * W)
tb3d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
     int nz, int ny, int nx, float c0, float c1, int zch, int tiles_x, int tiles_y, int remap)
{
    const int L = tb_logical_tile(remap);
    const int rest = L / tiles_x;
    // store modes measured per k at 1024^3 (see StoreMode); stage pinning where it pays (see
    // tb3d_tile): periodic k = 4 tiles of 2- and 4-site lanes
    tb3d_tile<T, P, S, V, R, W, K, K == 4 ? kStorePlanes : kStoreMasked,
              P && !S && K == 4 && V >= 2, tb3_one_exchange<P, S, V, W, K>()>(
        u, o, b, nz, ny, nx, c0, c1, zch, L % tiles_x, rest % tiles_y, rest / tiles_y);
}

template <typename T, bool P, bool S, int V, int K, int W, bool EW = false>
__global__ void __launch_bounds__(64 * W)
tb2d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
     int ny, int nx, float c0, float c1, int ych, int tiles_x, int remap)
{
    const int L = tb_logical_tile(remap);
    tb2d_tile<T, P, S, V, K, kStoreAll, EW>(u, o, b, ny, nx, c0, c1, ych,
                                            (L % tiles_x) * W + (int)(threadIdx.x >> 6), L / tiles_x);
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

// Default lane width for a k-step tile on a row of nx sites, starting from the widest allowed
// v: a wave writes 64*v - 2*KV sites, so the share of launched sites that are output sites,
// nx / (tiles * 64 * v), can drop for wide lanes on narrow rows (nx = 128, k = 3: 2-site lanes
// need two tiles of 124 output sites, 50%; 1-site lanes three of 58, 67%). Narrower lanes are
// taken when they launch over 10% fewer idle sites.
template <int K>
static int tb_fit_vec(int v, int nx) {
    auto share = [&](int w) {
        const int kv = ((K + w - 1) / w) * w;
        const int ox = 64 * w - 2 * kv;
        const long tiles = (nx + ox - 1) / ox;
        return (double)nx / ((double)tiles * 64.0 * w);
    };
    int best = v;
    for (int w = v >> 1; w >= 1; w >>= 1)
        if (share(w) > 1.1 * share(best)) best = w;
    return best;
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
    int zch = a.tile.zch;
    if (zch <= 0) {
        const long res = cached_resident<tb3d<T, P, S, V, R, W, K>>(64 * W);
        // a chunk fills a 2k-plane pipeline; grids too small to fill the device take chunks
        // down to one plane (more workgroups, each with a shorter chain of dependent steps:
        // 64^3 tb2 / tb3 / tb4 29 / 24 / 33 -> 44 / 39 / 56 GLUPS)
        zch = fill_chunk(a.nz, tiles, res, 16 * K, 2 * K, 1);
    }
    const long groups = tiles * ((a.nz + zch - 1) / zch);
    if (groups > 0x7fffffffL) { ph3host::set_error("tb3d: grid too large"); return PH3_EINVAL; }
    if (planned(a, PH3_TB, K, V, R, W, zch, 64 * V, W * R, G::OX, OY, 64 * W, groups))
        return PH3_OK;
    tb3d<T, P, S, V, R, W, K><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.nz, a.ny, a.nx, a.c0, a.c1, zch, tiles_x, tiles_y,
        groups % 8 == 0 ? 1 : 0);
    PH3_LAUNCH_CHECK("tb3d launch");
    return PH3_OK;
}

template <typename T, bool P, bool S, int V, int K, int W, bool EW = false>
static int tb2_launch(const SweepArgs& a) {
    using G = TbGeom<V, K>;
    const int strips = (a.nx + G::OX - 1) / G::OX;
    const int tiles_x = (strips + W - 1) / W;
    int ych = a.tile.zch;
    if (ych <= 0) {
        ych = tb_chunk(a.ny, tiles_x, K, 0);
        const long res = cached_resident<tb2d<T, P, S, V, K, W, EW>>(64 * W);
        if ((long)tiles_x * ((a.ny + ych - 1) / ych) < res)   // fewer workgroups than fit
            ych = fill_chunk(a.ny, tiles_x, res, ych, 2 * K, 2 * K);   // (1 row: 2048^2 tb4 -9%)
    }
    const long groups = (long)tiles_x * ((a.ny + ych - 1) / ych);
    if (groups > 0x7fffffffL) { ph3host::set_error("tb2d: grid too large"); return PH3_EINVAL; }
    if (planned(a, PH3_TB, K, V, 0, W, ych, 64 * V, 1, G::OX, 1, 64 * W, groups)) return PH3_OK;
    tb2d<T, P, S, V, K, W, EW><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.ny, a.nx, a.c0, a.c1, ych, tiles_x,
        groups % 8 == 0 ? 1 : 0);
    PH3_LAUNCH_CHECK("tb2d launch");
    return PH3_OK;
}

// 3D shapes (vec, rows per wave, waves): default (2, 4, 16); (1, 8, 8) for odd widths.
template <typename T, bool P, bool S, int K>
static int tb3_shape(const SweepArgs& a) {
    // Defaults measured on MI355X at 1024^3 float32 (vec, rows, waves): k=2 (4, 2, 16)
    // 1100 GLUPS; k=3 (2, 4, 16) 1620 GLUPS; k=4 (1, 4, 16) (register-bound). With a source
    // the k source planes held per lane push 16-wave shapes of k >= 3 and 2-4 site lanes past
    // 128 registers (spills: 377 vs 698 GLUPS for k=3 at 768^3), so those take 8 waves.
    const int vmax = K == 2 ? 4 : (K == 3 ? 2 : 1);
    int v = a.tile.vec ? a.tile.vec : tb_fit_vec<K>(tb_pick_vec(a, (int)sizeof(T), vmax), a.nx);
    int r = a.tile.ty ? a.tile.ty : (v == 4 ? 2 : 4);
    int w = a.tile.waves ? a.tile.waves : (S && K >= 3 && v > 1 ? 8 : 16);
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
    // Default lane width: the k float32 levels of a lane's ring grow with it (k = 8: 58 VGPRs
    // at 2 sites, 114 at 4, i.e. half the resident waves), so k = 8 takes 2 sites, others 4.
    int v = a.tile.vec ? a.tile.vec : tb_fit_vec<K>(tb_pick_vec(a, (int)sizeof(T), K >= 8 ? 2 : 4), a.nx);
    const int w = a.tile.waves ? a.tile.waves : 4;
    if (w != 4) { ph3host::set_error("tb2d: waves %d not built (4)", w); return PH3_EINVAL; }
    // float32 rows whose width or alignment allows no vector lanes: 2-site lanes loading and
    // storing site by site, where a 1-site strip would write fewer than 90% of the columns a
    // 2-site one writes (k = 8: 48 of 64 vs 112 of 128)
    constexpr bool kEwPays = TbGeom<1, K>::OX * 128 * 10 < 9 * TbGeom<2, K>::OX * 64;
    if constexpr (sizeof(T) == 4 && kEwPays) {
        if (!a.tile.vec && v == 1 && tb_fit_vec<K>(2, a.nx) == 2)
            return tb2_launch<T, P, S, 2, K, 4, true>(a);
    }
    if (a.nx % v) { ph3host::set_error("tb2d: vec %d does not divide nx %d", v, a.nx); return PH3_EINVAL; }
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
    // the tiles store (and float32 tiles and sources load) through buffer resources over one
    // plane (3D) or row (2D); a float32 source plane is 4 bytes a site whatever the storage
    const size_t elem = a.storage == kF32 || a.b ? 4 : 2;
    const size_t span = (a.dims == 3 ? (size_t)a.ny : 1) * (size_t)a.nx * elem;
    if (span > kMaxBufferBytes) {
        ph3host::set_error("tb: a %s of %zu bytes exceeds the %zu-byte store window",
                           a.dims == 3 ? "plane" : "row", span, (size_t)kMaxBufferBytes);
        return PH3_EUNSUP;
    }
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
