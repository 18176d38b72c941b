// Rung "zmarch": 2.5D blocking (counterpart of reference backends/stencil.cu:95-161 zmarch3d and
// of the K = 1 "ymarch" case of tb2d, stencil.cu:283-385), redesigned for wave64.
//
// 3D. A workgroup of W waves owns an xy tile of 64*V sites (x) by W*R rows (y) and marches a
// chunk of z planes. Wave w holds rows w*R .. w*R+R-1 of the tile; lane l holds the V
// contiguous sites x0 + l*V .. x0 + l*V + V-1 of each of them, so a tile row is one
// 64*V*sizeof(T)-byte access per wave (16 B per lane when V*sizeof(T) = 16). Three planes
// (z-1, z, z+1) sit in registers and plane z+2 is prefetched one iteration ahead.
// Neighbours:
//   x: inside the lane, else from lane l-1 / l+1 by a DPP wave shift, else (wave edges,
//      grid edges) from a halo value loaded ahead (with the plane's own rows, two planes
//      ahead, for float32 tiles of up to 4 rows: the neighbouring tile loads those cache
//      lines in the same step; one plane ahead otherwise);
//   y: rows of the same lane; across waves, each wave publishes its first and last row of
//      plane z in a double-buffered LDS slot (one barrier per plane); only the rows just above
//      and below the whole workgroup tile are read from memory (by the first / last wave);
//   z: the register queue.
// Workgroups are mapped XCD-aware: consecutive logical tiles run on the same XCD, so the halo
// rows shared by y-neighbouring workgroups hit one L2.
//
// 2D. A wave owns 64*V columns and marches a chunk of rows, R output rows per iteration from
// a window of R+2 rows in registers (the same scheme with the y rows playing z).
#include "march_tiles.h"

namespace ph3 {

template <typename T, bool P, bool S, int V, int R, int W>
__global__ void __launch_bounds__(64 * W)
zmarch3d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
         int nz, int ny, int nx, float c0, float c1, int zch, int tiles_x, int tiles_y,
         int remap, int nt)
{
    const int L = logical_tile(remap);
    const int tx = L % tiles_x;
    const int rest = L / tiles_x;
    zmarch3d_tile<T, P, S, V, R, W>(u, o, b, nz, ny, nx, c0, c1, zch, tx, rest % tiles_y,
                                    rest / tiles_y, nt);
}

template <typename T, bool P, bool S, int V, int R, int W>
__global__ void __launch_bounds__(64 * W)
ymarch2d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
         int ny, int nx, float c0, float c1, int ych, int tiles_x, int remap)
{
    const int L = logical_tile(remap);
    ymarch2d_tile<T, P, S, V, R, W>(u, o, b, ny, nx, c0, c1, ych, L % tiles_x, L / tiles_x);
}

static int pick_vec(const SweepArgs& a, int elem_bytes) {
    // widest lane vector up to 16 bytes (4 float32 or 8 16-bit sites) that divides nx and
    // keeps every row aligned
    const uintptr_t addr = (uintptr_t)a.u | (uintptr_t)a.o;
    for (int v = 16 / elem_bytes; v > 1; v >>= 1) {
        const bool aligned = (addr % (size_t)(v * elem_bytes)) == 0 &&
                             (!a.b || ((uintptr_t)a.b % (size_t)(v * 4)) == 0);
        if (a.nx % v == 0 && aligned) return v;
    }
    return 1;
}

// March chunk: enough workgroups for ~target per device, chunks of at least min_chunk.
static int pick_chunk(int n_march, long tiles, int min_chunk, int target_groups) {
    long want = (target_groups + tiles - 1) / tiles;
    if (want < 1) want = 1;
    int ch = (int)((n_march + want - 1) / want);
    if (ch < min_chunk) ch = min_chunk;
    return ch;
}

template <typename T, bool P, bool S, int V, int R, int W>
static int zmarch_launch3(const SweepArgs& a) {
    const int tiles_x = (a.nx + 64 * V - 1) / (64 * V);
    const int tiles_y = (a.ny + R * W - 1) / (R * W);
    const long tiles = (long)tiles_x * tiles_y;
    int zch = a.tile.zch;
    if (zch <= 0) {
        const long res = cached_resident<zmarch3d<T, P, S, V, R, W>>(64 * W);
        zch = fill_chunk(a.nz, tiles, res, 32, 2, 4);
    }
    const int chunks = (a.nz + zch - 1) / zch;
    const long groups = tiles * chunks;
    if (groups > 0x7fffffffL) { ph3host::set_error("zmarch: grid too large"); return PH3_EINVAL; }
    const int remap = (groups % 8) == 0 ? 1 : 0;
    const int nt = (a.tile.flags & PH3_TILE_NT_STORE) ? 1 : 0;
    if (planned(a, PH3_ZMARCH, 1, V, R, W, zch, 64 * V, R * W, 64 * V, R * W, 64 * W, groups))
        return PH3_OK;
    zmarch3d<T, P, S, V, R, W><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.nz, a.ny, a.nx, a.c0, a.c1, zch, tiles_x, tiles_y, remap,
        nt);
    PH3_LAUNCH_CHECK("zmarch3d launch");
    return PH3_OK;
}

template <typename T, bool P, bool S, int V, int R, int W>
static int ymarch_launch2(const SweepArgs& a) {
    const int strips = (a.nx + 64 * V - 1) / (64 * V);
    const int tiles_x = (strips + W - 1) / W;
    int ych = a.tile.zch;
    if (ych <= 0) {
        ych = pick_chunk(a.ny, tiles_x, 32, 2048);
        const long res = cached_resident<ymarch2d<T, P, S, V, R, W>>(64 * W);
        if ((long)tiles_x * ((a.ny + ych - 1) / ych) < res)   // fewer workgroups than fit
            ych = fill_chunk(a.ny, tiles_x, res, ych, 2, 4);
    }
    const int chunks = (a.ny + ych - 1) / ych;
    const long groups = (long)tiles_x * chunks;
    const int remap = (groups % 8) == 0 ? 1 : 0;
    if (planned(a, PH3_ZMARCH, 1, V, R, W, ych, 64 * V, 1, 64 * V, 1, 64 * W, groups))
        return PH3_OK;
    ymarch2d<T, P, S, V, R, W><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.ny, a.nx, a.c0, a.c1, ych, tiles_x, remap);
    PH3_LAUNCH_CHECK("ymarch2d launch");
    return PH3_OK;
}

// (rows per wave, waves per workgroup) shapes built for the 3D kernel
template <typename T, bool P, bool S, int V>
static int zmarch_shape3(const SweepArgs& a) {
    // Defaults measured on MI355X at 1024^3 (rows per wave, waves): float32 with 16-byte lanes
    // (4, 8) 731 GLUPS (halos loaded two planes ahead; (8, 8), which has no registers for
    // that, 716); 16-bit storage with 16-byte lanes (4, 4) 1300 GLUPS; narrower lanes (1, 16)
    // or (4, 4).
    int dr = 4, dw = 4;
    if (sizeof(T) == 4 && V == 4) { dr = 4; dw = 8; }
    else if ((sizeof(T) == 4 && V == 2) || (sizeof(T) == 2 && V == 4)) { dr = 1; dw = 16; }
    const int R = a.tile.ty ? a.tile.ty : dr;
    const int W = a.tile.waves ? a.tile.waves : dw;
#define PH3_ZM(r, w) if (R == r && W == w) return zmarch_launch3<T, P, S, V, r, w>(a)
    PH3_ZM(1, 16); PH3_ZM(2, 8); PH3_ZM(2, 4); PH3_ZM(4, 4); PH3_ZM(4, 8); PH3_ZM(8, 2);
    PH3_ZM(4, 2); PH3_ZM(8, 1); PH3_ZM(8, 8);
#undef PH3_ZM
    ph3host::set_error("zmarch3d: (rows %d, waves %d) not built; built: (1,16) (2,8) (2,4) "
                       "(4,4) (4,8) (8,2) (4,2) (8,1) (8,8)", R, W);
    return PH3_EINVAL;
}

template <typename T, bool P, bool S, int V>
static int ymarch_shape2(const SweepArgs& a) {
    const int R = a.tile.ty ? a.tile.ty : 4;
    switch (R) {
        case 1: return ymarch_launch2<T, P, S, V, 1, 4>(a);
        case 2: return ymarch_launch2<T, P, S, V, 2, 4>(a);
        case 4: return ymarch_launch2<T, P, S, V, 4, 4>(a);
        case 8: return ymarch_launch2<T, P, S, V, 8, 4>(a);
    }
    ph3host::set_error("ymarch2d: rows per step %d not in {1, 2, 4, 8}", R);
    return PH3_EINVAL;
}

template <typename T, bool P, bool S, int V>
static int zmarch_dims(const SweepArgs& a) {
    return a.dims == 3 ? zmarch_shape3<T, P, S, V>(a) : ymarch_shape2<T, P, S, V>(a);
}

int launch_zmarch(const SweepArgs& a) {
    if (a.tile.flags & ~PH3_TILE_NT_STORE) {
        ph3host::set_error("zmarch: unknown tile flags %#x", a.tile.flags);
        return PH3_EINVAL;
    }
    PH3_DISPATCH_TPS(a, {
        int v = a.tile.vec ? a.tile.vec : fit_vec(pick_vec(a, (int)sizeof(T)), a.nx);
        if (v < 1 || (v & (v - 1)) || v * (int)sizeof(T) > 16 || a.nx % v) {
            ph3host::set_error("zmarch: vec %d unsupported for nx %d and this storage", v, a.nx);
            return PH3_EINVAL;
        }
        if (v > 1 && (((uintptr_t)a.u | (uintptr_t)a.o) % (size_t)(v * sizeof(T)) ||
                      (a.b && (uintptr_t)a.b % (size_t)(v * 4)))) {
            ph3host::set_error("zmarch: buffers not aligned for vec %d", v);
            return PH3_EINVAL;
        }
        switch (v) {
            case 1: return zmarch_dims<T, P, S, 1>(a);
            case 2: return zmarch_dims<T, P, S, 2>(a);
            case 4: return zmarch_dims<T, P, S, 4>(a);
            case 8:
                if constexpr (sizeof(T) == 2) return zmarch_dims<T, P, S, 8>(a);
                break;
        }
        ph3host::set_error("zmarch: vec %d unsupported", v);
        return PH3_EINVAL;
    });
    return PH3_OK;
}

}  // namespace ph3
