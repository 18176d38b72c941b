// Rung "zmarch": 2.5D blocking (counterpart of reference backends/stencil.cu:95-161 zmarch3d and
// of the K = 1 "ymarch" case of tb2d, stencil.cu:283-385), redesigned for wave64.
//
// 3D. A workgroup of W waves owns an xy tile of 64*V sites (x) by W*R rows (y) and marches a
// chunk of z planes. Wave w holds rows w*R .. w*R+R-1 of the tile; lane l holds the V
// contiguous sites x0 + l*V .. x0 + l*V + V-1 of each of them, so a tile row is one
// 64*V*sizeof(T)-byte access per wave (16 B per lane when V*sizeof(T) = 16). Three planes
// (z-1, z, z+1) sit in registers and plane z+2 is prefetched one iteration ahead.
// Neighbours:
//   x: inside the lane, else from lane l-1 / l+1 by a cross-lane shuffle, else (wave edges,
//      grid edges) from a halo value loaded one plane ahead;
//   y: rows of the same lane; across waves, each wave publishes its first and last row of
//      plane z in a double-buffered LDS slot (one barrier per plane); only the rows just above
//      and below the whole workgroup tile are read from memory (by the first / last wave);
//   z: the register queue.
// Workgroups are mapped XCD-aware: consecutive logical tiles run on the same XCD, so the halo
// rows shared by y-neighbouring workgroups hit one L2.
//
// 2D. A wave owns 64*V columns and marches a chunk of rows, R output rows per iteration from
// a window of R+2 rows in registers (the same scheme with the y rows playing z).
#include "sweep_impl.h"
#include "vec_io.h"

namespace ph3 {

template <int V>
__device__ __forceinline__ void x_neighbours(const float (&c)[V], float hl, float hr, bool useL,
                                             bool useR, float (&xm)[V], float (&xp)[V]) {
    const float fromL = __shfl_up(c[V - 1], 1);
    const float fromR = __shfl_down(c[0], 1);
#pragma unroll
    for (int e = 0; e < V; ++e) {
        xm[e] = e > 0 ? c[e - 1] : (useL ? hl : fromL);
        xp[e] = e < V - 1 ? c[e + 1] : (useR ? hr : fromR);
    }
}

// Logical tile index of this workgroup: when the grid is a multiple of 8, workgroup L (which
// the dispatcher places on XCD L % 8) takes logical tile (L % 8) * (G / 8) + L / 8, so every
// XCD runs one contiguous range of logical tiles.
__device__ __forceinline__ int logical_tile(int remap) {
    const int L = blockIdx.x;
    if (!remap) return L;
    const int per = gridDim.x >> 3;
    return (L & 7) * per + (L >> 3);
}

template <typename T, int V>
__device__ __forceinline__ void store_out(T* __restrict__ o, size_t at, const float (&w)[V],
                                          bool nt) {
    if constexpr (V * sizeof(T) == 16 && sizeof(T) == 4) {
        if (nt) {
            typedef float f32x4 __attribute__((ext_vector_type(4)));
            f32x4 v = {w[0], w[1], w[2], w[3]};
            __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(o + at));
            return;
        }
    }
    store_vec<T, V>(o, at, w);
}

template <typename T, bool P, bool S, int V, int R, int W>
__global__ void __launch_bounds__(64 * W)
zmarch3d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
         int nz, int ny, int nx, float c0, float c1, int zch, int tiles_x, int tiles_y,
         int remap, int nt)
{
    __shared__ float edge[2][W][2][64 * V];  // [parity][wave][first/last row][site]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int L = logical_tile(remap);
    const int tx = L % tiles_x;
    const int rest = L / tiles_x;
    const int ty = rest % tiles_y;
    const int cz = rest / tiles_y;
    const int x = (tx * 64 + lane) * V;
    const int J0 = ty * W * R;          // first row of the workgroup tile
    const int j0 = J0 + wave * R;       // first row of this wave
    int z0 = cz * zch, z1 = min(z0 + zch, nz);
    if (!P) { z0 = max(z0, 1); z1 = min(z1, nz - 1); }
    if (z0 >= z1) return;               // uniform over the workgroup

    const bool xin = x < nx;
    const int xl = xin ? x : nx - V;    // load column (in bounds for idle lanes)
    const bool useL = lane == 0;
    const bool useR = lane == 63 || x + V >= nx;
    const int hxl = P ? (xl == 0 ? nx - 1 : xl - 1) : max(xl - 1, 0);
    const int hxr = P ? (xl + V >= nx ? xl + V - nx : xl + V) : min(xl + V, nx - 1);
    const size_t plane = (size_t)nx * ny;
    const bool top = wave == 0, bot = wave == W - 1;

    size_t roff[R];   // offset of row j0 + r within a plane (wrapped / clamped for loads)
    bool rstore[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        const int jl = P ? j % ny : min(j, ny - 1);
        roff[r] = (size_t)jl * nx;
        rstore[r] = j < ny && (P || (j > 0 && j < ny - 1));
    }
    // rows just above / below the workgroup tile (read by the first / last wave)
    const size_t hup = (size_t)(P ? (J0 + ny - 1) % ny : max(J0 - 1, 0)) * nx;
    const size_t hdn = (size_t)(P ? (J0 + W * R) % ny : min(J0 + W * R, ny - 1)) * nx;

    auto pidx = [&](int z) -> size_t {
        const int zz = P ? (z < 0 ? z + nz : (z >= nz ? z - nz : z)) : z;
        return (size_t)zz * plane;
    };

    float cm[R][V], cc[R][V], cn[R][V];
    float yu[V], yd[V];      // memory halo rows of plane z (first / last wave only)
    float xhl[R], xhr[R];    // x halos of plane z (used by edge lanes only)
    float bs[R][V];          // source of plane z

    {
        const size_t pm = pidx(z0 - 1), p0 = pidx(z0), p1 = pidx(z0 + 1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            load_vec<T, V>(u, pm + roff[r] + xl, cm[r]);
            load_vec<T, V>(u, p0 + roff[r] + xl, cc[r]);
            load_vec<T, V>(u, p1 + roff[r] + xl, cn[r]);
            xhl[r] = useL ? ld(u, p0 + roff[r] + hxl) : 0.f;
            xhr[r] = useR ? ld(u, p0 + roff[r] + hxr) : 0.f;
            if (S) load_vec<float, V>(b, p0 + roff[r] + xl, bs[r]);
        }
#pragma unroll
        for (int e = 0; e < V; ++e) yu[e] = yd[e] = 0.f;
        if (top) load_vec<T, V>(u, p0 + hup + xl, yu);
        if (bot) load_vec<T, V>(u, p0 + hdn + xl, yd);
    }

    for (int z = z0; z < z1; ++z) {
        // prefetch for plane z+1: its upper neighbour z+2, its halos and its source
        float nn[R][V], nyu[V], nyd[V], nxl[R], nxr[R], nbs[R][V];
        const bool more = z + 1 < z1;
        if (more) {
            const size_t p1 = pidx(z + 1), p2 = pidx(z + 2);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                load_vec<T, V>(u, p2 + roff[r] + xl, nn[r]);
                nxl[r] = useL ? ld(u, p1 + roff[r] + hxl) : 0.f;
                nxr[r] = useR ? ld(u, p1 + roff[r] + hxr) : 0.f;
                if (S) load_vec<float, V>(b, p1 + roff[r] + xl, nbs[r]);
            }
            if (top) load_vec<T, V>(u, p1 + hup + xl, nyu);
            if (bot) load_vec<T, V>(u, p1 + hdn + xl, nyd);
        }

        // exchange the first / last rows of plane z between the waves of the workgroup
        const int par = (z - z0) & 1;
        if (W > 1) {
#pragma unroll
            for (int e = 0; e < V; ++e) {
                edge[par][wave][0][lane * V + e] = cc[0][e];
                edge[par][wave][1][lane * V + e] = cc[R - 1][e];
            }
            __syncthreads();
            if (!top) {
#pragma unroll
                for (int e = 0; e < V; ++e) yu[e] = edge[par][wave - 1][1][lane * V + e];
            }
            if (!bot) {
#pragma unroll
                for (int e = 0; e < V; ++e) yd[e] = edge[par][wave + 1][0][lane * V + e];
            }
        }

        const size_t pz = (size_t)z * plane;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float xm[V], xp[V];
            x_neighbours<V>(cc[r], xhl[r], xhr[r], useL, useR, xm, xp);
            float w[V];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const float ym = r > 0 ? cc[r - 1][e] : yu[e];
                const float yp = r < R - 1 ? cc[r + 1][e] : yd[e];
                float s = xm[e] + xp[e];
                s = s + (ym + yp);
                s = s + (cm[r][e] + cn[r][e]);
                w[e] = update(c0, c1, cc[r][e], s);
                if (S) w[e] = w[e] + bs[r][e];
            }
            if (xin && rstore[r]) {
                const size_t at = pz + roff[r] + x;
                if (P || (x > 0 && x + V < nx)) {
                    store_out<T, V>(o, at, w, nt);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        if (x + e > 0 && x + e < nx - 1) st(o, at + e, w[e]);
                }
            }
        }

        if (more) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                xhl[r] = nxl[r];
                xhr[r] = nxr[r];
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    cm[r][e] = cc[r][e];
                    cc[r][e] = cn[r][e];
                    cn[r][e] = nn[r][e];
                    if (S) bs[r][e] = nbs[r][e];
                }
            }
            if (top) {
#pragma unroll
                for (int e = 0; e < V; ++e) yu[e] = nyu[e];
            }
            if (bot) {
#pragma unroll
                for (int e = 0; e < V; ++e) yd[e] = nyd[e];
            }
        }
    }
}

// 2D: wave = 64*V columns; window of R+2 rows (y0-1 .. y0+R) in registers; R outputs per step.
template <typename T, bool P, bool S, int V, int R, int W>
__global__ void __launch_bounds__(64 * W)
ymarch2d(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
         int ny, int nx, float c0, float c1, int ych, int tiles_x, int remap)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int L = logical_tile(remap);
    const int gx = L % (tiles_x);          // workgroup column group
    const int cy = L / (tiles_x);
    const int x = ((gx * W + wave) * 64 + lane) * V;
    if ((gx * W + wave) * 64 * V >= nx) return;
    int y0 = cy * ych, y1 = min(y0 + ych, ny);
    if (!P) { y0 = max(y0, 1); y1 = min(y1, ny - 1); }
    if (y0 >= y1) return;

    const bool xin = x < nx;
    const int xl = xin ? x : nx - V;
    const bool useL = lane == 0;
    const bool useR = lane == 63 || x + V >= nx;
    const int hxl = P ? (xl == 0 ? nx - 1 : xl - 1) : max(xl - 1, 0);
    const int hxr = P ? (xl + V >= nx ? xl + V - nx : xl + V) : min(xl + V, nx - 1);

    auto ridx = [&](int y) -> size_t {
        int yy = y;
        if (P) { yy = y % ny; if (yy < 0) yy += ny; }
        else yy = min(max(y, 0), ny - 1);
        return (size_t)yy * nx;
    };

    // win[q] holds row y0 - 1 + q, q = 0 .. R+1
    float win[R + 2][V];
    float xhl[R], xhr[R], bs[R][V];
#pragma unroll
    for (int q = 0; q < R + 2; ++q) load_vec<T, V>(u, ridx(y0 - 1 + q) + xl, win[q]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const size_t ro = ridx(y0 + r);
        xhl[r] = useL ? ld(u, ro + hxl) : 0.f;
        xhr[r] = useR ? ld(u, ro + hxr) : 0.f;
        if (S) load_vec<float, V>(b, ro + xl, bs[r]);
    }

    for (int y = y0; y < y1; y += R) {
        // prefetch the next window's R new rows (y+R+1 .. y+2R), halos and source of y+R ..
        float nw[R][V], nxl[R], nxr[R], nbs[R][V];
        const bool more = y + R < y1;
        if (more) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                load_vec<T, V>(u, ridx(y + R + 1 + r) + xl, nw[r]);
                const size_t ro = ridx(y + R + r);
                nxl[r] = useL ? ld(u, ro + hxl) : 0.f;
                nxr[r] = useR ? ld(u, ro + hxr) : 0.f;
                if (S) load_vec<float, V>(b, ro + xl, nbs[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int yr = y + r;
            float xm[V], xp[V];
            x_neighbours<V>(win[r + 1], xhl[r], xhr[r], useL, useR, xm, xp);
            float w[V];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                float s = xm[e] + xp[e];
                s = s + (win[r][e] + win[r + 2][e]);
                w[e] = update(c0, c1, win[r + 1][e], s);
                if (S) w[e] = w[e] + bs[r][e];
            }
            if (xin && yr < y1) {
                const size_t at = (size_t)yr * nx + x;
                if (P || (x > 0 && x + V < nx)) {
                    store_vec<T, V>(o, at, w);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        if (x + e > 0 && x + e < nx - 1) st(o, at + e, w[e]);
                }
            }
        }
        if (more) {
#pragma unroll
            for (int e = 0; e < V; ++e) {
                win[0][e] = win[R][e];
                win[1][e] = win[R + 1][e];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                xhl[r] = nxl[r];
                xhr[r] = nxr[r];
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    win[r + 2][e] = nw[r][e];
                    if (S) bs[r][e] = nbs[r][e];
                }
            }
        }
    }
}

static int pick_vec(const SweepArgs& a, int elem_bytes) {
    // widest access (<= 16 B) that divides nx and keeps every row aligned
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
    const int zch = a.tile.zch > 0 ? a.tile.zch : pick_chunk(a.nz, tiles, 64, 1024);
    const int chunks = (a.nz + zch - 1) / zch;
    const long groups = tiles * chunks;
    if (groups > 0x7fffffffL) { ph3host::set_error("zmarch: grid too large"); return PH3_EINVAL; }
    const int remap = (groups % 8) == 0 ? 1 : 0;
    const int nt = (a.tile.flags & PH3_TILE_NT_STORE) ? 1 : 0;
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
    const int ych = a.tile.zch > 0 ? a.tile.zch : pick_chunk(a.ny, tiles_x, 32, 2048);
    const int chunks = (a.ny + ych - 1) / ych;
    const long groups = (long)tiles_x * chunks;
    const int remap = (groups % 8) == 0 ? 1 : 0;
    ymarch2d<T, P, S, V, R, W><<<(unsigned)groups, 64 * W, 0, a.stream>>>(
        (const T*)a.u, (T*)a.o, a.b, a.ny, a.nx, a.c0, a.c1, ych, tiles_x, remap);
    PH3_LAUNCH_CHECK("ymarch2d launch");
    return PH3_OK;
}

// (rows per wave, waves per workgroup) shapes built for the 3D kernel
template <typename T, bool P, bool S, int V>
static int zmarch_shape3(const SweepArgs& a) {
    const int R = a.tile.ty ? a.tile.ty : 4;
    const int W = a.tile.waves ? a.tile.waves : 4;
#define PH3_ZM(r, w) if (R == r && W == w) return zmarch_launch3<T, P, S, V, r, w>(a)
    PH3_ZM(1, 16); PH3_ZM(2, 8); PH3_ZM(2, 4); PH3_ZM(4, 4); PH3_ZM(4, 8); PH3_ZM(8, 2);
    PH3_ZM(4, 2); PH3_ZM(8, 1);
#undef PH3_ZM
    ph3host::set_error("zmarch3d: (rows %d, waves %d) not built; built: (1,16) (2,8) (2,4) "
                       "(4,4) (4,8) (8,2) (4,2) (8,1)", R, W);
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
        int v = a.tile.vec ? a.tile.vec : pick_vec(a, (int)sizeof(T));
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
