// Tile bodies of the march rungs, shared by the single-grid kernels (sweep_zmarch.hip) and the
// batched varlen kernels (batch.hip) so that both apply the same operations to every site.
#pragma once
#include "sweep_impl.h"
#include "vec_io.h"

namespace ph3 {

template <int V>
__device__ __forceinline__ void x_neighbours(const float (&c)[V], float hl, float hr, bool useL,
                                             bool useR, float (&xm)[V], float (&xp)[V]) {
    const float fromL = from_prev_lane(c[V - 1]);
    const float fromR = from_next_lane(c[0]);
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

// One workgroup tile (tx, ty) of z chunk cz. Called by the single-grid kernel below and by the
// batched varlen kernel, so a grid's sites get the same operations in both. Returns early only
// uniformly over the workgroup. A caller that runs several tiles in a row must put a barrier
// between them (the LDS slot is reused).
template <typename T, bool P, bool S, int V, int R, int W>
__device__ __forceinline__ void
zmarch3d_tile(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
              int nz, int ny, int nx, float c0, float c1, int zch, int tx, int ty, int cz, int nt)
{
    __shared__ float edge[2][W][2][64 * V];  // [parity][wave][first/last row][site]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int x = (tx * 64 + lane) * V;
    const int J0 = ty * W * R;          // first row of the workgroup tile
    const int j0 = J0 + wave * R;       // first row of this wave
    // Dirichlet marches every plane too (plane indices of loads clamped, planes 0 and nz - 1
    // not stored): clamping the march range to 1 .. nz-2 instead doubled the registers the
    // compiler kept live (262 vs 150 VGPRs for 4 rows x 4 waves), spilling the taller tiles.
    const int z0 = cz * zch, z1 = min(z0 + zch, nz);
    if (z0 >= z1) return;               // uniform over the workgroup

    const bool xin = x < nx;
    const int xl = xin ? x : nx - V;    // load column (in bounds for idle lanes)
    const bool useL = lane == 0;
    const bool useR = lane == 63 || x + V >= nx;
    const int hxl = P ? (xl == 0 ? nx - 1 : xl - 1) : max(xl - 1, 0);
    const int hxr = P ? (xl + V >= nx ? xl + V - nx : xl + V) : min(xl + V, nx - 1);
    const bool fixL = x == 0, fixR = x + V == nx;   // lanes holding a fixed x column
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
        const int zz = P ? (z < 0 ? z + nz : (z >= nz ? z - nz : z)) : min(max(z, 0), nz - 1);
        return (size_t)zz * plane;
    };

    // Plane queue in four register slots: at step k of an unrolled group of four planes, slot
    // k holds plane z-1, k+1 plane z, k+2 plane z+1 and k+3 is free (slots mod 4). Rotating
    // slot roles instead of values removes the per-plane register moves; the source
    // double-buffers the same way with the parity of k.
    // Loads run one plane ahead. Float32 rows land in the free slot directly. Everything that
    // needs a conversion (16-bit rows, halo values) stays in its stored format until the step
    // that uses it, so that no conversion right after a load makes the wave wait there.
    constexpr bool kDirect = sizeof(T) == 4;
    float q[4][R][V];
    Packed<T, V> ahead[R];       // 16-bit storage: rows of the next plane, still packed
    // Halo timing. The x halo scalars and the rows above / below the tile sit in cache lines
    // that the neighbouring tiles load as their own rows two planes ahead. kAhead2: halos are
    // loaded in that same step (slot (z - z0) & 3 holds plane z's), so the lines are still in
    // L2; 1024^3 float32, 4 rows x 8 waves: reads 1.108 -> 1.037x the minimal bytes, 703 ->
    // 731 GLUPS. It costs 2R + 2 registers per lane, which spills 8-row float32 tiles and
    // slowed 16-bit storage (bf16 1358 -> 1174 GLUPS), so those load halos one plane ahead
    // (slot parity k & 1).
    constexpr bool kAhead2 = kDirect && R <= 4;
    constexpr int HS = kAhead2 ? 4 : 2;
    T xhl[HS][R], xhr[HS][R];      // x halos (used by edge lanes only)
    Packed<T, V> ryu[HS], ryd[HS]; // memory halo rows (first / last wave only)
    float bs[2][R][V];           // source of plane z

    {
        const size_t pm = pidx(z0 - 1), p0 = pidx(z0), p1 = pidx(z0 + 1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            load_vec<T, V>(u, pm + roff[r] + xl, q[0][r]);
            load_vec<T, V>(u, p0 + roff[r] + xl, q[1][r]);
            if constexpr (kDirect) load_vec<T, V>(u, p1 + roff[r] + xl, q[2][r]);
            else ahead[r] = load_packed<T, V>(u, p1 + roff[r] + xl);
            xhl[0][r] = useL ? u[p0 + roff[r] + hxl] : T{};
            xhr[0][r] = useR ? u[p0 + roff[r] + hxr] : T{};
            if constexpr (kAhead2) {
                xhl[1][r] = useL ? u[p1 + roff[r] + hxl] : T{};
                xhr[1][r] = useR ? u[p1 + roff[r] + hxr] : T{};
            }
            if (S) load_vec<float, V>(b, p0 + roff[r] + xl, bs[0][r]);
        }
#pragma unroll
        for (int h = 0; h < HS; ++h) ryu[h] = ryd[h] = Packed<T, V>{};
        if (top) ryu[0] = load_packed<T, V>(u, p0 + hup + xl);
        if (bot) ryd[0] = load_packed<T, V>(u, p0 + hdn + xl);
        if constexpr (kAhead2) {
            if (top) ryu[1] = load_packed<T, V>(u, p1 + hup + xl);
            if (bot) ryd[1] = load_packed<T, V>(u, p1 + hdn + xl);
        }
    }

    for (int zb = z0; zb < z1; zb += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int z = zb + k;
            if (z >= z1) break;                  // uniform over the workgroup
            if constexpr (!kDirect) {
                // plane z+1, loaded during the previous step
#pragma unroll
                for (int r = 0; r < R; ++r) unpack<T, V>(ahead[r], q[(k + 2) & 3][r]);
            }
            const float (&cm)[R][V] = q[k & 3];
            const float (&cc)[R][V] = q[(k + 1) & 3];
            const float (&cn)[R][V] = q[(k + 2) & 3];
            const int cur = k & 1, nxt = cur ^ 1;
            // halo slots of plane z and of the plane whose halos this step loads
            const int hc = kAhead2 ? (k & 3) : cur, hn = kAhead2 ? ((k + 2) & 3) : nxt;

            // prefetch: plane z+2's rows, the halos of plane z+2 (kAhead2) or z+1, and the
            // source of plane z+1
            if (z + 1 < z1) {
                const size_t p1 = pidx(z + 1), p2 = pidx(z + 2), ph = kAhead2 ? p2 : p1;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (kDirect) load_vec<T, V>(u, p2 + roff[r] + xl, q[(k + 3) & 3][r]);
                    else ahead[r] = load_packed<T, V>(u, p2 + roff[r] + xl);
                    xhl[hn][r] = useL ? u[ph + roff[r] + hxl] : T{};
                    xhr[hn][r] = useR ? u[ph + roff[r] + hxr] : T{};
                    if (S) load_vec<float, V>(b, p1 + roff[r] + xl, bs[nxt][r]);
                }
                if (top) ryu[hn] = load_packed<T, V>(u, ph + hup + xl);
                if (bot) ryd[hn] = load_packed<T, V>(u, ph + hdn + xl);
            }

            // rows above / below the wave: memory halo rows for the first / last wave, else
            // the neighbouring waves' first / last rows of plane z exchanged through LDS
            // (LDS parity (z - z0) & 1 == k & 1 since groups start at multiples of four)
            float yu[V], yd[V];
            unpack<T, V>(ryu[hc], yu);
            unpack<T, V>(ryd[hc], yd);
            if (W > 1) {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    edge[cur][wave][0][lane * V + e] = cc[0][e];
                    edge[cur][wave][1][lane * V + e] = cc[R - 1][e];
                }
                __syncthreads();
                if constexpr (kDirect) {
                    if (!top) {
#pragma unroll
                        for (int e = 0; e < V; ++e) yu[e] = edge[cur][wave - 1][1][lane * V + e];
                    }
                    if (!bot) {
#pragma unroll
                        for (int e = 0; e < V; ++e) yd[e] = edge[cur][wave + 1][0][lane * V + e];
                    }
                } else {
                    // 16-bit storage measured faster with a read by every wave (its own slot
                    // at the tile edges) and a select instead of a branch
                    const int wu = top ? 0 : wave - 1, wd = bot ? W - 1 : wave + 1;
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        const float eu = edge[cur][wu][1][lane * V + e];
                        const float ed = edge[cur][wd][0][lane * V + e];
                        yu[e] = top ? yu[e] : eu;
                        yd[e] = bot ? yd[e] : ed;
                    }
                }
            }

            const size_t pz = (size_t)z * plane;
            const bool zin = P || (z > 0 && z < nz - 1);   // Dirichlet: fixed planes stay
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float xm[V], xp[V];
                x_neighbours<V>(cc[r], Codec<T>::load(xhl[hc][r]), Codec<T>::load(xhr[hc][r]),
                                useL, useR, xm, xp);
                float w[V];
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const float ym = r > 0 ? cc[r - 1][e] : yu[e];
                    const float yp = r < R - 1 ? cc[r + 1][e] : yd[e];
                    float s = xm[e] + xp[e];
                    s = s + (ym + yp);
                    s = s + (cm[r][e] + cn[r][e]);
                    w[e] = update(c0, c1, cc[r][e], s);
                    if (S) w[e] = w[e] + bs[cur][r][e];
                }
                if (xin && rstore[r] && zin) {
                    // Dirichlet: the fixed sites x = 0 and x = nx - 1 keep their value, which
                    // the whole-vector store writes back unchanged (no per-site branch)
                    if (!P) {
                        if (fixL) w[0] = cc[r][0];
                        if (fixR) w[V - 1] = cc[r][V - 1];
                    }
                    store_out<T, V>(o, pz + roff[r] + x, w, nt);
                }
            }
        }
    }
}

// 2D: wave = 64*V columns; window of R+2 rows (y0-1 .. y0+R) in registers; R outputs per step.
// Workgroup column group gx of row chunk cy; waves are independent (no LDS, no barrier).
template <typename T, bool P, bool S, int V, int R, int W>
__device__ __forceinline__ void
ymarch2d_tile(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
              int ny, int nx, float c0, float c1, int ych, int gx, int cy)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
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
    const bool fixL = x == 0, fixR = x + V == nx;

    auto ridx = [&](int y) -> size_t {
        int yy = y;
        if (P) { yy = y % ny; if (yy < 0) yy += ny; }
        else yy = min(max(y, 0), ny - 1);
        return (size_t)yy * nx;
    };

    // win[q] holds row y0 - 1 + q, q = 0 .. R+1. Prefetched 16-bit rows and all halo values
    // stay in their stored format until the window advances (a conversion right after a load
    // would make the wave wait for it there).
    constexpr bool kDirect = sizeof(T) == 4;
    float win[R + 2][V];
    T xhl[R], xhr[R];
    float bs[R][V];
#pragma unroll
    for (int q = 0; q < R + 2; ++q) load_vec<T, V>(u, ridx(y0 - 1 + q) + xl, win[q]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const size_t ro = ridx(y0 + r);
        xhl[r] = useL ? u[ro + hxl] : T{};
        xhr[r] = useR ? u[ro + hxr] : T{};
        if (S) load_vec<float, V>(b, ro + xl, bs[r]);
    }

    for (int y = y0; y < y1; y += R) {
        // prefetch the next window's R new rows (y+R+1 .. y+2R), halos and source of y+R ..
        float nw[R][V], nbs[R][V];
        Packed<T, V> nwp[R];
        T nxl[R], nxr[R];
        const bool more = y + R < y1;
        if (more) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (kDirect) load_vec<T, V>(u, ridx(y + R + 1 + r) + xl, nw[r]);
                else nwp[r] = load_packed<T, V>(u, ridx(y + R + 1 + r) + xl);
                const size_t ro = ridx(y + R + r);
                nxl[r] = useL ? u[ro + hxl] : T{};
                nxr[r] = useR ? u[ro + hxr] : T{};
                if (S) load_vec<float, V>(b, ro + xl, nbs[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int yr = y + r;
            float xm[V], xp[V];
            x_neighbours<V>(win[r + 1], Codec<T>::load(xhl[r]), Codec<T>::load(xhr[r]), useL,
                            useR, xm, xp);
            float w[V];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                float s = xm[e] + xp[e];
                s = s + (win[r][e] + win[r + 2][e]);
                w[e] = update(c0, c1, win[r + 1][e], s);
                if (S) w[e] = w[e] + bs[r][e];
            }
            if (xin && yr < y1) {
                if (!P) {   // fixed x columns keep their value (written back unchanged)
                    if (fixL) w[0] = win[r + 1][0];
                    if (fixR) w[V - 1] = win[r + 1][V - 1];
                }
                store_vec<T, V>(o, (size_t)yr * nx + x, w);
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
                if constexpr (kDirect) {
#pragma unroll
                    for (int e = 0; e < V; ++e) win[r + 2][e] = nw[r][e];
                } else {
                    unpack<T, V>(nwp[r], win[r + 2]);
                }
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (S) bs[r][e] = nbs[r][e];
            }
        }
    }
}

}  // namespace ph3
