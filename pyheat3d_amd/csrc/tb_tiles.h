// Tile bodies of the temporal-blocking rung, shared by the single-grid kernels
// (sweep_tb.hip) and the batched varlen kernels (batch.hip).
#pragma once
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

// How a tile stores its output rows (measured on MI355X; the choice never changes the bits):
//  kStoreMasked  global stores behind exec-masked branches (lanes and planes that do not write
//                skip them). The compiler cannot count stores behind a branch, so the wait for
//                the next plane's prefetch also waits for them. Best for the batched varlen
//                passes, whose short tiles pay for the entry wait of the other modes (bench
//                batch mix: 970 vs 815-945 solves/s).
//  kStoreAll     a raw buffer store of every row on every step; lanes and planes that must not
//                write get the drop offset or a 0-byte buffer. The waits count the stores
//                exactly. 2D single grids at 16384^2: tb8 4047 -> 4245 GLUPS (vec 2), 2771 ->
//                3187 (vec 1); 3D tb2 / tb3 no better.
//  kStorePlanes  buffer stores on the planes that are stored (a uniform branch), drop offsets
//                for lanes. 3D tb4 at 1024^3: 1747 -> 1812 GLUPS periodic, 1537 -> 1597
//                Dirichlet.
// The buffer-store modes make the wave wait for the tile's first prefetch before its loop (see
// arrived()), so that the loop entry and the back edge look alike to the wait counting.
enum StoreMode : int { kStoreMasked = 0, kStoreAll = 1, kStorePlanes = 2 };

// One overlapped tile (tx, ty) of z chunk cz, shared with the batched varlen kernels. Every
// wave reaches every barrier (no early return). A caller running several tiles in a row must
// put a barrier between them (the LDS slot is reused).
//
// PIN: keep each stage's arithmetic before the next stage's LDS exchange. The exchanges of a
// plane step read only planes of the previous step, so without it the compiler hoists all k
// exchanges (and their barriers) to the top of the step and runs the k stages' arithmetic
// after them, holding the up / down rows of every stage at once. Pinned, the 2-site 4 x 8
// periodic tile of k = 4 drops from 132 to 125 registers (two workgroups per CU instead of
// one) and the 16-wave tiles of k = 4 stop spilling. Measured at 1024^3 periodic float32
// (GLUPS, best chunk): (vec 2, rows 4, waves 8) 1654 -> 1930, (2, 4, 16) 1366 -> 1578,
// (4, 2, 16) 799 -> 1231; bfloat16 (2, 4, 8) 1770 -> 2070. 1-site lanes lose 7-15% and k = 2,
// 3 tiles up to 15% (no occupancy to gain, the exchange latency is exposed instead), so those
// stay unpinned. Dirichlet tiles keep this order anyway (the per-stage select is a branch).
//
// X1: exchange the edge rows of all k stages at the top of the plane step, behind one barrier
// (k slots per parity), instead of one exchange and one barrier per stage. Each stage still
// reads its own slot just before its arithmetic. Costs k times the LDS of one exchange.
template <typename T, bool P, bool S, int V, int R, int W, int K, int SM = kStoreMasked,
          bool PIN = false, bool X1 = false>
__device__ __forceinline__ void
tb3d_tile(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
          int nz, int ny, int nx, float c0, float c1, int zch, int tx, int ty, int cz)
{
    using G = TbGeom<V, K>;
    constexpr int OY = W * R - 2 * K;  // output rows per tile
    constexpr int NXS = X1 ? K : 1;   // exchange slots per parity
    __shared__ float edge[2][NXS][W][2][64 * V];
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int z0 = cz * zch, z1 = min(z0 + zch, nz);
    const int gx = tx * G::OX - G::KV + lane * V;   // first x site of this lane (may be outside)
    const int gy0 = ty * OY - K + wave * R;         // first row of this wave (may be outside)
    const int lx = vec_col<V>(gx, nx, P);           // vector-aligned: nx % V == 0, gx % V == 0
    const size_t plane = (size_t)nx * ny;

    uint32_t roff[R];          // row offsets in a plane (a plane has fewer than 2^31 sites)
    uint32_t fixmask = 0;      // bit r*V+e: site (row r, element e) is a fixed Dirichlet site
    uint32_t wmask = 0;        // bit r*V+e: site is written (inside output region and grid)
    // Lanes are whole aligned V-site vectors and the row conditions are uniform per row, so a
    // row's V bits of wmask are all set or all clear. Fixed sites are written too: they keep
    // their value through every stage, so the store writes back what the grid already holds.
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gy = gy0 + r;
        roff[r] = (uint32_t)wrap_or_clamp(gy, ny, P) * (uint32_t)nx;
        const int row = wave * R + r;
        const bool row_out = row >= K && row < W * R - K && gy >= 0 && gy < ny;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int x = gx + e;
            const int col = lane * V + e;
            const bool fixed = !P && (x <= 0 || x >= nx - 1 || gy <= 0 || gy >= ny - 1);
            if (fixed) fixmask |= 1u << (r * V + e);
            if (row_out && col >= G::KV && col < 64 * V - G::KV && x >= 0 && x < nx)
                wmask |= 1u << (r * V + e);
        }
    }
    // Dirichlet: whether any lane of this wave holds a fixed site (uniform over the wave)
    const bool wave_fix = !P && __any(fixmask != 0u);
    // float32 rows and the float32 source are read through a buffer resource over the plane
    // (base in scalar registers), at byte offsets fixed for the whole march; the stores' byte
    // offsets are fixed too (lanes and rows that never write hold the drop offset). pbytes: one
    // plane of stored values. (Splitting each offset into a lane part and a wave-uniform row
    // part in a scalar register cut the registers of the Dirichlet tb4 tile from 131 to 117,
    // two workgroups per CU, but measured slower: 1024^3 periodic 1795 -> 1615 GLUPS,
    // Dirichlet 1579 -> 1279; only 512^3 Dirichlet gained, 1124 -> 1287.)
    constexpr bool kDirect = sizeof(T) == 4;   // float32 storage: rows read through buffers
    // Lanes and rows past the last site the tile's outputs depend on (k beyond its last output
    // column / row; the grid's far edge cuts the last tile of a row or column short) load
    // nothing: their float32 loads take the drop offset, which returns zeros and fetches no
    // bytes. Those zeros reach only sites that depend on them, none of which is stored.
    const int xneed = min(tx * G::OX + G::OX, nx) + K;   // first x site no output depends on
    const int yneed = min(ty * OY + OY, ny) + K;
    const bool lane_need = gx < xneed;
    const uint32_t pbytes = (uint32_t)(plane * sizeof(T));
    uint32_t loff[R], soff[R];
    constexpr uint32_t fullv = (1u << V) - 1u;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool need = !kDirect || (lane_need && gy0 + r < yneed);
        loff[r] = need ? (roff[r] + (uint32_t)lx) * (uint32_t)sizeof(T) : kDropOffset;
        soff[r] = ((wmask >> (r * V)) & fullv) == fullv ? (roff[r] + (uint32_t)gx) * (uint32_t)sizeof(T)
                                                       : kDropOffset;
    }

    // Plane queues in three register slots per level: at march step i (i = iz - z0 + K) level
    // s holds planes c-1, c, c+1 in slots i, i+1, i+2 (mod 3). A new value (the prefetched
    // plane at level 0, the stage result at level s >= 1) overwrites slot i+2, whose plane c-2
    // is no longer needed. Rotating slot roles over an unrolled group of three steps replaces
    // the per-step register moves of a shifted queue.
    float q[K][3][R][V];
    // the next plane, loaded one step ahead: float32 as values, 16-bit storage still packed
    // (a conversion right after the load would make the wave wait for it there)
    float pref[R][V];
    Packed<T, V> prefp[R];
    // float32 source: at march step iz, bq[s-1] holds plane iz-s (used by stage s); the planes
    // shift one slot per step and plane iz is prefetched into bnext, so each plane of b is read
    // once per pass instead of once per stage
    float bq[S ? K : 1][R][V];
    float bnext[R][V];
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < V; ++e) q[s][p][r][e] = 0.f;
    if constexpr (S) {
#pragma unroll
        for (int s = 0; s < K; ++s)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < V; ++e) bq[s][r][e] = 0.f;
    }
    {
        const size_t zo = (size_t)wrap_or_clamp(z0 - K, nz, P) * plane;
        const size_t zb = (size_t)wrap_or_clamp(z0 - K - 1, nz, P) * plane;
        const __amdgpu_buffer_rsrc_t ub = buffer_over(u + zo, pbytes);
        const __amdgpu_buffer_rsrc_t bb =
            buffer_over(b + (S ? zb : 0), S ? (uint32_t)(plane * sizeof(float)) : 0u);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (kDirect) load_vec_buf<V>(ub, loff[r], pref[r]);
            else prefp[r] = load_packed<T, V>(u, zo + roff[r] + lx);
            if constexpr (S)
                load_vec_buf<V>(bb, kDirect ? loff[r] : (roff[r] + (uint32_t)lx) * 4u, bnext[r]);
        }
        if constexpr (SM != kStoreMasked) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (kDirect) arrived(pref[r]);
                else arrived(prefp[r]);
                if constexpr (S) arrived(bnext[r]);
            }
        }
    }
    int par = 0;   // LDS slot parity, flipped once per stage (X1: once per plane step)
    for (int ib = z0 - K; ib < z1 + K; ib += 3) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int iz = ib + k;
        if (iz >= z1 + K) break;                      // uniform over the workgroup
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (kDirect) {
#pragma unroll
                for (int e = 0; e < V; ++e) q[0][(k + 2) % 3][r][e] = pref[r][e];
            } else {
                unpack<T, V>(prefp[r], q[0][(k + 2) % 3][r]);
            }
        }
        if (iz + 1 < z1 + K) {
            const size_t zo = (size_t)wrap_or_clamp(iz + 1, nz, P) * plane;
            const __amdgpu_buffer_rsrc_t ub = buffer_over(u + zo, pbytes);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (kDirect) load_vec_buf<V>(ub, loff[r], pref[r]);
                else prefp[r] = load_packed<T, V>(u, zo + roff[r] + lx);
            }
        }
        if constexpr (S) {
#pragma unroll
            for (int s = K - 1; s > 0; --s)
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < V; ++e) bq[s][r][e] = bq[s - 1][r][e];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int e = 0; e < V; ++e) bq[0][r][e] = bnext[r][e];
            const size_t zb = (size_t)wrap_or_clamp(iz, nz, P) * plane;
            const __amdgpu_buffer_rsrc_t bb = buffer_over(b + zb, (uint32_t)(plane * sizeof(float)));
#pragma unroll
            for (int r = 0; r < R; ++r)
                load_vec_buf<V>(bb, kDirect ? loff[r] : (roff[r] + (uint32_t)lx) * 4u, bnext[r]);
        }
        float res[R][V];
        if constexpr (X1 && W > 1) {
#pragma unroll
            for (int s = 1; s <= K; ++s)
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    edge[par][s - 1][wave][0][lane * V + e] = q[s - 1][(k + 1) % 3][0][e];
                    edge[par][s - 1][wave][1][lane * V + e] = q[s - 1][(k + 1) % 3][R - 1][e];
                }
            __syncthreads();
        }
#pragma unroll
        for (int s = 1; s <= K; ++s) {
            const int c = iz - s;                       // plane computed by this stage
            const bool zfixed = !P && (c <= 0 || c >= nz - 1);
            const float (&qm)[R][V] = q[s - 1][k];
            const float (&ctr)[R][V] = q[s - 1][(k + 1) % 3];
            const float (&qp)[R][V] = q[s - 1][(k + 2) % 3];
            float (&dst)[R][V] = s < K ? q[s < K ? s : 0][(k + 2) % 3] : res;
            // rows of the neighbouring waves (level s-1, plane c)
            float up[V], dn[V];
            if (W > 1) {
                if constexpr (!X1) {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        edge[par][0][wave][0][lane * V + e] = ctr[0][e];
                        edge[par][0][wave][1][lane * V + e] = ctr[R - 1][e];
                    }
                    __syncthreads();
                }
                // the first and last wave read their own rows: the tile's outer k rows are
                // never stored and a value there reaches only rows within the light cone of
                // the error, so any finite value does; an unconditional read costs no select
                // (per stage 4 moves and 4 branches fewer in the 2-site tiles)
                const int wu = wave > 0 ? wave - 1 : wave, wd = wave < W - 1 ? wave + 1 : wave;
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    up[e] = edge[par][X1 ? s - 1 : 0][wu][1][lane * V + e];
                    dn[e] = edge[par][X1 ? s - 1 : 0][wd][0][lane * V + e];
                }
                if constexpr (!X1) par ^= 1;
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) { up[e] = ctr[0][e]; dn[e] = ctr[R - 1][e]; }
            }
            const float (&bsrc)[R][V] = bq[S ? s - 1 : 0];   // source of plane c
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float fromL = from_prev_lane(ctr[r][V - 1]);
                const float fromR = from_next_lane(ctr[r][0]);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const float xm = e > 0 ? ctr[r][e - 1] : fromL;
                    const float xp = e < V - 1 ? ctr[r][e + 1] : fromR;
                    const float ym = r > 0 ? ctr[r - 1][e] : up[e];
                    const float yp = r < R - 1 ? ctr[r + 1][e] : dn[e];
                    float sum = xm + xp;
                    sum = sum + (ym + yp);
                    sum = sum + (qm[r][e] + qp[r][e]);
                    float v = update(c0, c1, ctr[r][e], sum);
                    if (S) v = v + bsrc[r][e];
                    dst[r][e] = v;
                }
            }
            // Dirichlet: fixed sites keep their value; only a fixed plane or a wave that holds
            // a fixed site runs the selects
            if (!P && (zfixed || wave_fix)) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        if (zfixed || ((fixmask >> (r * V + e)) & 1u)) dst[r][e] = ctr[r][e];
            }
            if constexpr (PIN) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int e = 0; e < V; ++e) asm volatile("" : "+v"(dst[r][e]));
            }
        }
        if constexpr (X1) par ^= 1;
        const int c = iz - K;
        const bool zok = c >= z0 && c < z1 && (P || (c > 0 && c < nz - 1));
        constexpr uint32_t full = (1u << V) - 1u;
        if constexpr (SM == kStoreMasked) {
            if (zok && wmask) {
                const size_t zo = (size_t)c * plane;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (((wmask >> (r * V)) & full) == full)
                        store_vec<T, V>(o, zo + roff[r] + gx, res[r]);
            }
        } else if (SM == kStoreAll || zok) {
            // kStoreAll: every step issues the same R stores (a plane outside the chunk, or a
            // fixed one, gets a 0-byte buffer); lanes outside the output region get the drop
            // offset
            const __amdgpu_buffer_rsrc_t ob = buffer_over(
                o + (zok ? (size_t)c * plane : 0), zok ? pbytes : 0u);
#pragma unroll
            for (int r = 0; r < R; ++r) store_vec_buf<T, V>(ob, soff[r], res[r]);
        }
    }
    }
}

// Strip `strip` (64*V extended columns) of row chunk cy, marched by one wave; waves are
// independent (no LDS, no barrier). SM: the StoreMode of tb3d_tile (kStorePlanes skips the
// stores of rows that are not stored). EW (float32 only): each of a lane's V sites is loaded
// and stored on its own (V dword accesses instead of one V-wide access), so V-site lanes run
// on rows of any width and alignment (odd widths would otherwise take 1-site lanes, whose
// strips write 48 of 64 columns at k = 8 instead of 112 of 128).
template <typename T, bool P, bool S, int V, int K, int SM = kStoreMasked, bool EW = false>
__device__ __forceinline__ void
tb2d_tile(const T* __restrict__ u, T* __restrict__ o, const float* __restrict__ b,
          int ny, int nx, float c0, float c1, int ych, int strip, int cy)
{
    static_assert(!EW || sizeof(T) == 4, "element-wise lanes are float32");
    using G = TbGeom<V, K>;
    const int lane = threadIdx.x & 63;
    const int y0 = cy * ych, y1 = min(y0 + ych, ny);
    const int gx = strip * G::OX - G::KV + lane * V;
    if (strip * G::OX - G::KV >= nx) return;          // whole wave past the grid (no barriers)
    const int lx = EW ? 0 : vec_col<V>(gx, nx, P);
    int lxe[V];                                       // EW: column of each site, wrapped / clamped
#pragma unroll
    for (int e = 0; e < V; ++e) lxe[e] = EW ? wrap_or_clamp(gx + e, nx, P) : 0;
    uint32_t fixmask = 0, wmask = 0;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int x = gx + e;
        const int col = lane * V + e;
        const bool fixed = !P && (x <= 0 || x >= nx - 1);
        if (fixed) fixmask |= 1u << e;
        // fixed sites included (written back unchanged), as in tb3d_tile; without EW a lane's
        // V bits are all set or all clear (nx % V == 0)
        if (col >= G::KV && col < 64 * V - G::KV && x >= 0 && x < nx) wmask |= 1u << e;
    }
    // one row of stored values (or of the source) at row offset ro, as a lane's V floats
    auto load_row = [&](const T* __restrict__ p, size_t ro, float (&out)[V]) {
        if constexpr (EW) {
#pragma unroll
            for (int e = 0; e < V; ++e) out[e] = Codec<T>::load(p[ro + lxe[e]]);
        } else {
            load_vec<T, V>(p, ro + lx, out);
        }
    };
    auto load_src = [&](size_t ro, float (&out)[V]) {
        if constexpr (EW) {
#pragma unroll
            for (int e = 0; e < V; ++e) out[e] = b[ro + lxe[e]];
        } else {
            load_vec<float, V>(b, ro + lx, out);
        }
    };
    const bool wave_fix = !P && __any(fixmask != 0u);
    // row queues rotated by slot as in tb3d_tile (slots i, i+1, i+2 mod 3 hold rows c-1, c,
    // c+1 of level s at march step i)
    float q[K][3][V];
    constexpr bool kDirect = sizeof(T) == 4;   // 16-bit rows stay packed until used
    float pref[V];
    Packed<T, V> prefp;
    float bq[S ? K : 1][V], bnext[V];   // source rows iy-1 .. iy-K, as in tb3d_tile
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int e = 0; e < V; ++e) q[s][p][e] = 0.f;
    if constexpr (S) {
#pragma unroll
        for (int s = 0; s < K; ++s)
#pragma unroll
            for (int e = 0; e < V; ++e) bq[s][e] = 0.f;
        load_src((size_t)wrap_or_clamp(y0 - K - 1, ny, P) * nx, bnext);
    }
    if constexpr (kDirect) load_row(u, (size_t)wrap_or_clamp(y0 - K, ny, P) * nx, pref);
    else prefp = load_packed<T, V>(u, (size_t)wrap_or_clamp(y0 - K, ny, P) * nx + lx);
    if constexpr (SM != kStoreMasked) {   // loop entry like the back edge (see arrived())
        if constexpr (kDirect) arrived(pref);
        else arrived(prefp);
        if constexpr (S) arrived(bnext);
    }
    for (int ib = y0 - K; ib < y1 + K; ib += 3) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int iy = ib + k;
        if (iy >= y1 + K) break;
        if constexpr (kDirect) {
#pragma unroll
            for (int e = 0; e < V; ++e) q[0][(k + 2) % 3][e] = pref[e];
        } else {
            unpack<T, V>(prefp, q[0][(k + 2) % 3]);
        }
        if (iy + 1 < y1 + K) {
            const size_t ro = (size_t)wrap_or_clamp(iy + 1, ny, P) * nx;
            if constexpr (kDirect) load_row(u, ro, pref);
            else prefp = load_packed<T, V>(u, ro + lx);
        }
        if constexpr (S) {
#pragma unroll
            for (int s = K - 1; s > 0; --s)
#pragma unroll
                for (int e = 0; e < V; ++e) bq[s][e] = bq[s - 1][e];
#pragma unroll
            for (int e = 0; e < V; ++e) bq[0][e] = bnext[e];
            load_src((size_t)wrap_or_clamp(iy, ny, P) * nx, bnext);
        }
        float res[V];
#pragma unroll
        for (int s = 1; s <= K; ++s) {
            const int c = iy - s;
            const bool yfixed = !P && (c <= 0 || c >= ny - 1);
            const float (&qm)[V] = q[s - 1][k];
            const float (&ctr)[V] = q[s - 1][(k + 1) % 3];
            const float (&qp)[V] = q[s - 1][(k + 2) % 3];
            float (&dst)[V] = s < K ? q[s < K ? s : 0][(k + 2) % 3] : res;
            const float (&bsrc)[V] = bq[S ? s - 1 : 0];   // source of row c
            const float fromL = from_prev_lane(ctr[V - 1]);
            const float fromR = from_next_lane(ctr[0]);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const float xm = e > 0 ? ctr[e - 1] : fromL;
                const float xp = e < V - 1 ? ctr[e + 1] : fromR;
                float sum = xm + xp;
                sum = sum + (qm[e] + qp[e]);
                float v = update(c0, c1, ctr[e], sum);
                if (S) v = v + bsrc[e];
                dst[e] = v;
            }
            if (!P && (yfixed || wave_fix)) {   // fixed sites keep their value
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (yfixed || ((fixmask >> e) & 1u)) dst[e] = ctr[e];
            }
        }
        const int c = iy - K;
        const bool yok = c >= y0 && c < y1 && (P || (c > 0 && c < ny - 1));
        if constexpr (SM == kStoreMasked) {
            if constexpr (EW) {
#pragma unroll
                for (int e = 0; e < V; ++e)
                    if (yok && ((wmask >> e) & 1u)) st(o, (size_t)c * nx + gx + e, res[e]);
            } else {
                if (wmask && yok) store_vec<T, V>(o, (size_t)c * nx + gx, res);
            }
        } else if (SM == kStoreAll || yok) {
            // kStoreAll: a store on every step, dropped outside the output rows and lanes
            const __amdgpu_buffer_rsrc_t ob =
                buffer_over(o + (yok ? (size_t)c * nx : 0), yok ? (uint32_t)(nx * sizeof(T)) : 0u);
            if constexpr (EW) {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const float one[1] = {res[e]};
                    store_vec_buf<T, 1>(ob, ((wmask >> e) & 1u) ? (uint32_t)((gx + e) * sizeof(T))
                                                                : kDropOffset, one);
                }
            } else {
                store_vec_buf<T, V>(ob, wmask ? (uint32_t)(gx * sizeof(T)) : kDropOffset, res);
            }
        }
    }
    }
}

}  // namespace ph3
