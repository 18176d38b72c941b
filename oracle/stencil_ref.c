/* Scalar per-site restatement of the pyheat3d stencil update. TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/ to pin the NumPy oracle (oracle/numpy_ref.py) bit for bit, and never linked
 * into or called by the pyheat3d_amd product path.
 *
 * Canonical order (reference src/pyheat3d/problem.py:3-10, backends/numpy_ref.py:30-47):
 *     s   = (u[x-1] + u[x+1]);  s = s + (u[y-1] + u[y+1]);  [3D] s = s + (u[z-1] + u[z+1])
 *     new = (c0 * u) + (c1 * s);  [source] new = new + b
 * Each line is one correctly rounded IEEE operation; this file must be compiled with
 * -ffp-contract=off so no multiply-add is fused.  Dirichlet: the outer layer is copied, not
 * updated.  Periodic: every site is updated with wrapped neighbours.
 *
 * Storage codecs (reference backends/stencil.cu:16-30, backends/cupy_backend.py:103-112,
 * validation.py:157-165) are written from the IEEE-754 binary16 definition and the bfloat16
 * bit formula, both round to nearest, ties to even.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <math.h>

#define DEFINE_STEP(NAME, T)                                                                   \
void NAME(const T* u, T* o, int dims, int nz, int ny, int nx, double c0d, double c1d,         \
          const T* b, int periodic)                                                            \
{                                                                                              \
    const T c0 = (T)c0d, c1 = (T)c1d;                                                          \
    if (dims == 2) nz = 1;                                                                     \
    const ptrdiff_t P = (ptrdiff_t)ny * nx;                                                    \
    for (int k = 0; k < nz; ++k)                                                               \
        for (int j = 0; j < ny; ++j)                                                           \
            for (int i = 0; i < nx; ++i) {                                                     \
                const ptrdiff_t c = (ptrdiff_t)k * P + (ptrdiff_t)j * nx + i;                  \
                const int edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 ||             \
                                 (dims == 3 && (k == 0 || k == nz - 1));                       \
                if (!periodic && edge) { o[c] = u[c]; continue; }                              \
                const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;          \
                const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;          \
                const ptrdiff_t row = (ptrdiff_t)k * P;                                        \
                T s = u[row + (ptrdiff_t)j * nx + im] + u[row + (ptrdiff_t)j * nx + ip];       \
                T t = u[row + (ptrdiff_t)jm * nx + i] + u[row + (ptrdiff_t)jp * nx + i];       \
                s = s + t;                                                                     \
                if (dims == 3) {                                                               \
                    const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;      \
                    T z = u[(ptrdiff_t)km * P + (ptrdiff_t)j * nx + i] +                       \
                          u[(ptrdiff_t)kp * P + (ptrdiff_t)j * nx + i];                        \
                    s = s + z;                                                                 \
                }                                                                              \
                T a = c0 * u[c];                                                               \
                T m = c1 * s;                                                                  \
                T v = a + m;                                                                   \
                if (b) v = v + b[c];                                                           \
                o[c] = v;                                                                      \
            }                                                                                  \
}

DEFINE_STEP(ref_step_f32, float)
DEFINE_STEP(ref_step_f64, double)

/* n steps of ref_step_f32 with ping-pong; result in out. work must hold as many values. */
void ref_run_f32(const float* u0, float* out, float* work, int dims, int nz, int ny, int nx,
                 double c0, double c1, const float* b, int periodic, int steps)
{
    const size_t n = (size_t)(dims == 2 ? 1 : nz) * ny * nx;
    float *a = out, *o = work;
    memcpy(a, u0, n * sizeof(float));
    for (int s = 0; s < steps; ++s) {
        ref_step_f32(a, o, dims, nz, ny, nx, c0, c1, b, periodic);
        float* t = a; a = o; o = t;
    }
    if (a != out) memcpy(out, a, n * sizeof(float));
}

/* float32 -> IEEE binary16 bits, round to nearest even (overflow -> inf, NaN kept quiet). */
static uint16_t f32_to_f16_bits(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ex = (x >> 23) & 0xFFu;
    uint32_t man = x & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u | (man >> 13) : 0u));
    const int e = (int)ex - 127;                 /* unbiased exponent */
    if (e > 15) return (uint16_t)(sign | 0x7C00u);
    if (e >= -14) {                              /* normal binary16 */
        uint32_t h = ((uint32_t)(e + 15) << 10) | (man >> 13);
        const uint32_t rest = man & 0x1FFFu;     /* 13 dropped bits */
        if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;  /* carry may reach inf */
        return (uint16_t)(sign | h);
    }
    if (e < -25) return (uint16_t)sign;          /* below half the smallest subnormal */
    /* subnormal: value = m * 2^-24 with m = (1.man) * 2^(e+24) */
    man |= 0x800000u;
    const int shift = -e - 14 + 13;              /* 14..24 */
    uint32_t h = man >> shift;
    const uint32_t rest = man & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rest > half || (rest == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

static float f16_bits_to_f32(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    uint32_t x;
    if (ex == 0x1Fu) x = sign | 0x7F800000u | (man << 13);
    else if (ex) x = sign | ((ex + 112u) << 23) | (man << 13);
    else if (!man) x = sign;
    else {                                       /* subnormal: normalise */
        int e = -14;
        uint32_t m = man;
        while (!(m & 0x400u)) { m <<= 1; --e; }
        x = sign | ((uint32_t)(e + 127) << 23) | ((m & 0x3FFu) << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static uint16_t f32_to_bf16_bits(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint64_t w = x;
    return (uint16_t)((w + 0x7FFFu + ((w >> 16) & 1u)) >> 16);
}

/* storage: 1 = float16, 2 = bfloat16. Writes the rounded bits and the widened values. */
void ref_round_storage(const float* in, uint16_t* bits, float* back, size_t n, int storage)
{
    for (size_t i = 0; i < n; ++i) {
        uint16_t h;
        float v;
        if (storage == 1) { h = f32_to_f16_bits(in[i]); v = f16_bits_to_f32(h); }
        else {
            h = f32_to_bf16_bits(in[i]);
            const uint32_t x = (uint32_t)h << 16;
            memcpy(&v, &x, 4);
        }
        if (bits) bits[i] = h;
        if (back) back[i] = v;
    }
}

/* Sum of squares of b + c0*u + c1*S(u) - u over updated sites, every value widened to double
 * and each operation done in double (the Jacobi defect; no counterpart in the reference). */
double ref_residual_f64(const float* u, const float* b, int dims, int nz, int ny, int nx,
                        double c0, double c1, int periodic)
{
    if (dims == 2) nz = 1;
    const ptrdiff_t P = (ptrdiff_t)ny * nx;
    const double c0f = (double)(float)c0, c1f = (double)(float)c1;
    double acc = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 ||
                                 (dims == 3 && (k == 0 || k == nz - 1));
                if (!periodic && edge) continue;
                const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
                const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
                const ptrdiff_t row = (ptrdiff_t)k * P, c = row + (ptrdiff_t)j * nx + i;
                double s = (double)u[row + (ptrdiff_t)j * nx + im] + (double)u[row + (ptrdiff_t)j * nx + ip];
                s += (double)u[row + (ptrdiff_t)jm * nx + i] + (double)u[row + (ptrdiff_t)jp * nx + i];
                if (dims == 3) {
                    const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
                    s += (double)u[(ptrdiff_t)km * P + (ptrdiff_t)j * nx + i] +
                         (double)u[(ptrdiff_t)kp * P + (ptrdiff_t)j * nx + i];
                }
                double r = c0f * (double)u[c] + c1f * s - (double)u[c];
                if (b) r += (double)b[c];
                acc += r * r;
            }
    return acc;
}
