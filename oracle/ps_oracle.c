/*
 * ps_oracle.c -- CPU oracle (restatement) of Puresoft3D's per-frame raster path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
 * library never links, loads or calls this file.
 *
 * Arithmetic contract: IEEE FP32 in source order, no FMA contraction
 * (build with -ffp-contract=off -fno-fast-math), correctly rounded division
 * and sqrtf.  rcpps / rsqrtss of the reference are replaced by 1.0f/x and
 * 1.0f/sqrtf(x) (the reference's bits for those depend on the CPU vendor).
 *
 * Parity status per part (what pins each function):
 *  - or_push_triangle (rasterizer.cpp:95-283): pinned bit-exact against the
 *    reference's own rasterizer.cpp compiled into oracle/_ref, via committed
 *    fixtures (tests/golden/raster_golden.npz) and live random cross-checks.
 *  - or_interp_setup / pixel loop (interp.cpp:26-92, fragthrd.cpp:208-252):
 *    pinned by closed forms (affine depth, cf2*(1/cf2) ~ 1, exact replay of
 *    the serial sums with the documented operation order) and tiny brute-force
 *    scenes; interp.cpp itself is unbuildable here (needs atlbase.h + 32-bit
 *    __asm), so parity there is against this restatement + those checks.
 *  - or_is_back_face (vertthrd.cpp:56-65): pinned against sign(cross.z) closed
 *    form and the NaN/inf special cases.
 *  - samplers (samplr2d.cpp, samplrcube.cpp, fbo.cpp:552-594): pinned by the
 *    closed-form index formulas, exhaustive over small textures and the face table.
 *  - shading (tex1light1.cpp etc. and the processors added here): pinned by
 *    closed forms on flat quads (light and camera at set angles to the
 *    normal, distinct B/G/R texels, turned TBN frames, turned models under a
 *    real perspective, skybox faces), and by an FP64 variant of the same
 *    formulas (shade_fp64=1) within 1 LSB.  shade_fp64=2 keeps FP32 but takes
 *    rcpps / rsqrtss where the reference's DEF01-DEF04 call div_3_4 and
 *    norm_3_4, to measure the colour against the reference's own arithmetic;
 *    3 and 4 put those reciprocals at the two ends of their error bound.
 *  - bilinear / trilinear / mip chain / log2 approximation (not in the
 *    reference): pinned by textbook identities (bilinear of affine data is
 *    exact, mip of a constant is constant, log2 of powers of two is exact,
 *    FP64 log2 within 1e-6).
 */
#define _GNU_SOURCE
#include "ps_oracle.h"

#include <emmintrin.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_NV 20

/* ------------------------------------------------------------------ */
/* x86 conversions                                                     */
/* ------------------------------------------------------------------ */

/* (int)f as compiled for x86 (cvttss2si): out-of-range and NaN give INT_MIN. */
static inline int cvt_i32(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

/* (unsigned int)f with gcc x86-64: 64-bit cvttss2si, low 32 bits kept
 * (samplr2d.cpp:6-7).  Out of int64 range / NaN -> 0x8000000000000000 -> 0. */
static inline uint32_t cvt_u32(float f)
{
    if (f >= -9223372036854775808.0f && f < 9223372036854775808.0f)
        return (uint32_t)(uint64_t)(int64_t)f;
    return 0u;
}

static inline float bits2f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t f2bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* ------------------------------------------------------------------ */
/* mcemaths restatement (src/mcemath/vector.cpp, matrix.cpp)           */
/* ------------------------------------------------------------------ */

/* mcemaths_transform_m4v4, matrix.cpp:515-557: ((x*C0 + y*C1) + z*C2) + w*C3 */
static void m4v4(float* r, const float* m, const float* v)
{
    float x = v[0], y = v[1], z = v[2], w = v[3];
    for (int i = 0; i < 4; i++) {
        float a = x * m[i], b = y * m[4 + i], c = z * m[8 + i], d = w * m[12 + i];
        r[i] = ((a + b) + c) + d;
    }
}

/* mcemaths_transform_m4m4, matrix.cpp:588-701: column j of d = l * (column j of r) */
static void m4m4(float* d, const float* l, const float* r)
{
    for (int j = 0; j < 4; j++) m4v4(d + 4 * j, l, r + 4 * j);
}

/* mcemaths_dot_3_4, vector.cpp:85-112: (a0b0 + a1b1) + (a2b2 + a3b3) */
static inline float dot4(const float* a, const float* b)
{
    float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2], p3 = a[3] * b[3];
    return (p0 + p1) + (p2 + p3);
}

/* mcemaths_norm_3_4, vector.cpp:225-250, rsqrtss -> 1/sqrtf */
static inline void norm4(float* a)
{
    float s = dot4(a, a);
    float r = 1.0f / sqrtf(s);
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}

/* div_3_4 and norm_3_4 with the reference's own approximations, as the x86
 * cores this runs on compute them: rcpps of the broadcast divisor
 * (vector.cpp:158-169) and rsqrtss of the haddps-summed squares
 * (vector.cpp:225-250).  Their bits are vendor-specific (|rel err| <=
 * 1.5 * 2^-12 by the SDM), so they serve to measure how far the exact
 * replacement above moves the colour, not as the parity reference. */
static inline void div4_x86(float* a, float f)
{
    float r = _mm_cvtss_f32(_mm_rcp_ps(_mm_set1_ps(f)));
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}
static inline void norm4_x86(float* a)
{
    float r = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(dot4(a, a))));
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}

/* the same with the SDM's error bound taken at one end: the exact
 * reciprocal (square root) scaled by 1 + g_approx_rel, for modes 3 (+) and
 * 4 (-) of shade_fp64 */
static const double kApproxRel = 1.5 / 4096.0;
static inline void div4_bound(float* a, float f, double rel)
{
    float r = (float)((1.0 / (double)f) * (1.0 + rel));
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}
static inline void norm4_bound(float* a, double rel)
{
    float r = (float)((1.0 / sqrt((double)dot4(a, a))) * (1.0 + rel));
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}
static inline void div4_hi(float* a, float f) { div4_bound(a, f, kApproxRel); }
static inline void div4_lo(float* a, float f) { div4_bound(a, f, -kApproxRel); }
static inline void norm4_hi(float* a) { norm4_bound(a, kApproxRel); }
static inline void norm4_lo(float* a) { norm4_bound(a, -kApproxRel); }

/* mcemaths_len_3_4, vector.cpp:196-223 */
static inline float len4(const float* a) { return sqrtf(dot4(a, a)); }

/* mcemaths_div_3_4, vector.cpp:158-169, rcpps -> 1/x */
static inline void div4(float* a, float f)
{
    float r = 1.0f / f;
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}

/* mcemaths_clamp_3_4, vector.cpp:370-383: maxps then minps (NaN -> lo) */
static inline float clampf_sse(float a, float lo, float hi)
{
    float t = (a > lo) ? a : lo;
    return (t < hi) ? t : hi;
}

/* mcemaths_cross_3, vector.cpp:114-144 */
static inline void cross4(float* r, const float* a, const float* b)
{
    float r0 = a[1] * b[2] - a[2] * b[1];
    float r1 = a[2] * b[0] - a[0] * b[2];
    float r2 = a[0] * b[1] - a[1] * b[0];
    float r3 = a[3] * b[3] - a[3] * b[3];
    r[0] = r0; r[1] = r1; r[2] = r2; r[3] = r3;
}

/* opt_pow, proc.h:73-86 */
static inline float opt_powf(float x, unsigned n)
{
    float pw = 1.0f;
    while (n > 0) {
        if (n & 1) pw *= x;
        x *= x;
        n >>= 1;
    }
    return pw;
}
static inline double opt_powd(double x, unsigned n)
{
    double pw = 1.0;
    while (n > 0) {
        if (n & 1) pw *= x;
        x *= x;
        n >>= 1;
    }
    return pw;
}

/* 3-lane helpers for the two processors without a reference counterpart */
static inline float dot3(const float* a, const float* b)
{
    float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2];
    return (p0 + p1) + p2;
}
static inline void norm3(float* a)
{
    float r = 1.0f / sqrtf(dot3(a, a));
    a[0] *= r; a[1] *= r; a[2] *= r;
}

uint32_t or_tri_hash(uint32_t t)
{
    /* 32-bit integer mixer (murmur3 finalizer) */
    t ^= t >> 16; t *= 0x85ebca6bu; t ^= t >> 13; t *= 0xc2b2ae35u; t ^= t >> 16;
    return t;
}

/* ------------------------------------------------------------------ */
/* textures / FBO addressing (fbo.cpp)                                 */
/* ------------------------------------------------------------------ */

static inline int lvl_w(const or_texture* t, int l) { int w = t->width >> l; return w < 1 ? 1 : w; }
static inline int lvl_h(const or_texture* t, int l) { int h = t->height >> l; return h < 1 ? 1 : h; }
static inline int lvl_scan(const or_texture* t, int l) { return l == 0 ? t->scanline : lvl_w(t, l) * t->elem_len; }

static size_t layer_bytes(const or_texture* t)
{
    size_t s = (size_t)t->scanline * (size_t)t->height;
    for (int l = 1; l < t->levels; l++) s += (size_t)lvl_w(t, l) * lvl_h(t, l) * t->elem_len;
    return s;
}

static const uint8_t* level_ptr(const or_texture* t, int layer, int level)
{
    size_t off = layer_bytes(t) * (size_t)layer;
    if (level > 0) {
        off += (size_t)t->scanline * (size_t)t->height;
        for (int l = 1; l < level; l++) off += (size_t)lvl_w(t, l) * lvl_h(t, l) * t->elem_len;
    }
    return t->data + off;
}

/* PuresoftFBO::clampCoord, fbo.cpp:552-594 */
static inline void clamp_coord(int wrap, int max_row, int max_col, int* row, int* col)
{
    if (!wrap) {
        if (*row > max_row) *row = max_row; else if (*row < 0) *row = 0;
        if (*col > max_col) *col = max_col; else if (*col < 0) *col = 0;
    } else {
        /* row % m_maxRow; a one-texel dimension would divide by zero in the
           reference, here it addresses texel 0 */
        if (max_row > 0) { *row = *row % max_row; if (*row < 0) *row += max_row; } else *row = 0;
        if (max_col > 0) { *col = *col % max_col; if (*col < 0) *col += max_col; } else *col = 0;
    }
}

static inline void fetch4(const or_texture* t, int layer, int level, int row, int col, uint8_t* out)
{
    int w = lvl_w(t, level), h = lvl_h(t, level);
    clamp_coord(t->wrap, h - 1, w - 1, &row, &col);
    const uint8_t* p = level_ptr(t, layer, level) + (size_t)row * lvl_scan(t, level) + (size_t)col * 4;
    memcpy(out, p, 4);
}

/* PuresoftSamplerProjection::get, samplrproj.cpp:3-37.  div_3_4 multiplies by
   the reciprocal of w (vector.cpp:158-169, rcpps as 1/x); the row offsets are
   POISSON_DISK_X[i]/500 and the column offsets POISSON_DISK_Y[i]/500, as there;
   the taps convert with (unsigned int) and read through directRead4
   (clampCoord, fbo.cpp:286-291), the four bytes taken as a float. */
float or_sample_projection(const or_texture* t, const float* proj)
{
    static const float PDX[4] = {-0.94201624f, 0.94558609f, -0.094184101f, 0.34495938f};
    static const float PDY[4] = {-0.39906216f, -0.76890725f, -0.92938870f, 0.29387760f};
    float r = 1.0f / proj[3];
    float tx = proj[0] * r, ty = proj[1] * r, tz = proj[2] * r;
    float f = 1.0f;
    for (int i = 0; i < 4; i++) {
        int row = (int)cvt_u32((float)t->height * (ty + PDX[i] / 500.0f));
        int col = (int)cvt_u32((float)t->width * (tx + PDY[i] / 500.0f));
        uint8_t b[4];
        fetch4(t, 0, 0, row, col, b);
        float d;
        memcpy(&d, b, 4);
        if (d < tz) f -= 0.2f;
    }
    return f;
}

/* PuresoftFBO::blend4, fbo.cpp:208-229: per byte, the saturating difference
   src - dst (subs_pu16) times the source alpha, low 16 bits (mullo_pi16),
   shifted right by 8, plus dst, packed with unsigned saturation */
void or_blend4(uint8_t* dst, const uint8_t* src)
{
    unsigned a = src[3];
    for (int k = 0; k < 4; k++) {
        unsigned s = src[k], d = dst[k];
        unsigned t = s > d ? s - d : 0u;
        unsigned v = (((t * a) & 0xffffu) >> 8) + d;
        dst[k] = (uint8_t)(v > 255u ? 255u : v);
    }
}

/* PuresoftSampler2D::get4, samplr2d.cpp:19-25 -> directRead4, fbo.cpp:286-291 */
void or_sample_nearest(const or_texture* t, int layer, float u, float v, uint8_t* out4)
{
    int row = (int)cvt_u32((float)t->height * v + 0.5f);
    int col = (int)cvt_u32((float)t->width * u + 0.5f);
    fetch4(t, layer, 0, row, col, out4);
}

/* texcoordFromDirection, samplrcube.cpp:29-97 (strict > comparisons) */
static int cube_face(const float* d, float* s, float* tt)
{
    float X = d[0], Y = d[1], Z = d[2];
    if (fabsf(X) > fabsf(Y)) {
        if (fabsf(X) > fabsf(Z)) {
            if (X > 0) { *s = (-Z / fabsf(X) + 1.0f) / 2.0f; *tt = (-Y / fabsf(X) + 1.0f) / 2.0f; return 0; }
            *s = (Z / fabsf(X) + 1.0f) / 2.0f; *tt = (-Y / fabsf(X) + 1.0f) / 2.0f; return 1;
        }
        if (Z > 0) { *s = (X / fabsf(Z) + 1.0f) / 2.0f; *tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; return 4; }
        *s = (-X / fabsf(Z) + 1.0f) / 2.0f; *tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; return 5;
    }
    if (fabsf(Z) > fabsf(Y)) {
        if (Z > 0) { *s = (X / fabsf(Z) + 1.0f) / 2.0f; *tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; return 4; }
        *s = (-X / fabsf(Z) + 1.0f) / 2.0f; *tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; return 5;
    }
    if (Z > 0) { *s = (X / fabsf(Y) + 1.0f) / 2.0f; *tt = (Z / fabsf(Y) + 1.0f) / 2.0f; return 2; }
    *s = (X / fabsf(Y) + 1.0f) / 2.0f; *tt = (-Z / fabsf(Y) + 1.0f) / 2.0f; return 3;
}

/* PuresoftSamplerCube::get4, samplrcube.cpp:119-127: row from S, column from T,
 * dimensions of layer 0 */
void or_sample_cube(const or_texture* t, const float* dir, uint8_t* out4)
{
    float s, tt;
    int face = cube_face(dir, &s, &tt);
    int row = (int)cvt_u32((float)t->height * s + 0.5f);
    int col = (int)cvt_u32((float)t->width * tt + 0.5f);
    if (face >= t->layers) face = 0;
    fetch4(t, face, 0, row, col, out4);
}

/* Bilinear filter (an addition: the reference has nearest only).
 * Texel j of a level covers u = j/W (the nearest sampler's rounding rule),
 * addressing follows clampCoord. */
void or_sample_bilinear(const or_texture* t, int layer, int level, float u, float v, float* out4)
{
    int w = lvl_w(t, level), h = lvl_h(t, level);
    float fx = (float)w * u, fy = (float)h * v;
    float x0 = floorf(fx), y0 = floorf(fy);
    float ax = fx - x0, ay = fy - y0;
    int ix = cvt_i32(x0), iy = cvt_i32(y0);
    int ix1 = (int)((unsigned)ix + 1u), iy1 = (int)((unsigned)iy + 1u);
    uint8_t a[4], b[4], c[4], d[4];
    fetch4(t, layer, level, iy, ix, a);
    fetch4(t, layer, level, iy, ix1, b);
    fetch4(t, layer, level, iy1, ix, c);
    fetch4(t, layer, level, iy1, ix1, d);
    for (int k = 0; k < 4; k++) {
        float fa = (float)a[k], fb = (float)b[k], fc = (float)c[k], fd = (float)d[k];
        float top = fa + (fb - fa) * ax;
        float bot = fc + (fd - fc) * ax;
        out4[k] = top + (bot - top) * ay;
    }
}

static void sample_bilinear_d(const or_texture* t, int layer, int level, double u, double v, double* out4)
{
    int w = lvl_w(t, level), h = lvl_h(t, level);
    double fx = (double)w * u, fy = (double)h * v;
    double x0 = floor(fx), y0 = floor(fy);
    double ax = fx - x0, ay = fy - y0;
    int ix = (x0 >= -2147483648.0 && x0 < 2147483648.0) ? (int)x0 : INT_MIN;
    int iy = (y0 >= -2147483648.0 && y0 < 2147483648.0) ? (int)y0 : INT_MIN;
    int ix1 = (int)((unsigned)ix + 1u), iy1 = (int)((unsigned)iy + 1u);
    uint8_t a[4], b[4], c[4], d[4];
    fetch4(t, layer, level, iy, ix, a);
    fetch4(t, layer, level, iy, ix1, b);
    fetch4(t, layer, level, iy1, ix, c);
    fetch4(t, layer, level, iy1, ix1, d);
    for (int k = 0; k < 4; k++) {
        double top = a[k] + (b[k] - (double)a[k]) * ax;
        double bot = c[k] + (d[k] - (double)c[k]) * ax;
        out4[k] = top + (bot - top) * ay;
    }
}

/* log2 by exponent split + atanh series (log m = 2 atanh((m-1)/(m+1))),
 * m reduced to [sqrt(1/2), sqrt(2)); truncation error < 1e-9. */
float or_log2_approx(float x)
{
    if (!(x > 0.0f)) return -INFINITY;
    if (x == INFINITY) return 128.0f;
    uint32_t b = f2bits(x);
    int e = (int)((b >> 23) & 0xffu);
    int bias = 127;
    if (e == 0) {               /* subnormal: scale by 2^23 */
        x = x * 8388608.0f;
        b = f2bits(x);
        e = (int)((b >> 23) & 0xffu);
        bias = 127 + 23;
    }
    int ex = e - bias;
    float m = bits2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; ex = ex + 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.32059889797532520f;
    p = 0.41219858311113240f + s2 * p;
    p = 0.57707801635558540f + s2 * p;
    p = 0.96179669392597560f + s2 * p;
    p = 2.88539008177792680f + s2 * p;
    return (float)ex + s * p;
}

/* mip level selection from the texel-space footprint (GL scale factor),
 * then trilinear blend of two bilinear levels */
static float lod_from_grad(const float* g, float u, float v, float r, int w0, int h0)
{
    /* g = {Qx, Qy, Ux, Uy, Vx, Vy}: screen gradients of rw, u*rw, v*rw */
    float dudx = (g[2] - u * g[0]) * r;
    float dudy = (g[3] - u * g[1]) * r;
    float dvdx = (g[4] - v * g[0]) * r;
    float dvdy = (g[5] - v * g[1]) * r;
    float sx = dudx * (float)w0, tx = dvdx * (float)h0;
    float sy = dudy * (float)w0, ty = dvdy * (float)h0;
    float px = sx * sx + tx * tx;
    float py = sy * sy + ty * ty;
    float rho2 = px > py ? px : py;
    if (!(rho2 > 0.0f)) return 0.0f;
    return 0.5f * or_log2_approx(rho2);
}

static double lod_from_grad_d(const float* g, double u, double v, double r, int w0, int h0)
{
    double dudx = (g[2] - u * g[0]) * r, dudy = (g[3] - u * g[1]) * r;
    double dvdx = (g[4] - v * g[0]) * r, dvdy = (g[5] - v * g[1]) * r;
    double sx = dudx * w0, tx = dvdx * h0, sy = dudy * w0, ty = dvdy * h0;
    double px = sx * sx + tx * tx, py = sy * sy + ty * ty;
    double rho2 = px > py ? px : py;
    if (!(rho2 > 0.0)) return 0.0;
    return 0.5 * log2(rho2);
}

static void sample_trilinear(const or_texture* t, float u, float v, float lod, float* out4)
{
    int maxl = t->levels - 1;
    if (!(lod > 0.0f) || maxl == 0) { or_sample_bilinear(t, 0, 0, u, v, out4); return; }
    if (!(lod < (float)maxl)) { or_sample_bilinear(t, 0, maxl, u, v, out4); return; }
    int l0 = (int)lod;
    float f = lod - (float)l0;
    float a[4], b[4];
    or_sample_bilinear(t, 0, l0, u, v, a);
    or_sample_bilinear(t, 0, l0 + 1, u, v, b);
    for (int k = 0; k < 4; k++) out4[k] = a[k] + (b[k] - a[k]) * f;
}

static void sample_trilinear_d(const or_texture* t, double u, double v, double lod, double* out4)
{
    int maxl = t->levels - 1;
    if (!(lod > 0.0) || maxl == 0) { sample_bilinear_d(t, 0, 0, u, v, out4); return; }
    if (!(lod < (double)maxl)) { sample_bilinear_d(t, 0, maxl, u, v, out4); return; }
    int l0 = (int)lod;
    double f = lod - l0;
    double a[4], b[4];
    sample_bilinear_d(t, 0, l0, u, v, a);
    sample_bilinear_d(t, 0, l0 + 1, u, v, b);
    for (int k = 0; k < 4; k++) out4[k] = a[k] + (b[k] - a[k]) * f;
}

void or_build_mips(uint8_t* data, int width, int height, int elem_len, int layers, int levels)
{
    or_texture t = {width, height, width * elem_len, elem_len, layers, 0, levels, data};
    for (int L = 0; L < layers; L++) {
        for (int l = 1; l < levels; l++) {
            int ws = lvl_w(&t, l - 1), hs = lvl_h(&t, l - 1), ss = lvl_scan(&t, l - 1);
            int wd = lvl_w(&t, l), hd = lvl_h(&t, l), sd = lvl_scan(&t, l);
            const uint8_t* src = level_ptr(&t, L, l - 1);
            uint8_t* dst = (uint8_t*)level_ptr(&t, L, l);
            for (int y = 0; y < hd; y++) {
                int y0 = 2 * y, y1 = 2 * y + 1 < hs ? 2 * y + 1 : hs - 1;
                for (int x = 0; x < wd; x++) {
                    int x0 = 2 * x, x1 = 2 * x + 1 < ws ? 2 * x + 1 : ws - 1;
                    for (int c = 0; c < elem_len; c++) {
                        int s = src[(size_t)y0 * ss + x0 * elem_len + c] + src[(size_t)y0 * ss + x1 * elem_len + c] +
                                src[(size_t)y1 * ss + x0 * elem_len + c] + src[(size_t)y1 * ss + x1 * elem_len + c];
                        dst[(size_t)y * sd + x * elem_len + c] = (uint8_t)((s + 2) >> 2);
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* rasterizer (rasterizer.cpp)                                         */
/* ------------------------------------------------------------------ */

typedef struct { int left, leftClamped, lv0, lv1, right, rightClamped, rv0, rv1; } or_row;

typedef struct { float dx, dy, x0, y0; int i0, i1; } or_edge;

/* LineSegment ctor, rasterizer.cpp:99-112 */
static void edge_init(or_edge* e, const float* vx, const float* vy, int i0, int i1)
{
    e->x0 = vx[i0];
    e->y0 = vy[i0];
    e->dx = vx[i1] - e->x0;
    e->dy = vy[i1] - e->y0;
    e->i0 = i0;
    e->i1 = i1;
    if (fabsf(e->dy) < 0.000001f) e->dy = FLT_MAX;
}

/* LineSegment::operator(), rasterizer.cpp:114-117 */
static inline float edge_eval(const or_edge* e, float y) { return e->dx * (y - e->y0) / e->dy + e->x0; }

typedef struct {
    int W, H, hw, hh;
    float vx[3], vy[3];
    int first, last;
    or_row* rows;
} or_raster;

/* processTriangle, rasterizer.cpp:142-167 */
static void process_half(or_raster* R, const or_edge* L, const or_edge* Rt, float yMin, float yMax)
{
    int iy = yMin < 0 ? 0 : cvt_i32(yMin);
    int lastRowIdx = R->H - 1;
    int iyMax = yMax > (float)lastRowIdx ? lastRowIdx : cvt_i32(yMax);
    for (; iy <= iyMax; iy++) {
        or_row* row = &R->rows[iy];
        row->left = cvt_i32(edge_eval(L, (float)iy) + 0.5f);
        row->leftClamped = row->left < 0 ? 0 : row->left;
        row->right = cvt_i32(edge_eval(Rt, (float)iy) + 0.5f);
        row->rightClamped = row->right >= R->W ? R->W - 1 : row->right;
        row->lv0 = L->i0; row->lv1 = L->i1;
        row->rv0 = Rt->i0; row->rv1 = Rt->i1;
    }
}

/* processStandingTriangle, rasterizer.cpp:170-189 */
static void process_standing(or_raster* R, int top, int bottom, int third)
{
    or_edge tt, tb, bt;
    edge_init(&tt, R->vx, R->vy, top, third);
    edge_init(&tb, R->vx, R->vy, top, bottom);
    edge_init(&bt, R->vx, R->vy, third, bottom);
    if (edge_eval(&tb, R->vy[third]) > edge_eval(&tt, R->vy[third])) {
        process_half(R, &tt, &tb, R->vy[third], R->vy[top]);
        process_half(R, &bt, &tb, R->vy[bottom], R->vy[third]);
    } else {
        process_half(R, &tb, &tt, R->vy[third], R->vy[top]);
        process_half(R, &tb, &bt, R->vy[bottom], R->vy[third]);
    }
}

/* pushVertex, rasterizer.cpp:122-139 */
static void push_vertex(or_raster* R, int i, const float* v)
{
    R->vy[i] = ((float)R->hh * v[1]) + (float)R->hh;
    R->vx[i] = ((float)R->hw * v[0]) + (float)R->hw;
    if (i == 0) R->first = R->last = cvt_i32(R->vy[0]);
    else if (R->vy[i] > (float)R->last) R->last = cvt_i32(R->vy[i]);
    else if (R->vy[i] < (float)R->first) R->first = cvt_i32(R->vy[i]);
}

/* pushTriangle, rasterizer.cpp:191-283 */
static int push_triangle(or_raster* R, const float* v0, const float* v1, const float* v2)
{
    push_vertex(R, 0, v0);
    push_vertex(R, 1, v1);
    push_vertex(R, 2, v2);
    if (R->first >= R->H || R->last < 0) { R->first = 0; R->last = -1; }
    if (R->first < 0) R->first = 0;
    if (R->last >= R->H) R->last = R->H - 1;
    if (R->first == R->last) return 0;

    const float* vx = R->vx;
    const float* vy = R->vy;
    or_edge a, b;
    if (vy[0] == vy[1]) {
        if (vx[0] < vx[1]) { edge_init(&a, vx, vy, 0, 2); edge_init(&b, vx, vy, 1, 2); }
        else { edge_init(&a, vx, vy, 1, 2); edge_init(&b, vx, vy, 0, 2); }
        process_half(R, &a, &b, (float)R->first, (float)R->last);
    } else if (vy[0] == vy[2]) {
        if (vx[0] < vx[2]) { edge_init(&a, vx, vy, 0, 1); edge_init(&b, vx, vy, 2, 1); }
        else { edge_init(&a, vx, vy, 2, 1); edge_init(&b, vx, vy, 0, 1); }
        process_half(R, &a, &b, (float)R->first, (float)R->last);
    } else if (vy[1] == vy[2]) {
        if (vx[1] < vx[2]) { edge_init(&a, vx, vy, 1, 0); edge_init(&b, vx, vy, 2, 0); }
        else { edge_init(&a, vx, vy, 2, 0); edge_init(&b, vx, vy, 1, 0); }
        process_half(R, &a, &b, (float)R->first, (float)R->last);
    } else {
        int top, bottom, third = 2;
        if (vy[0] > vy[1]) { top = 0; bottom = 1; } else { top = 1; bottom = 0; }
        if (vy[top] < vy[third]) { int t = top; top = third; third = t; }
        else if (vy[bottom] > vy[third]) { int t = bottom; bottom = third; third = t; }
        process_standing(R, top, bottom, third);
    }
    return 1;
}

int or_push_triangle(int width, int height, const float* v0, const float* v1, const float* v2,
                     int* first_row, int* last_row, int* rows_out, float* verts_out)
{
    or_raster R;
    R.W = width; R.H = height; R.hw = width / 2; R.hh = height / 2;
    R.rows = (or_row*)rows_out;
    int ok = push_triangle(&R, v0, v1, v2);
    *first_row = R.first;
    *last_row = R.last;
    for (int i = 0; i < 3; i++) { verts_out[2 * i] = R.vx[i]; verts_out[2 * i + 1] = R.vy[i]; }
    return ok;
}

/* ------------------------------------------------------------------ */
/* interpolater (interp.cpp)                                           */
/* ------------------------------------------------------------------ */

/* lineSegmentlinearInterpolate, interp.cpp:151-160 */
static void line_contrib(const float* verts, int v1, int v2, float x, float y, float* c)
{
    float p1x = verts[2 * v1], p1y = verts[2 * v1 + 1];
    float p2x = verts[2 * v2], p2y = verts[2 * v2 + 1];
    float dx = p1x - p2x, dy = p1y - p2y;
    c[v1] = fabsf(dx) > fabsf(dy) ? ((x - p2x) / dx) : ((y - p2y) / dy);
    c[v2] = 1.0f - c[v1];
    c[3 - v1 - v2] = 0;
}

typedef struct {
    float ccL[4], ccR[4];   /* corrected contributions (interp.cpp:36-41) */
    int step_count;         /* right - left (wrapping int arithmetic)     */
    int skip;               /* leftClamped - left                         */
} or_span_geom;

static void span_geom(const float* verts, const float* rw4, int row, int left, int right, int leftClamped,
                      int lv0, int lv1, int rv0, int rv1, or_span_geom* g)
{
    float cl[4], cr[4];
    cl[3] = cr[3] = 0;
    line_contrib(verts, lv0, lv1, (float)left, (float)row, cl);
    line_contrib(verts, rv0, rv1, (float)right, (float)row, cr);
    for (int i = 0; i < 4; i++) { g->ccL[i] = cl[i] * rw4[i]; g->ccR[i] = cr[i] * rw4[i]; }
    g->step_count = (int)((unsigned)right - (unsigned)left);
    g->skip = (int)((unsigned)leftClamped - (unsigned)left);
}

/* interpolateStartAndStep, interp.cpp:26-80 (depth and correction factor part) */
static void span_vis(const or_span_geom* g, const float* z4, float* cf2s, float* cf2st, float* pzs, float* pzst)
{
    float rlen = 1.0f / (float)g->step_count;
    float pz0 = dot4(g->ccL, z4);
    float pz1 = dot4(g->ccR, z4);
    float pzstep = (pz1 - pz0) * rlen;
    /* haddps x2: (c0 + c1) + (c2 + c3), interp.cpp:55-66 */
    float c0 = (g->ccL[0] + g->ccL[1]) + (g->ccL[2] + g->ccL[3]);
    float c1 = (g->ccR[0] + g->ccR[1]) + (g->ccR[2] + g->ccR[3]);
    float cstep = (c1 - c0) * rlen;
    if (g->skip > 0) {
        c0 += cstep * (float)g->skip;
        pz0 += pzstep * (float)g->skip;
    }
    *cf2s = c0; *cf2st = cstep; *pzs = pz0; *pzst = pzstep;
}

void or_interp_setup(const float* verts6, const float* rw3, const float* z3, int row, int left, int right,
                     int lv0, int lv1, int rv0, int rv1, float* cf2_start, float* cf2_step, float* pz_start,
                     float* pz_step)
{
    float rw4[4] = {rw3[0], rw3[1], rw3[2], 0.0f};
    float z4[4] = {z3[0], z3[1], z3[2], 0.0f};
    or_span_geom g;
    int lc = left < 0 ? 0 : left;
    span_geom(verts6, rw4, row, left, right, lc, lv0, lv1, rv0, rv1, &g);
    span_vis(&g, z4, cf2_start, cf2_step, pz_start, pz_step);
}

/* IP built-ins (tex1light1.cpp:63-138 and siblings): identical lane-wise
 * operations for every float of the varying block. */
static void span_varyings(const or_span_geom* g, const float (*user)[OR_MAX_NV], int nv, float* start, float* step)
{
    float end[OR_MAX_NV];
    for (int i = 0; i < nv; i++) {
        /* interpolateByContributes: ((a0*c0) + (a1*c1)) + (a2*c2) */
        float a0 = user[0][i] * g->ccL[0], a1 = user[1][i] * g->ccL[1], a2 = user[2][i] * g->ccL[2];
        start[i] = (a0 + a1) + a2;
        float b0 = user[0][i] * g->ccR[0], b1 = user[1][i] * g->ccR[1], b2 = user[2][i] * g->ccR[2];
        end[i] = (b0 + b1) + b2;
    }
    /* calcStep: (end - start) * (stepCount == 0 ? 1 : 1/stepCount) */
    float rsc = g->step_count == 0 ? 1.0f : 1.0f / (float)g->step_count;
    for (int i = 0; i < nv; i++) step[i] = (end[i] - start[i]) * rsc;
    if (g->skip > 0) {
        /* stepForward(start, step, skip): n==1 add, else start + step*(float)n */
        if (g->skip == 1) {
            for (int i = 0; i < nv; i++) start[i] = start[i] + step[i];
        } else {
            float n = (float)g->skip;
            for (int i = 0; i < nv; i++) start[i] = start[i] + step[i] * n;
        }
    }
}

/* isBackFace, vertthrd.cpp:56-65 */
int or_is_back_face(const float* v0, const float* v1, const float* v2)
{
    float a[4], b[4], c[4];
    static const float test[4] = {0, 0, 1.0f, 0};
    for (int i = 0; i < 4; i++) { a[i] = v1[i] - v0[i]; b[i] = v2[i] - v1[i]; }
    cross4(c, a, b);
    norm4(c);
    return dot4(test, c) < 0;
}

/* ------------------------------------------------------------------ */
/* processors                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    int prog, nv, needs_grad;
    const float *PV, *M, *Mrot, *P, *V, *lightPos, *cameraPos, *lightDir;
    const or_texture *tex0, *tex1, *tex2;
    const float* shadowPV;  /* SHADOWED, CLOUD */
    unsigned spec_exp;
    float invView[16];  /* DEF04 */
    float PVM[16];      /* DEF05 */
} or_prog;

static int prog_nv(int prog)
{
    switch (prog) {
    case OR_PROG_DEF01: return 12;
    case OR_PROG_DEF02: return 12;
    case OR_PROG_DEF03: return 20;
    case OR_PROG_DEF04: return 4;
    case OR_PROG_DEF05: return 0;
    case OR_PROG_LAMBERT: return 5;
    case OR_PROG_BLINN: return 14;
    case OR_PROG_SHADOWED: return 18;
    case OR_PROG_CLOUD: return 14;
    case OR_PROG_CLOUD_SHADOW: return 2;
    default: return -1;
    }
}

static const or_texture* tex_from_uniform(const or_draw* d, int slot)
{
    const int* idx = (const int*)d->uniforms[slot];
    if (!idx || *idx < 0 || *idx >= d->n_textures) return NULL;
    const or_texture* t = &d->textures[*idx];
    if (t->elem_len != 4 || !t->data) return NULL;
    return t;
}

/* the preprocess() step of each programme: resolve uniform slots */
static int prog_prepare(const or_draw* d, or_prog* p)
{
    memset(p, 0, sizeof(*p));
    p->prog = d->prog;
    p->nv = prog_nv(d->prog);
    if (p->nv < 0) return -2;
    const void* const* u = d->uniforms;
    switch (d->prog) {
    case OR_PROG_DEF01: case OR_PROG_DEF02: case OR_PROG_DEF03:
        /* tex1light1.cpp:15-20,145-150; colr1light1.cpp; tex1bump1light1.cpp:15-20,172-178 */
        p->PV = u[3]; p->M = u[4]; p->Mrot = u[5]; p->lightPos = u[7]; p->cameraPos = u[8];
        if (!p->PV || !p->M || !p->Mrot || !p->lightPos || !p->cameraPos) return -3;
        if (d->prog != OR_PROG_DEF02) { p->tex0 = tex_from_uniform(d, 9); if (!p->tex0) return -4; }
        if (d->prog == OR_PROG_DEF03) { p->tex1 = tex_from_uniform(d, 10); if (!p->tex1) return -4; }
        break;
    case OR_PROG_DEF04: {
        /* skybox.cpp:17-21,122-125 */
        p->P = u[0]; p->V = u[1];
        if (!p->V) return -3;
        p->tex0 = tex_from_uniform(d, 2);
        if (!p->tex0) return -4;
        /* mat4cpy + mat4transpose + zero the last column (skybox.cpp:34-37) */
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) p->invView[c * 4 + r] = p->V[r * 4 + c];
        p->invView[12] = p->invView[13] = p->invView[14] = p->invView[15] = 0.0f;
        break;
    }
    case OR_PROG_DEF05: case OR_PROG_CLOUD_SHADOW:
        /* shadow.cpp:15-29; VP_Shadow, test2/testproc.cpp:505-520 */
        p->PV = u[3]; p->M = u[4];
        if (!p->PV || !p->M) return -3;
        m4m4(p->PVM, p->PV, p->M);
        if (d->prog == OR_PROG_CLOUD_SHADOW) { p->tex0 = tex_from_uniform(d, 9); if (!p->tex0) return -4; }
        break;
    case OR_PROG_CLOUD:   /* VP_Cloud / FP_Cloud::preprocess, test/testproc.cpp:384-390, 526-532 */
        p->PV = u[3]; p->M = u[4]; p->Mrot = u[5]; p->lightPos = u[7]; p->shadowPV = u[16];
        if (!p->PV || !p->M || !p->Mrot || !p->lightPos || !p->shadowPV) return -3;
        p->tex0 = tex_from_uniform(d, 9);
        p->tex2 = tex_from_uniform(d, 15);
        if (!p->tex0 || !p->tex2) return -4;
        break;
    case OR_PROG_LAMBERT:
        p->PV = u[3]; p->M = u[4]; p->Mrot = u[5]; p->lightDir = u[7];
        if (!p->PV || !p->M || !p->Mrot || !p->lightDir) return -3;
        p->tex0 = tex_from_uniform(d, 9);
        if (!p->tex0) return -4;
        break;
    case OR_PROG_BLINN: case OR_PROG_SHADOWED:
        p->PV = u[3]; p->M = u[4]; p->Mrot = u[5]; p->lightPos = u[7]; p->cameraPos = u[8];
        if (!p->PV || !p->M || !p->Mrot || !p->lightPos || !p->cameraPos) return -3;
        p->tex0 = tex_from_uniform(d, 9);
        p->tex1 = tex_from_uniform(d, 10);
        if (!p->tex0 || !p->tex1) return -4;
        p->spec_exp = u[11] ? (unsigned)*(const int*)u[11] : 32u;
        p->needs_grad = 1;
        if (d->prog == OR_PROG_SHADOWED) {
            p->shadowPV = u[12];
            p->tex2 = tex_from_uniform(d, 13);
            if (!p->shadowPV) return -3;
            if (!p->tex2) return -4;
        }
        break;
    }
    return 0;
}

static inline const float* slot_elem(const or_draw* d, int slot, int64_t i)
{
    return (const float*)((const uint8_t*)d->slot_data[slot] + (size_t)i * (size_t)d->slot_unit_bytes[slot]);
}

/* VP::process for vertex i of the draw; pos gets clip coordinates. */
static void run_vp(const or_draw* d, const or_prog* p, int64_t i, float* pos, float* user)
{
    float wp[4];
    memset(user, 0, sizeof(float) * OR_MAX_NV);
    switch (p->prog) {
    case OR_PROG_DEF01: {   /* tex1light1.cpp:28-41 */
        m4v4(wp, p->M, slot_elem(d, 0, i));
        memcpy(pos, wp, 16);
        wp[3] = 0;
        float t[4]; memcpy(t, pos, 16); m4v4(pos, p->PV, t);
        m4v4(user + 0, p->Mrot, slot_elem(d, 3, i));
        memcpy(user + 4, wp, 16);
        const float* uv = slot_elem(d, 4, i);
        user[8] = uv[0]; user[9] = uv[1];
        break;
    }
    case OR_PROG_DEF02: {   /* colr1light1.cpp:22-40 */
        m4v4(wp, p->M, slot_elem(d, 0, i));
        memcpy(pos, wp, 16);
        wp[3] = 0;
        float t[4]; memcpy(t, pos, 16); m4v4(pos, p->PV, t);
        m4v4(user + 0, p->Mrot, slot_elem(d, 1, i));
        memcpy(user + 4, wp, 16);
        memcpy(user + 8, slot_elem(d, 2, i), 16);
        break;
    }
    case OR_PROG_DEF03: {   /* tex1bump1light1.cpp:22-45 */
        m4v4(wp, p->M, slot_elem(d, 0, i));
        memcpy(pos, wp, 16);
        wp[3] = 0;
        float t[4]; memcpy(t, pos, 16); m4v4(pos, p->PV, t);
        m4v4(user + 0, p->Mrot, slot_elem(d, 1, i));
        m4v4(user + 4, p->Mrot, slot_elem(d, 2, i));
        m4v4(user + 8, p->Mrot, slot_elem(d, 3, i));
        memcpy(user + 12, wp, 16);
        const float* uv = slot_elem(d, 4, i);
        user[16] = uv[0]; user[17] = uv[1];
        break;
    }
    case OR_PROG_DEF04: {   /* skybox.cpp:23-44 */
        const float* src = slot_elem(d, 0, i);
        static const float observer[4] = {0, 0, 1.0f, 0};
        float v[4];
        for (int k = 0; k < 4; k++) v[k] = src[k] - observer[k];
        v[3] = 0;
        switch (d->shade_fp64) {
        case 2: norm4_x86(v); break;
        case 3: norm4_hi(v); break;
        case 4: norm4_lo(v); break;
        default: norm4(v);
        }
        m4v4(user, p->invView, v);
        memcpy(pos, src, 16);
        break;
    }
    case OR_PROG_DEF05: case OR_PROG_CLOUD_SHADOW: {
        /* shadow.cpp:21-29 (0.9); VP_CloudShadow, test/testproc.cpp:595-609 (0.95) */
        float pa[4], f = p->prog == OR_PROG_CLOUD_SHADOW ? 0.95f : 0.9f;
        memcpy(pa, slot_elem(d, 0, i), 16);
        pa[0] *= f; pa[1] *= f; pa[2] *= f;
        m4v4(pos, p->PVM, pa);
        if (p->prog == OR_PROG_CLOUD_SHADOW) {
            const float* uv = slot_elem(d, 4, i);
            user[0] = uv[0]; user[1] = uv[1];
        }
        break;
    }
    case OR_PROG_CLOUD: {   /* VP_Cloud::process, test/testproc.cpp:392-420 */
        float sc[4], n[4];
        m4v4(wp, p->M, slot_elem(d, 0, i));
        m4v4(sc, p->shadowPV, wp);
        m4v4(pos, p->PV, wp);
        m4v4(n, p->Mrot, slot_elem(d, 3, i));
        const float* uv = slot_elem(d, 4, i);
        /* PROCDATA_CLOUD: normal, worldPos (w = 0), texcoord, shadowcoord */
        for (int k = 0; k < 4; k++) { user[k] = n[k]; user[10 + k] = sc[k]; }
        user[4] = wp[0]; user[5] = wp[1]; user[6] = wp[2]; user[7] = 0.0f;
        user[8] = uv[0]; user[9] = uv[1];
        break;
    }
    case OR_PROG_LAMBERT: {
        m4v4(wp, p->M, slot_elem(d, 0, i));
        m4v4(pos, p->PV, wp);
        float n[4];
        m4v4(n, p->Mrot, slot_elem(d, 3, i));
        const float* uv = slot_elem(d, 4, i);
        user[0] = n[0]; user[1] = n[1]; user[2] = n[2]; user[3] = uv[0]; user[4] = uv[1];
        break;
    }
    case OR_PROG_BLINN: case OR_PROG_SHADOWED: {
        m4v4(wp, p->M, slot_elem(d, 0, i));
        m4v4(pos, p->PV, wp);
        if (p->prog == OR_PROG_SHADOWED) {   /* world position in shadow-map clip space */
            float sc[4];
            m4v4(sc, p->shadowPV, wp);
            for (int k = 0; k < 4; k++) user[14 + k] = sc[k];
        }
        float t[4], b[4], n[4];
        m4v4(t, p->Mrot, slot_elem(d, 1, i));
        m4v4(b, p->Mrot, slot_elem(d, 2, i));
        m4v4(n, p->Mrot, slot_elem(d, 3, i));
        const float* uv = slot_elem(d, 4, i);
        for (int k = 0; k < 3; k++) { user[k] = t[k]; user[3 + k] = b[k]; user[6 + k] = n[k]; user[9 + k] = wp[k]; }
        user[12] = uv[0]; user[13] = uv[1];
        break;
    }
    }
}

static inline uint8_t u8_of(float v) { return (uint8_t)(int)v; }

/* cvtps2dq, packusdw, packuswb of one lane (testproc.cpp:562-568): cvtps2dq
   itself (MXCSR round-to-nearest-even, INT_MIN for NaN and out of range),
   then the two saturating packs, the second reading its words as signed */
uint8_t or_sse_pack_byte(float f)
{
    int v = _mm_cvtsi128_si32(_mm_cvtps_epi32(_mm_set1_ps(f)));
    unsigned w = v < 0 ? 0u : (v > 65535 ? 65535u : (unsigned)v);
    return (uint8_t)((int16_t)(uint16_t)w < 0 ? 0u : (w > 255u ? 255u : w));
}
static inline uint8_t u8_of_d(double v) { return (uint8_t)(int)v; }

/* returns 1 when a colour is written to *out, 2 when the fragment is
   discarded (FBOBridge::discard, fragthrd.cpp:21-24: no depth update) */
static int run_fp(const or_prog* p, const float* in, float r, const float* grad, int fp64, uint8_t* out)
{
    switch (p->prog) {
    case OR_PROG_CLOUD: {   /* FP_Cloud::process, test/testproc.cpp:534-574 */
        uint8_t t[4];
        or_sample_nearest(p->tex0, 0, in[8], in[9], t);
        float sf = or_sample_projection(p->tex2, in + 10);
        float L[4];
        for (int k = 0; k < 4; k++) L[k] = p->lightPos[k] - in[4 + k];
        float dist = len4(L);
        div4(L, dist);
        /* E and H (testproc.cpp:553-558) feed nothing */
        float lambert = 2.0f * dot4(L, in);
        float cc = 255.0f * (lambert * sf);
        uint8_t b = or_sse_pack_byte(cc);
        out[0] = out[1] = out[2] = b;
        out[3] = or_sse_pack_byte((float)t[2]);
        return 1;
    }
    case OR_PROG_CLOUD_SHADOW: {   /* FP_CloudShadow, test/testproc.cpp:677-686 */
        uint8_t c[4];
        or_sample_nearest(p->tex0, 0, in[0], in[1], c);
        return c[2] < 150 ? 2 : 0;
    }
    case OR_PROG_DEF01: case OR_PROG_DEF02: case OR_PROG_DEF03: {
        /* tex1light1.cpp:152-193, colr1light1.cpp:144-190, tex1bump1light1.cpp:180-239 */
        const float* wp; const float* nrm;
        float oc[4], bump[4];
        if (p->prog == OR_PROG_DEF03) { wp = in + 12; nrm = NULL; } else { wp = in + 4; nrm = in; }
        if (p->prog == OR_PROG_DEF02) {
            memcpy(oc, in + 8, 16);
        } else {
            const float* tc = p->prog == OR_PROG_DEF03 ? in + 16 : in + 8;
            uint8_t c[4];
            or_sample_nearest(p->tex0, 0, tc[0], tc[1], c);
            oc[0] = c[0]; oc[1] = c[1]; oc[2] = c[2]; oc[3] = c[3];
            if (p->prog == OR_PROG_DEF03) {
                uint8_t bb[4];
                or_sample_nearest(p->tex1, 0, tc[0], tc[1], bb);
                bump[0] = bb[2]; bump[1] = bb[1]; bump[2] = bb[0]; bump[3] = 0;
            }
        }
        if (fp64 != 1) {
            /* modes 2-4: the reference's rcpps / rsqrtss in div_3_4 and
               norm_3_4, as this CPU computes them or at either end of their
               error bound */
            void (*divf)(float*, float) = fp64 == 2 ? div4_x86 : fp64 == 3 ? div4_hi : fp64 == 4 ? div4_lo : div4;
            void (*normf)(float*) = fp64 == 2 ? norm4_x86 : fp64 == 3 ? norm4_hi : fp64 == 4 ? norm4_lo : norm4;
            if (p->prog == OR_PROG_DEF03) {
                divf(bump, 255.0f);
                for (int k = 0; k < 4; k++) bump[k] *= 2.0f;
                for (int k = 0; k < 4; k++) bump[k] -= 1.0f;
                float tbn[16];
                memcpy(tbn, in, 16); memcpy(tbn + 4, in + 4, 16); memcpy(tbn + 8, in + 8, 16);
                tbn[12] = tbn[13] = tbn[14] = tbn[15] = 0.0f;
                float t[4]; memcpy(t, bump, 16);
                m4v4(bump, tbn, t);
                normf(bump);
                nrm = bump;
            }
            float L[4], E[4], H[4];
            for (int k = 0; k < 4; k++) L[k] = p->lightPos[k] - wp[k];
            float dist = len4(L);
            divf(L, dist);
            for (int k = 0; k < 4; k++) E[k] = p->cameraPos[k] - wp[k];
            normf(E);
            for (int k = 0; k < 4; k++) H[k] = E[k] + L[k];
            normf(H);
            float lambert = dot4(L, nrm);
            float spec = dot4(H, nrm);
            spec = spec < 0 ? 0 : spec;
            spec = opt_powf(spec, 50);
            float s255 = 255.0f * spec;
            for (int k = 0; k < 4; k++) oc[k] = oc[k] + s255;
            for (int k = 0; k < 4; k++) oc[k] = oc[k] * lambert;
            for (int k = 0; k < 4; k++) oc[k] = clampf_sse(oc[k], 0, 255.0f);
            out[0] = u8_of(oc[0]); out[1] = u8_of(oc[1]); out[2] = u8_of(oc[2]); out[3] = 0;
        } else {
            double n[4], L[4], E[4], H[4], o[4], nb[4];
            for (int k = 0; k < 4; k++) o[k] = oc[k];
            if (p->prog == OR_PROG_DEF03) {
                double b[4];
                for (int k = 0; k < 4; k++) b[k] = bump[k] * (2.0 / 255.0) - 1.0;
                b[3] = -1.0;
                for (int k = 0; k < 4; k++) nb[k] = b[0] * in[k] + b[1] * in[4 + k] + b[2] * in[8 + k];
                double s = sqrt(nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2] + nb[3] * nb[3]);
                for (int k = 0; k < 4; k++) n[k] = nb[k] / s;
            } else {
                for (int k = 0; k < 4; k++) n[k] = nrm[k];
            }
            for (int k = 0; k < 4; k++) L[k] = (double)p->lightPos[k] - wp[k];
            double dist = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2] + L[3] * L[3]);
            for (int k = 0; k < 4; k++) L[k] /= dist;
            for (int k = 0; k < 4; k++) E[k] = (double)p->cameraPos[k] - wp[k];
            double el = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2] + E[3] * E[3]);
            for (int k = 0; k < 4; k++) E[k] /= el;
            for (int k = 0; k < 4; k++) H[k] = E[k] + L[k];
            double hl = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2] + H[3] * H[3]);
            for (int k = 0; k < 4; k++) H[k] /= hl;
            double lambert = L[0] * n[0] + L[1] * n[1] + L[2] * n[2] + L[3] * n[3];
            double spec = H[0] * n[0] + H[1] * n[1] + H[2] * n[2] + H[3] * n[3];
            spec = spec < 0 ? 0 : spec;
            spec = opt_powd(spec, 50);
            for (int k = 0; k < 3; k++) {
                double v = (o[k] + 255.0 * spec) * lambert;
                v = v > 0 ? v : 0; v = v < 255.0 ? v : 255.0;
                out[k] = u8_of_d(v);
            }
            out[3] = 0;
        }
        return 1;
    }
    case OR_PROG_DEF04: {   /* skybox.cpp:127-134 */
        or_sample_cube(p->tex0, in, out);
        return 1;
    }
    case OR_PROG_DEF05:     /* shadow.cpp:76-78: writes nothing */
        return 0;
    case OR_PROG_LAMBERT: {
        if (fp64 != 1) {
            float t[4];
            or_sample_bilinear(p->tex0, 0, 0, in[3], in[4], t);
            float n[3] = {in[0], in[1], in[2]};
            norm3(n);
            float lam = dot3(n, p->lightDir);
            lam = lam > 0.0f ? lam : 0.0f;
            for (int k = 0; k < 3; k++) out[k] = u8_of(clampf_sse(t[k] * lam, 0.0f, 255.0f));
        } else {
            double t[4];
            sample_bilinear_d(p->tex0, 0, 0, in[3], in[4], t);
            double n0 = in[0], n1 = in[1], n2 = in[2];
            double s = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
            double lam = (n0 * p->lightDir[0] + n1 * p->lightDir[1] + n2 * p->lightDir[2]) / s;
            lam = lam > 0 ? lam : 0;
            for (int k = 0; k < 3; k++) {
                double v = t[k] * lam;
                v = v > 0 ? v : 0; v = v < 255.0 ? v : 255.0;
                out[k] = u8_of_d(v);
            }
        }
        out[3] = 0;
        return 1;
    }
    case OR_PROG_BLINN: case OR_PROG_SHADOWED: {
        const float *T = in, *B = in + 3, *N = in + 6, *wp = in + 9;
        float u = in[12], v = in[13];
        /* the shadow factor is a discrete lookup: both variants take it in FP32 */
        float sf = p->prog == OR_PROG_SHADOWED ? or_sample_projection(p->tex2, in + 14) : 1.0f;
        if (fp64 != 1) {
            float lod = lod_from_grad(grad, u, v, r, p->tex0->width, p->tex0->height);
            float alb[4], nm[4];
            sample_trilinear(p->tex0, u, v, lod, alb);
            sample_trilinear(p->tex1, u, v, lod, nm);
            const float k2 = 2.0f / 255.0f;
            float bx = nm[2] * k2 - 1.0f, by = nm[1] * k2 - 1.0f, bz = nm[0] * k2 - 1.0f;
            float n[3];
            for (int k = 0; k < 3; k++) {
                float a = bx * T[k], b = by * B[k], c = bz * N[k];
                n[k] = (a + b) + c;
            }
            norm3(n);
            float L[3], E[3], H[3];
            for (int k = 0; k < 3; k++) L[k] = p->lightPos[k] - wp[k];
            float dist = sqrtf(dot3(L, L));
            float rd = 1.0f / dist;
            for (int k = 0; k < 3; k++) L[k] *= rd;
            for (int k = 0; k < 3; k++) E[k] = p->cameraPos[k] - wp[k];
            norm3(E);
            for (int k = 0; k < 3; k++) H[k] = E[k] + L[k];
            norm3(H);
            float lam = dot3(L, n);
            lam = lam * sf;
            float spec = dot3(H, n);
            spec = spec < 0 ? 0 : spec;
            spec = opt_powf(spec, p->spec_exp);
            float s255 = 255.0f * spec;
            for (int k = 0; k < 3; k++) out[k] = u8_of(clampf_sse((alb[k] + s255) * lam, 0.0f, 255.0f));
        } else {
            double lod = lod_from_grad_d(grad, u, v, r, p->tex0->width, p->tex0->height);
            double alb[4], nm[4];
            sample_trilinear_d(p->tex0, u, v, lod, alb);
            sample_trilinear_d(p->tex1, u, v, lod, nm);
            double bx = nm[2] * (2.0 / 255.0) - 1.0, by = nm[1] * (2.0 / 255.0) - 1.0, bz = nm[0] * (2.0 / 255.0) - 1.0;
            double n[3], L[3], E[3], H[3];
            for (int k = 0; k < 3; k++) n[k] = bx * T[k] + by * B[k] + bz * N[k];
            double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; k++) n[k] /= s;
            for (int k = 0; k < 3; k++) L[k] = (double)p->lightPos[k] - wp[k];
            s = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
            for (int k = 0; k < 3; k++) L[k] /= s;
            for (int k = 0; k < 3; k++) E[k] = (double)p->cameraPos[k] - wp[k];
            s = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
            for (int k = 0; k < 3; k++) E[k] /= s;
            for (int k = 0; k < 3; k++) H[k] = E[k] + L[k];
            s = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
            for (int k = 0; k < 3; k++) H[k] /= s;
            double lam = (L[0] * n[0] + L[1] * n[1] + L[2] * n[2]) * sf;
            double spec = H[0] * n[0] + H[1] * n[1] + H[2] * n[2];
            spec = spec < 0 ? 0 : spec;
            spec = opt_powd(spec, p->spec_exp);
            for (int k = 0; k < 3; k++) {
                double c = (alb[k] + 255.0 * spec) * lam;
                c = c > 0 ? c : 0; c = c < 255.0 ? c : 255.0;
                out[k] = u8_of_d(c);
            }
        }
        out[3] = 0;
        return 1;
    }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* fragment loop (fragthrd.cpp:160-252)                                */
/* ------------------------------------------------------------------ */

typedef struct {
    int x1, x2, y, tri;
    float cf2, cf2_step, pz, pz_step;
    float vs[OR_MAX_NV], vt[OR_MAX_NV];
    float grad[6];
    int kind;   /* 0 draw, 1 end-of-draw, 2 quit */
} or_task;

typedef struct {
    int64_t frags_rasterised, frags_shaded;
} or_wstats;

static void run_task(const or_draw* d, const or_prog* p, or_target* t, const or_task* k, or_wstats* ws)
{
    int nv = p->nv;
    float cf2 = k->cf2, pz = k->pz;
    float vs[OR_MAX_NV];
    memcpy(vs, k->vs, sizeof(float) * nv);
    int W = t->width, H = t->height, y = k->y;
    float* drow = t->depth + (size_t)y * (size_t)t->depth_scanline;
    uint8_t* crow = t->colour + (size_t)(t->colour_top_down ? (H - 1 - y) : y) * (size_t)t->colour_scanline;
    int beh = d->behavior;
    uint32_t th = or_tri_hash((uint32_t)k->tri);
    for (int x = k->x1; x <= k->x2; x++) {
        /* interpolateNextStep, interp.cpp:82-92 */
        float r = 1.0f / cf2;
        cf2 += k->cf2_step;
        float in[OR_MAX_NV];
        for (int i = 0; i < nv; i++) in[i] = vs[i] * r;
        for (int i = 0; i < nv; i++) vs[i] = vs[i] + k->vt[i];
        float z = pz * r;
        pz += k->pz_step;
        float cur = (beh & OR_TEST_DEPTH) ? drow[x] : 1.0f;
        ws->frags_rasterised++;
        size_t pi = (size_t)y * W + x;
        if (t->dbg_raster_count) { t->dbg_raster_count[pi]++; t->dbg_raster_hash[pi] ^= th; }
        if (-1.0f < z && (z - cur < -0.0001f)) {
            uint8_t c[4];
            int fr = run_fp(p, in, r, k->grad, d->shade_fp64, c);
            if (fr == 1) {   /* FBOBridge::write4: blend4 under ALPHABLEND (fragthrd.cpp:70-82) */
                if (beh & OR_ALPHABLEND) or_blend4(crow + (size_t)x * 4, c);
                else memcpy(crow + (size_t)x * 4, c, 4);
            }
            if (fr != 2 && (beh & OR_UPDATE_DEPTH)) drow[x] = z;
            ws->frags_shaded++;
            if (t->dbg_pass_count) {
                t->dbg_pass_count[pi]++; t->dbg_pass_hash[pi] ^= th; t->dbg_last_pass[pi] = k->tri;
            }
        }
    }
    (void)W;
}

/* ------------------------------------------------------------------ */
/* SPSC ring queues and worker threads (rinque.h, fragthrd.cpp)        */
/* ------------------------------------------------------------------ */

#define OR_QLEN 512   /* config.h:9 MAX_FRAGTASKS */

typedef struct {
    _Alignas(64) atomic_uint head;   /* producer */
    _Alignas(64) atomic_uint tail;   /* consumer */
    or_task* slots;
} or_queue;

typedef struct or_pool or_pool;
typedef struct {
    or_pool* pool;
    int index;
    pthread_t th;
    or_wstats ws;
} or_worker;

struct or_pool {
    int n_workers;
    or_queue* q;
    or_worker* w;
    /* draw context, published before tasks are pushed */
    const or_draw* d;
    const or_prog* p;
    or_target* t;
};

static void q_push(or_queue* q, const or_task* k)
{
    unsigned h = atomic_load_explicit(&q->head, memory_order_relaxed);
    while (h - atomic_load_explicit(&q->tail, memory_order_acquire) >= OR_QLEN) sched_yield();
    q->slots[h % OR_QLEN] = *k;
    atomic_store_explicit(&q->head, h + 1, memory_order_release);
}

static void* worker_main(void* arg)
{
    or_worker* w = (or_worker*)arg;
    or_queue* q = &w->pool->q[w->index];
    for (;;) {
        unsigned tl = atomic_load_explicit(&q->tail, memory_order_relaxed);
        while (atomic_load_explicit(&q->head, memory_order_acquire) == tl) sched_yield();
        or_task* k = &q->slots[tl % OR_QLEN];
        if (k->kind == 2) { atomic_store_explicit(&q->tail, tl + 1, memory_order_release); break; }
        if (k->kind == 0) run_task(w->pool->d, w->pool->p, w->pool->t, k, &w->ws);
        atomic_store_explicit(&q->tail, tl + 1, memory_order_release);
    }
    return NULL;
}

static or_pool* g_pool = NULL;
static pthread_mutex_t g_pool_lock = PTHREAD_MUTEX_INITIALIZER;

static void pool_destroy(or_pool* P)
{
    or_task quit;
    memset(&quit, 0, sizeof(quit));
    quit.kind = 2;
    for (int i = 0; i < P->n_workers; i++) q_push(&P->q[i], &quit);
    for (int i = 0; i < P->n_workers; i++) pthread_join(P->w[i].th, NULL);
    for (int i = 0; i < P->n_workers; i++) free(P->q[i].slots);
    free(P->q);
    free(P->w);
    free(P);
}

static or_pool* pool_get(int n_workers)
{
    if (g_pool && g_pool->n_workers == n_workers) return g_pool;
    if (g_pool) { pool_destroy(g_pool); g_pool = NULL; }
    or_pool* P = (or_pool*)calloc(1, sizeof(or_pool));
    P->n_workers = n_workers;
    P->q = (or_queue*)aligned_alloc(64, sizeof(or_queue) * (size_t)n_workers);
    P->w = (or_worker*)calloc((size_t)n_workers, sizeof(or_worker));
    for (int i = 0; i < n_workers; i++) {
        atomic_init(&P->q[i].head, 0);
        atomic_init(&P->q[i].tail, 0);
        P->q[i].slots = (or_task*)malloc(sizeof(or_task) * OR_QLEN);
    }
    for (int i = 0; i < n_workers; i++) {
        P->w[i].pool = P;
        P->w[i].index = i;
        pthread_create(&P->w[i].th, NULL, worker_main, &P->w[i]);
    }
    g_pool = P;
    return P;
}

/* ------------------------------------------------------------------ */
/* drawVAO (drawvao.cpp:3-133)                                         */
/* ------------------------------------------------------------------ */

/* per-triangle gradients of rw, u*rw, v*rw in screen space (for mip selection) */
static void tri_grad(const float* verts, const float* rw4, const float (*user)[OR_MAX_NV], int iu, float* g)
{
    float e1x = verts[2] - verts[0], e1y = verts[3] - verts[1];
    float e2x = verts[4] - verts[0], e2y = verts[5] - verts[1];
    float det = e1x * e2y - e2x * e1y;
    float inv = 1.0f / det;
    float f[3][3];
    for (int i = 0; i < 3; i++) {
        f[0][i] = rw4[i];
        f[1][i] = user[i][iu] * rw4[i];
        f[2][i] = user[i][iu + 1] * rw4[i];
    }
    for (int a = 0; a < 3; a++) {
        float d1 = f[a][1] - f[a][0], d2 = f[a][2] - f[a][0];
        g[2 * a] = (d1 * e2y - d2 * e1y) * inv;
        g[2 * a + 1] = (d2 * e1x - d1 * e2x) * inv;
    }
}

int or_draw_vao(const or_draw* d, or_target* t, or_stats* st, int n_threads)
{
    or_prog p;
    int rc = prog_prepare(d, &p);
    if (rc) return rc;
    int W = t->width, H = t->height;
    if (W < 1 || H < 1) return -5;

    /* processVertices stops when any attached VBO runs dry (vertthrd.cpp:14-51) */
    int64_t nverts = -1;
    for (int s = 0; s < OR_MAX_SLOTS; s++)
        if (d->slot_count[s] > 0 && (nverts < 0 || d->slot_count[s] < nverts)) nverts = d->slot_count[s];
    if (nverts < 0) nverts = 0;
    int64_t ntri = nverts / 3;

    or_raster R;
    R.W = W; R.H = H; R.hw = W / 2; R.hh = H / 2;
    R.rows = (or_row*)malloc(sizeof(or_row) * (size_t)H);
    memset(R.rows, 0, sizeof(or_row) * (size_t)H);

    or_pool* P = NULL;
    int nworkers = n_threads >= 2 ? n_threads - 1 : 0;
    if (nworkers) {
        pthread_mutex_lock(&g_pool_lock);
        P = pool_get(nworkers);
        P->d = d; P->p = &p; P->t = t;
        for (int i = 0; i < nworkers; i++) memset(&P->w[i].ws, 0, sizeof(or_wstats));
    }
    or_wstats local = {0, 0};
    int64_t culled = 0, zrej = 0, dropped = 0, spans = 0;

    float rw4[4] = {0, 0, 0, 0}, z4[4] = {0, 0, 0, 0};
    float user[3][OR_MAX_NV];
    float pos[3][4];
    or_task task;
    memset(&task, 0, sizeof(task));

    for (int64_t tri = 0; tri < ntri; tri++) {
        for (int v = 0; v < 3; v++) {
            run_vp(d, &p, tri * 3 + v, pos[v], user[v]);
            float rw = 1.0f / pos[v][3];
            pos[v][0] *= rw; pos[v][1] *= rw; pos[v][2] *= rw; pos[v][3] *= rw;
            z4[v] = pos[v][2];
            rw4[v] = rw;
        }
        if ((d->behavior & OR_FACE_CULLING) && or_is_back_face(pos[0], pos[1], pos[2])) { culled++; continue; }
        int zr = 0;
        for (int v = 0; v < 3; v++) if (pos[v][2] < -1.0f || pos[v][2] > 1.0f) zr = 1;
        if (zr) { zrej++; continue; }
        /* Screen y outside the int range (or NaN) makes the reference read rows
           it never wrote (or crash); such triangles are dropped here. */
        int bad = 0;
        for (int v = 0; v < 3; v++) {
            float sy = ((float)R.hh * pos[v][1]) + (float)R.hh;
            if (!(sy >= -2147483648.0f && sy < 2147483648.0f)) bad = 1;
        }
        if (bad) { dropped++; continue; }
        if (!push_triangle(&R, pos[0], pos[1], pos[2])) { dropped++; continue; }

        float verts[6] = {R.vx[0], R.vy[0], R.vx[1], R.vy[1], R.vx[2], R.vy[2]};
        float grad[6] = {0, 0, 0, 0, 0, 0};
        if (p.needs_grad) tri_grad(verts, rw4, (const float(*)[OR_MAX_NV])user, 12, grad);

        for (int row = R.first; row <= R.last; row++) {
            or_row* rr = &R.rows[row];
            if (rr->left == rr->right) continue;   /* drawvao.cpp:72-75 */
            or_span_geom g;
            span_geom(verts, rw4, row, rr->left, rr->right, rr->leftClamped, rr->lv0, rr->lv1, rr->rv0, rr->rv1, &g);
            task.kind = 0;
            task.x1 = rr->leftClamped;
            task.x2 = rr->rightClamped;
            task.y = row;
            task.tri = d->tri_base + (int)tri;
            span_vis(&g, z4, &task.cf2, &task.cf2_step, &task.pz, &task.pz_step);
            span_varyings(&g, (const float(*)[OR_MAX_NV])user, p.nv, task.vs, task.vt);
            memcpy(task.grad, grad, sizeof(grad));
            spans++;
            if (nworkers) q_push(&P->q[row % nworkers], &task);
            else run_task(d, &p, t, &task, &local);
        }
    }

    if (nworkers) {
        /* ENDOFDRAW + pollEmpty (drawvao.cpp:121-132) */
        or_task end;
        memset(&end, 0, sizeof(end));
        end.kind = 1;
        for (int i = 0; i < nworkers; i++) q_push(&P->q[i], &end);
        for (int i = 0; i < nworkers; i++) {
            or_queue* q = &P->q[i];
            while (atomic_load_explicit(&q->tail, memory_order_acquire) != atomic_load_explicit(&q->head, memory_order_relaxed))
                sched_yield();
            local.frags_rasterised += P->w[i].ws.frags_rasterised;
            local.frags_shaded += P->w[i].ws.frags_shaded;
        }
        pthread_mutex_unlock(&g_pool_lock);
    }
    free(R.rows);
    if (st) {
        st->tris_in += ntri;
        st->tris_culled += culled;
        st->tris_zrejected += zrej;
        st->tris_dropped += dropped;
        st->spans += spans;
        st->frags_rasterised += local.frags_rasterised;
        st->frags_shaded += local.frags_shaded;
    }
    return 0;
}

void or_pool_shutdown(void)
{
    pthread_mutex_lock(&g_pool_lock);
    if (g_pool) { pool_destroy(g_pool); g_pool = NULL; }
    pthread_mutex_unlock(&g_pool_lock);
}

void or_clear_depth(or_target* t, float value)
{
    size_t n = (size_t)t->depth_scanline * (size_t)t->height;
    for (size_t i = 0; i < n; i++) t->depth[i] = value;
}

void or_clear_colour(or_target* t, uint32_t bgra)
{
    for (int y = 0; y < t->height - 1; y++) {
        uint8_t* row = t->colour + (size_t)y * (size_t)t->colour_scanline;
        for (int x = 0; x < t->width; x++) memcpy(row + (size_t)x * 4, &bgra, 4);
    }
}

/* ------------------------------------------------------------------ */
/* differentiable shading on a geometry buffer, FP64                    */
/* ------------------------------------------------------------------ */
typedef struct {
    int covered;
    double lam, s, spec, u, v;
} or_geo;

static void norm3d(double* a)
{
    double r = 1.0 / sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    a[0] *= r; a[1] *= r; a[2] *= r;
}

static or_geo diff_geometry(const float* g, double e, double ks, const float* light, const float* camera)
{
    or_geo o;
    memset(&o, 0, sizeof(o));
    o.covered = g[3] != 0.0f;
    if (!o.covered) return o;
    double n[3] = {g[0], g[1], g[2]}, L[3], E[3], Hh[3];
    norm3d(n);
    for (int k = 0; k < 3; k++) { L[k] = (double)light[k] - g[4 + k]; E[k] = (double)camera[k] - g[4 + k]; }
    norm3d(L);
    norm3d(E);
    for (int k = 0; k < 3; k++) Hh[k] = L[k] + E[k];
    norm3d(Hh);
    double lam = n[0] * L[0] + n[1] * L[1] + n[2] * L[2];
    double s = n[0] * Hh[0] + n[1] * Hh[1] + n[2] * Hh[2];
    o.lam = lam > 0 ? lam : 0;
    o.s = s > 0 ? s : 0;
    o.spec = o.s > 0 ? ks * pow(o.s, e) : 0;
    o.u = g[8];
    o.v = g[9];
    return o;
}

static void diff_footprint(int tw, int th, double u, double v, int64_t* idx, double* w)
{
    /* the texel footprint follows the FP32 coordinate the kernel forms */
    float fx = (float)u * (float)tw, fy = (float)v * (float)th;
    if (!(fabsf(fx) < 1073741824.0f)) fx = 0.0f;
    if (!(fabsf(fy) < 1073741824.0f)) fy = 0.0f;
    double x0 = floor((double)fx), y0 = floor((double)fy);
    double ax = (double)fx - x0, ay = (double)fy - y0;
    long ix = (long)x0 % tw, iy = (long)y0 % th;
    if (ix < 0) ix += tw;
    if (iy < 0) iy += th;
    long ix1 = (ix + 1) % tw, iy1 = (iy + 1) % th;
    idx[0] = iy * tw + ix; idx[1] = iy * tw + ix1; idx[2] = iy1 * tw + ix; idx[3] = iy1 * tw + ix1;
    w[0] = (1 - ax) * (1 - ay); w[1] = ax * (1 - ay); w[2] = (1 - ax) * ay; w[3] = ax * ay;
}

void or_diff_forward_f64(int W, int H, const float* gbuf, int tw, int th, const float* tex, double albedo_scale,
                         double spec_exp, double ks, const float* light, const float* camera, double* out)
{
    for (int64_t p = 0; p < (int64_t)W * H; p++) {
        or_geo g = diff_geometry(gbuf + p * 12, spec_exp, ks, light, camera);
        for (int c = 0; c < 3; c++) out[p * 3 + c] = 0;
        if (!g.covered) continue;
        int64_t idx[4];
        double w[4];
        diff_footprint(tw, th, g.u, g.v, idx, w);
        for (int c = 0; c < 3; c++) {
            double t = 0;
            for (int k = 0; k < 4; k++) t += w[k] * tex[idx[k] * 3 + c];
            out[p * 3 + c] = albedo_scale * t * g.lam + g.spec;
        }
    }
}

void or_diff_backward_f64(int W, int H, const float* gbuf, int tw, int th, const float* tex, double albedo_scale,
                          double spec_exp, double ks, const float* light, const float* camera, const double* gout,
                          double* gtex, double* gparams)
{
    for (int64_t p = 0; p < (int64_t)W * H; p++) {
        or_geo g = diff_geometry(gbuf + p * 12, spec_exp, ks, light, camera);
        if (!g.covered) continue;
        int64_t idx[4];
        double w[4];
        diff_footprint(tw, th, g.u, g.v, idx, w);
        const double* go = gout + p * 3;
        for (int c = 0; c < 3; c++) {
            double t = 0;
            for (int k = 0; k < 4; k++) t += w[k] * tex[idx[k] * 3 + c];
            gparams[0] += go[c] * t * g.lam;
            for (int k = 0; k < 4; k++) gtex[idx[k] * 3 + c] += go[c] * albedo_scale * g.lam * w[k];
        }
        if (g.s > 0) gparams[1] += (go[0] + go[1] + go[2]) * g.spec * log(g.s);
    }
}
