// ps_device.h -- gfx950 device arithmetic of the raster path.
//
// Every function restates a reference operation with its exact FP32
// operation order (the library is built with -ffp-contract=off, no fast-math,
// correctly rounded division and sqrt, denormals kept).  rcpps and rsqrtss
// of the reference (vector.cpp:158-250) are 1.0f/x and 1.0f/sqrtf(x) here.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <stdint.h>

#include "ps_types.h"
#include "../../include/psgpu.h"

#define PS_DEV __device__ __forceinline__

// ---- x86 conversions ---------------------------------------------------
// (int)f compiled for x86 (cvttss2si): NaN and out-of-range give INT_MIN.
// v_cvt_i32_f32 saturates instead, so the range is checked first.
PS_DEV int ps_cvt_i32(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

// (unsigned)f on x86-64 gcc: 64-bit cvttss2si, low 32 bits (samplr2d.cpp:6-7).
PS_DEV uint32_t ps_cvt_u32(float f)
{
    if (f >= -2147483648.0f && f < 2147483648.0f) return (uint32_t)(int)f;
    if (f >= -9223372036854775808.0f && f < 9223372036854775808.0f) return (uint32_t)(uint64_t)(int64_t)f;
    return 0u;
}

// ---- mcemaths (vector.cpp, matrix.cpp) --------------------------------
// transform_m4v4, matrix.cpp:515-557
PS_DEV void ps_m4v4(float* r, const float* m, const float* v)
{
    const float x = v[0], y = v[1], z = v[2], w = v[3];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float a = x * m[i], b = y * m[4 + i], c = z * m[8 + i], d = w * m[12 + i];
        r[i] = ((a + b) + c) + d;
    }
}

// dot_3_4, vector.cpp:85-112
PS_DEV float ps_dot4(const float* a, const float* b)
{
    float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2], p3 = a[3] * b[3];
    return (p0 + p1) + (p2 + p3);
}

PS_DEV void ps_norm4(float* a)   // norm_3_4
{
    float r = 1.0f / sqrtf(ps_dot4(a, a));
    a[0] *= r; a[1] *= r; a[2] *= r; a[3] *= r;
}

PS_DEV float ps_dot3(const float* a, const float* b)
{
    float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2];
    return (p0 + p1) + p2;
}

PS_DEV void ps_norm3(float* a)
{
    float r = 1.0f / sqrtf(ps_dot3(a, a));
    a[0] *= r; a[1] *= r; a[2] *= r;
}

// clamp_3_4, vector.cpp:370-383 (maxps then minps; NaN -> lo)
PS_DEV float ps_clamp(float a, float lo, float hi)
{
    float t = (a > lo) ? a : lo;
    return (t < hi) ? t : hi;
}

// opt_pow, proc.h:73-86
PS_DEV float ps_opt_pow(float x, unsigned n)
{
    float pw = 1.0f;
    while (n > 0) {
        if (n & 1) pw *= x;
        x *= x;
        n >>= 1;
    }
    return pw;
}

// ---- isBackFace, vertthrd.cpp:56-65 ----------------------------------
PS_DEV bool ps_back_face(const float* v0, const float* v1, const float* v2)
{
    float a[4], b[4], c[4];
#pragma unroll
    for (int i = 0; i < 4; i++) { a[i] = v1[i] - v0[i]; b[i] = v2[i] - v1[i]; }
    // cross_3, vector.cpp:114-144
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    c[3] = a[3] * b[3] - a[3] * b[3];
    ps_norm4(c);
    const float test[4] = {0.0f, 0.0f, 1.0f, 0.0f};
    return ps_dot4(test, c) < 0.0f;
}

// element i (0..2) of a 3-array by selects: a run-time index into a small
// local array would otherwise send the array to LDS or scratch memory
PS_DEV float ps_sel3(const float* a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

// ---- rasterizer edges (rasterizer.cpp:95-120) --------------------------
struct PsEdge { float dx, dy, x0, y0; int i0, i1; };

PS_DEV PsEdge ps_edge(const float* vx, const float* vy, int i0, int i1)
{
    PsEdge e;
    e.x0 = ps_sel3(vx, i0);
    e.y0 = ps_sel3(vy, i0);
    e.dx = ps_sel3(vx, i1) - e.x0;
    e.dy = ps_sel3(vy, i1) - e.y0;
    e.i0 = i0;
    e.i1 = i1;
    if (fabsf(e.dy) < 0.000001f) e.dy = FLT_MAX;
    return e;
}

PS_DEV float ps_edge_eval(const PsEdge& e, float y) { return e.dx * (y - e.y0) / e.dy + e.x0; }

// Left/right edges of row `row` of a set-up triangle (processTriangle calls of
// pushTriangle / processStandingTriangle, rasterizer.cpp:170-280).  The case
// analysis picks vertex indices only (selects, no branches), and each edge
// is then built once: with an edge built per case the compiler emitted
// divergent branches that moved every field of every case.
PS_DEV void ps_row_edges(const PsTri& t, int row, PsEdge& L, PsEdge& R)
{
    const uint32_t c = t.code;
    const int a = c & 3, b = (c >> 2) & 3, th = (c >> 4) & 3;
    const bool standing = (c & 64u) != 0;
    const bool left = (c & 128u) != 0;    // standing: a = top, b = bottom, th = third
    const bool lower = row <= t.split;    // standing: the lower half, written last
    // flat top/bottom: (a, apex) and (b, apex)
    // standing, lower half: left ? (th, b) | (a, b) : (a, b) | (th, b)
    // standing, upper half: left ? (a, th) | (a, b) : (a, b) | (a, th)
    int l0 = a, l1 = th, r0 = b, r1 = th;
    if (standing) {
        l0 = lower ? (left ? th : a) : a;
        l1 = lower ? b : (left ? th : b);
        r0 = lower ? (left ? a : th) : a;
        r1 = lower ? b : (left ? b : th);
    }
    L = ps_edge(t.vx, t.vy, l0, l1);
    R = ps_edge(t.vx, t.vy, r0, r1);
}

// ---- interpolater (interp.cpp) ----------------------------------------
struct PsSpanGeom {
    float ccL[4], ccR[4];
    int step_count;   // right - left, wrapping
    int skip;         // leftClamped - left, wrapping
};

// lineSegmentlinearInterpolate, interp.cpp:151-160
PS_DEV void ps_line_contrib(const float* vx, const float* vy, int v1, int v2, float x, float y, float* c)
{
    float x1 = ps_sel3(vx, v1), x2 = ps_sel3(vx, v2), y1 = ps_sel3(vy, v1), y2 = ps_sel3(vy, v2);
    float dx = x1 - x2, dy = y1 - y2;
    // one division of the chosen operands
    const bool by_x = fabsf(dx) > fabsf(dy);
    float cv = (by_x ? (x - x2) : (y - y2)) / (by_x ? dx : dy);
    float o = 1.0f - cv;
    // c[v1] = cv, c[v2] = 1 - cv, c[other] = 0
    c[0] = (v1 == 0) ? cv : ((v2 == 0) ? o : 0.0f);
    c[1] = (v1 == 1) ? cv : ((v2 == 1) ? o : 0.0f);
    c[2] = (v1 == 2) ? cv : ((v2 == 2) ? o : 0.0f);
    c[3] = 0.0f;
}

// interp.cpp:26-41: contributions at the (rounded) span ends, times 1/w
PS_DEV void ps_span_geom(const PsTri& t, int row, int left, int right, const PsEdge& L, const PsEdge& R, PsSpanGeom& g)
{
    float cl[4], cr[4];
    ps_line_contrib(t.vx, t.vy, L.i0, L.i1, (float)left, (float)row, cl);
    ps_line_contrib(t.vx, t.vy, R.i0, R.i1, (float)right, (float)row, cr);
    const float rw4[4] = {t.rw[0], t.rw[1], t.rw[2], 0.0f};
#pragma unroll
    for (int i = 0; i < 4; i++) { g.ccL[i] = cl[i] * rw4[i]; g.ccR[i] = cr[i] * rw4[i]; }
    int lc = left < 0 ? 0 : left;
    g.step_count = (int)((unsigned)right - (unsigned)left);
    g.skip = (int)((unsigned)lc - (unsigned)left);
}

// interp.cpp:45-79, depth and correction factor
PS_DEV void ps_span_vis(const PsTri& t, const PsSpanGeom& g, float& cf2, float& cf2st, float& pz, float& pzst)
{
    const float z4[4] = {t.z[0], t.z[1], t.z[2], 0.0f};
    float rlen = 1.0f / (float)g.step_count;
    float pz0 = ps_dot4(g.ccL, z4);
    float pz1 = ps_dot4(g.ccR, z4);
    float pstep = (pz1 - pz0) * rlen;
    float c0 = (g.ccL[0] + g.ccL[1]) + (g.ccL[2] + g.ccL[3]);
    float c1 = (g.ccR[0] + g.ccR[1]) + (g.ccR[2] + g.ccR[3]);
    float cstep = (c1 - c0) * rlen;
    if (g.skip > 0) {
        c0 += cstep * (float)g.skip;
        pz0 += pstep * (float)g.skip;
    }
    cf2 = c0; cf2st = cstep; pz = pz0; pzst = pstep;
}

// IP built-ins (tex1light1.cpp:63-138 and siblings), lane-wise for NV floats
template <int NV>
PS_DEV void ps_span_varyings(const PsSpanGeom& g, const float* u0, const float* u1, const float* u2,
                             float* start, float* step)
{
    float rsc = g.step_count == 0 ? 1.0f : 1.0f / (float)g.step_count;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        float a0 = u0[i] * g.ccL[0], a1 = u1[i] * g.ccL[1], a2 = u2[i] * g.ccL[2];
        float s = (a0 + a1) + a2;
        float b0 = u0[i] * g.ccR[0], b1 = u1[i] * g.ccR[1], b2 = u2[i] * g.ccR[2];
        float e = (b0 + b1) + b2;
        start[i] = s;
        step[i] = (e - s) * rsc;
    }
    if (g.skip > 0) {
        if (g.skip == 1) {
#pragma unroll
            for (int i = 0; i < NV; i++) start[i] = start[i] + step[i];
        } else {
            float n = (float)g.skip;
#pragma unroll
            for (int i = 0; i < NV; i++) start[i] = start[i] + step[i] * n;
        }
    }
}

// the same for a run-time number of varying floats (nv <= NMAX)
template <int NMAX = PS_MAX_NV>
PS_DEV void ps_span_varyings_n(const PsSpanGeom& g, const float* u0, const float* u1, const float* u2, float* start,
                               float* step, int nv)
{
    float rsc = g.step_count == 0 ? 1.0f : 1.0f / (float)g.step_count;
#pragma unroll
    for (int i = 0; i < NMAX; i++) {
        if (i >= nv) break;
        float a0 = u0[i] * g.ccL[0], a1 = u1[i] * g.ccL[1], a2 = u2[i] * g.ccL[2];
        float s = (a0 + a1) + a2;
        float b0 = u0[i] * g.ccR[0], b1 = u1[i] * g.ccR[1], b2 = u2[i] * g.ccR[2];
        float e = (b0 + b1) + b2;
        start[i] = s;
        step[i] = (e - s) * rsc;
    }
    if (g.skip > 0) {
        if (g.skip == 1) {
#pragma unroll
            for (int i = 0; i < NMAX; i++)
                if (i < nv) start[i] = start[i] + step[i];
        } else {
            float n = (float)g.skip;
#pragma unroll
            for (int i = 0; i < NMAX; i++)
                if (i < nv) start[i] = start[i] + step[i] * n;
        }
    }
}

// ---- textures (fbo.cpp:552-594, samplr2d.cpp, samplrcube.cpp) ----------
PS_DEV int ps_lvl_w(const PsTex& t, int l) { int w = t.width >> l; return w < 1 ? 1 : w; }
PS_DEV int ps_lvl_h(const PsTex& t, int l) { int h = t.height >> l; return h < 1 ? 1 : h; }

// floor modulo v - m*floor(v/m), in [0, m), for m >= 1: what clampCoord's
// "v % m, plus m when negative" yields (fbo.cpp:552-594).  For |v| < 2^21 the
// quotient comes from the 1-ulp hardware reciprocal: |v * rcp(m) - v/m| <
// 2^21 * 1.5 * 2^-23 < 1, so the floored estimate is within one of the true
// quotient and a single correction each way makes the remainder exact.
// Larger |v| (and INT_MIN, the conversion of NaN) take the integer remainder.
// Textures have at most 2^22 texels a side (psgpu_texture_create), so m and
// |q| fit the 24-bit multiplier, whose low 32 bits equal the 32-bit product.
PS_DEV int ps_floor_mod(int v, int m, float rm)
{
    if ((unsigned)v + (1u << 21) < (2u << 21)) {
        int q = (int)floorf((float)v * rm);
        int r = v - __mul24(q, m);
        r = r < 0 ? r + m : r;
        return r >= m ? r - m : r;
    }
    int r = v % m;
    return r < 0 ? r + m : r;
}

PS_DEV float ps_rcp_approx(int m) { return __builtin_amdgcn_rcpf((float)m); }

PS_DEV void ps_clamp_coord(int wrap, int max_row, int max_col, int& row, int& col)
{
    if (!wrap) {
        if (row > max_row) row = max_row; else if (row < 0) row = 0;
        if (col > max_col) col = max_col; else if (col < 0) col = 0;
    } else {
        row = max_row > 0 ? ps_floor_mod(row, max_row, ps_rcp_approx(max_row)) : 0;
        col = max_col > 0 ? ps_floor_mod(col, max_col, ps_rcp_approx(max_col)) : 0;
    }
}

// one texel of texture memory (device global memory): a global load, not a
// flat one, addressed as base + 32-bit offset
typedef __attribute__((address_space(1))) const uint8_t ps_gbyte;
typedef __attribute__((address_space(1))) const uint32_t ps_gword;
PS_DEV uint32_t ps_texel(const uint8_t* base, uint32_t off)
{
    return *(ps_gword*)((ps_gbyte*)base + off);
}

PS_DEV uint32_t ps_fetch(const PsTex& t, int layer, int level, int row, int col)
{
    int w = ps_lvl_w(t, level), h = ps_lvl_h(t, level);
    ps_clamp_coord(t.wrap, h - 1, w - 1, row, col);
    int scan = level == 0 ? t.scanline : w * t.elem_len;
    const uint8_t* p = t.data + t.layer_bytes * layer + t.level_off[level] + (int64_t)row * scan + (int64_t)col * 4;
    return *(const uint32_t*)p;
}

// PuresoftSampler2D::get4, samplr2d.cpp:19-25
PS_DEV uint32_t ps_sample_nearest(const PsTex& t, float u, float v)
{
    int row = (int)ps_cvt_u32((float)t.height * v + 0.5f);
    int col = (int)ps_cvt_u32((float)t.width * u + 0.5f);
    return ps_fetch(t, 0, 0, row, col);
}

// PuresoftSamplerCube::get4 + texcoordFromDirection, samplrcube.cpp:29-127
PS_DEV uint32_t ps_sample_cube(const PsTex& t, const float* d)
{
    float X = d[0], Y = d[1], Z = d[2], s, tt;
    int face;
    if (fabsf(X) > fabsf(Y)) {
        if (fabsf(X) > fabsf(Z)) {
            if (X > 0) { s = (-Z / fabsf(X) + 1.0f) / 2.0f; tt = (-Y / fabsf(X) + 1.0f) / 2.0f; face = 0; }
            else { s = (Z / fabsf(X) + 1.0f) / 2.0f; tt = (-Y / fabsf(X) + 1.0f) / 2.0f; face = 1; }
        } else if (Z > 0) { s = (X / fabsf(Z) + 1.0f) / 2.0f; tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; face = 4; }
        else { s = (-X / fabsf(Z) + 1.0f) / 2.0f; tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; face = 5; }
    } else if (fabsf(Z) > fabsf(Y)) {
        if (Z > 0) { s = (X / fabsf(Z) + 1.0f) / 2.0f; tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; face = 4; }
        else { s = (-X / fabsf(Z) + 1.0f) / 2.0f; tt = (-Y / fabsf(Z) + 1.0f) / 2.0f; face = 5; }
    } else if (Z > 0) { s = (X / fabsf(Y) + 1.0f) / 2.0f; tt = (Z / fabsf(Y) + 1.0f) / 2.0f; face = 2; }
    else { s = (X / fabsf(Y) + 1.0f) / 2.0f; tt = (-Z / fabsf(Y) + 1.0f) / 2.0f; face = 3; }
    int row = (int)ps_cvt_u32((float)t.height * s + 0.5f);
    int col = (int)ps_cvt_u32((float)t.width * tt + 0.5f);
    if (face >= t.layers) face = 0;
    return ps_fetch(t, face, 0, row, col);
}

// PuresoftSamplerProjection::get (samplrproj.cpp:20-37): the projected
// coordinate divided by w (div_3_4: times the reciprocal), then four taps
// offset by the Poisson disk / 500, each darkening by 0.2 where the stored
// depth is below the coordinate's z.  The row offsets come from
// POISSON_DISK_X and the column offsets from POISSON_DISK_Y, as there; the
// taps convert with (unsigned int) and read through clampCoord (directRead4).
PS_DEV float ps_sample_projection(const PsTex& t, const float* proj)
{
    const float DX0 = -0.94201624f, DX1 = 0.94558609f, DX2 = -0.094184101f, DX3 = 0.34495938f;
    const float DY0 = -0.39906216f, DY1 = -0.76890725f, DY2 = -0.92938870f, DY3 = 0.29387760f;
    float r = 1.0f / proj[3];
    float tx = proj[0] * r, ty = proj[1] * r, tz = proj[2] * r;
    float f = 1.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float dx = i == 0 ? DX0 : i == 1 ? DX1 : i == 2 ? DX2 : DX3;
        float dy = i == 0 ? DY0 : i == 1 ? DY1 : i == 2 ? DY2 : DY3;
        int row = (int)ps_cvt_u32((float)t.height * (ty + dx / 500.0f));
        int col = (int)ps_cvt_u32((float)t.width * (tx + dy / 500.0f));
        float d = __uint_as_float(ps_fetch(t, 0, 0, row, col));
        if (d < tz) f -= 0.2f;
    }
    return f;
}

// PuresoftFBO::blend4 (fbo.cpp:208-229) of src over dst: per byte, the
// saturating difference src - dst (subs_pu16) times the source alpha (low 16
// bits, mullo_pi16), shifted right by 8, plus dst, packed with unsigned
// saturation (packs_pu16)
PS_DEV uint32_t ps_blend4(uint32_t dst, uint32_t src)
{
    uint32_t a = src >> 24, out = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t s = (src >> (8 * k)) & 255u, d = (dst >> (8 * k)) & 255u;
        uint32_t t = s > d ? s - d : 0u;
        uint32_t v = (((t * a) & 0xffffu) >> 8) + d;
        out |= (v > 255u ? 255u : v) << (8 * k);
    }
    return out;
}

// clampCoord of the two neighbouring coordinates v and v + 1 (the second
// wraps around at INT_MAX like the reference's int arithmetic).  With wrap, the
// second is the first's successor modulo maxv unless v + 1 overflows.
PS_DEV void ps_clamp_pair(int wrap, int maxv, int v, int& a, int& b)
{
    int v1 = (int)((unsigned)v + 1u);
    if (!wrap) {
        a = v > maxv ? maxv : (v < 0 ? 0 : v);
        b = v1 > maxv ? maxv : (v1 < 0 ? 0 : v1);
        return;
    }
    if (maxv <= 0) { a = b = 0; return; }
    float rm = ps_rcp_approx(maxv);
    a = ps_floor_mod(v, maxv, rm);
    if (v != 0x7fffffff) b = a + 1 == maxv ? 0 : a + 1;
    else b = ps_floor_mod(v1, maxv, rm);
}

// bilinear filter: texel j covers u = j/W (the nearest sampler's rounding rule).
// Rows and columns are addressed once each; the four texels share them.  The
// built-in programmes use the B, G, R channels only (NCH = 3).
template <int NCH = 3>
PS_DEV void ps_sample_bilinear(const PsTex& t, int level, float u, float v, float* out)
{
    int w = ps_lvl_w(t, level), h = ps_lvl_h(t, level);
    float fx = (float)w * u, fy = (float)h * v;
    float x0 = floorf(fx), y0 = floorf(fy);
    float ax = fx - x0, ay = fy - y0;
    int ix = ps_cvt_i32(x0), iy = ps_cvt_i32(y0);
    int r0, r1, c0, c1;
    ps_clamp_pair(t.wrap, h - 1, iy, r0, r1);
    ps_clamp_pair(t.wrap, w - 1, ix, c0, c1);
    int scan = level == 0 ? t.scanline : w * t.elem_len;
    // a layer stays below 4 GiB (psgpu_texture_create): 32-bit texel offsets
    const uint32_t lo = (uint32_t)t.level_off[level];
    // rows below 2^22 and rows of fewer than 2^24 bytes (psgpu_texture_create): 24-bit multiplies
    const uint32_t q0 = lo + __umul24((uint32_t)r0, (uint32_t)scan), q1 = lo + __umul24((uint32_t)r1, (uint32_t)scan);
    uint32_t a = ps_texel(t.data, q0 + (uint32_t)c0 * 4u);
    uint32_t b = ps_texel(t.data, q0 + (uint32_t)c1 * 4u);
    uint32_t c = ps_texel(t.data, q1 + (uint32_t)c0 * 4u);
    uint32_t d = ps_texel(t.data, q1 + (uint32_t)c1 * 4u);
#pragma unroll
    for (int k = 0; k < NCH; k++) {
        float fa = (float)((a >> (8 * k)) & 255u), fb = (float)((b >> (8 * k)) & 255u);
        float fc = (float)((c >> (8 * k)) & 255u), fd = (float)((d >> (8 * k)) & 255u);
        float top = fa + (fb - fa) * ax;
        float bot = fc + (fd - fc) * ax;
        out[k] = top + (bot - top) * ay;
    }
}

// log2 via exponent split and the atanh series of log((1+s)/(1-s))
PS_DEV float ps_log2(float x)
{
    if (!(x > 0.0f)) return -INFINITY;
    if (x == INFINITY) return 128.0f;
    uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xffu);
    int bias = 127;
    if (e == 0) {
        x = x * 8388608.0f;
        b = __float_as_uint(x);
        e = (int)((b >> 23) & 0xffu);
        bias = 127 + 23;
    }
    int ex = e - bias;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
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

// mip level from the texel footprint (GL scale factor)
PS_DEV float ps_lod(const float* g, float u, float v, float r, int w0, int h0)
{
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
    return 0.5f * ps_log2(rho2);
}

template <int NCH = 3>
PS_DEV void ps_sample_trilinear(const PsTex& t, float u, float v, float lod, float* out)
{
    int maxl = t.levels - 1;
    if (!(lod > 0.0f) || maxl == 0) { ps_sample_bilinear<NCH>(t, 0, u, v, out); return; }
    if (!(lod < (float)maxl)) { ps_sample_bilinear<NCH>(t, maxl, u, v, out); return; }
    int l0 = (int)lod;
    float f = lod - (float)l0;
    float a[4], b[4];
    ps_sample_bilinear<NCH>(t, l0, u, v, a);
    ps_sample_bilinear<NCH>(t, l0 + 1, u, v, b);
#pragma unroll
    for (int k = 0; k < NCH; k++) out[k] = a[k] + (b[k] - a[k]) * f;
}

// Two textures of one geometry (size, row pitch, texel size, levels, wrap):
// texel coordinates, wrap and addresses are computed once and both are
// fetched and filtered with them; each result is exactly what
// ps_sample_trilinear gives for that texture alone.
PS_DEV bool ps_same_geometry(const PsTex& a, const PsTex& b)
{
    return a.width == b.width && a.height == b.height && a.scanline == b.scanline && a.elem_len == b.elem_len &&
           a.levels == b.levels && a.wrap == b.wrap && a.level_off[1] == b.level_off[1];
}

template <int NCH = 3>
PS_DEV void ps_sample_bilinear_pair(const PsTex& t, const uint8_t* data1, int level, float u, float v, float* out0,
                                    float* out1)
{
    int w = ps_lvl_w(t, level), h = ps_lvl_h(t, level);
    float fx = (float)w * u, fy = (float)h * v;
    float x0 = floorf(fx), y0 = floorf(fy);
    float ax = fx - x0, ay = fy - y0;
    int ix = ps_cvt_i32(x0), iy = ps_cvt_i32(y0);
    int r0, r1, c0, c1;
    ps_clamp_pair(t.wrap, h - 1, iy, r0, r1);
    ps_clamp_pair(t.wrap, w - 1, ix, c0, c1);
    int scan = level == 0 ? t.scanline : w * t.elem_len;
    // a layer stays below 4 GiB (psgpu_texture_create): 32-bit texel offsets
    const uint32_t lo = (uint32_t)t.level_off[level];
    // rows below 2^22 and rows of fewer than 2^24 bytes (psgpu_texture_create): 24-bit multiplies
    const uint32_t q0 = lo + __umul24((uint32_t)r0, (uint32_t)scan), q1 = lo + __umul24((uint32_t)r1, (uint32_t)scan);
    const uint32_t o00 = q0 + (uint32_t)c0 * 4u, o01 = q0 + (uint32_t)c1 * 4u;
    const uint32_t o10 = q1 + (uint32_t)c0 * 4u, o11 = q1 + (uint32_t)c1 * 4u;
    const uint8_t* ds[2] = {t.data, data1};
    float* outs[2] = {out0, out1};
#pragma unroll
    for (int q = 0; q < 2; q++) {
        uint32_t a = ps_texel(ds[q], o00);
        uint32_t b = ps_texel(ds[q], o01);
        uint32_t c = ps_texel(ds[q], o10);
        uint32_t d = ps_texel(ds[q], o11);
#pragma unroll
        for (int k = 0; k < NCH; k++) {
            float fa = (float)((a >> (8 * k)) & 255u), fb = (float)((b >> (8 * k)) & 255u);
            float fc = (float)((c >> (8 * k)) & 255u), fd = (float)((d >> (8 * k)) & 255u);
            float top = fa + (fb - fa) * ax;
            float bot = fc + (fd - fc) * ax;
            outs[q][k] = top + (bot - top) * ay;
        }
    }
}

template <int NCH = 3>
PS_DEV void ps_sample_trilinear_pair(const PsTex& t, const uint8_t* data1, float u, float v, float lod, float* out0,
                                     float* out1)
{
    int maxl = t.levels - 1;
    if (!(lod > 0.0f) || maxl == 0) { ps_sample_bilinear_pair<NCH>(t, data1, 0, u, v, out0, out1); return; }
    if (!(lod < (float)maxl)) { ps_sample_bilinear_pair<NCH>(t, data1, maxl, u, v, out0, out1); return; }
    int l0 = (int)lod;
    float f = lod - (float)l0;
    float a0[4], b0[4], a1[4], b1[4];
    ps_sample_bilinear_pair<NCH>(t, data1, l0, u, v, a0, a1);
    ps_sample_bilinear_pair<NCH>(t, data1, l0 + 1, u, v, b0, b1);
#pragma unroll
    for (int k = 0; k < NCH; k++) {
        out0[k] = a0[k] + (b0[k] - a0[k]) * f;
        out1[k] = a1[k] + (b1[k] - a1[k]) * f;
    }
}

// screen gradients of rw, u*rw, v*rw over a triangle (mip selection)
PS_DEV void ps_tri_grad(const PsTri& t, const float* u0, const float* u1, const float* u2, int iu, float* g)
{
    float e1x = t.vx[1] - t.vx[0], e1y = t.vy[1] - t.vy[0];
    float e2x = t.vx[2] - t.vx[0], e2y = t.vy[2] - t.vy[0];
    float det = e1x * e2y - e2x * e1y;
    float inv = 1.0f / det;
    const float* us[3] = {u0, u1, u2};
    float f[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f[0][i] = t.rw[i];
        f[1][i] = us[i][iu] * t.rw[i];
        f[2][i] = us[i][iu + 1] * t.rw[i];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float d1 = f[a][1] - f[a][0], d2 = f[a][2] - f[a][0];
        g[2 * a] = (d1 * e2y - d2 * e1y) * inv;
        g[2 * a + 1] = (d2 * e1x - d1 * e2x) * inv;
    }
}

// one lane of cvtps2dq, packusdw, packuswb (testproc.cpp:562-568): round to
// nearest even (out of range or NaN -> INT_MIN), saturate to an unsigned word,
// then pack that word *as a signed one* to an unsigned byte, so words of
// 0x8000 and above come out 0.
PS_DEV uint32_t ps_sse_pack_byte(float f)
{
    int v = ps_cvt_i32(rintf(f));
    uint32_t w = v < 0 ? 0u : (v > 65535 ? 65535u : (uint32_t)v);
    return w >= 0x8000u ? 0u : (w > 255u ? 255u : w);
}

PS_DEV uint32_t ps_pack_bgr(float b, float g, float r)
{
    uint32_t ub = (uint32_t)(int)b, ug = (uint32_t)(int)g, ur = (uint32_t)(int)r;
    return ub | (ug << 8) | (ur << 16);
}
