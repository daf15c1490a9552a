// ps_procs.h -- GPU twins of the reference's shader processors.
//
// Each twin is a VP (vertex processor), the varying layout its IP moves along
// a span (NV floats, lane-wise linear like every built-in IP), and an FP.
// Processor sources: tex1light1.cpp (DEF01), colr1light1.cpp (DEF02),
// tex1bump1light1.cpp (DEF03), skybox.cpp (DEF04), shadow.cpp (DEF05); LAMBERT
// and BLINN are two processors without a reference counterpart, added for the
// benchmark scenes; SHADOWED is BLINN lit through the projective shadow
// sampler (samplrproj.cpp); CLOUD and CLOUD_SHADOW are the demo's cloud pair
// (test/testproc.cpp:384-686), shaded inside the raster walk.
#pragma once
#include "ps_device.h"

template <int PROG> struct PsProc;

static_assert(ps_shares_varyings(PSGPU_PROC_DEF01) && ps_shares_varyings(PSGPU_PROC_DEF02) &&
                  ps_shares_varyings(PSGPU_PROC_DEF03) && ps_shares_varyings(PSGPU_PROC_LAMBERT) &&
                  ps_shares_varyings(PSGPU_PROC_BLINN) && ps_shares_varyings(PSGPU_PROC_SHADOWED) &&
                  ps_shares_varyings(PSGPU_PROC_CLOUD) && !ps_shares_varyings(PSGPU_PROC_DEF04) &&
                  !ps_shares_varyings(PSGPU_PROC_DEF05) && !ps_shares_varyings(PSGPU_PROC_CLOUD_SHADOW),
              "the programmes that share varyings across views");


PS_DEV const float* ps_slot(const PsDraw& d, int s, int64_t i)
{
    return (const float*)(d.slot[s] + i * (int64_t)d.unit[s]);
}

PS_DEV void ps_load4(float* r, const float* p)
{
    r[0] = p[0]; r[1] = p[1]; r[2] = p[2]; r[3] = p[3];
}

// the clip position of a sharing programme, the same operations as its VP,
// plus the texcoord its mip-level gradients need (slot 4 into u[iu], u[iu + 1])
PS_DEV void ps_vp_pos(const PsDraw& d, int64_t i, float* pos, float* u, int iu)
{
    float p[4], wp[4];
    ps_load4(p, ps_slot(d, 0, i));
    ps_m4v4(wp, d.M, p);
    ps_m4v4(pos, d.PV, wp);
    if (iu >= 0) {
        const float* tc = ps_slot(d, 4, i);
        u[iu] = tc[0]; u[iu + 1] = tc[1];
    }
}

// common lighting tail of DEF01/02/03 (tex1light1.cpp:321-349)
PS_DEV uint32_t ps_light_def(const PsDraw& d, float* oc, const float* wp, const float* n)
{
    float L[4], E[4], H[4];
#pragma unroll
    for (int k = 0; k < 4; k++) L[k] = d.lightPos[k] - wp[k];
    float dist = sqrtf(ps_dot4(L, L));
    float rd = 1.0f / dist;
#pragma unroll
    for (int k = 0; k < 4; k++) L[k] *= rd;
#pragma unroll
    for (int k = 0; k < 4; k++) E[k] = d.cameraPos[k] - wp[k];
    ps_norm4(E);
#pragma unroll
    for (int k = 0; k < 4; k++) H[k] = E[k] + L[k];
    ps_norm4(H);
    float lambert = ps_dot4(L, n);
    float spec = ps_dot4(H, n);
    spec = spec < 0 ? 0 : spec;
    spec = ps_opt_pow(spec, 50);
    float s255 = 255.0f * spec;
#pragma unroll
    for (int k = 0; k < 4; k++) oc[k] = oc[k] + s255;
#pragma unroll
    for (int k = 0; k < 4; k++) oc[k] = oc[k] * lambert;
#pragma unroll
    for (int k = 0; k < 4; k++) oc[k] = ps_clamp(oc[k], 0.0f, 255.0f);
    return ps_pack_bgr(oc[0], oc[1], oc[2]);
}

PS_DEV void ps_unpack4(uint32_t c, float* o)
{
    o[0] = (float)(c & 255u); o[1] = (float)((c >> 8) & 255u);
    o[2] = (float)((c >> 16) & 255u); o[3] = (float)(c >> 24);
}

// VP part shared by DEF01/02/03: worldPos = M p; clip = PV (M p); worldPos.w = 0
PS_DEV void ps_vp_world(const PsDraw& d, int64_t i, float* pos, float* wp)
{
    float p[4];
    ps_load4(p, ps_slot(d, 0, i));
    ps_m4v4(wp, d.M, p);
    float t[4] = {wp[0], wp[1], wp[2], wp[3]};
    wp[3] = 0.0f;
    ps_m4v4(pos, d.PV, t);
}

// ---- DEF01: tex1light1 -------------------------------------------------
// varyings: normal[4], worldPos[4], texcoord[2]
template <> struct PsProc<PSGPU_PROC_DEF01> {
    static constexpr int NV = 10;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float wp[4], n[4], nn[4];
        ps_vp_world(d, i, pos, wp);
        ps_load4(n, ps_slot(d, 3, i));
        ps_m4v4(nn, d.Mrot, n);
        const float* tc = ps_slot(d, 4, i);
#pragma unroll
        for (int k = 0; k < 4; k++) { u[k] = nn[k]; u[4 + k] = wp[k]; }
        u[8] = tc[0]; u[9] = tc[1];
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        float oc[4];
        ps_unpack4(ps_sample_nearest(d.tex0, in[8], in[9]), oc);
        out = ps_light_def(d, oc, in + 4, in);
        return true;
    }
};

// ---- DEF02: colr1light1 ------------------------------------------------
// varyings: normal[4], worldPos[4], colour[4]
template <> struct PsProc<PSGPU_PROC_DEF02> {
    static constexpr int NV = 12;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float wp[4], n[4], nn[4], c[4];
        ps_vp_world(d, i, pos, wp);
        ps_load4(n, ps_slot(d, 1, i));
        ps_m4v4(nn, d.Mrot, n);
        ps_load4(c, ps_slot(d, 2, i));
#pragma unroll
        for (int k = 0; k < 4; k++) { u[k] = nn[k]; u[4 + k] = wp[k]; u[8 + k] = c[k]; }
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        float oc[4] = {in[8], in[9], in[10], in[11]};
        out = ps_light_def(d, oc, in + 4, in);
        return true;
    }
};

// ---- DEF03: tex1bump1light1 --------------------------------------------
// varyings: tangent[4], binormal[4], normal[4], worldPos[4], texcoord[2]
template <> struct PsProc<PSGPU_PROC_DEF03> {
    static constexpr int NV = 18;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float wp[4], a[4], r[4];
        ps_vp_world(d, i, pos, wp);
#pragma unroll
        for (int s = 0; s < 3; s++) {
            ps_load4(a, ps_slot(d, 1 + s, i));
            ps_m4v4(r, d.Mrot, a);
#pragma unroll
            for (int k = 0; k < 4; k++) u[4 * s + k] = r[k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) u[12 + k] = wp[k];
        const float* tc = ps_slot(d, 4, i);
        u[16] = tc[0]; u[17] = tc[1];
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        float oc[4];
        ps_unpack4(ps_sample_nearest(d.tex0, in[16], in[17]), oc);
        uint32_t bb = ps_sample_nearest(d.tex1, in[16], in[17]);
        float bump[4] = {(float)((bb >> 16) & 255u), (float)((bb >> 8) & 255u), (float)(bb & 255u), 0.0f};
        float rc = 1.0f / 255.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) bump[k] *= rc;
#pragma unroll
        for (int k = 0; k < 4; k++) bump[k] *= 2.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) bump[k] -= 1.0f;
        float tbn[16];
#pragma unroll
        for (int k = 0; k < 12; k++) tbn[k] = in[k];
        tbn[12] = tbn[13] = tbn[14] = tbn[15] = 0.0f;
        float n[4];
        ps_m4v4(n, tbn, bump);
        ps_norm4(n);
        out = ps_light_def(d, oc, in + 12, n);
        return true;
    }
};

// ---- DEF04: skybox -------------------------------------------------------
// varyings: direction[3] (the FP reads x, y, z)
template <> struct PsProc<PSGPU_PROC_DEF04> {
    static constexpr int NV = 3;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float p[4], v[4], r[4];
        ps_load4(p, ps_slot(d, 0, i));
        v[0] = p[0] - 0.0f; v[1] = p[1] - 0.0f; v[2] = p[2] - 1.0f; v[3] = p[3] - 0.0f;
        v[3] = 0.0f;
        ps_norm4(v);
        ps_m4v4(r, d.invView, v);
        u[0] = r[0]; u[1] = r[1]; u[2] = r[2];
        pos[0] = p[0]; pos[1] = p[1]; pos[2] = p[2]; pos[3] = p[3];
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        out = ps_sample_cube(d.tex0, in);
        return true;
    }
};

// ---- DEF05: shadow (depth only) ----------------------------------------
template <> struct PsProc<PSGPU_PROC_DEF05> {
    static constexpr int NV = 0;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float p[4];
        ps_load4(p, ps_slot(d, 0, i));
        p[0] *= 0.9f; p[1] *= 0.9f; p[2] *= 0.9f;
        ps_m4v4(pos, d.PVM, p);
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out) { return false; }
};

// ---- LAMBERT: bilinear texture x Lambert, directional light ------------
// varyings: normal[3], texcoord[2]
template <> struct PsProc<PSGPU_PROC_LAMBERT> {
    static constexpr int NV = 5;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float p[4], wp[4], n[4], nn[4];
        ps_load4(p, ps_slot(d, 0, i));
        ps_m4v4(wp, d.M, p);
        ps_m4v4(pos, d.PV, wp);
        ps_load4(n, ps_slot(d, 3, i));
        ps_m4v4(nn, d.Mrot, n);
        const float* tc = ps_slot(d, 4, i);
        u[0] = nn[0]; u[1] = nn[1]; u[2] = nn[2]; u[3] = tc[0]; u[4] = tc[1];
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        float t[4];
        ps_sample_bilinear(d.tex0, 0, in[3], in[4], t);
        float n[3] = {in[0], in[1], in[2]};
        ps_norm3(n);
        float lam = ps_dot3(n, d.lightDir);
        lam = lam > 0.0f ? lam : 0.0f;
        out = ps_pack_bgr(ps_clamp(t[0] * lam, 0.0f, 255.0f), ps_clamp(t[1] * lam, 0.0f, 255.0f),
                          ps_clamp(t[2] * lam, 0.0f, 255.0f));
        return true;
    }
};

// ---- BLINN: Blinn-Phong + tangent-space normal map, trilinear ----------
// varyings: tangent[3], binormal[3], normal[3], worldPos[3], texcoord[2]
// (SHADOWED adds the world position in shadow-map clip space[4])
PS_DEV void ps_blinn_vp(const PsDraw& d, int64_t i, float* pos, float* u, float* wp)
{
    float p[4], a[4], r[4];
    ps_load4(p, ps_slot(d, 0, i));
    ps_m4v4(wp, d.M, p);
    ps_m4v4(pos, d.PV, wp);
#pragma unroll
    for (int s = 0; s < 3; s++) {
        ps_load4(a, ps_slot(d, 1 + s, i));
        ps_m4v4(r, d.Mrot, a);
        u[3 * s] = r[0]; u[3 * s + 1] = r[1]; u[3 * s + 2] = r[2];
    }
    u[9] = wp[0]; u[10] = wp[1]; u[11] = wp[2];
    const float* tc = ps_slot(d, 4, i);
    u[12] = tc[0]; u[13] = tc[1];
}

// the Blinn-Phong colour, the Lambert term scaled by `shadow`
PS_DEV bool ps_blinn_fp(const PsDraw& d, const float* in, float r, const float* g, float shadow, uint32_t& out)
{
    const float* T = in;
    const float* B = in + 3;
    const float* N = in + 6;
    const float* wp = in + 9;
    float u = in[12], v = in[13];
    float lod = ps_lod(g, u, v, r, d.tex0.width, d.tex0.height);
    float alb[4], nm[4];
    if (ps_same_geometry(d.tex0, d.tex1)) {
        ps_sample_trilinear_pair(d.tex0, d.tex1.data, u, v, lod, alb, nm);
    } else {
        ps_sample_trilinear(d.tex0, u, v, lod, alb);
        ps_sample_trilinear(d.tex1, u, v, lod, nm);
    }
    const float k2 = 2.0f / 255.0f;
    float bx = nm[2] * k2 - 1.0f, by = nm[1] * k2 - 1.0f, bz = nm[0] * k2 - 1.0f;
    float n[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float a = bx * T[k], b = by * B[k], c = bz * N[k];
        n[k] = (a + b) + c;
    }
    ps_norm3(n);
    float L[3], E[3], H[3];
#pragma unroll
    for (int k = 0; k < 3; k++) L[k] = d.lightPos[k] - wp[k];
    float dist = sqrtf(ps_dot3(L, L));
    float rd = 1.0f / dist;
#pragma unroll
    for (int k = 0; k < 3; k++) L[k] *= rd;
#pragma unroll
    for (int k = 0; k < 3; k++) E[k] = d.cameraPos[k] - wp[k];
    ps_norm3(E);
#pragma unroll
    for (int k = 0; k < 3; k++) H[k] = E[k] + L[k];
    ps_norm3(H);
    float lam = ps_dot3(L, n);
    lam = lam * shadow;   // x 1.0f (Blinn-Phong) is exact
    float spec = ps_dot3(H, n);
    spec = spec < 0 ? 0 : spec;
    spec = ps_opt_pow(spec, (unsigned)d.spec_exp);
    float s255 = 255.0f * spec;
    out = ps_pack_bgr(ps_clamp((alb[0] + s255) * lam, 0.0f, 255.0f), ps_clamp((alb[1] + s255) * lam, 0.0f, 255.0f),
                      ps_clamp((alb[2] + s255) * lam, 0.0f, 255.0f));
    return true;
}

template <> struct PsProc<PSGPU_PROC_BLINN> {
    static constexpr int NV = 14;
    static constexpr bool GRAD = true;
    static constexpr int IU = 12;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float wp[4];
        ps_blinn_vp(d, i, pos, u, wp);
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        return ps_blinn_fp(d, in, r, g, 1.0f, out);
    }
};

// ---- SHADOWED: BLINN lit through a projective shadow map ---------------
template <> struct PsProc<PSGPU_PROC_SHADOWED> {
    static constexpr int NV = 18;
    static constexpr bool GRAD = true;
    static constexpr int IU = 12;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float wp[4], sc[4];
        ps_blinn_vp(d, i, pos, u, wp);
        ps_m4v4(sc, d.shadowPV, wp);
        u[14] = sc[0]; u[15] = sc[1]; u[16] = sc[2]; u[17] = sc[3];
    }
    PS_DEV static bool fp(const PsDraw& d, const float* in, float r, const float* g, uint32_t& out)
    {
        return ps_blinn_fp(d, in, r, g, ps_sample_projection(d.tex2, in + 14), out);
    }
};

// ---- programmes shaded inside the raster walk (blending, discard) ------
// fpi() returns 1 with a colour to write, 2 to discard the fragment (no depth
// update, FBOBridge::discard), 0 for neither.

// CLOUD: VP_Cloud + FP_Cloud (test/testproc.cpp:384-420, 534-574): white
// with the texel's red as alpha, times 2 (L . n) and the shadow factor,
// packed the way the reference's SSE tail packs it.
// varyings: normal[4], worldPos[4] (w = 0), texcoord[2], shadowcoord[4]
template <> struct PsProc<PSGPU_PROC_CLOUD> {
    static constexpr int NV = 14;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float p[4], wp[4], sc[4], n[4];
        ps_load4(p, ps_slot(d, 0, i));
        ps_m4v4(wp, d.M, p);
        ps_m4v4(sc, d.shadowPV, wp);
        ps_m4v4(pos, d.PV, wp);
        ps_load4(p, ps_slot(d, 3, i));
        ps_m4v4(n, d.Mrot, p);
        const float* tc = ps_slot(d, 4, i);
#pragma unroll
        for (int k = 0; k < 4; k++) { u[k] = n[k]; u[10 + k] = sc[k]; }
        u[4] = wp[0]; u[5] = wp[1]; u[6] = wp[2]; u[7] = 0.0f;
        u[8] = tc[0]; u[9] = tc[1];
    }
    PS_DEV static int fpi(const PsDraw& d, const float* in, uint32_t& out)
    {
        uint32_t t = ps_sample_nearest(d.tex0, in[8], in[9]);
        float shadow = ps_sample_projection(d.tex2, in + 10);
        float L[4];
#pragma unroll
        for (int k = 0; k < 4; k++) L[k] = d.lightPos[k] - in[4 + k];
        float rd = 1.0f / sqrtf(ps_dot4(L, L));
#pragma unroll
        for (int k = 0; k < 4; k++) L[k] *= rd;
        float lambert = 2.0f * ps_dot4(L, in);
        uint32_t c = ps_sse_pack_byte(255.0f * (lambert * shadow));
        out = c | (c << 8) | (c << 16) | (((t >> 16) & 255u) << 24);
        return 1;
    }
};

// CLOUD_SHADOW: VP_CloudShadow + FP_CloudShadow (test/testproc.cpp:595-609, 677-686)
template <> struct PsProc<PSGPU_PROC_CLOUD_SHADOW> {
    static constexpr int NV = 2;
    static constexpr bool GRAD = false;
    PS_DEV static void vp(const PsDraw& d, int64_t i, float* pos, float* u)
    {
        float p[4];
        ps_load4(p, ps_slot(d, 0, i));
        p[0] *= 0.95f; p[1] *= 0.95f; p[2] *= 0.95f;
        ps_m4v4(pos, d.PVM, p);
        const float* tc = ps_slot(d, 4, i);
        u[0] = tc[0]; u[1] = tc[1];
    }
    PS_DEV static int fpi(const PsDraw& d, const float* in, uint32_t& out)
    {
        uint32_t t = ps_sample_nearest(d.tex0, in[0], in[1]);
        return ((t >> 16) & 255u) < 150u ? 2 : 0;
    }
};
