/*
 * ps_oracle.h -- CPU oracle for the Puresoft3D per-frame raster path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product path (libpsgpu.so) never links or calls it.
 *
 * It restates, in plain C with source-order IEEE FP32 arithmetic and no FMA
 * contraction, the algorithm of the reference files
 *   src/puresoft3d/drawvao.cpp, vertthrd.cpp, rasterizer.cpp, interp.cpp,
 *   fragthrd.cpp, fbo.cpp, samplr2d.cpp, samplrcube.cpp, pipeline.cpp, proc.h,
 *   tex1light1.cpp, colr1light1.cpp, tex1bump1light1.cpp, skybox.cpp, shadow.cpp
 *   src/mcemath/vector.cpp, matrix.cpp
 * Every function in ps_oracle.c cites the file:line it follows.
 *
 * Two deliberate, documented substitutions (the reference's own results
 * depend on the CPU vendor there):
 *   rcpps(x)   -> 1.0f / x          (vector.cpp:158-194, matrix.cpp:473-476)
 *   rsqrtss(x) -> 1.0f / sqrtf(x)   (vector.cpp:225-250)
 *
 * Parity status of each function is listed in the header comment of
 * ps_oracle.c ("pinned by" lines).
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_SLOTS     16    /* config.h:4  MAX_VBOS      */
#define OR_MAX_UNIFORMS  1024  /* config.h:5  MAX_UNIFORMS  */

/* behaviour bits, pipeline.h:17-20 */
#define OR_UPDATE_DEPTH  0x1
#define OR_TEST_DEPTH    0x2
#define OR_FACE_CULLING  0x4
#define OR_ALPHABLEND    0x8

/* programme kinds (one VP/IP/FP triple each) */
#define OR_PROG_DEF01 1   /* tex1light1.cpp      */
#define OR_PROG_DEF02 2   /* colr1light1.cpp     */
#define OR_PROG_DEF03 3   /* tex1bump1light1.cpp */
#define OR_PROG_DEF04 4   /* skybox.cpp          */
#define OR_PROG_DEF05 5   /* shadow.cpp          */
#define OR_PROG_LAMBERT 6 /* Lambert x bilinear texture, directional light  */
#define OR_PROG_BLINN   7 /* Blinn-Phong + normal map, trilinear, point light */
#define OR_PROG_CLOUD   9 /* VP_Cloud + FP_Cloud (test/testproc.cpp:384-420,526-574): white, alpha = the nearest
                             texel's red, times 2 (L . n) and the projective shadow factor (uniforms 15, 16);
                             blend4 under OR_ALPHABLEND */
#define OR_PROG_CLOUD_SHADOW 10 /* VP_CloudShadow + FP_CloudShadow (test/testproc.cpp:589-609,672-686):
                                  depth only (0.95 shrink), discarded where the nearest texel's red < 150 */
#define OR_PROG_SHADOWED 8 /* OR_PROG_BLINN with the Lambert term times the projective shadow sampler
                              (uniform 12: the light's PV, 16 floats; uniform 13: FP32 shadow-map texture) */

typedef struct {
    int width, height;     /* level-0 texel dims                         */
    int scanline;          /* level-0 bytes per row                      */
    int elem_len;          /* 1 or 4 (fbo.cpp:21-24)                     */
    int layers;            /* 1 + extraLayers                            */
    int wrap;              /* 0 CLAMP, 1 WRAP (fbo.h:18)                 */
    int levels;            /* mip levels, >= 1                           */
    const uint8_t* data;   /* layer-major; levels in order inside a layer;
                              levels >= 1 are tightly packed             */
} or_texture;

typedef struct {
    int width, height;             /* viewport == target size            */
    uint8_t* colour;               /* BGRA8                              */
    int colour_scanline;           /* bytes                              */
    int colour_top_down;           /* 1 for the display FBO (pipeline.cpp:48) */
    float* depth;
    int depth_scanline;            /* in floats                          */
    /* optional per-pixel debug planes, index y*width+x (display rows), or NULL */
    int32_t*  dbg_last_pass;       /* global triangle id of last depth-test pass, -1 none */
    int32_t*  dbg_raster_count;    /* fragments visited by the pixel loop */
    uint32_t* dbg_raster_hash;     /* xor of or_tri_hash(tri) over visited fragments */
    int32_t*  dbg_pass_count;      /* depth-test passes (= FP invocations) */
    uint32_t* dbg_pass_hash;       /* xor of or_tri_hash(tri) over passes */
} or_target;

typedef struct {
    int prog;                                   /* OR_PROG_*             */
    int behavior;                               /* OR_* bits             */
    const void* slot_data[OR_MAX_SLOTS];        /* VBO contents          */
    int slot_unit_bytes[OR_MAX_SLOTS];
    int slot_count[OR_MAX_SLOTS];               /* 0 = slot not attached */
    const void* uniforms[OR_MAX_UNIFORMS];      /* uniform slot bytes    */
    const or_texture* textures;                 /* texture pool          */
    int n_textures;
    int tri_base;                               /* id of this draw's first triangle */
    int shade_fp64;                             /* 1: FP64 shading variant; 2: FP32 with x86 rcpps/rsqrtss
                                                   in DEF01-DEF04's div_3_4 and norm_3_4; 3, 4: those
                                                   reciprocals at +/- 1.5 * 2^-12 relative error */
} or_draw;

typedef struct {
    int64_t tris_in, tris_culled, tris_zrejected, tris_dropped;
    int64_t spans, frags_rasterised, frags_shaded;
} or_stats;

/* Full draw (drawvao.cpp:3-133).  n_threads >= 2 uses the reference's
 * threading shape: the caller thread produces row tasks, n_threads-1 workers
 * consume them in per-row FIFO order (drawvao.cpp:77-85, fragthrd.cpp).
 * n_threads <= 1 runs everything on the caller thread.  Results do not depend
 * on n_threads.  Returns 0 or a negative error code. */
int or_draw_vao(const or_draw* d, or_target* t, or_stats* st, int n_threads);

/* stop the worker threads of the threaded mode (they spin while idle, like
 * the reference's fragment threads) */
void or_pool_shutdown(void);

/* clearDepth -> clear16 over the whole buffer (pipeline.cpp:334-338, fbo.cpp:348-371) */
void or_clear_depth(or_target* t, float value);
/* clearColour -> clear4, which skips the last memory row (pipeline.cpp:340-343, fbo.cpp:332-346) */
void or_clear_colour(or_target* t, uint32_t bgra);

/* Rasterizer restatement for one triangle (rasterizer.cpp:142-234).
 * Inputs are NDC float4 positions (after the perspective divide).
 * rows_out must hold height*8 ints: per row {left, leftClamped, lv0, lv1,
 * right, rightClamped, rv0, rv1}.  Only rows written by this triangle are
 * touched; callers pre-fill with a sentinel.  verts_out gets 6 floats.
 * Returns 1 when pushTriangle returns true, 0 otherwise. */
int or_push_triangle(int width, int height, const float* v0, const float* v1,
                     const float* v2, int* first_row, int* last_row,
                     int* rows_out, float* verts_out);

/* Span interpolation setup for cf2/projZ (interp.cpp:26-80), exposed for tests. */
void or_interp_setup(const float* verts6, const float* rw3, const float* z3, int row,
                     int left, int right, int lv0, int lv1, int rv0, int rv1,
                     float* cf2_start, float* cf2_step, float* pz_start, float* pz_step);

/* Back-face test (vertthrd.cpp:56-65) on NDC float4 positions. */
int or_is_back_face(const float* v0, const float* v1, const float* v2);

/* 2x2 box-filter mip chain for a BGRA8 / R8 texture with a tight level-0
 * scanline: fills levels 1..levels-1 of `data` in place (rounded mean). */
void or_build_mips(uint8_t* data, int width, int height, int elem_len, int layers, int levels);

/* Samplers, exposed for tests.  out4 receives 4 floats (b,g,r,a). */
void or_sample_nearest(const or_texture* tex, int layer, float u, float v, uint8_t* out4);
void or_sample_cube(const or_texture* tex, const float* dir, uint8_t* out4);
void or_sample_bilinear(const or_texture* tex, int layer, int level, float u, float v, float* out4);
float or_log2_approx(float x);
/* PuresoftSamplerProjection::get (samplrproj.cpp:20-37) over an FP32 texture:
 * 1 minus 0.2 per Poisson tap whose stored depth is below the projected z */
float or_sample_projection(const or_texture* tex, const float* proj4);
/* PuresoftFBO::blend4 (fbo.cpp:208-229) of a BGRA texel onto dst, in place */
void or_blend4(uint8_t* dst, const uint8_t* src);
/* one lane of FP_Cloud's cvtps2dq/packusdw/packuswb tail (testproc.cpp:562-568) */
uint8_t or_sse_pack_byte(float f);

/* Hash used by the debug planes. */
uint32_t or_tri_hash(uint32_t tri);

/* Differentiable shading on a geometry buffer (12 floats per pixel: normal
 * xyz, covered flag | world position xyz, 0 | u, v, 0, 0), in FP64: the
 * reference of psgpu_diff_forward / psgpu_diff_backward (no counterpart in
 * Puresoft3D).  tex: th rows of tw RGB float texels, bilinear with texel j at
 * u = j/tw, wrapping.  out: W*H*3.  gtex: tw*th*3 and gparams[2] are
 * accumulated (dLoss/dtex, dLoss/d{albedo_scale, spec_exp}) from gout. */
void or_diff_forward_f64(int W, int H, const float* gbuf, int tw, int th, const float* tex, double albedo_scale,
                         double spec_exp, double ks, const float* light, const float* camera, double* out);
void or_diff_backward_f64(int W, int H, const float* gbuf, int tw, int th, const float* tex, double albedo_scale,
                          double spec_exp, double ks, const float* light, const float* camera, const double* gout,
                          double* gtex, double* gparams);

#ifdef __cplusplus
}
#endif
#endif
