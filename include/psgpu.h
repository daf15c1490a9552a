/*
 * psgpu.h -- C ABI of the MI355X (gfx950) raster hot path for Puresoft3D.
 *
 * Every entry point replaces one member of the reference's C++ pipeline API
 * (src/puresoft3d/pipeline.h:30-66 and the plugin types in proc.h, vbo.h,
 * vao.h, fbo.h).  The reference throws C++ exceptions; here every call
 * returns PSGPU_OK (0) or a negative status, and psgpu_last_error() gives
 * the message.  Handles are small integers, as in the reference's pools.
 *
 * Ordering: calls on one context are stream-ordered.  Draws are queued and
 * executed in submission order when the queue is flushed (explicitly by
 * psgpu_flush/psgpu_sync, or implicitly before any call that reads or
 * replaces a buffer the queued draws use).  Reads (psgpu_read_*) return
 * results identical to the reference's synchronous drawVAO (drawvao.cpp:3-133).
 * A context is used by one host thread at a time.
 */
#ifndef PSGPU_H
#define PSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PSGPU_OK                 0
#define PSGPU_ERR_INVALID       -1   /* std::invalid_argument in the reference */
#define PSGPU_ERR_RANGE         -2   /* std::out_of_range                      */
#define PSGPU_ERR_NOMEM         -3   /* std::bad_alloc                         */
#define PSGPU_ERR_DEVICE        -4   /* HIP runtime error                      */
#define PSGPU_ERR_STATE         -5   /* call not valid in the current state    */
#define PSGPU_ERR_UNSUPPORTED   -6   /* processor without a GPU twin, etc.     */

/* ---- behaviour bits (pipeline.h:17-20) ---- */
#define PSGPU_UPDATE_DEPTH  0x1
#define PSGPU_TEST_DEPTH    0x2
#define PSGPU_FACE_CULLING  0x4
#define PSGPU_ALPHABLEND    0x8

/* ---- processor kinds with a GPU twin (one VP/IP/FP triple each) ---- */
#define PSGPU_PROC_DEF01    1   /* tex1light1.cpp      */
#define PSGPU_PROC_DEF02    2   /* colr1light1.cpp     */
#define PSGPU_PROC_DEF03    3   /* tex1bump1light1.cpp */
#define PSGPU_PROC_DEF04    4   /* skybox.cpp          */
#define PSGPU_PROC_DEF05    5   /* shadow.cpp          */
#define PSGPU_PROC_LAMBERT  6   /* Lambert x bilinear texture, directional light (slot 7 = light dir, 9 = tex) */
#define PSGPU_PROC_BLINN    7   /* Blinn-Phong + normal map, trilinear (7 light, 8 camera, 9 albedo, 10 normal map, 11 exponent) */
#define PSGPU_PROC_CLOUD    9   /* VP_Cloud + FP_Cloud (test/testproc.cpp:384-420, 526-574): white with the
                                   nearest texel's red as alpha, times 2 (L . n) and the projective shadow
                                   factor, blended by blend4 (fbo.cpp:208-229) under PSGPU_ALPHABLEND
                                   (3 PV, 4 M, 5 Mrot, 7 light, 9 texture, 15 shadow map, 16 the light's PV) */
#define PSGPU_PROC_CLOUD_SHADOW 10 /* VP_CloudShadow + FP_CloudShadow (test/testproc.cpp:589-609, 672-686):
                                   depth only, the position shrunk by 0.95, the fragment discarded (no depth
                                   write) where the nearest texel's red is below 150 (3 PV, 4 M, 9 texture) */
#define PSGPU_PROC_SHADOWED 8   /* PSGPU_PROC_BLINN lit through a shadow map: the Lambert term times
                                   PuresoftSamplerProjection::get (samplrproj.cpp:20-37) of the world position
                                   projected by uniform 12 (16 floats, the light's PV) into the FP32 depth
                                   texture named by uniform 13 (rendered by a PSGPU_PROC_DEF05 pass) */

/* ---- wrap modes (fbo.h:18) ---- */
#define PSGPU_CLAMP 0
#define PSGPU_WRAP  1

/* ---- render-target selectors ---- */
#define PSGPU_DISPLAY  (-1)     /* setFBO(0, NULL): the display buffer (pipeline.cpp:267-270) */
#define PSGPU_DEFAULT_DEPTH (-1) /* setDepth(-1) (pipeline.cpp:218-223)                       */

typedef struct psgpu_ctx psgpu_ctx;

typedef struct {
    int64_t tris_in, tris_culled, tris_zrejected, tris_dropped;
    int64_t spans, frags_rasterised, frags_shaded;
    int64_t segments;           /* span pieces after splitting at bin-chunk borders */
    int64_t flushes, retries;   /* pipeline executions, capacity-growth re-runs      */
    int64_t splits;             /* flushes split in two for the per-flush span limit */
} psgpu_stats;

/* tile configuration of the raster kernel; searched by the autotuner */
typedef struct {
    int chunk_width;    /* pixels per lane along a row: 1, 2, 4, 8, 16, 32 or 64 */
    int tile_rows;      /* rows per wave tile (tile_rows * lanes-per-row = 64): 1..64, power of two */
    int waves_per_group;/* wavefronts per workgroup of the raster kernel: 1..8 */
} psgpu_tiling;

/* ---- context: PuresoftPipeline ctor/dtor (pipeline.cpp:24-116) ---- */
/* flags: PSGPU_CREATE_TINY_SCRATCH starts the intermediate buffers at a few
 * hundred entries, so the first flushes overflow and are replayed (tests). */
#define PSGPU_CREATE_TINY_SCRATCH 0x1
/* PSGPU_CREATE_SMALL_FLUSHES lowers the most spans one flush may hold from
 * 2^26 to 4,096 (tests): a flush that needs more is split into flushes of
 * about half its triangles, re-run in order, with the same result. */
#define PSGPU_CREATE_SMALL_FLUSHES 0x2
/* Display and viewport sizes are limited to PSGPU_MAX_VIEW_WIDTH x
 * PSGPU_MAX_VIEW_HEIGHT (intermediate records keep a column in 16 bits and a
 * row in 24); larger ones give PSGPU_ERR_RANGE. */
#define PSGPU_MAX_VIEW_WIDTH  65535
#define PSGPU_MAX_VIEW_HEIGHT 16777215
int psgpu_create(int device, int width, int height, unsigned flags, psgpu_ctx** out);
int psgpu_destroy(psgpu_ctx* ctx);
const char* psgpu_last_error(psgpu_ctx* ctx);
/* the HIP stream the context runs on (hipStream_t), for callers that share it */
void* psgpu_stream(psgpu_ctx* ctx);

/* ---- textures: createTexture / destroyTexture / getTexture (tex.cpp:4-59) ---- */
/* pixels may be NULL (contents then undefined, as with image->pixels == NULL).
 * mip_levels >= 1; levels beyond 0 are filled by psgpu_texture_generate_mips.
 * width and height at most 2^22 texels and scanline below 2^24 bytes
 * (PSGPU_ERR_RANGE beyond: texel addressing uses 24-bit multiplies). */
int psgpu_texture_create(psgpu_ctx* ctx, int width, int scanline, int height, int elem_len,
                         int extra_layers, int wrap_mode, int mip_levels, const void* pixels, int* out_idx);
int psgpu_texture_upload(psgpu_ctx* ctx, int tex, int layer, int level, const void* pixels, size_t bytes);
int psgpu_texture_download(psgpu_ctx* ctx, int tex, int layer, int level, void* pixels, size_t bytes);
int psgpu_texture_generate_mips(psgpu_ctx* ctx, int tex);
int psgpu_texture_destroy(psgpu_ctx* ctx, int tex);
/* width, scanline, height, elem_len of one layer/level (getTexture) */
int psgpu_texture_info(psgpu_ctx* ctx, int tex, int level, int* width, int* scanline, int* height, int* elem_len);

/* ---- vertex streams: PuresoftVBO / PuresoftVAO (vbo.cpp, vao.cpp, pipeline.cpp:118-205) ---- */
int psgpu_vbo_create(psgpu_ctx* ctx, size_t unit_bytes, size_t unit_count, int* out_vbo);
int psgpu_vbo_update(psgpu_ctx* ctx, int vbo, const void* src);            /* updateContent */
int psgpu_vbo_destroy(psgpu_ctx* ctx, int vbo);
int psgpu_vao_create(psgpu_ctx* ctx, int* out_vao);
int psgpu_vao_attach(psgpu_ctx* ctx, int vao, int slot, int vbo, int* out_previous);  /* -1: none */
int psgpu_vao_detach(psgpu_ctx* ctx, int vao, int slot, int* out_previous);
int psgpu_vao_get(psgpu_ctx* ctx, int vao, int slot, int* out_vbo);
int psgpu_vao_destroy(psgpu_ctx* ctx, int vao);   /* also destroys attached VBOs (pipeline.cpp:183-205) */

/* ---- programmes: createProgramme / useProgramme (prog.cpp:43-132) ---- */
int psgpu_program_create(psgpu_ctx* ctx, int vp_kind, int ip_kind, int fp_kind, int* out_prog);
int psgpu_program_destroy(psgpu_ctx* ctx, int prog);
int psgpu_program_use(psgpu_ctx* ctx, int prog);

/* ---- state (pipeline.cpp:207-343) ---- */
int psgpu_uniform_set(psgpu_ctx* ctx, int slot, const void* data, size_t len);   /* NULL clears */
int psgpu_behavior_enable(psgpu_ctx* ctx, int bits);
int psgpu_behavior_disable(psgpu_ctx* ctx, int bits);
int psgpu_behavior_get(psgpu_ctx* ctx, int* bits);
int psgpu_viewport(psgpu_ctx* ctx, int width, int height);
/* setFBO (pipeline.cpp:260-275): slot 0 takes a texture or PSGPU_DISPLAY; slots
 * 1..15 take a texture or PSGPU_DISPLAY for none, and are only recorded: every
 * processor (here and in the reference) writes slot 0 */
int psgpu_target_set(psgpu_ctx* ctx, int fbo_slot, int tex_or_display);
int psgpu_depth_set(psgpu_ctx* ctx, int tex_or_default);                   /* setDepth */
/* geometry buffer of the visible fragments (no counterpart in the reference):
 * while set, Blinn-Phong draws also write, for every pixel whose colour they
 * produce, 12 floats at device_ptr + (y*vpW + x)*48 bytes (rows bottom-up):
 * interpolated normal xyz, 1 | world position xyz, 0 | u, v, 0, 0.  Pixels
 * no draw colours are left as they were.  NULL stops it.  The adjoint entry
 * points below read this buffer: it fixes visibility for the backward pass. */
int psgpu_target_gbuffer(psgpu_ctx* ctx, void* device_ptr, size_t bytes);
int psgpu_clear_depth(psgpu_ctx* ctx, float furthest);                     /* clear16 */
int psgpu_clear_colour(psgpu_ctx* ctx, uint32_t bgra);                     /* clear4 of the display */
/* clears every row of the current colour target (FBO slot 0); the reference
 * can only clear the display (clearColour), this serves off-screen views */
int psgpu_clear_target(psgpu_ctx* ctx, uint32_t bgra);

/* ---- drawing ---- */
/* Draws are queued per view (render targets + viewport).  Changing targets or
 * viewport between draws starts another view of the same queued flush, so
 * many views render in one pass of the kernels; a view whose buffers an
 * earlier queued view already uses makes the queue flush first. */
int psgpu_draw(psgpu_ctx* ctx, int vao);          /* drawVAO: queued */
int psgpu_flush(psgpu_ctx* ctx);                  /* execute queued draws (async) */
int psgpu_sync(psgpu_ctx* ctx);                   /* flush + wait; re-runs on capacity growth */
/* display: top-down BGRA8 rows (the renderer's back buffer); depth: default depth buffer rows */
int psgpu_read_colour(psgpu_ctx* ctx, void* host, size_t bytes);
int psgpu_read_depth(psgpu_ctx* ctx, void* host, size_t bytes);
/* the display and default depth buffer written from host memory (same layouts
 * as the reads): the reference's are host FBOs that a post-processor may
 * write in place (PuresoftFBO::getBuffer, post.cpp:3-18); synchronises first */
int psgpu_write_colour(psgpu_ctx* ctx, const void* host, size_t bytes);
int psgpu_write_depth(psgpu_ctx* ctx, const void* host, size_t bytes);
int psgpu_display_device_ptr(psgpu_ctx* ctx, void** ptr);
int psgpu_depth_device_ptr(psgpu_ctx* ctx, void** ptr, int* scanline_floats);
int psgpu_get_stats(psgpu_ctx* ctx, psgpu_stats* out);
int psgpu_reset_stats(psgpu_ctx* ctx);
int psgpu_swap_buffers(psgpu_ctx* ctx);           /* swapBuffers: flush; display keeps its buffer */

/* ---- post-processing: PuresoftPipeline::postProcess (post.cpp:3-18, proc.h:89-94) ----
 * The reference hands its post-processor the colour target of slot 0 and the
 * depth buffer once the draws before it are done.  Here fn runs in stream
 * order: queued draws are flushed, then fn is called with the device
 * addresses of the current colour target and depth buffer and the context
 * stream, and enqueues its own work (kernels, copies) on that stream; later
 * draws and reads come after that work.  fn sees finished frames only: it is
 * called once the work before it has run, and when a flush before it
 * overflowed its scratch buffers the call waits for the log replay that
 * re-runs that flush, so fn runs exactly once per call and may change the
 * frame in place.  fn must only enqueue stream work.  Inside a recorded
 * command list (psgpu_capture_begin/end) its work becomes part of the graph
 * and runs once per replay.  fn returns 0 or a negative status, which the
 * call reports. */
typedef struct {
    void* colour;          /* BGRA8 rows of colour_pitch bytes; display rows are top-down */
    int colour_pitch, colour_top_down;
    void* depth;           /* FP32 rows of depth_pitch bytes */
    int depth_pitch;
    int width, height;     /* the viewport */
} psgpu_post_frame;
typedef int (*psgpu_post_fn)(void* user, const psgpu_post_frame* frame, void* stream);
int psgpu_post_process(psgpu_ctx* ctx, psgpu_post_fn fn, void* user);

/* ---- adjoint of texture sampling and shading (no counterpart in the reference) ----
 * A differentiable shading model evaluated on a geometry buffer (above):
 *   c = albedo_scale * T(u,v) * max(N.L, 0) + ks * max(N.H, 0)^spec_exp   (RGB)
 * with N the normalised interpolated normal, L and E the normalised
 * directions to light and camera, H = normalize(L + E), and T a bilinear
 * sample of an FP32 RGB texture (th rows of tw texels, row 0 at v = 0,
 * texel j at u = j/tw, wrapping).  Uncovered pixels give 0.
 * All buffers are device memory; params = {albedo_scale, spec_exp}.
 * backward adds dLoss/dT into grad_tex (atomics) and dLoss/dparams into
 * grad_params, given grad_out = dLoss/dc; visibility stays that of the
 * geometry buffer.  stream: the hipStream_t to run on (NULL is the null
 * stream; pass psgpu_stream(ctx) for the context's own). */
typedef struct {
    int width, height;          /* geometry buffer / image size */
    int tex_w, tex_h;           /* learnable texture size */
    float light[4], camera[4];  /* world positions */
    float ks;                   /* specular weight */
} psgpu_diff_desc;
int psgpu_diff_forward(psgpu_ctx* ctx, const psgpu_diff_desc* desc, const float* gbuf, const float* tex,
                       const float* params, float* out_rgb, void* stream);
int psgpu_diff_backward(psgpu_ctx* ctx, const psgpu_diff_desc* desc, const float* gbuf, const float* tex,
                        const float* params, const float* grad_out, float* grad_tex, float* grad_params, void* stream);

/* ---- asynchronous readback and fences (serving) ----
 * No counterpart in the reference, whose drawVAO returns with the frame done
 * (drawvao.cpp:122-132).  A server queues many views, then reads each one
 * back without waiting: psgpu_texture_read_async queues a device-to-host
 * copy of a region (memory rows y..y+h-1, columns x..x+w-1 of one layer;
 * PSGPU_DISPLAY for the display) into host memory, after everything queued
 * before it.  Copies run on a second stream, so they overlap later flushes.
 * Host memory should come from psgpu_host_alloc (pinned).  psgpu_fence marks
 * the current point; psgpu_fence_query reports done = 1 once every command
 * before it (copies included) has finished with correct results.  A command
 * that overflowed its scratch buffers is re-run by the query itself (a full
 * sync), so results are never returned from an overflowed run.  Host buffers
 * must stay valid until the fence after their copy is done. */
int psgpu_host_alloc(psgpu_ctx* ctx, size_t bytes, void** ptr);
int psgpu_host_free(psgpu_ctx* ctx, void* ptr);   /* syncs first */
int psgpu_texture_read_async(psgpu_ctx* ctx, int tex, int layer, int x, int y, int w, int h, void* host,
                             size_t host_pitch);
int psgpu_fence(psgpu_ctx* ctx, uint64_t* out_fence);
int psgpu_fence_query(psgpu_ctx* ctx, uint64_t fence, int* done);
int psgpu_fence_wait(psgpu_ctx* ctx, uint64_t fence);

/* ---- debug / parity planes (per pixel, display rows bottom-up: index y*W+x) ---- */
/* enable=1 makes every flush record: last depth-test-passing triangle id,
 * fragments rasterised, xor-hash of rasterised triangle ids, depth-test passes,
 * xor-hash of passing triangle ids.  Triangle ids count across the draws since
 * the last psgpu_debug_reset. */
int psgpu_debug_enable(psgpu_ctx* ctx, int enable);
int psgpu_debug_reset(psgpu_ctx* ctx);
int psgpu_debug_read(psgpu_ctx* ctx, int32_t* last_pass, int32_t* raster_count, uint32_t* raster_hash,
                     int32_t* pass_count, uint32_t* pass_hash);

/* ---- tuning and timing ---- */
int psgpu_tiling_set(psgpu_ctx* ctx, const psgpu_tiling* t);
int psgpu_tiling_get(psgpu_ctx* ctx, psgpu_tiling* t);
/* per-kernel timing with HIP events on the context stream; names[i] valid until destroy */
int psgpu_profile_enable(psgpu_ctx* ctx, int enable);
int psgpu_profile_read(psgpu_ctx* ctx, int max, const char** names, double* total_ms, int64_t* launches, int* count);
int psgpu_profile_reset(psgpu_ctx* ctx);

/* ---- recorded command lists (hipGraph replay) ---- */
/* Between capture_begin and capture_end, clears, draws and flushes are recorded
 * instead of run.  capture_end executes the recording once (which sizes the
 * scratch buffers), then captures its kernels into a hipGraph.  graph_launch
 * replays it `count` times, stream-ordered; the snapshot of programme,
 * uniforms, behaviour and targets taken at record time is reused.  Resource
 * updates (textures, VBOs) are refused while recording.  Graphs are captured
 * again automatically when scratch buffers have been reallocated or a VBO
 * they draw was updated (new geometry may need more scratch): the first
 * launch after that runs the commands once eagerly, which counts as one of
 * the `count` replays, and captures them again.  No extern
 * counterpart in the reference: it replaces the per-frame drawVAO calls of a
 * static scene by one launch. */
int psgpu_capture_begin(psgpu_ctx* ctx);
int psgpu_capture_end(psgpu_ctx* ctx, int* out_graph);
int psgpu_graph_launch(psgpu_ctx* ctx, int graph, int count);
int psgpu_graph_destroy(psgpu_ctx* ctx, int graph);

/* ---- synchronising without allocating ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory while they render (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): psgpu_sync_noalloc runs the
 * queued clears and draws, as psgpu_sync would, with every per-flush buffer
 * (triangle records, varyings, spans, pieces, sort keys, the winner plane,
 * the draw records themselves) carved from a fixed array of
 * PSGPU_STATIC_SCRATCH_BYTES in the device's memory, and waits for them.  It
 * calls no hipMalloc, hipFree or other allocating routine.  Each GPU the
 * library runs on holds one such array from the library's load until the
 * process ends; all contexts on a GPU share it, and calls are serialised by a
 * lock inside the library (the array keeps the last call's contents; nothing
 * in it needs to start at zero).  The queued work must be one flush: if
 * anything was submitted since the last sync (psgpu_flush, a flush psgpu_draw
 * forces, a post-process, a graph launch) the call returns PSGPU_ERR_STATE.
 * Work too large for the array returns PSGPU_ERR_NOMEM: before anything runs
 * (triangle, view and pixel counts) with the queue left intact for
 * psgpu_sync, or after the run when its spans or pieces overflowed (the
 * clears were applied, the draws were not).  The 1080p benchmark frame
 * (200,480 triangles) needs about 120 MB. */
#define PSGPU_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int psgpu_sync_noalloc(psgpu_ctx* ctx);

/* ---- arithmetic self-check: out[i] for op in {0:a/b, 1:sqrtf(a), 2:1/a, 3:(int)x86(a), 4:floorf(a),
 *      5:(int)(unsigned)x86(a), 6:log2 approximation of the trilinear LOD, 7:a (a plain copy kernel,
 *      the calibration point for HBM traffic counters), 8 and 9: the wrapped texel coordinates of v and
 *      v + 1 modulo m (v, m and the result are int32 bit patterns in the float arrays a, b, out),
 *      10: one lane of FP_Cloud's cvtps2dq/packusdw/packuswb tail (testproc.cpp:562-568) as a float} ---- */
int psgpu_selftest_math(psgpu_ctx* ctx, int op, const float* a, const float* b, float* out, int n);

#ifdef __cplusplus
}
#endif
#endif
