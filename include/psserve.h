/*
 * psserve.h -- render-serving loop over libpsgpu (continuous batching of view
 * requests against resident scenes, with HBM-budget admission).
 *
 * No counterpart in the reference, which renders one view per drawVAO call
 * on the caller's thread (drawvao.cpp:3-133).  This is the loop that keeps
 * many such views in flight: scenes and textures stay resident on one
 * psgpu context; each request names a scene, a resolution and a camera.
 *
 * Policy (one GPU, one host thread driving one or more contexts, each a
 * render stream with its own copy of the scenes):
 *  - arrivals join a FIFO; the head is admitted when its render targets
 *    (colour + depth, rounded up to a size class) plus a per-pixel scratch
 *    estimate fit in the HBM budget, and fewer than max_batch_views requests
 *    are admitted but not yet launched;
 *  - whenever fewer than max_inflight batches are on the GPU, every admitted
 *    request not yet launched (up to max_batch_views / max_batch_pixels) is
 *    queued as one multi-view flush, followed by asynchronous copies of each
 *    view into pinned host memory and a fence;
 *  - finished batches release their targets, which admits more requests.
 *    Requests arriving while batches run are admitted and launched as soon as
 *    targets and a batch slot free up, so the GPU always has the next batch
 *    queued behind the current one and readback copies overlap it.
 * With several render streams, admitted requests go to the stream with the
 * fewest batches in flight, so a small batch need not wait behind a large
 * one and the streams' kernels overlap on the GPU.
 * max_batch_views = 1 with max_inflight = 1 and one render stream is the
 * plain one-view-per-launch path.  A view renders to the same pixels alone
 * or in any batch, on any stream.
 */
#ifndef PSSERVE_H
#define PSSERVE_H

#include <stddef.h>
#include <stdint.h>

#include "psgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* a resident scene: one VAO drawn with the Blinn-Phong programme */
typedef struct {
    int vao;
    int albedo_tex;     /* uniform slot 9  */
    int normal_tex;     /* uniform slot 10 */
} psserve_scene;

typedef struct {
    int scene;
    int width, height;
    float pv[16];       /* slot 3: projection x view, column-major */
    float eye[4];       /* slot 8 */
    float light[4];     /* slot 7 */
    int spec_exp;       /* slot 11 */
    double arrival_s;   /* seconds after the start of the run */
} psserve_request;

typedef struct {
    int program;                /* psgpu programme used for every view */
    size_t hbm_budget;          /* bytes for render targets + scratch estimate */
    int scratch_bytes_per_px;   /* scratch estimate per pixel in flight */
    int max_batch_views;
    int64_t max_batch_pixels;
    int max_inflight;           /* batches queued on the GPU at once */
    uint32_t clear_bgra;
    int deliver_to_host;        /* 1: each image is copied to pinned host memory before the
                                 * request completes (PCIe-inclusive); 0: images stay in HBM */
    int render_streams;         /* contexts used, 1 .. the server's (0: all of them) */
} psserve_config;

typedef struct {
    double arrival_s, admitted_s, launched_s, done_s;
    int batch;
} psserve_result;

typedef struct {
    int64_t batches, views, pixels;
    int64_t target_allocs, target_evictions;
    int64_t peak_bytes;          /* render targets allocated at once (pooled ones included) */
    int max_batch_views_seen;
    double wall_s;
    double host_submit_s;        /* host time spent queueing batches (draws, reads, fence) */
    double host_flush_s;         /* part of it spent in the flush the first read triggers */
} psserve_stats;

/* A server: the scene list and configuration, plus the pooled render
 * targets and pinned buffers it keeps between runs. */
typedef struct psserve_server psserve_server;

int psserve_create(psgpu_ctx* ctx, const psserve_scene* scenes, int n_scenes, const psserve_config* cfg,
                   psserve_server** out);
/* The same over n_ctx contexts of one device (render streams): scenes holds
 * n_ctx * n_scenes entries, scene k of context c at [c * n_scenes + k]; the
 * configuration's programme id must name the same programme on each. */
int psserve_create_multi(psgpu_ctx* const* ctxs, int n_ctx, const psserve_scene* scenes, int n_scenes,
                         const psserve_config* cfg, psserve_server** out);
int psserve_set_config(psserve_server* server, const psserve_config* cfg);
/* frees the pooled targets (syncs the context) */
int psserve_destroy(psserve_server* server);

/* Runs the request list in real time (arrivals are waited for) and returns
 * per-request timestamps.  images: NULL, or one host pointer per request
 * (NULL entries allowed) that receives height rows of width*4 bytes (the
 * colour target's memory rows); only filled with deliver_to_host. */
int psserve_run(psserve_server* server, const psserve_request* reqs, int n, psserve_result* results,
                uint8_t* const* images, psserve_stats* stats);

/* message of the last failed psserve_* call on this thread */
const char* psserve_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
