// psserve.cpp -- the render-serving loop declared in include/psserve.h.
//
// One host thread drives one or more psgpu contexts (render streams).
// Requests move through
//   waiting (FIFO) -> admitted (targets held on a stream) -> in flight (in a
//   batch) -> done.
// Render targets are pooled per stream and size class and reused; a target
// class is allocated lazily while the HBM budget (shared by the streams)
// allows, and idle targets of other classes are evicted when it does not.
#include "../../include/psserve.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// sizes of the target classes: sqrt(2) steps, rounded to multiples of 16
int class_side(int v)
{
    static const int ladder[] = {128, 176, 256, 368, 512, 736, 1024, 1456, 2048, 2896, 4096};
    for (int s : ladder)
        if (v <= s) return s;
    return (v + 63) / 64 * 64;
}

struct Slot {
    int ctx = 0;               // the render stream (context) holding it
    int colour = -1, depth = -1;
    uint8_t* host = nullptr;   // pinned, cw*ch*4 bytes
    int cw = 0, ch = 0;
    size_t bytes() const { return (size_t)cw * ch * 8; }
};

struct Batch {
    uint64_t fence = 0;
    std::vector<int> reqs;
};

}  // namespace

struct psserve_server {
    std::vector<psgpu_ctx*> ctxs;
    std::vector<psserve_scene> scene_list;   // scene k of context c at [c * n_scenes + k]
    int n_scenes;
    const psserve_request* reqs;
    int n;
    psserve_config cfg;
    psserve_result* res;
    uint8_t* const* images;
    psserve_stats st;

    // idle targets by (stream, class width, class height)
    std::map<std::tuple<int, int, int>, std::vector<Slot>> idle;
    std::vector<Slot> held;                                  // per request
    size_t target_bytes = 0;       // all allocated targets (held + idle)
    size_t inflight_scratch = 0;   // scratch estimate of admitted requests

    int psgpu(int rc, int c)
    {
        if (rc) g_err = std::string("psgpu: ") + psgpu_last_error(ctxs[c]);
        return rc;
    }

    int streams() const
    {
        const int n = (int)ctxs.size();
        return cfg.render_streams >= 1 && cfg.render_streams < n ? cfg.render_streams : n;
    }

    int free_slot(Slot& s)
    {
        int rc;
        psgpu_ctx* ctx = ctxs[s.ctx];
        if ((rc = psgpu(psgpu_texture_destroy(ctx, s.colour), s.ctx)) ||
            (rc = psgpu(psgpu_texture_destroy(ctx, s.depth), s.ctx)) || (rc = psgpu(psgpu_host_free(ctx, s.host), s.ctx)))
            return rc;
        target_bytes -= s.bytes();
        st.target_evictions++;
        return PSGPU_OK;
    }

    // a target for request i on stream c, within the budget; *ok = 0 when it must wait
    int acquire(int i, int c, int* ok)
    {
        const psserve_request& r = reqs[i];
        std::tuple<int, int, int> key(c, class_side(r.width), class_side(r.height));
        size_t scratch = (size_t)r.width * r.height * (size_t)cfg.scratch_bytes_per_px;
        *ok = 0;
        auto& pool = idle[key];
        if (!pool.empty()) {
            if (inflight_scratch + scratch + target_bytes > cfg.hbm_budget && inflight_scratch > 0) return PSGPU_OK;
            held[i] = pool.back();
            pool.pop_back();
            inflight_scratch += scratch;
            *ok = 1;
            return PSGPU_OK;
        }
        Slot s;
        s.ctx = c;
        s.cw = std::get<1>(key);
        s.ch = std::get<2>(key);
        size_t need = s.bytes() + scratch;
        // evict idle targets of other classes while over budget
        while (target_bytes + inflight_scratch + need > cfg.hbm_budget) {
            bool evicted = false;
            for (auto& kv : idle)
                if (!kv.second.empty()) {
                    int rc = free_slot(kv.second.back());
                    if (rc) return rc;
                    kv.second.pop_back();
                    evicted = true;
                    break;
                }
            if (!evicted) break;
        }
        // a request always fits an idle GPU, whatever its size
        bool empty = target_bytes == 0 && inflight_scratch == 0;
        if (target_bytes + inflight_scratch + need > cfg.hbm_budget && !empty) return PSGPU_OK;
        int rc;
        psgpu_ctx* ctx = ctxs[c];
        if ((rc = psgpu(psgpu_texture_create(ctx, s.cw, s.cw * 4, s.ch, 4, 0, PSGPU_CLAMP, 1, nullptr, &s.colour), c)) ||
            (rc = psgpu(psgpu_texture_create(ctx, s.cw, s.cw * 4, s.ch, 4, 0, PSGPU_CLAMP, 1, nullptr, &s.depth), c)) ||
            (rc = psgpu(psgpu_host_alloc(ctx, (size_t)s.cw * s.ch * 4, (void**)&s.host), c)))
            return rc;
        target_bytes += s.bytes();
        st.target_allocs++;
        st.peak_bytes = std::max<int64_t>(st.peak_bytes, (int64_t)target_bytes);
        held[i] = s;
        inflight_scratch += scratch;
        *ok = 1;
        return PSGPU_OK;
    }

    void release(int i)
    {
        const psserve_request& r = reqs[i];
        Slot s = held[i];
        held[i] = Slot();
        idle[std::make_tuple(s.ctx, s.cw, s.ch)].push_back(s);
        inflight_scratch -= (size_t)r.width * r.height * (size_t)cfg.scratch_bytes_per_px;
    }

    static double secs()
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

    // the batch ids on stream c
    int launch(int c, const std::vector<int>& ids, uint64_t* fence)
    {
        int rc;
        double t0 = secs();
        psgpu_ctx* ctx = ctxs[c];
        auto psgpu = [&](int e) { return this->psgpu(e, c); };
        if ((rc = psgpu(psgpu_program_use(ctx, cfg.program)))) return rc;
        for (int i : ids) {
            const psserve_request& r = reqs[i];
            const psserve_scene& sc = scene_list[(size_t)c * n_scenes + r.scene];
            const Slot& s = held[i];
            static const float I4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
            if ((rc = psgpu(psgpu_target_set(ctx, 0, s.colour))) || (rc = psgpu(psgpu_depth_set(ctx, s.depth))) ||
                (rc = psgpu(psgpu_viewport(ctx, r.width, r.height))) || (rc = psgpu(psgpu_clear_depth(ctx, 1.0f))) ||
                (rc = psgpu(psgpu_clear_target(ctx, cfg.clear_bgra))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 3, r.pv, sizeof(r.pv)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 4, I4, sizeof(I4)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 5, I4, sizeof(I4)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 7, r.light, sizeof(r.light)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 8, r.eye, sizeof(r.eye)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 9, &sc.albedo_tex, sizeof(int)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 10, &sc.normal_tex, sizeof(int)))) ||
                (rc = psgpu(psgpu_uniform_set(ctx, 11, &r.spec_exp, sizeof(int)))) || (rc = psgpu(psgpu_draw(ctx, sc.vao))))
                return rc;
        }
        // all views of the batch are queued; the first read flushes them as one pass
        double t1 = secs();
        if ((rc = psgpu(psgpu_flush(ctx)))) return rc;
        st.host_flush_s += secs() - t1;
        if (cfg.deliver_to_host) {
            for (int i : ids) {
                const psserve_request& r = reqs[i];
                const Slot& s = held[i];
                if ((rc = psgpu(psgpu_texture_read_async(ctx, s.colour, 0, 0, 0, r.width, r.height, s.host,
                                                         (size_t)r.width * 4))))
                    return rc;
            }
        }
        rc = psgpu(psgpu_fence(ctx, fence));
        st.host_submit_s += secs() - t0;
        return rc;
    }

    int run()
    {
        using clk = std::chrono::steady_clock;
        held.assign(n, Slot());
        const int ns = streams();
        std::deque<int> waiting;
        std::vector<std::vector<int>> admitted(ns);   // per stream: holding targets, not launched yet
        std::vector<std::deque<Batch>> inflight(ns);
        int next = 0, done = 0;
        auto t0 = clk::now();
        auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
        int rc;
        while (done < n) {
            bool progress = false;
            // completions, in launch order on each stream
            for (int c = 0; c < ns; c++)
                while (!inflight[c].empty()) {
                    int fin = 0;
                    if ((rc = psgpu(psgpu_fence_query(ctxs[c], inflight[c].front().fence, &fin), c))) return rc;
                    if (!fin) break;
                    double t = now();
                    for (int i : inflight[c].front().reqs) {
                        const psserve_request& r = reqs[i];
                        res[i].done_s = t;
                        if (images && images[i] && cfg.deliver_to_host)
                            memcpy(images[i], held[i].host, (size_t)r.width * r.height * 4);
                        release(i);
                        done++;
                    }
                    inflight[c].pop_front();
                    progress = true;
                }
            double t = now();
            while (next < n && reqs[next].arrival_s <= t) {
                res[next].arrival_s = reqs[next].arrival_s;
                waiting.push_back(next++);
            }
            // targets are held only by requests the next batch of a stream can
            // take; the head of the queue goes to the stream with the fewest
            // batches in flight (then the fewest admitted)
            while (!waiting.empty()) {
                int c = -1;
                for (int k = 0; k < ns; k++) {
                    if ((int)admitted[k].size() >= cfg.max_batch_views) continue;
                    if (c < 0 || inflight[k].size() < inflight[c].size() ||
                        (inflight[k].size() == inflight[c].size() && admitted[k].size() < admitted[c].size()))
                        c = k;
                }
                if (c < 0) break;
                int ok = 0;
                if ((rc = acquire(waiting.front(), c, &ok))) return rc;
                if (!ok) break;
                res[waiting.front()].admitted_s = t;
                admitted[c].push_back(waiting.front());
                waiting.pop_front();
                progress = true;
            }
            for (int c = 0; c < ns; c++) {
                if (admitted[c].empty() || (int)inflight[c].size() >= cfg.max_inflight) continue;
                Batch b;
                int64_t px = 0;
                size_t k = 0;
                for (; k < admitted[c].size() && (int)b.reqs.size() < cfg.max_batch_views; k++) {
                    int64_t p = (int64_t)reqs[admitted[c][k]].width * reqs[admitted[c][k]].height;
                    if (!b.reqs.empty() && px + p > cfg.max_batch_pixels) break;
                    b.reqs.push_back(admitted[c][k]);
                    px += p;
                }
                admitted[c].erase(admitted[c].begin(), admitted[c].begin() + k);
                if ((rc = launch(c, b.reqs, &b.fence))) return rc;
                double tl = now();
                for (int i : b.reqs) {
                    res[i].launched_s = tl;
                    res[i].batch = (int)st.batches;
                }
                st.batches++;
                st.views += (int64_t)b.reqs.size();
                st.pixels += px;
                st.max_batch_views_seen = std::max<int>(st.max_batch_views_seen, (int)b.reqs.size());
                inflight[c].push_back(std::move(b));
                progress = true;
            }
            if (!progress) {
                // idle until the next arrival or a completion; short sleeps only
                // when nothing is on the GPU and the next arrival is far
                bool busy = false;
                for (int c = 0; c < ns; c++) busy = busy || !inflight[c].empty();
                if (!busy && next < n && reqs[next].arrival_s - now() > 200e-6)
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                else
                    std::this_thread::yield();
            }
        }
        st.wall_s = now();
        return PSGPU_OK;
    }

    // after a run: targets still held (a failed run) go back to the pools
    int settle()
    {
        int rc = PSGPU_OK;
        for (int c = 0; c < (int)ctxs.size(); c++) {
            int e = psgpu(psgpu_sync(ctxs[c]), c);
            if (!rc) rc = e;
        }
        for (Slot& s : held)
            if (s.colour >= 0) idle[std::make_tuple(s.ctx, s.cw, s.ch)].push_back(s);
        held.clear();
        inflight_scratch = 0;
        return rc;
    }

    // frees every pooled target
    int release_all()
    {
        int rc = settle();
        for (auto& kv : idle)
            for (Slot& s : kv.second)
                if (!rc) rc = free_slot(s);
        idle.clear();
        return rc;
    }
};

namespace {

int check_config(const psserve_config* cfg)
{
    if (!cfg) return fail(PSGPU_ERR_INVALID, "null configuration");
    if (cfg->max_batch_views < 1 || cfg->max_inflight < 1 || cfg->max_batch_pixels < 1 || cfg->scratch_bytes_per_px < 0 ||
        cfg->render_streams < 0)
        return fail(PSGPU_ERR_INVALID, "bad configuration");
    return PSGPU_OK;
}

}  // namespace

extern "C" {

int psserve_create_multi(psgpu_ctx* const* ctxs, int n_ctx, const psserve_scene* scenes, int n_scenes,
                         const psserve_config* cfg, psserve_server** out)
{
    if (!out) return fail(PSGPU_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!ctxs || n_ctx < 1 || !scenes || n_scenes < 1) return fail(PSGPU_ERR_INVALID, "null argument");
    for (int c = 0; c < n_ctx; c++)
        if (!ctxs[c]) return fail(PSGPU_ERR_INVALID, "null context %d", c);
    int rc = check_config(cfg);
    if (rc) return rc;
    psserve_server* s = new psserve_server();
    s->ctxs.assign(ctxs, ctxs + n_ctx);
    s->scene_list.assign(scenes, scenes + (size_t)n_ctx * n_scenes);
    s->n_scenes = n_scenes;
    s->cfg = *cfg;
    *out = s;
    return PSGPU_OK;
}

int psserve_create(psgpu_ctx* ctx, const psserve_scene* scenes, int n_scenes, const psserve_config* cfg,
                   psserve_server** out)
{
    return psserve_create_multi(&ctx, 1, scenes, n_scenes, cfg, out);
}

int psserve_set_config(psserve_server* s, const psserve_config* cfg)
{
    if (!s) return fail(PSGPU_ERR_INVALID, "null server");
    int rc = check_config(cfg);
    if (rc) return rc;
    s->cfg = *cfg;
    return PSGPU_OK;
}

int psserve_run(psserve_server* s, const psserve_request* reqs, int n, psserve_result* results, uint8_t* const* images,
                psserve_stats* stats)
{
    if (!s || !reqs || !results || n < 0) return fail(PSGPU_ERR_INVALID, "null argument");
    for (int i = 0; i < n; i++) {
        const psserve_request& r = reqs[i];
        if (r.scene < 0 || r.scene >= s->n_scenes)
            return fail(PSGPU_ERR_RANGE, "request %d: scene %d out of range", i, r.scene);
        if (r.width < 1 || r.height < 1 || r.width > 16384 || r.height > 16384)
            return fail(PSGPU_ERR_INVALID, "request %d: bad size %dx%d", i, r.width, r.height);
        if (i && r.arrival_s < reqs[i - 1].arrival_s) return fail(PSGPU_ERR_INVALID, "requests must be in arrival order");
    }
    s->reqs = reqs;
    s->n = n;
    s->res = results;
    s->images = images;
    memset(&s->st, 0, sizeof(s->st));
    s->st.peak_bytes = (int64_t)s->target_bytes;   // pooled targets count from the start
    memset(results, 0, sizeof(psserve_result) * (size_t)n);
    int rc = s->run();
    int rc2 = s->settle();
    if (stats) *stats = s->st;
    s->reqs = nullptr;
    s->res = nullptr;
    s->images = nullptr;
    if (rc) return rc;
    return rc2;
}

int psserve_destroy(psserve_server* s)
{
    if (!s) return PSGPU_OK;
    int rc = s->release_all();
    delete s;
    return rc;
}

const char* psserve_last_error(void) { return g_err.c_str(); }

}  // extern "C"
