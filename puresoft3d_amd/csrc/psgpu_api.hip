// psgpu_api.hip -- C ABI of the raster hot path (include/psgpu.h).
//
// Host side of libpsgpu: object pools with the reference's semantics
// (pipeline.cpp, tex.cpp, prog.cpp, vbo.cpp, vao.cpp), a per-context command
// log of flushes and clears, and the orchestration of the kernels in
// ps_kernels.hip on one HIP stream.
//
// Capacity handling: intermediate buffers (spans, bin pairs) are sized from
// the previous frames.  Kernels that would overrun a buffer set an overflow
// flag instead; every later kernel of the log then skips its work, and
// psgpu_sync() grows the buffers and replays the whole log since the last
// sync.  Clears and flushes only write depth/colour when the flag is clear,
// so a replay starts from the state the log started from.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psgpu.h"
#include "ps_kernels.h"

// Per-flush scratch of psgpu_sync_noalloc: a fixed array in every GPU's
// memory, held from the library's load until the process ends.  Host code
// never uses this name as an address: hipGetSymbolAddress gives each
// device's copy (static_scratch_init), and that address is what kernels get.
static __device__ __align__(256) uint8_t scratch[PSGPU_STATIC_SCRATCH_BYTES];

namespace {

const int kMaxSlots = 16;      // config.h:4
const int kMaxUniforms = 1024; // config.h:5
const int64_t kMaxFlushTris = (int64_t)8 << 20;  // triangles queued before a flush is forced
// piece records hold a column in 16 bits and span records a row in 24 bits
const int kMaxViewW = PSGPU_MAX_VIEW_WIDTH, kMaxViewH = PSGPU_MAX_VIEW_HEIGHT;

struct Tex {
    bool alive = false;
    uint8_t* d = nullptr;
    int w = 0, h = 0, scan = 0, elem = 0, layers = 1, wrap = 0, levels = 1;
    int64_t layer_bytes = 0;
    int64_t level_off[PS_MAX_LEVELS] = {0};
};

struct Vbo {
    bool alive = false;
    uint8_t* d = nullptr;
    size_t unit = 0, count = 0;
};

struct Vao {
    bool alive = false;
    int slot[kMaxSlots];
};

struct Prog {
    bool alive = false;
    int vp = 0, ip = 0, fp = 0;
};

struct Targets {
    uint32_t* colour; int colour_scan; int colour_top_down; int colour_w; int colour_h;
    float* depth; int depth_scan; int depth_w; int depth_h;
    int vpw, vph;
    float* gbuf;
    bool same(const Targets& o) const
    {
        return colour == o.colour && depth == o.depth && vpw == o.vpw && vph == o.vph && colour_scan == o.colour_scan &&
               depth_scan == o.depth_scan && gbuf == o.gbuf;
    }
};

// pending draws of one view (render targets + viewport) of the next flush,
// with the clears of its buffers that the flush applies itself
struct Segment {
    Targets tg;
    std::vector<PsDraw> draws;
    bool dclear = false;
    float dclear_value = 0.0f;
    int cclear_rows = 0;
    uint32_t cclear_value = 0;
};

enum CmdType { CMD_FLUSH = 0, CMD_CLEARS = 1, CMD_GRAPH = 3, CMD_COPY = 4, CMD_POST = 5 };

// a device-to-host copy of a 2D region, run on the copy stream after the
// commands before it
struct CopyRegion {
    const uint8_t* src = nullptr;
    size_t src_pitch = 0;
    void* dst = nullptr;
    size_t dst_pitch = 0, width_bytes = 0, rows = 0;
};

struct Cmd {
    CmdType type;
    std::vector<Segment> views;   // CMD_FLUSH
    std::vector<PsFill> fills;    // CMD_CLEARS
    int graph = -1, count = 0;    // CMD_GRAPH
    CopyRegion copy;              // CMD_COPY
    psgpu_post_fn post = nullptr; // CMD_POST: callback, its argument and the frame it sees
    void* post_user = nullptr;
    psgpu_post_frame post_frame = {};
};

// a point in the command stream: done when both streams passed it and no
// command before it overflowed its scratch buffers
struct Fence {
    uint64_t id;
    hipEvent_t main_ev, copy_ev;
    size_t log_end;   // commands [0, log_end) of the log come before it
};

// A recorded command list replayed as one hipGraph (stream capture of the
// kernels of every command).  Each command's device records (draws, views,
// fills) live in a buffer owned by the graph, so replays copy nothing.
struct Graph {
    bool alive = false;
    std::vector<Cmd> cmds;
    std::vector<uint8_t*> blobs;
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
    int64_t span_cap = 0, seg_cap = 0;   // the capacities its kernels were captured with
    bool stale = false;                  // a vertex stream it draws was updated
    int64_t flushes = 0;
};

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
};

// The per-flush scratch of run_flush: everything it reads or writes besides
// the command records, the render targets and the counters.  span_cap and
// seg_cap are the span and piece capacities the kernels are told.
struct FlushBufs {
    DevBuf<PsTri> tris;
    DevBuf<float> vary;
    DevBuf<float> grad;   // per triangle: screen gradients for the mip level (8 floats)
    DevBuf<uint32_t> nrows, rowoff, partial;
    DevBuf<uint32_t> bstart;   // triangle of every 64th row item, written by the row scan
    DevBuf<PsSpan> spans;
    // tile lists: per-tile piece and run counts (zero between flushes: the
    // piece kernel counts them back down), their scan, the scan's block sums,
    // and the run records (two words each) with scratch for reordering them
    DevBuf<unsigned long long> tcnt, tstart, tpartial;
    DevBuf<uint32_t> runs, runs2;
    DevBuf<uint32_t> heavy;   // count and list of the heavy tiles, rastered first
    DevBuf<int> winner;
    DevBuf<PsSeg> segs;
    DevBuf<float> ck;
    int64_t span_cap = 0, seg_cap = 0;
};

struct ProfRec { const char* name; hipEvent_t a, b; };

// device image of one command: draws, then views, then fills
struct Blob {
    std::vector<uint8_t> bytes;
    size_t off_views = 0, off_fills = 0;
    int nd = 0, nv = 0, nf = 0;
    bool shares[PS_N_PROGS] = {false};   // some draws of the programme share varyings
};

}  // namespace

struct psgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // device-to-host copies overlap the next flushes
    hipEvent_t copy_ev = nullptr;
    std::vector<Fence> fences;
    std::vector<hipEvent_t> fence_ev_pool;
    uint64_t next_fence = 1;             // a completed sync clears the log and retires every fence
    std::string err;
    int dev_w = 0, dev_h = 0;
    int vp_w = 0, vp_h = 0;
    uint32_t* display = nullptr;
    float* depth = nullptr;
    int depth_scan = 0;
    std::vector<Tex> tex;
    std::vector<Vbo> vbo;
    std::vector<Vao> vao;
    std::vector<Prog> prog;
    std::vector<std::vector<uint8_t>> uniforms;
    int behavior = PSGPU_UPDATE_DEPTH | PSGPU_TEST_DEPTH | PSGPU_FACE_CULLING;  // pipeline.cpp:34
    int cur_prog = -1;
    int fbo0 = PSGPU_DISPLAY;
    int fbo_more[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};   // slots 1..15
    int depth_sel = PSGPU_DEFAULT_DEPTH;
    float* gbuf = nullptr;         // caller's geometry buffer (device memory), or none
    size_t gbuf_bytes = 0;
    std::vector<Segment> queue;        // pending views with draws
    std::vector<PsFill> pending_fills; // pending clears (run before the queued draws)
    int64_t queue_tris = 0;
    int queue_imm = 0;        // programme of the queued draws when they shade in the raster walk, else 0
    std::vector<Cmd> log;
    bool capturing = false;        // between psgpu_capture_begin and _end
    bool in_stream_capture = false;
    std::vector<Cmd> capture_list;
    std::vector<Graph> graphs;
    uint64_t buf_gen = 1;          // bumped whenever a scratch buffer moves
    psgpu_tiling tiling = {4, 8, 1};      // 8 x 32-pixel tiles: short per-tile lists for dense and sparse scenes alike
    // scratch
    DevBuf<uint8_t> d_blob;
    FlushBufs fb;
    bool fixed_scratch = false;   // fb is carved from the static array: never reallocate it
    // most spans and pieces one flush may hold (pieces keep their span's index
    // in 26 bits, run records the index of its wave of 64 below 2^20 - 1, the
    // winner plane a piece's index in 31 bits, all ones meaning none); a
    // flush that needs more is split in two and re-run
    int64_t max_spans = ((int64_t)1 << 26) - 64, max_segs = ((int64_t)1 << 31) - 64;
    uint32_t* d_sizes = nullptr;
    uint32_t* h_sizes = nullptr;
    // per logged command: the overflow flag as it stood after the command ran
    static const int kFlagCap = 16384;
    int* h_flags = nullptr;
    PsCounters* d_cnt = nullptr;
    PsCounters* h_cnt = nullptr;
    // command staging (pinned), ring of slots
    static const int kStage = 4;
    uint8_t* h_stage[kStage] = {nullptr};
    size_t h_stage_cap[kStage] = {0};
    hipEvent_t stage_ev[kStage] = {nullptr};
    int stage_next = 0;
    // debug planes
    bool debug = false;
    int32_t* dbg_last = nullptr; int32_t* dbg_rcount = nullptr; uint32_t* dbg_rhash = nullptr;
    int32_t* dbg_pcount = nullptr; uint32_t* dbg_phash = nullptr;
    int dbg_w = 0, dbg_h = 0;
    int64_t dbg_tri_counter = 0;
    // stats and profiling
    psgpu_stats stats;
    bool profiling = false;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, std::pair<double, int64_t>> prof;
    std::vector<std::string> prof_names;
};

namespace {

int fail(psgpu_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, x)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return fail((c), PSGPU_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <typename T> int ensure(psgpu_ctx* c, DevBuf<T>& b, size_t n)
{
    if (n <= b.cap && b.p) return PSGPU_OK;
    if (c->fixed_scratch)
        return fail(c, PSGPU_ERR_NOMEM, "the flush needs %zu bytes in a buffer of %zu in the static scratch array",
                    n * sizeof(T), b.cap * sizeof(T));
    size_t cap = std::max(n, (size_t)(b.cap * 3 / 2));
    if (cap < 64) cap = 64;
    if (c->in_stream_capture) return fail(c, PSGPU_ERR_STATE, "scratch buffer too small during graph capture");
    if (b.p) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc((void**)&b.p, cap * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();   // the failure is reported here, not by the next launch check
        b.p = nullptr;
        return fail(c, e == hipErrorOutOfMemory ? PSGPU_ERR_NOMEM : PSGPU_ERR_DEVICE,
                    "hipMalloc of %zu bytes (span capacity %lld, piece capacity %lld): %s", cap * sizeof(T),
                    (long long)c->fb.span_cap, (long long)c->fb.seg_cap, hipGetErrorString(e));
    }
    b.cap = cap;
    c->buf_gen++;
    return PSGPU_OK;
}

hipEvent_t ev_get(psgpu_ctx* c)
{
    if (!c->ev_pool.empty()) { hipEvent_t e = c->ev_pool.back(); c->ev_pool.pop_back(); return e; }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

struct ProfScope {
    psgpu_ctx* c; const char* name; hipEvent_t a = nullptr;
    ProfScope(psgpu_ctx* c_, const char* n) : c(c_), name(n)
    {
        if (c->profiling && !c->in_stream_capture && !c->fixed_scratch) { a = ev_get(c); hipEventRecord(a, c->stream); }
    }
    ~ProfScope()
    {
        if (a) {
            hipEvent_t b = ev_get(c);
            hipEventRecord(b, c->stream);
            c->prof_pending.push_back({name, a, b});
        }
    }
};

void prof_collect(psgpu_ctx* c)
{
    for (auto& r : c->prof_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& p = c->prof[r.name];
            p.first += ms;
            p.second += 1;
        }
        c->ev_pool.push_back(r.a);
        c->ev_pool.push_back(r.b);
    }
    c->prof_pending.clear();
}

bool valid_tex(psgpu_ctx* c, int i) { return i >= 0 && i < (int)c->tex.size() && c->tex[i].alive; }

int prog_nv(int kind)
{
    switch (kind) {
    case PSGPU_PROC_DEF01: return 10;
    case PSGPU_PROC_DEF02: return 12;
    case PSGPU_PROC_DEF03: return 18;
    case PSGPU_PROC_DEF04: return 3;
    case PSGPU_PROC_DEF05: return 0;
    case PSGPU_PROC_LAMBERT: return 5;
    case PSGPU_PROC_BLINN: return 14;
    case PSGPU_PROC_SHADOWED: return 18;
    case PSGPU_PROC_CLOUD: return 14;
    case PSGPU_PROC_CLOUD_SHADOW: return 2;
    default: return -1;
    }
}

// host copy of mcemaths_transform_m4m4 (matrix.cpp:588-701), same operation order
void host_m4m4(float* d, const float* l, const float* r)
{
    for (int j = 0; j < 4; j++) {
        const float* v = r + 4 * j;
        for (int i = 0; i < 4; i++) {
            float a = v[0] * l[i], b = v[1] * l[4 + i], cc = v[2] * l[8 + i], dd = v[3] * l[12 + i];
            d[4 * j + i] = ((a + b) + cc) + dd;
        }
    }
}

int current_targets(psgpu_ctx* c, Targets& t)
{
    t.vpw = c->vp_w;
    t.vph = c->vp_h;
    t.gbuf = c->gbuf;
    if (c->fbo0 == PSGPU_DISPLAY) {
        t.colour = c->display; t.colour_scan = c->dev_w; t.colour_top_down = 1; t.colour_w = c->dev_w; t.colour_h = c->dev_h;
    } else {
        if (!valid_tex(c, c->fbo0)) return fail(c, PSGPU_ERR_STATE, "colour target texture %d was destroyed", c->fbo0);
        Tex& x = c->tex[c->fbo0];
        if (x.elem != 4 || x.scan % 4) return fail(c, PSGPU_ERR_INVALID, "colour target must have 4-byte texels");
        t.colour = (uint32_t*)x.d; t.colour_scan = x.scan / 4; t.colour_top_down = 0; t.colour_w = x.w; t.colour_h = x.h;
    }
    if (c->depth_sel == PSGPU_DEFAULT_DEPTH) {
        t.depth = c->depth; t.depth_scan = c->depth_scan; t.depth_w = c->dev_w; t.depth_h = c->dev_h;
    } else {
        if (!valid_tex(c, c->depth_sel)) return fail(c, PSGPU_ERR_STATE, "depth texture %d was destroyed", c->depth_sel);
        Tex& x = c->tex[c->depth_sel];
        t.depth = (float*)x.d; t.depth_scan = x.scan / 4; t.depth_w = x.w; t.depth_h = x.h;
    }
    return PSGPU_OK;
}

PsTex tex_desc(const Tex& x)
{
    PsTex d;
    memset(&d, 0, sizeof(d));
    d.data = x.d; d.width = x.w; d.height = x.h; d.scanline = x.scan; d.elem_len = x.elem;
    d.layers = x.layers; d.wrap = x.wrap; d.levels = x.levels; d.layer_bytes = x.layer_bytes;
    for (int l = 0; l < PS_MAX_LEVELS; l++) d.level_off[l] = x.level_off[l];
    return d;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Views of one scene share its varyings: a draw whose programme, vertex
// streams and model transforms (M, Mrot, shadowPV) equal those of an earlier
// draw of the same command reads that draw's varyings, which then writes them
// for all its triangles, and computes clip positions only.
static bool same_varyings(const PsDraw& a, const PsDraw& b)
{
    return a.prog == b.prog && a.tri_count == b.tri_count && !memcmp(a.slot, b.slot, sizeof(a.slot)) &&
           !memcmp(a.unit, b.unit, sizeof(a.unit)) && !memcmp(a.M, b.M, sizeof(a.M)) &&
           !memcmp(a.Mrot, b.Mrot, sizeof(a.Mrot)) && !memcmp(a.shadowPV, b.shadowPV, sizeof(a.shadowPV));
}

static void share_varyings(std::vector<PsDraw>& draws)
{
    std::multimap<uint64_t, size_t> seen;
    for (size_t i = 0; i < draws.size(); i++) {
        PsDraw& d = draws[i];
        d.vary_delta = 0;
        d.vary_all = d.vary_skip = 0;
        if (!ps_shares_varyings(d.prog)) continue;
        // FNV-1a over the fields same_varyings compares
        uint64_t h = 1469598103934665603ull;
        auto mix = [&h](const void* p, size_t n) {
            for (size_t k = 0; k < n; k++) h = (h ^ ((const uint8_t*)p)[k]) * 1099511628211ull;
        };
        mix(&d.prog, sizeof(d.prog)); mix(&d.tri_count, sizeof(d.tri_count)); mix(d.slot, sizeof(d.slot));
        mix(d.M, sizeof(d.M)); mix(d.Mrot, sizeof(d.Mrot));
        bool found = false;
        for (auto r = seen.equal_range(h); r.first != r.second; ++r.first) {
            PsDraw& s = draws[r.first->second];
            if (same_varyings(s, d)) {
                d.vary_delta = (int)(s.tri_start - d.tri_start);
                d.vary_skip = 1;
                s.vary_all = 1;
                found = true;
                break;
            }
        }
        if (!found) seen.emplace(h, i);
    }
}

// lay out a command's device records: draws (view-tagged, in submission
// order), views (tiling and scratch offsets are filled in here), fills
int build_blob(psgpu_ctx* c, const Cmd& cmd, Blob& b, PsFrame& f, int64_t& T, int& nvs, bool* progs, bool* shade_progs,
               int64_t& n_pix)
{
    std::vector<PsDraw> draws;
    std::vector<PsView> views;
    T = 0;
    nvs = 1;
    n_pix = 0;
    memset(&f, 0, sizeof(f));
    f.chunk_w = c->tiling.chunk_width;
    f.tile_rows = c->tiling.tile_rows;
    f.tile_cols = 64 / f.tile_rows;
    int64_t tile_base = 0, stile_base = 0;
    for (size_t v = 0; v < cmd.views.size(); v++) {
        const Segment& sg = cmd.views[v];
        PsView pv;
        memset(&pv, 0, sizeof(pv));
        pv.W = sg.tg.vpw;
        pv.H = sg.tg.vph;
        pv.n_chunks = (pv.W + f.chunk_w - 1) / f.chunk_w;
        pv.n_tiles_x = (pv.n_chunks + f.tile_cols - 1) / f.tile_cols;
        pv.n_tiles_y = (pv.H + f.tile_rows - 1) / f.tile_rows;
        pv.n_stiles_x = (pv.W + 15) / 16;
        pv.tile_base = tile_base;
        pv.stile_base = stile_base;
        tile_base += (int64_t)pv.n_tiles_x * pv.n_tiles_y;
        stile_base += (int64_t)pv.n_stiles_x * ((pv.H + 15) / 16);
        pv.colour = sg.tg.colour; pv.colour_scan = sg.tg.colour_scan; pv.colour_top_down = sg.tg.colour_top_down;
        pv.colour_h = sg.tg.colour_h;
        pv.depth = sg.tg.depth; pv.depth_scan = sg.tg.depth_scan;
        pv.gbuf = sg.tg.gbuf;
        pv.dclear = sg.dclear ? 1 : 0;
        pv.dclear_value = sg.dclear_value;
        pv.cclear_rows = sg.cclear_rows;
        pv.cclear_value = sg.cclear_value;
        pv.winner = (int*)(intptr_t)n_pix;   // offset; rebased once the winner buffer is known
        n_pix += (int64_t)pv.W * pv.H;
        views.push_back(pv);
        for (PsDraw d : sg.draws) {
            d.view = (int)v;
            d.tri_start = T;
            T += d.tri_count;
            nvs = std::max(nvs, d.nv);
            progs[d.prog] = true;
            if (d.writes_colour) shade_progs[d.prog] = true;
            draws.push_back(d);
        }
    }
    share_varyings(draws);
    for (const PsDraw& d : draws)
        if (d.vary_all) b.shares[d.prog] = true;
    f.n_views = (int)views.size();
    f.n_tiles = tile_base;
    f.n_stiles = stile_base;
    int rc = ensure(c, c->fb.winner, (size_t)std::max<int64_t>(n_pix, 1));
    if (rc) return rc;
    for (PsView& pv : views) pv.winner = c->fb.winner.p + (intptr_t)pv.winner;
    b.nd = (int)draws.size();
    b.nv = (int)views.size();
    b.nf = (int)cmd.fills.size();
    b.off_views = align256(draws.size() * sizeof(PsDraw));
    b.off_fills = b.off_views + align256(views.size() * sizeof(PsView));
    b.bytes.assign(b.off_fills + cmd.fills.size() * sizeof(PsFill) + 16, 0);
    if (!draws.empty()) memcpy(b.bytes.data(), draws.data(), draws.size() * sizeof(PsDraw));
    if (!views.empty()) memcpy(b.bytes.data() + b.off_views, views.data(), views.size() * sizeof(PsView));
    if (!cmd.fills.empty()) memcpy(b.bytes.data() + b.off_fills, cmd.fills.data(), cmd.fills.size() * sizeof(PsFill));
    return PSGPU_OK;
}

// copy a blob to the device through the pinned staging ring
int stage_blob(psgpu_ctx* c, const Blob& b, uint8_t** out)
{
    int s = c->stage_next;
    c->stage_next = (s + 1) % psgpu_ctx::kStage;
    if (c->stage_ev[s]) HIPCHK(c, hipEventSynchronize(c->stage_ev[s]));
    if (c->h_stage_cap[s] < b.bytes.size()) {
        if (c->h_stage[s]) HIPCHK(c, hipHostFree(c->h_stage[s]));
        size_t cap = std::max(b.bytes.size(), (size_t)4096);
        HIPCHK(c, hipHostMalloc((void**)&c->h_stage[s], cap));
        c->h_stage_cap[s] = cap;
    }
    memcpy(c->h_stage[s], b.bytes.data(), b.bytes.size());
    // one device blob per staging slot, so copies in flight never alias
    int rc = ensure(c, c->d_blob, psgpu_ctx::kStage * align256(std::max(b.bytes.size(), (size_t)1)) + 256);
    if (rc) return rc;
    size_t slot_bytes = c->d_blob.cap / psgpu_ctx::kStage / 256 * 256;
    if (slot_bytes < b.bytes.size()) {
        if ((rc = ensure(c, c->d_blob, psgpu_ctx::kStage * align256(b.bytes.size()) * 2))) return rc;
        slot_bytes = c->d_blob.cap / psgpu_ctx::kStage / 256 * 256;
    }
    uint8_t* dst = c->d_blob.p + (size_t)s * slot_bytes;
    HIPCHK(c, hipMemcpyAsync(dst, c->h_stage[s], b.bytes.size(), hipMemcpyHostToDevice, c->stream));
    if (!c->stage_ev[s]) HIPCHK(c, hipEventCreateWithFlags(&c->stage_ev[s], hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(c->stage_ev[s], c->stream));
    *out = dst;
    return PSGPU_OK;
}

// partial sums of the row scan: one per scan block, or one per setup block
int64_t row_partials(int64_t T)
{
    const int sb = psk_setup_block();
    return std::max<int64_t>(psk_scan_blocks(T), (T + sb - 1) / sb) + 1;
}

// size the per-flush scratch for a flush of T triangles, nvs varying floats
// per vertex and n_tiles raster tiles at the current span and piece capacities
int size_flush_scratch(psgpu_ctx* c, int64_t T, int nvs, int64_t n_tiles)
{
    FlushBufs& b = c->fb;
    int rc;
    if ((rc = ensure(c, b.tris, T)) || (rc = ensure(c, b.vary, (size_t)T * 3 * nvs)) ||
        (rc = ensure(c, b.grad, (size_t)T * 8)) || (rc = ensure(c, b.nrows, T)) || (rc = ensure(c, b.rowoff, T + 1)))
        return rc;
    if (b.span_cap == 0) {
        b.span_cap = std::min<int64_t>(std::max<int64_t>(1 << 20, 6 * T), c->max_spans);
        b.seg_cap = std::min<int64_t>(b.span_cap * 3 / 2, c->max_segs);
    }
    if (b.span_cap > c->max_spans || b.seg_cap > c->max_segs)
        return fail(c, PSGPU_ERR_NOMEM, "more than %lld spans or %lld pieces in one flush", (long long)c->max_spans,
                    (long long)c->max_segs);
    if ((rc = ensure(c, b.partial, row_partials(T))) ||
        (rc = ensure(c, b.bstart, b.span_cap / 64 + 2)) || (rc = ensure(c, b.spans, b.span_cap)) ||
        (rc = ensure(c, b.segs, b.seg_cap)) || (rc = ensure(c, b.ck, (size_t)b.seg_cap * nvs)) ||
        (rc = ensure(c, b.runs, (size_t)b.seg_cap * 2)) || (rc = ensure(c, b.runs2, (size_t)b.seg_cap * 2)) ||
        (rc = ensure(c, b.tstart, n_tiles + 1)) || (rc = ensure(c, b.tpartial, psk_scan_blocks(n_tiles) + 1)) ||
        (rc = ensure(c, b.heavy, psk_heavy_cap() + 1)))
        return rc;
    // a new tile-count array starts at zero (a reallocation may return the
    // old address, so the test is whether one happened)
    const uint64_t gen = c->buf_gen;
    if ((rc = ensure(c, b.tcnt, n_tiles + 1))) return rc;
    if (c->buf_gen != gen) HIPCHK(c, hipMemsetAsync(b.tcnt.p, 0, b.tcnt.cap * sizeof(unsigned long long), c->stream));
    return PSGPU_OK;
}

// run one flush: the whole pipeline over the draws of all its views
int run_flush(psgpu_ctx* c, const Cmd& cmd, uint8_t* staged)
{
    Blob blob;
    PsFrame f;
    int64_t T, n_pix;
    int nvs;
    bool progs[PS_N_PROGS] = {false}, shade_progs[PS_N_PROGS] = {false};
    int rc = build_blob(c, cmd, blob, f, T, nvs, progs, shade_progs, n_pix);
    if (rc) return rc;
    if (T == 0) return PSGPU_OK;
    uint8_t* dev = staged;
    if (!dev && (rc = stage_blob(c, blob, &dev))) return rc;
    PsDraw* dd = (PsDraw*)dev;
    f.views = (const PsView*)(dev + blob.off_views);
    if ((rc = size_flush_scratch(c, T, nvs, f.n_tiles))) return rc;
    f.ck_stride = nvs;
    // debug planes record the flushes of the display pass: one view of the
    // debug viewport into the default depth buffer (a shadow-map pass into a
    // depth texture is not recorded)
    bool dbg = false;
    if (c->debug) {
        const Targets& tg0 = cmd.views[0].tg;
        bool display_pass = tg0.depth == c->depth;
        if (display_pass && (f.n_views != 1 || c->dbg_w != tg0.vpw || c->dbg_h != tg0.vph))
            return fail(c, PSGPU_ERR_STATE, "debug planes need a single view of the debug viewport");
        if (display_pass) {
            dbg = true;
            f.dbg_last = c->dbg_last; f.dbg_rcount = c->dbg_rcount; f.dbg_rhash = c->dbg_rhash;
            f.dbg_pcount = c->dbg_pcount; f.dbg_phash = c->dbg_phash;
        }
    }

    hipStream_t s = c->stream;
    int nd = blob.nd;
    const PsView* dv = f.views;
    // a flush of one programme has one setup launch, which also leaves the
    // row totals of its blocks for the row scan
    int n_progs = 0;
    for (int k = 1; k < PS_N_PROGS; k++) n_progs += progs[k] ? 1 : 0;
    uint32_t* row_part = n_progs == 1 ? c->fb.partial.p : nullptr;
    {
        ProfScope p(c, "setup");
        for (int k = 1; k < PS_N_PROGS; k++)
            if (progs[k])
                psk_setup(s, k, blob.shares[k], dd, nd, T, c->fb.tris.p, c->fb.vary.p, nvs, c->fb.grad.p, c->fb.nrows.p,
                          row_part, c->d_cnt, dv);
    }
    {
        ProfScope p(c, "scan_rows");
        // row counts never depend on buffer capacities: always scanned, so the
        // largest row total of the log is known even after an overflow
        const int sb = psk_setup_block();
        if (row_part)
            psk_scan_parts(s, c->fb.nrows.p, T, c->fb.rowoff.p, row_part, sb, c->d_sizes + 0, c->fb.bstart.p,
                           c->fb.span_cap / 64 + 2);
        else
            psk_scan_rows(s, c->fb.nrows.p, T, c->fb.rowoff.p, c->fb.partial.p, c->d_sizes + 0, c->fb.bstart.p,
                          c->fb.span_cap / 64 + 2);
    }
    // k_span finds each row's triangle by a search in the row offsets,
    // started from the row scan's table of every 64th item's triangle and
    // bounded by the next entry, so runs of triangles without rows (culled,
    // off screen or sub-pixel) cost nothing
    int imm = 0;   // a flush of one programme that shades in the raster walk
    {
        const PsDraw* hd = (const PsDraw*)blob.bytes.data();
        if (nd > 0 && (hd[0].prog == PSGPU_PROC_CLOUD || hd[0].prog == PSGPU_PROC_CLOUD_SHADOW)) imm = hd[0].prog;
    }
    // the deferred raster takes the tiles with the longest lists first
    uint32_t* heavy = imm ? nullptr : c->fb.heavy.p;
    {
        ProfScope p(c, "span");
        psk_span(s, c->fb.tris.p, c->fb.rowoff.p, c->fb.bstart.p, T, c->fb.spans.p, c->fb.tcnt.p, dv,
                 f.n_views, f.chunk_w, f.tile_rows, c->fb.span_cap, heavy, c->d_cnt);
    }
    {
        ProfScope p(c, "scan_tiles");
        psk_scan_tiles(s, c->fb.tcnt.p, f.n_tiles, c->fb.tstart.p, c->fb.tpartial.p, c->d_sizes + 1, c->d_cnt, heavy);
    }
    f.tstart = c->fb.tstart.p;
    f.runs = c->fb.runs.p;
    f.runs2 = c->fb.runs2.p;
    f.heavy = heavy;
    {
        ProfScope p(c, "segments");
        psk_segments(s, c->fb.spans.p, c->fb.rowoff.p, T, f, c->fb.tris.p, dd, c->fb.vary.p, nvs, c->fb.tcnt.p,
                     c->fb.segs.p, c->fb.ck.p, c->fb.span_cap, c->fb.seg_cap, c->d_cnt);
    }
    if (imm) {
        ProfScope p(c, "raster");
        psk_raster_imm(s, imm, dbg, c->fb.spans.p, c->fb.segs.p, c->fb.tris.p, dd, f, c->tiling.waves_per_group,
                       c->d_cnt, c->fb.vary.p, nvs, c->fb.ck.p);
        HIPCHK(c, hipGetLastError());
        c->stats.flushes++;
        return PSGPU_OK;
    }
    {
        ProfScope p(c, "raster");
        psk_raster(s, dbg, c->fb.spans.p, c->fb.segs.p, c->fb.tris.p, dd, f, c->tiling.waves_per_group, c->d_cnt);
    }
    {
        ProfScope p(c, "shade");
        bool clears = true;   // the first shading launch also writes the fused colour clears
        for (int k = 1; k < PS_N_PROGS; k++)
            if (shade_progs[k]) {
                psk_shade(s, k, c->fb.spans.p, c->fb.segs.p, c->fb.ck.p, c->fb.tris.p, c->fb.vary.p, nvs, c->fb.grad.p, dd, f,
                          c->d_cnt, clears);
                clears = false;
            }
    }
    HIPCHK(c, hipGetLastError());
    c->stats.flushes++;
    return PSGPU_OK;
}

int run_clears(psgpu_ctx* c, const Cmd& cmd, uint8_t* staged)
{
    if (cmd.fills.empty()) return PSGPU_OK;
    Blob blob;
    blob.off_views = blob.off_fills = 0;
    blob.nf = (int)cmd.fills.size();
    blob.bytes.assign(cmd.fills.size() * sizeof(PsFill), 0);
    memcpy(blob.bytes.data(), cmd.fills.data(), blob.bytes.size());
    uint8_t* dev = staged;
    int rc;
    if (!dev && (rc = stage_blob(c, blob, &dev))) return rc;
    int64_t mx = 1;
    for (const PsFill& fl : cmd.fills) mx = std::max<int64_t>(mx, psk_fill_pieces(fl.rows, fl.width));
    ProfScope p(c, "clear");
    for (size_t i = 0; i < cmd.fills.size(); i += 65535)
        psk_fills(c->stream, (const PsFill*)dev + i, (int)std::min<size_t>(65535, cmd.fills.size() - i), mx, c->d_cnt);
    HIPCHK(c, hipGetLastError());
    return PSGPU_OK;
}

int graph_launch_n(psgpu_ctx* c, int id, int count);

int run_copy(psgpu_ctx* c, const CopyRegion& r)
{
    HIPCHK(c, hipEventRecord(c->copy_ev, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->copy_stream, c->copy_ev, 0));
    HIPCHK(c, hipMemcpy2DAsync(r.dst, r.dst_pitch, r.src, r.src_pitch, r.width_bytes, r.rows, hipMemcpyDeviceToHost,
                               c->copy_stream));
    return PSGPU_OK;
}

int run_cmd(psgpu_ctx* c, const Cmd& cmd, uint8_t* staged = nullptr)
{
    switch (cmd.type) {
    case CMD_FLUSH: return run_flush(c, cmd, staged);
    case CMD_CLEARS: return run_clears(c, cmd, staged);
    case CMD_GRAPH: return graph_launch_n(c, cmd.graph, cmd.count);
    case CMD_COPY: return run_copy(c, cmd.copy);
    case CMD_POST: {
        // The post-processor sees a finished frame only: after an overflow
        // earlier in the log the frame lacks that flush's draws, and the log
        // replay that follows runs this command again.  So it waits for the
        // work before it and is skipped while the overflow flag is set; every
        // logged post-process then runs its callback exactly once.  (Inside a
        // graph capture its work is recorded instead; a prepared graph fits
        // its scratch buffers.)
        if (!c->in_stream_capture) {
            int flag = 0;
            HIPCHK(c, hipMemcpyAsync(&flag, &c->d_cnt->overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (flag) return PSGPU_OK;
        }
        int rc = cmd.post(cmd.post_user, &cmd.post_frame, (void*)c->stream);
        if (rc) return fail(c, rc < 0 ? rc : PSGPU_ERR_DEVICE, "post-processor returned %d", rc);
        HIPCHK(c, hipGetLastError());
        return PSGPU_OK;
    }
    }
    return PSGPU_OK;
}

int do_sync(psgpu_ctx* c);

// Split the flush at log[idx] into two flushes of about half its triangles
// each, in submission order (the draw at the midpoint is cut into two draws
// of the same programme and state, its vertex streams offset by the first
// part's triangles).  Every pixel still sees its fragments in submission
// order, so the result is unchanged.  False when there is nothing to split.
bool split_flush(psgpu_ctx* c, size_t idx)
{
    if (idx >= c->log.size() || c->log[idx].type != CMD_FLUSH) return false;
    const Cmd& cmd = c->log[idx];
    int64_t T = 0;
    for (const Segment& sg : cmd.views)
        for (const PsDraw& d : sg.draws) T += d.tri_count;
    if (T < 2) return false;
    const int64_t half = T / 2;
    int64_t acc = 0;
    Cmd a, b;
    a.type = b.type = CMD_FLUSH;
    // clears the flush applied itself become a clear command before both parts
    Cmd clears;
    clears.type = CMD_CLEARS;
    for (const Segment& sg : cmd.views) {
        PsFill fl;
        memset(&fl, 0, sizeof(fl));
        if (sg.dclear) {
            fl.ptr = (uint32_t*)sg.tg.depth;
            fl.rows = 1;
            fl.width = (int64_t)sg.tg.depth_scan * sg.tg.depth_h;
            memcpy(&fl.value, &sg.dclear_value, 4);
            clears.fills.push_back(fl);
        }
        if (sg.cclear_rows > 0) {
            fl.ptr = sg.tg.colour;
            fl.rows = sg.cclear_rows;
            fl.width = sg.tg.colour_w;
            fl.scan = sg.tg.colour_scan;
            fl.value = sg.cclear_value;
            clears.fills.push_back(fl);
        }
    }
    for (const Segment& sg : cmd.views) {
        Segment sa, sb;
        sa.tg = sb.tg = sg.tg;
        for (const PsDraw& d : sg.draws) {
            if (acc + d.tri_count <= half) {
                sa.draws.push_back(d);
            } else if (acc >= half) {
                sb.draws.push_back(d);
            } else {
                const int64_t k = half - acc;
                PsDraw d1 = d, d2 = d;
                d1.tri_count = k;
                d2.tri_count = d.tri_count - k;
                for (int s = 0; s < 5; s++)
                    if (d2.slot[s]) d2.slot[s] += 3 * k * (int64_t)d2.unit[s];
                d2.id_base += k;
                sa.draws.push_back(d1);
                sb.draws.push_back(d2);
            }
            acc += d.tri_count;
        }
        if (!sa.draws.empty()) a.views.push_back(std::move(sa));
        if (!sb.draws.empty()) b.views.push_back(std::move(sb));
    }
    c->log[idx] = std::move(a);
    c->log.insert(c->log.begin() + idx + 1, std::move(b));
    if (!clears.fills.empty()) c->log.insert(c->log.begin() + idx, std::move(clears));
    return true;
}

// remember whether the overflow flag was already set when command i finished:
// the first command that finished with it set is where a replay starts.
// Only an eager flush can set it (clears, copies and post-processes have no
// intermediate buffers, and a graph replay fits the buffers it was sized
// for), so the first command found set is always a flush and the others are
// marked clear without a copy.
int mark_done(psgpu_ctx* c, int i)
{
    if (c->log[i].type != CMD_FLUSH) {
        c->h_flags[i] = 0;
        return PSGPU_OK;
    }
    HIPCHK(c, hipMemcpyAsync(c->h_flags + i, &c->d_cnt->overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    return PSGPU_OK;
}

// a command is either recorded (between capture begin/end) or logged and run
int submit(psgpu_ctx* c, Cmd&& cmd)
{
    if (c->capturing) {
        c->capture_list.push_back(std::move(cmd));
        return PSGPU_OK;
    }
    if ((int)c->log.size() >= psgpu_ctx::kFlagCap - 1) {
        int rc = do_sync(c);
        if (rc) return rc;
    }
    c->log.push_back(std::move(cmd));
    int rc = run_cmd(c, c->log.back());
    if (rc) return rc;
    return mark_done(c, (int)c->log.size() - 1);
}

// Clears that a flush can apply itself: a clear of a view's whole depth
// buffer (when the buffer is exactly the viewport, without padding), taken
// into the raster kernel's depth tiles, and a clear of its colour target's
// rows (when the target is exactly the viewport), written by the shading
// kernel to the pixels no fragment colours.  Only for flushes that shade
// deferred (the raster walk that blends reads the colour beneath), and only
// when no other pending clear touches the buffer.  The fills taken are
// removed from `fills`.
void fuse_clears(std::vector<PsFill>& fills, std::vector<Segment>& views, bool shades)
{
    if (fills.empty()) return;
    std::vector<bool> taken(fills.size(), false);
    auto only = [&](const void* p) {
        int found = -1;
        for (size_t i = 0; i < fills.size(); i++)
            if ((const void*)fills[i].ptr == p) {
                if (found >= 0) return -1;
                found = (int)i;
            }
        return found;
    };
    for (Segment& sg : views) {
        const Targets& tg = sg.tg;
        int i = only(tg.depth);
        if (i >= 0 && tg.depth_scan == tg.vpw && tg.depth_w == tg.vpw && tg.depth_h == tg.vph &&
            fills[i].rows == 1 && fills[i].width == (int64_t)tg.depth_scan * tg.depth_h) {
            sg.dclear = true;
            memcpy(&sg.dclear_value, &fills[i].value, 4);
            taken[i] = true;
        }
        i = only(tg.colour);
        if (shades && i >= 0 && tg.colour_w == tg.vpw && tg.colour_h == tg.vph && fills[i].width == tg.colour_w &&
            fills[i].scan == tg.colour_scan && fills[i].rows <= tg.colour_h && fills[i].rows > 0) {
            sg.cclear_rows = (int)fills[i].rows;
            sg.cclear_value = fills[i].value;
            taken[i] = true;
        }
    }
    size_t k = 0;
    for (size_t i = 0; i < fills.size(); i++)
        if (!taken[i]) fills[k++] = fills[i];
    fills.resize(k);
}

// whether a flush of these views shades in the deferred kernel
bool flush_shades(const std::vector<Segment>& views, int imm)
{
    if (imm) return false;
    for (const Segment& sg : views)
        for (const PsDraw& d : sg.draws)
            if (d.writes_colour) return true;
    return false;
}

// pending clears first, then the pending draws of every view as one flush
int queue_to_log(psgpu_ctx* c)
{
    int rc;
    if (!c->queue.empty() && !c->queue_imm) fuse_clears(c->pending_fills, c->queue, flush_shades(c->queue, 0));
    if (!c->pending_fills.empty()) {
        Cmd cmd;
        cmd.type = CMD_CLEARS;
        cmd.fills.swap(c->pending_fills);
        if ((rc = submit(c, std::move(cmd)))) return rc;
    }
    if (!c->queue.empty()) {
        Cmd cmd;
        cmd.type = CMD_FLUSH;
        cmd.views.swap(c->queue);
        c->queue_tris = 0;
        c->queue_imm = 0;
        if ((rc = submit(c, std::move(cmd)))) return rc;
    }
    return PSGPU_OK;
}

bool queue_uses(psgpu_ctx* c, const void* buf)
{
    for (const Segment& sg : c->queue)
        if ((const void*)sg.tg.colour == buf || (const void*)sg.tg.depth == buf) return true;
    return false;
}

int add_fill(psgpu_ctx* c, void* ptr, int64_t rows, int64_t width, int64_t scan, uint32_t value)
{
    if (queue_uses(c, ptr)) {       // draws already queued for this buffer come first
        int rc = queue_to_log(c);
        if (rc) return rc;
    }
    if (rows <= 0 || width <= 0) return PSGPU_OK;
    PsFill fl;
    memset(&fl, 0, sizeof(fl));
    fl.ptr = (uint32_t*)ptr;
    fl.rows = rows;
    fl.width = width;
    fl.scan = scan;
    fl.value = value;
    c->pending_fills.push_back(fl);
    return PSGPU_OK;
}

int do_sync(psgpu_ctx* c)
{
    if (c->capturing) return fail(c, PSGPU_ERR_STATE, "not allowed while recording a command list");
    int rc = queue_to_log(c);
    if (rc) return rc;
    const int64_t span_cap0 = c->fb.span_cap, seg_cap0 = c->fb.seg_cap;
    for (int attempt = 0; attempt < 64; attempt++) {
        HIPCHK(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(PsCounters), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->h_sizes, c->d_sizes, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamSynchronize(c->copy_stream));
        if (c->profiling) prof_collect(c);
        if (!c->h_cnt->overflow) {
            int64_t sum[PS_ST_N] = {0};
            for (int k = 0; k < PS_STRIPES; k++)
                for (int f = 0; f < PS_ST_N; f++) sum[f] += (int64_t)c->h_cnt->stat[k][f];
            c->stats.tris_culled += sum[PS_ST_CULLED];
            c->stats.tris_zrejected += sum[PS_ST_ZREJ];
            c->stats.tris_dropped += sum[PS_ST_DROPPED];
            c->stats.spans += sum[PS_ST_SPANS];
            c->stats.frags_rasterised += sum[PS_ST_FRAGS];
            c->stats.frags_shaded += sum[PS_ST_SHADED];
            c->stats.segments += sum[PS_ST_SEGS];
            c->log.clear();
            for (const Fence& fe : c->fences) {
                c->fence_ev_pool.push_back(fe.main_ev);
                c->fence_ev_pool.push_back(fe.copy_ev);
            }
            c->fences.clear();
            HIPCHK(c, hipMemsetAsync(c->d_cnt, 0, sizeof(PsCounters), c->stream));
            HIPCHK(c, hipMemsetAsync(c->d_sizes, 0, 4 * sizeof(uint32_t), c->stream));
            return PSGPU_OK;
        }
        if (c->h_cnt->overflow & 2)
            return fail(c, PSGPU_ERR_STATE, "internal: a tile list did not match its counts");
        // grow and replay the log; a flush that needs more than one flush may
        // hold is split in two (and again, if a half still does)
        int64_t need_spans = (int64_t)c->h_sizes[2], need_segs = (int64_t)c->h_sizes[3];
        size_t first = 0;
        while (first < c->log.size() && !c->h_flags[first]) first++;
        // only eager flushes carry a copied flag; the others cannot overflow
        if (first == c->log.size())
            return fail(c, PSGPU_ERR_STATE, "internal: the scratch buffers overflowed outside an eager flush");
        bool split = false;
        if (need_spans > c->max_spans || need_segs > c->max_segs) {
            if (c->log.size() + 2 >= (size_t)psgpu_ctx::kFlagCap || !split_flush(c, first)) break;
            split = true;
            c->stats.splits++;
        }
        if (need_spans > c->fb.span_cap) {
            c->fb.span_cap = std::min<int64_t>(need_spans + need_spans / 4 + 1024, c->max_spans);
            c->fb.seg_cap = std::min<int64_t>(std::max<int64_t>(c->fb.seg_cap, c->fb.span_cap * 3 / 2), c->max_segs);
        } else if (need_segs > c->fb.seg_cap) {
            c->fb.seg_cap = std::min<int64_t>(need_segs + need_segs / 4 + 1024, c->max_segs);
        } else if (!split) {
            c->fb.span_cap = std::min<int64_t>(c->fb.span_cap * 2, c->max_spans);
            c->fb.seg_cap = std::min<int64_t>(c->fb.seg_cap * 2, c->max_segs);
        }
        c->stats.retries++;
        // commands that finished before the overflow are complete (their
        // statistics are committed and stay); replay the rest
        c->log.erase(c->log.begin(), c->log.begin() + first);
        HIPCHK(c, hipMemsetAsync(&c->d_cnt->overflow, 0, sizeof(int), c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_cnt->pend, 0, sizeof(c->d_cnt->pend), c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_sizes, 0, 4 * sizeof(uint32_t), c->stream));
        // a flush cut short leaves tile counts behind
        if (c->fb.tcnt.p)
            HIPCHK(c, hipMemsetAsync(c->fb.tcnt.p, 0, c->fb.tcnt.cap * sizeof(unsigned long long), c->stream));
        for (size_t i = 0; i < c->log.size(); i++) {
            if ((rc = run_cmd(c, c->log[i]))) return rc;
            if ((rc = mark_done(c, (int)i))) return rc;
        }
    }
    // give up with the capacities as they were, so later flushes stay usable
    c->fb.span_cap = span_cap0;
    c->fb.seg_cap = seg_cap0;
    return fail(c, PSGPU_ERR_NOMEM, "intermediate buffers kept overflowing");
}

int graph_capture(psgpu_ctx* c, Graph& g)
{
    if (g.exec) { HIPCHK(c, hipGraphExecDestroy(g.exec)); g.exec = nullptr; }
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->in_stream_capture = true;
    int rc = PSGPU_OK;
    for (size_t i = 0; i < g.cmds.size() && !rc; i++) rc = run_cmd(c, g.cmds[i], g.blobs[i]);
    c->in_stream_capture = false;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(c, PSGPU_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(c, PSGPU_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    g.gen = c->buf_gen;
    g.span_cap = c->fb.span_cap;
    g.seg_cap = c->fb.seg_cap;
    g.stale = false;
    return PSGPU_OK;
}

// size the scratch buffers for every flush of a graph at the current
// capacities, then upload its commands' device records (their winner planes
// point into the scratch) and capture it
int graph_rebuild(psgpu_ctx* c, Graph& g)
{
    for (int pass = 0; pass < 2; pass++) {   // sizing may move the winner plane: records come after it
        for (size_t i = 0; i < g.cmds.size(); i++) {
            if (g.cmds[i].type != CMD_FLUSH || !g.blobs[i]) continue;
            Blob b;
            PsFrame f;
            int64_t T, npx;
            int nvs;
            bool p1[PS_N_PROGS] = {false}, p2[PS_N_PROGS] = {false};
            int rc = build_blob(c, g.cmds[i], b, f, T, nvs, p1, p2, npx);
            if (rc) return rc;
            if (pass == 0) {
                if (T > 0 && (rc = size_flush_scratch(c, T, nvs, f.n_tiles))) return rc;
                continue;
            }
            HIPCHK(c, hipMemcpyAsync(g.blobs[i], b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
    }
    return graph_capture(c, g);
}

// a recorded flush that had to be split cannot be replayed as captured
int graph_too_large(psgpu_ctx* c)
{
    return fail(c, PSGPU_ERR_NOMEM, "a flush of the recorded command list needs more than %lld spans or %lld pieces; "
                "record it with fewer triangles per flush (psgpu_flush between draws)", (long long)c->max_spans,
                (long long)c->max_segs);
}

// A graph whose capture is stale is first run once eagerly: that sizes the
// scratch buffers for the current tiling and geometry (overflow replays
// included).  The eager run counts as the first of the launches asked for.
int graph_prepare(psgpu_ctx* c, Graph& g)
{
    int rc = do_sync(c);
    if (rc) return rc;
    const int64_t splits = c->stats.splits;
    for (const Cmd& cmd : g.cmds) {
        Cmd copy = cmd;
        c->log.push_back(std::move(copy));
        if ((rc = run_cmd(c, c->log.back()))) return rc;
        if ((rc = mark_done(c, (int)c->log.size() - 1))) return rc;
    }
    if ((rc = do_sync(c))) return rc;
    if (c->stats.splits != splits) return graph_too_large(c);
    return graph_rebuild(c, g);
}

// whether a draw of the graph reads the vertex stream at ptr
bool graph_reads(const Graph& g, const uint8_t* ptr)
{
    for (const Cmd& cmd : g.cmds)
        for (const Segment& sg : cmd.views)
            for (const PsDraw& d : sg.draws)
                for (int s = 0; s < 5; s++)
                    if (d.slot[s] == ptr) return true;
    return false;
}

// a graph is current when it was captured against the scratch buffers and
// capacities in use and no vertex stream it reads changed since
bool graph_current(psgpu_ctx* c, const Graph& g)
{
    return g.exec && !g.stale && g.gen == c->buf_gen && g.span_cap == c->fb.span_cap && g.seg_cap == c->fb.seg_cap;
}

int graph_launch_n(psgpu_ctx* c, int id, int count)
{
    Graph& g = c->graphs[id];
    if (!graph_current(c, g)) {
        // buffers or capacities only grew since this launch was logged (a
        // replay of the log after an overflow): records and capture suffice
        int rc = graph_rebuild(c, g);
        if (rc) return rc;
    }
    for (int i = 0; i < count; i++) HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
    c->stats.flushes += g.flushes * count;
    return PSGPU_OK;
}

int need_ctx(psgpu_ctx* c) { return c ? PSGPU_OK : PSGPU_ERR_INVALID; }

// ---- the static scratch array ----------------------------------------------
// Each device holds its own copy of `scratch`; its address is looked up once
// per device, with that device current, when the first context on it is
// created.  Calls of psgpu_sync_noalloc share the array of their device, so
// they are serialised by g_static_lock.
const int kMaxStaticDevices = 64;
std::mutex g_static_lock;
uint8_t* g_static_ptr[kMaxStaticDevices] = {nullptr};
hipError_t g_static_err[kMaxStaticDevices] = {hipSuccess};

void static_scratch_init(int device)
{
    if (device < 0 || device >= kMaxStaticDevices) return;
    std::lock_guard<std::mutex> g(g_static_lock);
    if (g_static_ptr[device]) return;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_static_ptr[device] = (uint8_t*)p;
    else g_static_err[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
}

// bump carving of the static array
struct Arena {
    uint8_t* base;
    size_t cap, used = 0;
    template <typename T> bool take(DevBuf<T>& b, size_t n)
    {
        size_t bytes = align256(std::max<size_t>(n, 1) * sizeof(T));
        if (used + bytes > cap) return false;
        b.p = (T*)(base + used);
        b.cap = std::max<size_t>(n, 1);
        used += bytes;
        return true;
    }
    uint8_t* take_bytes(size_t n)
    {
        size_t bytes = align256(std::max<size_t>(n, 1));
        if (used + bytes > cap) return nullptr;
        uint8_t* p = base + used;
        used += bytes;
        return p;
    }
};

// the quantities a flush's scratch depends on, as build_blob lays them out
struct FlushShape { int64_t T = 0, n_pix = 0, n_tiles = 0; int nvs = 1; };

FlushShape flush_shape(psgpu_ctx* c, const Cmd& cmd)
{
    FlushShape sh;
    const int cw = c->tiling.chunk_width, tr = c->tiling.tile_rows, tc = 64 / tr;
    for (const Segment& sg : cmd.views) {
        int n_chunks = (sg.tg.vpw + cw - 1) / cw;
        sh.n_tiles += (int64_t)((n_chunks + tc - 1) / tc) * ((sg.tg.vph + tr - 1) / tr);
        sh.n_pix += (int64_t)sg.tg.vpw * sg.tg.vph;
        for (const PsDraw& d : sg.draws) {
            sh.T += d.tri_count;
            sh.nvs = std::max(sh.nvs, d.nv);
        }
    }
    return sh;
}

// carve every buffer of run_flush for a flush of this shape; the span and
// piece capacities take what is left (pieces at 1.5 per span, as the
// allocating path starts them)
bool carve_flush(Arena& a, const FlushShape& sh, FlushBufs& b)
{
    const int64_t T = sh.T;
    if (!a.take(b.winner, sh.n_pix) || !a.take(b.tris, T) || !a.take(b.vary, (size_t)T * 3 * sh.nvs) ||
        !a.take(b.grad, (size_t)T * 8) || !a.take(b.nrows, T) || !a.take(b.rowoff, T + 1) ||
        !a.take(b.partial, row_partials(T)) || !a.take(b.tcnt, sh.n_tiles + 1) ||
        !a.take(b.tstart, sh.n_tiles + 1) || !a.take(b.tpartial, psk_scan_blocks(sh.n_tiles) + 1) ||
        !a.take(b.heavy, psk_heavy_cap() + 1))
        return false;
    // bytes per span: record and row-start table entry, and 1.5 pieces of
    // record, checkpoint and two run records
    const double per_span = 32 + 4.0 / 64 + 1.5 * (sizeof(PsSeg) + 4.0 * sh.nvs + 16);
    size_t left = a.cap - a.used;
    int64_t spans = (int64_t)((double)left / per_span);
    spans = std::min<int64_t>(spans, (int64_t)1 << 26);
    for (; spans >= 1024; spans -= spans / 16 + 1) {
        Arena t = a;
        FlushBufs x = b;
        x.span_cap = spans;
        x.seg_cap = spans * 3 / 2;
        if (t.take(x.bstart, x.span_cap / 64 + 2) &&
            t.take(x.spans, x.span_cap) && t.take(x.segs, x.seg_cap) &&
            t.take(x.ck, (size_t)x.seg_cap * sh.nvs) && t.take(x.runs, (size_t)x.seg_cap * 2) &&
            t.take(x.runs2, (size_t)x.seg_cap * 2)) {
            a = t;
            b = x;
            return true;
        }
    }
    return false;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int psgpu_create(int device, int width, int height, unsigned flags, psgpu_ctx** out)
{
    if (!out) return PSGPU_ERR_INVALID;
    *out = nullptr;
    if (width < 1 || height < 1) return PSGPU_ERR_INVALID;
    if (width > kMaxViewW || height > kMaxViewH) return PSGPU_ERR_RANGE;
    // default depth scanline (pipeline.cpp:31); it is shorter than the row when
    // width % 4 == 1, which makes the reference alias the row ends
    int dscan = (int)(width / 4.0f + 0.5f) * 4;
    if (dscan < width) return PSGPU_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PSGPU_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PSGPU_ERR_DEVICE;
    static_scratch_init(device);
    psgpu_ctx* c = new psgpu_ctx();
    c->device = device;
    c->dev_w = c->vp_w = width;
    c->dev_h = c->vp_h = height;
    c->depth_scan = dscan;
    if (flags & PSGPU_CREATE_TINY_SCRATCH) {   // start tiny: exercises growth and log replay
        c->fb.span_cap = 256;
        c->fb.seg_cap = 256;
    }
    if (flags & PSGPU_CREATE_SMALL_FLUSHES) {  // a low ceiling per flush: exercises the splitting of flushes
        c->max_spans = 4096;
        c->max_segs = 6144;
    }
    c->uniforms.resize(kMaxUniforms);
    memset(&c->stats, 0, sizeof(c->stats));
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&c->copy_ev, hipEventDisableTiming) == hipSuccess &&
              hipMalloc((void**)&c->display, (size_t)width * height * 4) == hipSuccess &&
              hipMalloc((void**)&c->depth, (size_t)dscan * height * 4) == hipSuccess &&
              hipMalloc((void**)&c->d_cnt, sizeof(PsCounters)) == hipSuccess &&
              hipHostMalloc((void**)&c->h_cnt, sizeof(PsCounters)) == hipSuccess &&
              hipMalloc((void**)&c->d_sizes, 4 * sizeof(uint32_t)) == hipSuccess &&
              hipHostMalloc((void**)&c->h_sizes, 4 * sizeof(uint32_t)) == hipSuccess &&
              hipHostMalloc((void**)&c->h_flags, psgpu_ctx::kFlagCap * sizeof(int)) == hipSuccess &&
              hipMemsetAsync(c->display, 0, (size_t)width * height * 4, c->stream) == hipSuccess &&
              hipMemsetAsync(c->depth, 0, (size_t)dscan * height * 4, c->stream) == hipSuccess &&
              hipMemsetAsync(c->d_cnt, 0, sizeof(PsCounters), c->stream) == hipSuccess &&
              hipMemsetAsync(c->d_sizes, 0, 4 * sizeof(uint32_t), c->stream) == hipSuccess &&
              hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) {
        psgpu_destroy(c);
        return PSGPU_ERR_DEVICE;
    }
    memset(c->h_cnt, 0, sizeof(PsCounters));
    *out = c;
    return PSGPU_OK;
}

int psgpu_destroy(psgpu_ctx* c)
{
    if (!c) return PSGPU_ERR_INVALID;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->copy_stream) hipStreamSynchronize(c->copy_stream);
    auto fr = [](void* p) { if (p) hipFree(p); };
    fr(c->display); fr(c->depth); fr(c->d_cnt); fr(c->d_sizes);
    if (c->h_cnt) hipHostFree(c->h_cnt);
    if (c->h_sizes) hipHostFree(c->h_sizes);
    if (c->h_flags) hipHostFree(c->h_flags);
    for (auto& t : c->tex) fr(t.d);
    for (auto& v : c->vbo) fr(v.d);
    fr(c->d_blob.p); fr(c->fb.tris.p); fr(c->fb.vary.p); fr(c->fb.nrows.p); fr(c->fb.rowoff.p); fr(c->fb.partial.p);
    fr(c->fb.bstart.p); fr(c->fb.spans.p); fr(c->fb.tcnt.p); fr(c->fb.tstart.p); fr(c->fb.tpartial.p);
    fr(c->fb.runs.p); fr(c->fb.runs2.p); fr(c->fb.heavy.p); fr(c->fb.winner.p);
    fr(c->fb.segs.p); fr(c->fb.ck.p); fr(c->fb.grad.p);
    fr(c->dbg_last); fr(c->dbg_rcount); fr(c->dbg_rhash); fr(c->dbg_pcount); fr(c->dbg_phash);
    for (int i = 0; i < psgpu_ctx::kStage; i++) {
        if (c->h_stage[i]) hipHostFree(c->h_stage[i]);
        if (c->stage_ev[i]) hipEventDestroy(c->stage_ev[i]);
    }
    for (auto& g : c->graphs) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        for (auto p : g.blobs) if (p) hipFree(p);
    }
    for (auto& r : c->prof_pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    for (auto& fe : c->fences) { hipEventDestroy(fe.main_ev); hipEventDestroy(fe.copy_ev); }
    for (auto e : c->fence_ev_pool) hipEventDestroy(e);
    if (c->copy_ev) hipEventDestroy(c->copy_ev);
    if (c->copy_stream) hipStreamDestroy(c->copy_stream);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return PSGPU_OK;
}

const char* psgpu_last_error(psgpu_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* psgpu_stream(psgpu_ctx* c) { return c ? (void*)c->stream : nullptr; }

// ---- textures --------------------------------------------------------------
int psgpu_texture_create(psgpu_ctx* c, int width, int scanline, int height, int elem_len, int extra_layers,
                         int wrap_mode, int mip_levels, const void* pixels, int* out_idx)
{
    if (need_ctx(c) || !out_idx) return PSGPU_ERR_INVALID;
    if (elem_len != 1 && elem_len != 4) return fail(c, PSGPU_ERR_INVALID, "elemLen must be 1 or 4 (fbo.cpp:21-24)");
    if (width < 1 || height < 1 || scanline < width * elem_len) return fail(c, PSGPU_ERR_INVALID, "bad texture size");
    if (width > (1 << 22) || height > (1 << 22) || scanline >= (1 << 24))
        return fail(c, PSGPU_ERR_RANGE, "texture of %dx%d texels, %d-byte rows: at most 2^22 texels a side and rows "
                    "below 2^24 bytes", width, height, scanline);
    if (extra_layers < 0 || extra_layers > PS_MAX_LAYERS - 1) return fail(c, PSGPU_ERR_RANGE, "extraLayers out of range");
    if (wrap_mode != PSGPU_CLAMP && wrap_mode != PSGPU_WRAP) return fail(c, PSGPU_ERR_INVALID, "bad wrap mode");
    int maxl = 1;
    while (((width >> maxl) > 0 || (height >> maxl) > 0) && maxl < PS_MAX_LEVELS) maxl++;
    if (mip_levels < 1 || mip_levels > maxl) return fail(c, PSGPU_ERR_RANGE, "mip levels out of range");
    if (mip_levels > 1 && scanline != width * elem_len) return fail(c, PSGPU_ERR_INVALID, "mipmapped textures need a tight scanline");
    Tex t;
    t.w = width; t.h = height; t.scan = scanline; t.elem = elem_len; t.layers = 1 + extra_layers; t.wrap = wrap_mode;
    t.levels = mip_levels;
    int64_t off = (int64_t)scanline * height;
    t.level_off[0] = 0;
    for (int l = 1; l < mip_levels; l++) {
        t.level_off[l] = off;
        int w = std::max(1, width >> l), h = std::max(1, height >> l);
        off += (int64_t)w * h * elem_len;
    }
    t.layer_bytes = (off + 255) / 256 * 256;
    // the samplers address texels of a layer with 32-bit offsets
    if (t.layer_bytes > (int64_t)0xffffffffu) return fail(c, PSGPU_ERR_INVALID, "a texture layer must stay below 4 GiB");
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMalloc((void**)&t.d, (size_t)t.layer_bytes * t.layers));
    HIPCHK(c, hipMemsetAsync(t.d, 0, (size_t)t.layer_bytes * t.layers, c->stream));
    if (pixels) HIPCHK(c, hipMemcpyAsync(t.d, pixels, (size_t)scanline * height, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    t.alive = true;
    size_t i = 0;
    for (; i < c->tex.size(); i++) if (!c->tex[i].alive) break;   // tex.cpp:6-16: first vacant slot
    if (i == c->tex.size()) c->tex.push_back(t); else c->tex[i] = t;
    *out_idx = (int)i;
    return PSGPU_OK;
}

static int level_bytes(const Tex& t, int level, int* w, int* h, int* scan)
{
    *w = std::max(1, t.w >> level);
    *h = std::max(1, t.h >> level);
    *scan = level == 0 ? t.scan : *w * t.elem;
    return PSGPU_OK;
}

int psgpu_texture_upload(psgpu_ctx* c, int tex, int layer, int level, const void* pixels, size_t bytes)
{
    if (need_ctx(c) || !pixels) return PSGPU_ERR_INVALID;
    if (!valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
    Tex& t = c->tex[tex];
    if (layer < 0 || layer >= t.layers || level < 0 || level >= t.levels) return fail(c, PSGPU_ERR_RANGE, "layer/level out of range");
    int w, h, scan;
    level_bytes(t, level, &w, &h, &scan);
    if (bytes != (size_t)scan * h) return fail(c, PSGPU_ERR_INVALID, "expected %zu bytes", (size_t)scan * h);
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(t.d + t.layer_bytes * layer + t.level_off[level], pixels, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSGPU_OK;
}

int psgpu_texture_download(psgpu_ctx* c, int tex, int layer, int level, void* pixels, size_t bytes)
{
    if (need_ctx(c) || !pixels) return PSGPU_ERR_INVALID;
    if (!valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
    Tex& t = c->tex[tex];
    if (layer < 0 || layer >= t.layers || level < 0 || level >= t.levels) return fail(c, PSGPU_ERR_RANGE, "layer/level out of range");
    int w, h, scan;
    level_bytes(t, level, &w, &h, &scan);
    if (bytes != (size_t)scan * h) return fail(c, PSGPU_ERR_INVALID, "expected %zu bytes", (size_t)scan * h);
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(pixels, t.d + t.layer_bytes * layer + t.level_off[level], bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSGPU_OK;
}

int psgpu_texture_generate_mips(psgpu_ctx* c, int tex)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (!valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
    int rc = do_sync(c);
    if (rc) return rc;
    Tex& t = c->tex[tex];
    for (int L = 0; L < t.layers; L++)
        for (int l = 1; l < t.levels; l++) {
            int ws, hs, ss, wd, hd, sd;
            level_bytes(t, l - 1, &ws, &hs, &ss);
            level_bytes(t, l, &wd, &hd, &sd);
            psk_mip(c->stream, t.d + t.layer_bytes * L + t.level_off[l - 1], ws, hs, ss, t.d + t.layer_bytes * L + t.level_off[l],
                    wd, hd, t.elem);
        }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSGPU_OK;
}

int psgpu_texture_destroy(psgpu_ctx* c, int tex)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (tex < 0 || tex >= (int)c->tex.size()) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
    if (!c->tex[tex].alive) return PSGPU_OK;
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipFree(c->tex[tex].d));
    c->tex[tex] = Tex();
    return PSGPU_OK;
}

int psgpu_texture_info(psgpu_ctx* c, int tex, int level, int* width, int* scanline, int* height, int* elem_len)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (!valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
    Tex& t = c->tex[tex];
    if (level < 0 || level >= t.levels) return fail(c, PSGPU_ERR_RANGE, "level out of range");
    int w, h, s;
    level_bytes(t, level, &w, &h, &s);
    if (width) *width = w;
    if (scanline) *scanline = s;
    if (height) *height = h;
    if (elem_len) *elem_len = t.elem;
    return PSGPU_OK;
}

// ---- vertex streams ---------------------------------------------------------
int psgpu_vbo_create(psgpu_ctx* c, size_t unit_bytes, size_t unit_count, int* out_vbo)
{
    if (need_ctx(c) || !out_vbo) return PSGPU_ERR_INVALID;
    if (unit_bytes == 0 || unit_count == 0) return fail(c, PSGPU_ERR_INVALID, "empty VBO");
    Vbo v;
    v.unit = unit_bytes;
    v.count = unit_count;
    if (hipMalloc((void**)&v.d, unit_bytes * unit_count) != hipSuccess) return fail(c, PSGPU_ERR_NOMEM, "VBO allocation failed");
    v.alive = true;
    size_t i = 0;
    for (; i < c->vbo.size(); i++) if (!c->vbo[i].alive) break;
    if (i == c->vbo.size()) c->vbo.push_back(v); else c->vbo[i] = v;
    *out_vbo = (int)i;
    return PSGPU_OK;
}

int psgpu_vbo_update(psgpu_ctx* c, int vbo, const void* src)
{
    if (need_ctx(c) || !src) return PSGPU_ERR_INVALID;
    if (vbo < 0 || vbo >= (int)c->vbo.size() || !c->vbo[vbo].alive) return fail(c, PSGPU_ERR_RANGE, "VBO out of range");
    int rc = do_sync(c);
    if (rc) return rc;
    Vbo& v = c->vbo[vbo];
    HIPCHK(c, hipMemcpyAsync(v.d, src, v.unit * v.count, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // new geometry can need more spans and pieces than a graph that draws it
    // was captured with: such graphs are run eagerly and captured again
    for (Graph& g : c->graphs)
        if (g.alive && graph_reads(g, v.d)) g.stale = true;
    return PSGPU_OK;
}

int psgpu_vbo_destroy(psgpu_ctx* c, int vbo)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (vbo < 0 || vbo >= (int)c->vbo.size()) return fail(c, PSGPU_ERR_RANGE, "VBO out of range");
    if (!c->vbo[vbo].alive) return PSGPU_OK;
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipFree(c->vbo[vbo].d));
    c->vbo[vbo] = Vbo();
    for (auto& a : c->vao)
        if (a.alive)
            for (int s = 0; s < kMaxSlots; s++) if (a.slot[s] == vbo) a.slot[s] = -1;
    return PSGPU_OK;
}

int psgpu_vao_create(psgpu_ctx* c, int* out_vao)
{
    if (need_ctx(c) || !out_vao) return PSGPU_ERR_INVALID;
    Vao a;
    a.alive = true;
    for (int s = 0; s < kMaxSlots; s++) a.slot[s] = -1;
    size_t i = 0;
    for (; i < c->vao.size(); i++) if (!c->vao[i].alive) break;   // pipeline.cpp:118-136
    if (i == c->vao.size()) c->vao.push_back(a); else c->vao[i] = a;
    *out_vao = (int)i;
    return PSGPU_OK;
}

static int check_vao_slot(psgpu_ctx* c, int vao, int slot)
{
    if (vao < 0 || vao >= (int)c->vao.size() || !c->vao[vao].alive) return fail(c, PSGPU_ERR_RANGE, "vao out of range");
    if (slot < 0 || slot >= kMaxSlots) return fail(c, PSGPU_ERR_RANGE, "VBO slot out of range");
    return PSGPU_OK;
}

int psgpu_vao_attach(psgpu_ctx* c, int vao, int slot, int vbo, int* out_previous)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    int rc = check_vao_slot(c, vao, slot);
    if (rc) return rc;
    if (vbo != -1 && (vbo < 0 || vbo >= (int)c->vbo.size() || !c->vbo[vbo].alive)) return fail(c, PSGPU_ERR_RANGE, "VBO out of range");
    if (out_previous) *out_previous = c->vao[vao].slot[slot];
    c->vao[vao].slot[slot] = vbo;
    return PSGPU_OK;
}

int psgpu_vao_detach(psgpu_ctx* c, int vao, int slot, int* out_previous) { return psgpu_vao_attach(c, vao, slot, -1, out_previous); }

int psgpu_vao_get(psgpu_ctx* c, int vao, int slot, int* out_vbo)
{
    if (need_ctx(c) || !out_vbo) return PSGPU_ERR_INVALID;
    int rc = check_vao_slot(c, vao, slot);
    if (rc) return rc;
    *out_vbo = c->vao[vao].slot[slot];
    return PSGPU_OK;
}

int psgpu_vao_destroy(psgpu_ctx* c, int vao)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (vao < 0 || vao >= (int)c->vao.size()) return fail(c, PSGPU_ERR_RANGE, "vao out of range");
    if (!c->vao[vao].alive) return PSGPU_OK;
    int slots[kMaxSlots];
    memcpy(slots, c->vao[vao].slot, sizeof(slots));
    c->vao[vao].alive = false;
    for (int s = 0; s < kMaxSlots; s++)
        if (slots[s] >= 0) {
            int rc = psgpu_vbo_destroy(c, slots[s]);
            if (rc) return rc;
        }
    return PSGPU_OK;
}

// ---- programmes -------------------------------------------------------------
int psgpu_program_create(psgpu_ctx* c, int vp, int ip, int fp, int* out_prog)
{
    if (need_ctx(c) || !out_prog) return PSGPU_ERR_INVALID;
    if (prog_nv(vp) < 0 || prog_nv(ip) < 0 || prog_nv(fp) < 0)
        return fail(c, PSGPU_ERR_UNSUPPORTED, "processor has no GPU twin");
    if (vp != ip || ip != fp)
        return fail(c, PSGPU_ERR_UNSUPPORTED, "VP/IP/FP must belong to one built-in programme (varying layouts differ)");
    Prog p;
    p.alive = true; p.vp = vp; p.ip = ip; p.fp = fp;
    size_t i = 0;
    for (; i < c->prog.size(); i++) if (!c->prog[i].alive) break;   // prog.cpp:52-60
    if (i == c->prog.size()) c->prog.push_back(p); else c->prog[i] = p;
    *out_prog = (int)i;
    return PSGPU_OK;
}

int psgpu_program_destroy(psgpu_ctx* c, int prog)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (prog < 0 || prog >= (int)c->prog.size()) return fail(c, PSGPU_ERR_RANGE, "programme out of range");
    c->prog[prog].alive = false;
    if (c->cur_prog == prog) c->cur_prog = -1;
    return PSGPU_OK;
}

int psgpu_program_use(psgpu_ctx* c, int prog)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (prog < 0 || prog >= (int)c->prog.size()) return fail(c, PSGPU_ERR_RANGE, "programme out of range");
    c->cur_prog = c->prog[prog].alive ? prog : -1;
    return PSGPU_OK;
}

// ---- state ------------------------------------------------------------------
int psgpu_uniform_set(psgpu_ctx* c, int slot, const void* data, size_t len)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (slot < 0 || slot >= kMaxUniforms) return fail(c, PSGPU_ERR_RANGE, "uniform slot out of range");
    auto& u = c->uniforms[slot];
    if (!data) { u.clear(); u.shrink_to_fit(); return PSGPU_OK; }
    u.assign((const uint8_t*)data, (const uint8_t*)data + len);
    return PSGPU_OK;
}

int psgpu_behavior_enable(psgpu_ctx* c, int bits) { if (need_ctx(c)) return PSGPU_ERR_INVALID; c->behavior |= bits; return PSGPU_OK; }
int psgpu_behavior_disable(psgpu_ctx* c, int bits) { if (need_ctx(c)) return PSGPU_ERR_INVALID; c->behavior &= ~bits; return PSGPU_OK; }
int psgpu_behavior_get(psgpu_ctx* c, int* bits)
{
    if (need_ctx(c) || !bits) return PSGPU_ERR_INVALID;
    *bits = c->behavior;
    return PSGPU_OK;
}

int psgpu_viewport(psgpu_ctx* c, int width, int height)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (width < 1 || height < 1) return fail(c, PSGPU_ERR_INVALID, "bad viewport");
    if (width > kMaxViewW || height > kMaxViewH)
        return fail(c, PSGPU_ERR_RANGE, "viewport %d x %d beyond %d x %d", width, height, kMaxViewW, kMaxViewH);
    c->vp_w = width;
    c->vp_h = height;
    return PSGPU_OK;
}

int psgpu_target_set(psgpu_ctx* c, int fbo_slot, int tex)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (fbo_slot < 0 || fbo_slot >= 16) return fail(c, PSGPU_ERR_RANGE, "FBO slot out of range");
    if (tex != PSGPU_DISPLAY && !valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
    // slots 1..15 are bound as the reference binds them (pipeline.cpp:260-275;
    // PSGPU_DISPLAY stands for NULL there); no processor twin writes them, as
    // no processor of the reference does
    if (fbo_slot != 0) c->fbo_more[fbo_slot] = tex;
    else c->fbo0 = tex;
    return PSGPU_OK;
}

int psgpu_target_gbuffer(psgpu_ctx* c, void* device_ptr, size_t bytes)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (device_ptr && ((uintptr_t)device_ptr % 16)) return fail(c, PSGPU_ERR_INVALID, "geometry buffer must be 16-byte aligned");
    c->gbuf = (float*)device_ptr;
    c->gbuf_bytes = device_ptr ? bytes : 0;
    return PSGPU_OK;
}

int psgpu_depth_set(psgpu_ctx* c, int tex)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (tex != PSGPU_DEFAULT_DEPTH) {
        if (!valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "PuresoftPipeline::setDepth");
        // pipeline.cpp:231-235
        if (c->tex[tex].elem != 4 || c->tex[tex].scan % 4) return fail(c, PSGPU_ERR_INVALID, "PuresoftPipeline::setDepth");
    }
    c->depth_sel = tex;
    return PSGPU_OK;
}

int psgpu_clear_depth(psgpu_ctx* c, float furthest)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    Targets tg;
    int rc = current_targets(c, tg);
    if (rc) return rc;
    uint32_t v;
    memcpy(&v, &furthest, 4);
    // clear16 fills m_bytes = scanline * height (fbo.cpp:348-371)
    return add_fill(c, tg.depth, 1, (int64_t)tg.depth_scan * tg.depth_h, 0, v);
}

int psgpu_clear_colour(psgpu_ctx* c, uint32_t bgra)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    // clearColour clears the display with clear4, which skips the last memory row
    // (pipeline.cpp:340-343, fbo.cpp:332-346)
    return add_fill(c, c->display, c->dev_h - 1, c->dev_w, c->dev_w, bgra);
}

int psgpu_clear_target(psgpu_ctx* c, uint32_t bgra)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    Targets tg;
    int rc = current_targets(c, tg);
    if (rc) return rc;
    return add_fill(c, tg.colour, tg.colour_h, tg.colour_w, tg.colour_scan, bgra);
}

// ---- drawing ------------------------------------------------------------------
static int uniform_bytes(psgpu_ctx* c, int slot, size_t need, const uint8_t** out)
{
    const auto& u = c->uniforms[slot];
    if (u.size() < need) return fail(c, PSGPU_ERR_STATE, "uniform slot %d not set (needs %zu bytes)", slot, need);
    *out = u.data();
    return PSGPU_OK;
}

static int uniform_tex(psgpu_ctx* c, int slot, PsTex* out)
{
    const uint8_t* p;
    int rc = uniform_bytes(c, slot, 4, &p);
    if (rc) return rc;
    int idx;
    memcpy(&idx, p, 4);
    if (!valid_tex(c, idx)) return fail(c, PSGPU_ERR_STATE, "uniform %d names texture %d, which does not exist", slot, idx);
    if (c->tex[idx].elem != 4) return fail(c, PSGPU_ERR_STATE, "sampled textures need 4-byte texels");
    *out = tex_desc(c->tex[idx]);
    return PSGPU_OK;
}

int psgpu_draw(psgpu_ctx* c, int vao)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    // drawvao.cpp:12-15: silently nothing without a programme or with a bad VAO
    if (c->cur_prog < 0 || vao < 0 || vao >= (int)c->vao.size() || !c->vao[vao].alive) return PSGPU_OK;
    const Prog& pg = c->prog[c->cur_prog];
    // blending changes nothing for the depth-only programmes (the demo leaves it
    // on for its cloud's shadow pass, testobjs.cpp:93-112)
    if ((c->behavior & PSGPU_ALPHABLEND) && pg.vp != PSGPU_PROC_CLOUD && pg.vp != PSGPU_PROC_CLOUD_SHADOW &&
        pg.vp != PSGPU_PROC_DEF05)
        return fail(c, PSGPU_ERR_UNSUPPORTED, "alpha blending is supported for PSGPU_PROC_CLOUD only");
    const Vao& a = c->vao[vao];
    PsDraw d;
    memset(&d, 0, sizeof(d));
    d.prog = pg.vp;
    d.behavior = c->behavior;
    d.nv = prog_nv(pg.vp);
    d.writes_colour = pg.vp != PSGPU_PROC_DEF05 && pg.vp != PSGPU_PROC_CLOUD_SHADOW;
    // processVertices stops at the first empty VBO (vertthrd.cpp:14-51)
    int64_t nverts = -1;
    for (int s = 0; s < kMaxSlots; s++)
        if (a.slot[s] >= 0) {
            int64_t n = (int64_t)c->vbo[a.slot[s]].count;
            if (nverts < 0 || n < nverts) nverts = n;
        }
    if (nverts < 0) return PSGPU_OK;
    d.tri_count = nverts / 3;
    // vertex slots read by the VP twins, with their element sizes
    int need[5] = {16, 0, 0, 0, 0};
    switch (d.prog) {
    case PSGPU_PROC_DEF01: need[3] = 16; need[4] = 8; break;
    case PSGPU_PROC_DEF02: need[1] = 16; need[2] = 16; break;
    case PSGPU_PROC_DEF03: need[1] = need[2] = need[3] = 16; need[4] = 8; break;
    case PSGPU_PROC_LAMBERT: need[3] = 16; need[4] = 8; break;
    case PSGPU_PROC_BLINN: case PSGPU_PROC_SHADOWED: need[1] = need[2] = need[3] = 16; need[4] = 8; break;
    case PSGPU_PROC_CLOUD: need[3] = 16; need[4] = 8; break;
    case PSGPU_PROC_CLOUD_SHADOW: need[4] = 8; break;
    default: break;
    }
    for (int s = 0; s < 5; s++) {
        if (!need[s]) continue;
        if (a.slot[s] < 0) return fail(c, PSGPU_ERR_STATE, "VBO slot %d is not attached", s);
        const Vbo& v = c->vbo[a.slot[s]];
        if ((int)v.unit < need[s]) return fail(c, PSGPU_ERR_STATE, "VBO slot %d elements must hold %d bytes", s, need[s]);
        d.slot[s] = v.d;
        d.unit[s] = (int)v.unit;
    }
    const uint8_t* p;
    int rc;
    switch (d.prog) {
    case PSGPU_PROC_DEF01: case PSGPU_PROC_DEF02: case PSGPU_PROC_DEF03: case PSGPU_PROC_BLINN: case PSGPU_PROC_SHADOWED:
        if ((rc = uniform_bytes(c, 3, 64, &p))) return rc; memcpy(d.PV, p, 64);
        if ((rc = uniform_bytes(c, 4, 64, &p))) return rc; memcpy(d.M, p, 64);
        if ((rc = uniform_bytes(c, 5, 64, &p))) return rc; memcpy(d.Mrot, p, 64);
        if ((rc = uniform_bytes(c, 7, 16, &p))) return rc; memcpy(d.lightPos, p, 16);
        if ((rc = uniform_bytes(c, 8, 16, &p))) return rc; memcpy(d.cameraPos, p, 16);
        if (d.prog != PSGPU_PROC_DEF02 && (rc = uniform_tex(c, 9, &d.tex0))) return rc;
        if ((d.prog == PSGPU_PROC_DEF03 || d.prog == PSGPU_PROC_BLINN || d.prog == PSGPU_PROC_SHADOWED) &&
            (rc = uniform_tex(c, 10, &d.tex1)))
            return rc;
        d.spec_exp = 32;
        if ((d.prog == PSGPU_PROC_BLINN || d.prog == PSGPU_PROC_SHADOWED) && c->uniforms[11].size() >= 4)
            memcpy(&d.spec_exp, c->uniforms[11].data(), 4);
        if (d.prog == PSGPU_PROC_SHADOWED) {
            if ((rc = uniform_bytes(c, 12, 64, &p))) return rc;
            memcpy(d.shadowPV, p, 64);
            if ((rc = uniform_tex(c, 13, &d.tex2))) return rc;
        }
        if (d.spec_exp < 0) return fail(c, PSGPU_ERR_INVALID, "specular exponent must be >= 0");
        break;
    case PSGPU_PROC_DEF04: {
        if ((rc = uniform_bytes(c, 1, 64, &p))) return rc;
        float V[16];
        memcpy(V, p, 64);
        // skybox.cpp:34-37: transpose, zero the last column
        for (int r = 0; r < 4; r++)
            for (int cc = 0; cc < 4; cc++) d.invView[cc * 4 + r] = V[r * 4 + cc];
        d.invView[12] = d.invView[13] = d.invView[14] = d.invView[15] = 0.0f;
        if ((rc = uniform_tex(c, 2, &d.tex0))) return rc;
        break;
    }
    case PSGPU_PROC_DEF05: case PSGPU_PROC_CLOUD_SHADOW: {
        float PV[16], M[16];
        if ((rc = uniform_bytes(c, 3, 64, &p))) return rc; memcpy(PV, p, 64);
        if ((rc = uniform_bytes(c, 4, 64, &p))) return rc; memcpy(M, p, 64);
        host_m4m4(d.PVM, PV, M);
        if (d.prog == PSGPU_PROC_CLOUD_SHADOW && (rc = uniform_tex(c, 9, &d.tex0))) return rc;
        break;
    }
    case PSGPU_PROC_CLOUD:   // VP_Cloud / FP_Cloud::preprocess, test/testproc.cpp:384-390, 526-532
        if ((rc = uniform_bytes(c, 3, 64, &p))) return rc; memcpy(d.PV, p, 64);
        if ((rc = uniform_bytes(c, 4, 64, &p))) return rc; memcpy(d.M, p, 64);
        if ((rc = uniform_bytes(c, 5, 64, &p))) return rc; memcpy(d.Mrot, p, 64);
        if ((rc = uniform_bytes(c, 7, 16, &p))) return rc; memcpy(d.lightPos, p, 16);
        if ((rc = uniform_bytes(c, 16, 64, &p))) return rc; memcpy(d.shadowPV, p, 64);
        if ((rc = uniform_tex(c, 9, &d.tex0))) return rc;
        if ((rc = uniform_tex(c, 15, &d.tex2))) return rc;
        break;
    case PSGPU_PROC_LAMBERT:
        if ((rc = uniform_bytes(c, 3, 64, &p))) return rc; memcpy(d.PV, p, 64);
        if ((rc = uniform_bytes(c, 4, 64, &p))) return rc; memcpy(d.M, p, 64);
        if ((rc = uniform_bytes(c, 5, 64, &p))) return rc; memcpy(d.Mrot, p, 64);
        if ((rc = uniform_bytes(c, 7, 12, &p))) return rc; memcpy(d.lightDir, p, 12);
        if ((rc = uniform_tex(c, 9, &d.tex0))) return rc;
        break;
    }
    // a texture that a pending view renders into (a shadow map, say) is
    // sampled only after that view's flush: the shading of one flush runs
    // after all of its rasterisation, but colour is written while shading
    for (const PsTex* t : {&d.tex0, &d.tex1, &d.tex2})
        if (t->data && queue_uses(c, t->data) && (rc = queue_to_log(c))) return rc;
    // targets must cover the viewport
    Targets tg;
    if ((rc = current_targets(c, tg))) return rc;
    if (tg.colour_w < c->vp_w || tg.colour_h < c->vp_h || tg.depth_w < c->vp_w || tg.depth_h < c->vp_h)
        return fail(c, PSGPU_ERR_STATE, "render targets are smaller than the viewport");
    if (tg.gbuf) {
        if (d.prog != PSGPU_PROC_BLINN)
            return fail(c, PSGPU_ERR_UNSUPPORTED, "geometry buffers are written by the Blinn-Phong programme only");
        if (c->gbuf_bytes < (size_t)c->vp_w * c->vp_h * 48)
            return fail(c, PSGPU_ERR_STATE, "geometry buffer smaller than the viewport (48 bytes per pixel)");
    }
    if (d.tri_count == 0) return PSGPU_OK;
    // draws with the targets and viewport of the last pending view join it;
    // otherwise they open a new view of the same flush, unless a pending view
    // already renders into one of these buffers (then that work goes first)
    if (c->queue_tris + d.tri_count > kMaxFlushTris && !c->queue.empty()) {
        if ((rc = queue_to_log(c))) return rc;     // bound the scratch of one flush
    }
    // programmes that shade in the raster walk (blending, discard) get flushes
    // of their own, one programme each
    const int imm = (d.prog == PSGPU_PROC_CLOUD || d.prog == PSGPU_PROC_CLOUD_SHADOW) ? d.prog : 0;
    if (!c->queue.empty() && c->queue_imm != imm && (rc = queue_to_log(c))) return rc;
    c->queue_imm = imm;
    if (c->queue.empty() || !c->queue.back().tg.same(tg)) {
        if (queue_uses(c, tg.colour) || queue_uses(c, tg.depth)) {
            if ((rc = queue_to_log(c))) return rc;
        }
        Segment sg;
        sg.tg = tg;
        c->queue.push_back(std::move(sg));
    }
    d.tri_start = c->queue_tris;
    d.id_base = c->dbg_tri_counter;
    c->dbg_tri_counter += d.tri_count;
    c->queue_tris += d.tri_count;
    c->stats.tris_in += d.tri_count;
    c->queue.back().draws.push_back(d);
    return PSGPU_OK;
}

int psgpu_flush(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    return queue_to_log(c);
}

int psgpu_sync(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    return do_sync(c);
}

int psgpu_swap_buffers(psgpu_ctx* c) { return psgpu_flush(c); }

int psgpu_post_process(psgpu_ctx* c, psgpu_post_fn fn, void* user)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (!fn) return fail(c, PSGPU_ERR_INVALID, "no post-processor");
    Targets tg;
    int rc = current_targets(c, tg);
    if (rc) return rc;
    if ((rc = queue_to_log(c))) return rc;   // the draws before it (post.cpp: after the queued rows)
    Cmd cmd;
    cmd.type = CMD_POST;
    cmd.post = fn;
    cmd.post_user = user;
    cmd.post_frame.colour = tg.colour;
    cmd.post_frame.colour_pitch = tg.colour_scan * 4;
    cmd.post_frame.colour_top_down = tg.colour_top_down;
    cmd.post_frame.depth = tg.depth;
    cmd.post_frame.depth_pitch = tg.depth_scan * 4;
    cmd.post_frame.width = tg.vpw;
    cmd.post_frame.height = tg.vph;
    return submit(c, std::move(cmd));
}

int psgpu_read_colour(psgpu_ctx* c, void* host, size_t bytes)
{
    if (need_ctx(c) || !host) return PSGPU_ERR_INVALID;
    if (bytes != (size_t)c->dev_w * c->dev_h * 4) return fail(c, PSGPU_ERR_INVALID, "display is %d x %d BGRA8", c->dev_w, c->dev_h);
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(host, c->display, bytes, hipMemcpyDeviceToHost));
    return PSGPU_OK;
}

int psgpu_read_depth(psgpu_ctx* c, void* host, size_t bytes)
{
    if (need_ctx(c) || !host) return PSGPU_ERR_INVALID;
    if (bytes != (size_t)c->depth_scan * c->dev_h * 4) return fail(c, PSGPU_ERR_INVALID, "depth is %d x %d floats", c->depth_scan, c->dev_h);
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(host, c->depth, bytes, hipMemcpyDeviceToHost));
    return PSGPU_OK;
}

// the reference's display and default depth buffer are host memory that a
// post-processor may write (post.cpp:3-18, test2/testpost.cpp:9-44): the
// host copies are written back through these
int psgpu_write_colour(psgpu_ctx* c, const void* host, size_t bytes)
{
    if (need_ctx(c) || !host) return PSGPU_ERR_INVALID;
    if (bytes != (size_t)c->dev_w * c->dev_h * 4) return fail(c, PSGPU_ERR_INVALID, "display is %d x %d BGRA8", c->dev_w, c->dev_h);
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(c->display, host, bytes, hipMemcpyHostToDevice));
    return PSGPU_OK;
}

int psgpu_write_depth(psgpu_ctx* c, const void* host, size_t bytes)
{
    if (need_ctx(c) || !host) return PSGPU_ERR_INVALID;
    if (bytes != (size_t)c->depth_scan * c->dev_h * 4) return fail(c, PSGPU_ERR_INVALID, "depth is %d x %d floats", c->depth_scan, c->dev_h);
    int rc = do_sync(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(c->depth, host, bytes, hipMemcpyHostToDevice));
    return PSGPU_OK;
}

// ---- asynchronous readback and fences --------------------------------------
int psgpu_host_alloc(psgpu_ctx* c, size_t bytes, void** ptr)
{
    if (need_ctx(c) || !ptr || !bytes) return PSGPU_ERR_INVALID;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, bytes, 0) != hipSuccess) return fail(c, PSGPU_ERR_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
    return PSGPU_OK;
}

int psgpu_host_free(psgpu_ctx* c, void* ptr)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (!ptr) return PSGPU_OK;
    int rc = do_sync(c);   // no queued copy may still target it
    if (rc) return rc;
    HIPCHK(c, hipHostFree(ptr));
    return PSGPU_OK;
}

int psgpu_texture_read_async(psgpu_ctx* c, int tex, int layer, int x, int y, int w, int h, void* host, size_t host_pitch)
{
    if (need_ctx(c) || !host) return PSGPU_ERR_INVALID;
    if (c->capturing) return fail(c, PSGPU_ERR_STATE, "asynchronous reads are not recorded into command lists");
    const uint8_t* base;
    int tw, th, elem;
    size_t pitch;
    if (tex == PSGPU_DISPLAY) {
        if (layer != 0) return fail(c, PSGPU_ERR_RANGE, "the display has one layer");
        base = (const uint8_t*)c->display;
        tw = c->dev_w; th = c->dev_h; elem = 4; pitch = (size_t)c->dev_w * 4;
    } else {
        if (!valid_tex(c, tex)) return fail(c, PSGPU_ERR_RANGE, "texture index out of range");
        const Tex& t = c->tex[tex];
        if (layer < 0 || layer >= t.layers) return fail(c, PSGPU_ERR_RANGE, "layer out of range");
        base = t.d + t.layer_bytes * layer;
        tw = t.w; th = t.h; elem = t.elem; pitch = (size_t)t.scan;
    }
    if (x < 0 || y < 0 || w < 1 || h < 1 || x + w > tw || y + h > th) return fail(c, PSGPU_ERR_RANGE, "region outside the buffer");
    if (host_pitch < (size_t)w * elem) return fail(c, PSGPU_ERR_INVALID, "host pitch shorter than a row");
    int rc = queue_to_log(c);   // the draws and clears queued so far come first
    if (rc) return rc;
    Cmd cmd;
    cmd.type = CMD_COPY;
    cmd.copy.src = base + (size_t)y * pitch + (size_t)x * elem;
    cmd.copy.src_pitch = pitch;
    cmd.copy.dst = host;
    cmd.copy.dst_pitch = host_pitch;
    cmd.copy.width_bytes = (size_t)w * elem;
    cmd.copy.rows = (size_t)h;
    return submit(c, std::move(cmd));
}

int psgpu_fence(psgpu_ctx* c, uint64_t* out)
{
    if (need_ctx(c) || !out) return PSGPU_ERR_INVALID;
    if (c->capturing) return fail(c, PSGPU_ERR_STATE, "not allowed while recording a command list");
    int rc = queue_to_log(c);
    if (rc) return rc;
    Fence fe;
    for (hipEvent_t* e : {&fe.main_ev, &fe.copy_ev}) {
        if (!c->fence_ev_pool.empty()) {
            *e = c->fence_ev_pool.back();
            c->fence_ev_pool.pop_back();
        } else {
            HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
    }
    HIPCHK(c, hipEventRecord(fe.main_ev, c->stream));
    HIPCHK(c, hipEventRecord(fe.copy_ev, c->copy_stream));
    fe.log_end = c->log.size();
    fe.id = c->next_fence++;
    c->fences.push_back(fe);
    *out = fe.id;
    return PSGPU_OK;
}

static int fence_finish(psgpu_ctx* c, size_t k, int* done)
{
    Fence fe = c->fences[k];
    // a command before the fence that overflowed is replayed (with the
    // commands after it) by a full sync, which retires every fence
    for (size_t i = 0; i < fe.log_end && i < c->log.size(); i++)
        if (c->h_flags[i]) {
            int rc = do_sync(c);
            if (rc) return rc;
            *done = 1;
            return PSGPU_OK;
        }
    c->fence_ev_pool.push_back(fe.main_ev);
    c->fence_ev_pool.push_back(fe.copy_ev);
    c->fences.erase(c->fences.begin() + k);
    *done = 1;
    return PSGPU_OK;
}

static int fence_find(psgpu_ctx* c, uint64_t id)
{
    for (size_t k = 0; k < c->fences.size(); k++)
        if (c->fences[k].id == id) return (int)k;
    return -1;
}

int psgpu_fence_query(psgpu_ctx* c, uint64_t id, int* done)
{
    if (need_ctx(c) || !done) return PSGPU_ERR_INVALID;
    if (id == 0 || id >= c->next_fence) return fail(c, PSGPU_ERR_RANGE, "unknown fence");
    *done = 0;
    int k = fence_find(c, id);
    if (k < 0) { *done = 1; return PSGPU_OK; }   // retired by a sync or an earlier query
    for (hipEvent_t e : {c->fences[k].main_ev, c->fences[k].copy_ev}) {
        hipError_t q = hipEventQuery(e);
        if (q == hipErrorNotReady) return PSGPU_OK;
        if (q != hipSuccess) return fail(c, PSGPU_ERR_DEVICE, "hipEventQuery: %s", hipGetErrorString(q));
    }
    return fence_finish(c, (size_t)k, done);
}

int psgpu_fence_wait(psgpu_ctx* c, uint64_t id)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (id == 0 || id >= c->next_fence) return fail(c, PSGPU_ERR_RANGE, "unknown fence");
    int k = fence_find(c, id);
    if (k < 0) return PSGPU_OK;
    HIPCHK(c, hipEventSynchronize(c->fences[k].main_ev));
    HIPCHK(c, hipEventSynchronize(c->fences[k].copy_ev));
    int done = 0;
    return fence_finish(c, (size_t)k, &done);
}

int psgpu_display_device_ptr(psgpu_ctx* c, void** ptr)
{
    if (need_ctx(c) || !ptr) return PSGPU_ERR_INVALID;
    *ptr = c->display;
    return PSGPU_OK;
}

int psgpu_depth_device_ptr(psgpu_ctx* c, void** ptr, int* scan)
{
    if (need_ctx(c) || !ptr) return PSGPU_ERR_INVALID;
    *ptr = c->depth;
    if (scan) *scan = c->depth_scan;
    return PSGPU_OK;
}

int psgpu_get_stats(psgpu_ctx* c, psgpu_stats* out)
{
    if (need_ctx(c) || !out) return PSGPU_ERR_INVALID;
    *out = c->stats;
    return PSGPU_OK;
}

int psgpu_reset_stats(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    int rc = do_sync(c);
    memset(&c->stats, 0, sizeof(c->stats));
    return rc;
}

// ---- debug planes -----------------------------------------------------------
int psgpu_debug_enable(psgpu_ctx* c, int enable)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    int rc = queue_to_log(c);
    if (rc) return rc;
    c->debug = enable != 0;
    if (c->debug) return psgpu_debug_reset(c);
    return PSGPU_OK;
}

int psgpu_debug_reset(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    int rc = do_sync(c);
    if (rc) return rc;
    size_t n = (size_t)c->vp_w * c->vp_h;
    if (c->dbg_w != c->vp_w || c->dbg_h != c->vp_h || !c->dbg_last) {
        for (void* p : {(void*)c->dbg_last, (void*)c->dbg_rcount, (void*)c->dbg_rhash, (void*)c->dbg_pcount, (void*)c->dbg_phash})
            if (p) hipFree(p);
        HIPCHK(c, hipMalloc((void**)&c->dbg_last, n * 4));
        HIPCHK(c, hipMalloc((void**)&c->dbg_rcount, n * 4));
        HIPCHK(c, hipMalloc((void**)&c->dbg_rhash, n * 4));
        HIPCHK(c, hipMalloc((void**)&c->dbg_pcount, n * 4));
        HIPCHK(c, hipMalloc((void**)&c->dbg_phash, n * 4));
        c->dbg_w = c->vp_w;
        c->dbg_h = c->vp_h;
    }
    psk_fill_u32(c->stream, (uint32_t*)c->dbg_last, (int64_t)n, 0xffffffffu, nullptr);
    psk_fill_u32(c->stream, (uint32_t*)c->dbg_rcount, (int64_t)n, 0, nullptr);
    psk_fill_u32(c->stream, c->dbg_rhash, (int64_t)n, 0, nullptr);
    psk_fill_u32(c->stream, (uint32_t*)c->dbg_pcount, (int64_t)n, 0, nullptr);
    psk_fill_u32(c->stream, c->dbg_phash, (int64_t)n, 0, nullptr);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dbg_tri_counter = 0;
    return PSGPU_OK;
}

int psgpu_debug_read(psgpu_ctx* c, int32_t* last_pass, int32_t* raster_count, uint32_t* raster_hash, int32_t* pass_count,
                     uint32_t* pass_hash)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (!c->dbg_last) return fail(c, PSGPU_ERR_STATE, "debug planes are not enabled");
    int rc = do_sync(c);
    if (rc) return rc;
    size_t n = (size_t)c->dbg_w * c->dbg_h * 4;
    if (last_pass) HIPCHK(c, hipMemcpy(last_pass, c->dbg_last, n, hipMemcpyDeviceToHost));
    if (raster_count) HIPCHK(c, hipMemcpy(raster_count, c->dbg_rcount, n, hipMemcpyDeviceToHost));
    if (raster_hash) HIPCHK(c, hipMemcpy(raster_hash, c->dbg_rhash, n, hipMemcpyDeviceToHost));
    if (pass_count) HIPCHK(c, hipMemcpy(pass_count, c->dbg_pcount, n, hipMemcpyDeviceToHost));
    if (pass_hash) HIPCHK(c, hipMemcpy(pass_hash, c->dbg_phash, n, hipMemcpyDeviceToHost));
    return PSGPU_OK;
}

// ---- tuning and timing ------------------------------------------------------
int psgpu_tiling_set(psgpu_ctx* c, const psgpu_tiling* t)
{
    if (need_ctx(c) || !t) return PSGPU_ERR_INVALID;
    int cw = t->chunk_width, tr = t->tile_rows, w = t->waves_per_group;
    bool okc = cw >= 1 && cw <= 64 && (cw & (cw - 1)) == 0;
    bool okr = tr >= 1 && tr <= 64 && (tr & (tr - 1)) == 0;
    if (!okc || !okr || w < 1 || w > 8) return fail(c, PSGPU_ERR_INVALID, "unsupported tiling");
    if (psk_raster_lds(cw, w) > 65536) return fail(c, PSGPU_ERR_INVALID, "tiling needs more than 64 KiB of LDS");
    int rc = do_sync(c);    // logged commands are replayed with the tiling they ran with
    if (rc) return rc;
    c->tiling = *t;
    c->buf_gen++;     // graphs bake the tiling in: capture them again
    return PSGPU_OK;
}

int psgpu_tiling_get(psgpu_ctx* c, psgpu_tiling* t)
{
    if (need_ctx(c) || !t) return PSGPU_ERR_INVALID;
    *t = c->tiling;
    return PSGPU_OK;
}

int psgpu_profile_enable(psgpu_ctx* c, int enable)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    int rc = do_sync(c);
    c->profiling = enable != 0;
    return rc;
}

int psgpu_profile_read(psgpu_ctx* c, int max, const char** names, double* total_ms, int64_t* launches, int* count)
{
    if (need_ctx(c) || !count) return PSGPU_ERR_INVALID;
    int rc = do_sync(c);
    if (rc) return rc;
    c->prof_names.clear();
    for (auto& kv : c->prof) c->prof_names.push_back(kv.first);
    int i = 0;
    for (auto& kv : c->prof) {
        if (i >= max) break;
        if (names) names[i] = c->prof_names[i].c_str();
        if (total_ms) total_ms[i] = kv.second.first;
        if (launches) launches[i] = kv.second.second;
        i++;
    }
    *count = i;
    return PSGPU_OK;
}

int psgpu_profile_reset(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    int rc = do_sync(c);
    c->prof.clear();
    return rc;
}

// ---- recorded command lists replayed as hipGraphs ---------------------------
int psgpu_capture_begin(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (c->capturing) return fail(c, PSGPU_ERR_STATE, "already recording");
    int rc = queue_to_log(c);
    if (rc) return rc;
    c->capturing = true;
    c->capture_list.clear();
    return PSGPU_OK;
}

int psgpu_capture_end(psgpu_ctx* c, int* out_graph)
{
    if (need_ctx(c) || !out_graph) return PSGPU_ERR_INVALID;
    if (!c->capturing) return fail(c, PSGPU_ERR_STATE, "not recording");
    int rc = queue_to_log(c);      // pending draws join the recording
    c->capturing = false;
    if (rc) return rc;
    Graph g;
    g.alive = true;
    g.cmds.swap(c->capture_list);
    // first execution runs eagerly: it sizes every scratch buffer
    const int64_t splits = c->stats.splits;
    for (const Cmd& cmd : g.cmds) {
        Cmd copy = cmd;
        c->log.push_back(std::move(copy));
        if ((rc = run_cmd(c, c->log.back()))) return rc;
        if ((rc = mark_done(c, (int)c->log.size() - 1))) return rc;
    }
    if ((rc = do_sync(c))) return rc;
    if (c->stats.splits != splits) return graph_too_large(c);
    for (const Cmd& cmd : g.cmds) {
        uint8_t* p = nullptr;
        Blob b;
        if (cmd.type == CMD_FLUSH) {
            PsFrame f;
            int64_t T, npx;
            int nvs;
            bool p1[PS_N_PROGS] = {false}, p2[PS_N_PROGS] = {false};
            if ((rc = build_blob(c, cmd, b, f, T, nvs, p1, p2, npx))) return rc;
            g.flushes++;
        } else if (cmd.type == CMD_CLEARS) {
            b.bytes.assign(cmd.fills.size() * sizeof(PsFill), 0);
            memcpy(b.bytes.data(), cmd.fills.data(), b.bytes.size());
        }
        if (!b.bytes.empty()) {
            HIPCHK(c, hipMalloc((void**)&p, b.bytes.size()));
            HIPCHK(c, hipMemcpy(p, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice));
        }
        g.blobs.push_back(p);
    }
    if ((rc = graph_capture(c, g))) {
        for (auto p : g.blobs) if (p) hipFree(p);
        return rc;
    }
    size_t i = 0;
    for (; i < c->graphs.size(); i++) if (!c->graphs[i].alive) break;
    if (i == c->graphs.size()) c->graphs.push_back(std::move(g)); else c->graphs[i] = std::move(g);
    *out_graph = (int)i;
    return PSGPU_OK;
}

int psgpu_graph_launch(psgpu_ctx* c, int graph, int count)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (graph < 0 || graph >= (int)c->graphs.size() || !c->graphs[graph].alive) return fail(c, PSGPU_ERR_RANGE, "graph out of range");
    if (c->capturing) return fail(c, PSGPU_ERR_STATE, "not allowed while recording a command list");
    if (count <= 0) return PSGPU_OK;
    int rc = queue_to_log(c);
    if (rc) return rc;
    Graph& g = c->graphs[graph];
    if (!graph_current(c, g)) {
        if ((rc = graph_prepare(c, g))) return rc;
        if (--count == 0) return PSGPU_OK;   // the eager run was the first launch
    }
    // a prepared graph fits the scratch buffers, which only grow, and its
    // vertex streams are those it was sized for (psgpu_vbo_update marks the
    // graphs that read a stream stale): a replay cannot overflow, so a launch
    // of `count` replays is one log entry
    Cmd cmd;
    cmd.type = CMD_GRAPH;
    cmd.graph = graph;
    cmd.count = count;
    return submit(c, std::move(cmd));
}

int psgpu_graph_destroy(psgpu_ctx* c, int graph)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (graph < 0 || graph >= (int)c->graphs.size()) return fail(c, PSGPU_ERR_RANGE, "graph out of range");
    if (!c->graphs[graph].alive) return PSGPU_OK;
    int rc = do_sync(c);
    if (rc) return rc;
    Graph& g = c->graphs[graph];
    if (g.exec) hipGraphExecDestroy(g.exec);
    for (auto p : g.blobs) if (p) hipFree(p);
    c->graphs[graph] = Graph();
    return PSGPU_OK;
}

int psgpu_sync_noalloc(psgpu_ctx* c)
{
    if (need_ctx(c)) return PSGPU_ERR_INVALID;
    if (c->capturing) return fail(c, PSGPU_ERR_STATE, "not allowed while recording a command list");
    if (!c->log.empty())
        return fail(c, PSGPU_ERR_STATE, "work submitted since the last sync (psgpu_flush, a flush forced by psgpu_draw, a "
                                        "post-process or a graph launch) must be synchronised with psgpu_sync first");
    if (c->pending_fills.empty() && c->queue.empty()) return PSGPU_OK;
    uint8_t* base = (c->device >= 0 && c->device < kMaxStaticDevices) ? g_static_ptr[c->device] : nullptr;
    if (!base)
        return fail(c, PSGPU_ERR_DEVICE, "static scratch array of device %d unavailable: %s", c->device,
                    hipGetErrorString(c->device < kMaxStaticDevices ? g_static_err[c->device] : hipErrorInvalidDevice));
    Cmd clears, flush;
    clears.type = CMD_CLEARS;
    clears.fills = c->pending_fills;
    flush.type = CMD_FLUSH;
    flush.views = c->queue;
    if (!c->queue_imm) fuse_clears(clears.fills, flush.views, flush_shades(flush.views, 0));
    std::lock_guard<std::mutex> guard(g_static_lock);
    Arena arena{base, PSGPU_STATIC_SCRATCH_BYTES};
    uint8_t* fill_dev = nullptr;
    if (!clears.fills.empty() && !(fill_dev = arena.take_bytes(clears.fills.size() * sizeof(PsFill))))
        return fail(c, PSGPU_ERR_NOMEM, "%zu clears exceed the static scratch array", clears.fills.size());
    FlushShape sh = flush_shape(c, flush);
    FlushBufs sb;
    Blob blob;
    uint8_t* blob_dev = nullptr;
    // the flush's records go first, then its buffers; the blob's size does
    // not depend on where the buffers land
    {
        size_t draws = 0;
        for (const Segment& sg : flush.views) draws += sg.draws.size();
        size_t bytes = align256(draws * sizeof(PsDraw)) + align256(flush.views.size() * sizeof(PsView)) + 16;
        if (sh.T > 0 && !(blob_dev = arena.take_bytes(bytes)))
            return fail(c, PSGPU_ERR_NOMEM, "the draw records exceed the static scratch array");
    }
    if (sh.T > 0 && !carve_flush(arena, sh, sb))
        return fail(c, PSGPU_ERR_NOMEM,
                    "a flush of %lld triangles over %lld pixels needs more than the %zu-byte static scratch array",
                    (long long)sh.T, (long long)sh.n_pix, (size_t)PSGPU_STATIC_SCRATCH_BYTES);
    // from here on the queued work is consumed, as psgpu_flush consumes it
    c->pending_fills.clear();
    c->queue.clear();
    c->queue_tris = 0;
    c->queue_imm = 0;
    std::swap(c->fb, sb);
    c->fixed_scratch = true;
    int rc = PSGPU_OK;
    if (fill_dev) {
        hipError_t e = hipMemcpyAsync(fill_dev, clears.fills.data(), clears.fills.size() * sizeof(PsFill),
                                      hipMemcpyHostToDevice, c->stream);
        rc = e != hipSuccess ? fail(c, PSGPU_ERR_DEVICE, "hipMemcpyAsync: %s", hipGetErrorString(e))
                             : run_clears(c, clears, fill_dev);
    }
    // the tile counts start at zero (the array may have held anything)
    if (!rc && blob_dev) {
        hipError_t e = hipMemsetAsync(c->fb.tcnt.p, 0, c->fb.tcnt.cap * sizeof(unsigned long long), c->stream);
        if (e != hipSuccess) rc = fail(c, PSGPU_ERR_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
    }
    if (!rc && blob_dev) {
        PsFrame f;
        int64_t T, n_pix;
        int nvs;
        bool p1[PS_N_PROGS] = {false}, p2[PS_N_PROGS] = {false};
        rc = build_blob(c, flush, blob, f, T, nvs, p1, p2, n_pix);
        if (!rc && (T != sh.T || n_pix != sh.n_pix || f.n_tiles != sh.n_tiles))
            rc = fail(c, PSGPU_ERR_STATE, "internal: flush shape mismatch");
        if (!rc) {
            hipError_t e = hipMemcpyAsync(blob_dev, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice, c->stream);
            rc = e != hipSuccess ? fail(c, PSGPU_ERR_DEVICE, "hipMemcpyAsync: %s", hipGetErrorString(e))
                                 : run_flush(c, flush, blob_dev);
        }
    }
    // wait, then read and reset the counters as psgpu_sync does
    hipError_t e = hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(PsCounters), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    std::swap(c->fb, sb);
    c->fixed_scratch = false;
    if (rc) return rc;
    if (e != hipSuccess) return fail(c, PSGPU_ERR_DEVICE, "%s", hipGetErrorString(e));
    bool overflow = c->h_cnt->overflow != 0;
    if (!overflow) {
        int64_t sum[PS_ST_N] = {0};
        for (int k = 0; k < PS_STRIPES; k++)
            for (int f = 0; f < PS_ST_N; f++) sum[f] += (int64_t)c->h_cnt->stat[k][f];
        c->stats.tris_culled += sum[PS_ST_CULLED];
        c->stats.tris_zrejected += sum[PS_ST_ZREJ];
        c->stats.tris_dropped += sum[PS_ST_DROPPED];
        c->stats.spans += sum[PS_ST_SPANS];
        c->stats.frags_rasterised += sum[PS_ST_FRAGS];
        c->stats.frags_shaded += sum[PS_ST_SHADED];
        c->stats.segments += sum[PS_ST_SEGS];
    }
    HIPCHK(c, hipMemsetAsync(c->d_cnt, 0, sizeof(PsCounters), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_sizes, 0, 4 * sizeof(uint32_t), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (overflow)
        return fail(c, PSGPU_ERR_NOMEM,
                    "the flush's spans or pieces exceed what the %zu-byte static scratch array holds (%lld spans, "
                    "%lld pieces); its draws were not rendered (the clears were)",
                    (size_t)PSGPU_STATIC_SCRATCH_BYTES, (long long)sb.span_cap, (long long)sb.seg_cap);
    return PSGPU_OK;
}

int psgpu_selftest_math(psgpu_ctx* c, int op, const float* a, const float* b, float* out, int n)
{
    if (need_ctx(c) || !a || !out || n <= 0) return PSGPU_ERR_INVALID;
    float *da = nullptr, *db = nullptr, *dout = nullptr;
    HIPCHK(c, hipMalloc((void**)&da, (size_t)n * 4));
    HIPCHK(c, hipMalloc((void**)&db, (size_t)n * 4));
    HIPCHK(c, hipMalloc((void**)&dout, (size_t)n * 4));
    HIPCHK(c, hipMemcpy(da, a, (size_t)n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(db, b ? b : a, (size_t)n * 4, hipMemcpyHostToDevice));
    psk_selftest(c->stream, op, da, db, dout, n);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost));
    hipFree(da); hipFree(db); hipFree(dout);
    return PSGPU_OK;
}

}  // extern "C"
