// ps_kernels.hip -- gfx950 kernels of the sort-middle raster path.
//
//  setup   : one thread per triangle: vertex twin, 1/w divide, back-face cull,
//            z reject, pushVertex/pushTriangle row-range and edge choice
//            (drawvao.cpp:36-63, vertthrd.cpp, rasterizer.cpp:122-283)
//  rows    : exclusive scan of the row counts, and the triangle of every
//            64th row item
//  span    : one thread per (triangle,row), its triangle searched from that
//            table: span ends, clamping and the
//            depth / correction-factor start and step (drawvao.cpp:66-110,
//            interp.cpp:26-80); counts, per tile, the pieces (bin chunks) of
//            the spans and the waves that have pieces there
//  tiles   : exclusive scan of the per-tile counts: each tile's piece and
//            run ranges
//  segment : one thread per span walks it and writes its pieces; a wave's
//            pieces of one tile form one run of consecutive pieces, placed by
//            one atomic on the tile's counter
//  raster  : one wave per tile of 64 bins, lane = bin (a row piece of
//            chunk_w pixels); the wave puts its tile's runs in wave order,
//            which is submission order, and each lane walks its bin's spans
//            in triangle order, keeping depth and the winning span in LDS
//            (fragthrd.cpp:208-252); depth test -1 < z && z - d < -1e-4
//  shade   : one thread per pixel whose last colour-writing depth-test pass
//            is known: re-derives that span's varyings, walks them serially
//            to the pixel, runs the FP twin and writes BGRA8.
//  The depth test is order dependent, so per pixel the fragments are applied
//  strictly in submission order; no atomics touch depth or colour.
#include "ps_kernels.h"
#include "ps_procs.h"

#define PS_WAVE 64

__device__ __forceinline__ bool ps_poisoned(const PsCounters* c)
{
    return *((volatile const int*)&c->overflow) != 0;
}
// Kernels read the flag together with their first loads that stay in bounds
// whatever an overflow left behind, and test it before the first store or
// the first load whose address depends on the flush's device-side sizes:
// the flag's round trip then overlaps those loads instead of preceding them.

__device__ __forceinline__ uint64_t ps_lt_mask()
{
    return __lanemask_lt();
}

// the flush got past every overflow check: its pending statistics become
// committed (blocks 0..63 of the raster launch take the stripes between them)
__device__ __forceinline__ void ps_commit_pending(PsCounters* c)
{
    if (blockIdx.x >= PS_STRIPES || threadIdx.x >= PS_ST_N) return;
    for (int s = blockIdx.x; s < PS_STRIPES; s += gridDim.x) {
        unsigned long long v = c->pend[s][threadIdx.x];
        if (v) {
            atomicAdd(&c->stat[s][threadIdx.x], v);
            c->pend[s][threadIdx.x] = 0;
        }
    }
}

__device__ __forceinline__ int ps_view_of_tile(const PsFrame& f, int64_t t)
{
    int lo = 0, hi = f.n_views - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (f.views[mid].tile_base <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int ps_view_of_stile(const PsFrame& f, int64_t t)
{
    int lo = 0, hi = f.n_views - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (f.views[mid].stile_base <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  This maps workgroups to logical
// indices so that each XCD gets runs of G consecutive indices: neighbouring
// screen tiles then share texels and triangle records in one L2, while the
// runs still alternate over the XCDs (contiguous eighths leave XCDs idle
// where the screen is empty).  A bijection on [0, n) whatever the dispatch:
// whole groups of 8G are permuted within, the tail keeps its order.
template <int G>
__device__ __forceinline__ int64_t ps_xcd_remap(int64_t b, int64_t n)
{
    const int64_t s = b / (8 * G), j = b - s * (8 * G);
    if ((s + 1) * (8 * G) > n) return b;
    return s * (8 * G) + (j & 7) * G + (j >> 3);
}

// where triangle tri of draw d finds its three vertices' varyings
__device__ __forceinline__ int ps_vary_index(const PsDraw& d, int tri) { return tri + d.vary_delta; }

// the first of the two texcoord varyings the mip-level gradients use (-1: none)
template <class P> __device__ constexpr int ps_grad_iu()
{
    if constexpr (P::GRAD) return P::IU;
    else return -1;
}

// f(draw record): when every active lane of the wave uses the same draw, the
// record is addressed through a wave-uniform index, so its fields come in
// with scalar loads instead of one vector load per lane
template <typename F>
__device__ __forceinline__ void ps_with_draw(const PsDraw* __restrict__ draws, int di, F&& f)
{
    int d0 = __builtin_amdgcn_readfirstlane(di);
    if (__all(di == d0)) f(draws[d0]);
    else f(draws[di]);
}

// Wave-wide min / max / inclusive sum over 64 lanes with DPP row operations
// (no LDS crossbar round trips).  Control codes of GFX9 DPP: quad_perm
// 0x00-0xff, row_shr:n 0x110+n, row_mirror 0x140, row_half_mirror 0x141,
// row_bcast:15 0x142, row_bcast:31 0x143.
#define PS_DPP(v, ctrl, rmask, bound) \
    __builtin_amdgcn_update_dpp(0u, (v), (ctrl), (rmask), 0xf, (bound))

// the minimum, the same in every lane
// a tile start without the heavy-tile mark (bit 63)
#define PS_TS(x) ((x) & 0x7fffffffffffffffull)

__device__ __forceinline__ uint32_t ps_wave_min(uint32_t v)
{
    // within rows of 16: quad swaps, then the half-row and row mirrors
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false));   // quad_perm [1,0,3,2]
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false));   // quad_perm [2,3,0,1]
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xf, 0xf, false));  // row_half_mirror
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xf, 0xf, false));  // row_mirror
    // the four row minima
    const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return min(min(r0, r1), min(r2, r3));
}

__device__ __forceinline__ uint32_t ps_wave_max(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xf, 0xf, false));
    const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return max(max(r0, r1), max(r2, r3));
}

// inclusive prefix sum over the lanes (every lane of the wave active)
__device__ __forceinline__ uint32_t ps_wave_incl_scan(uint32_t v)
{
    v += PS_DPP(v, 0x111, 0xf, true);   // row_shr:1, zero from outside the row
    v += PS_DPP(v, 0x112, 0xf, true);   // row_shr:2
    v += PS_DPP(v, 0x114, 0xf, true);   // row_shr:4
    v += PS_DPP(v, 0x118, 0xf, true);   // row_shr:8
    v += PS_DPP(v, 0x142, 0xa, false);  // row_bcast:15 into rows 1 and 3
    v += PS_DPP(v, 0x143, 0xc, false);  // row_bcast:31 into rows 2 and 3
    return v;
}

// a wave's statistic (lane values whose wave total fits 32 bits, every lane
// active): into the committed counters (pending = false) or the flush's
// pending ones; one atomic per wave, on a per-block stripe
__device__ __forceinline__ void ps_wave_add(PsCounters* c, int field, uint32_t v, bool pending = false)
{
    const uint32_t tot = __builtin_amdgcn_readlane(ps_wave_incl_scan(v), 63);
    if ((threadIdx.x & (PS_WAVE - 1)) == 0 && tot)
        atomicAdd(pending ? &c->pend[blockIdx.x & (PS_STRIPES - 1)][field] : &c->stat[blockIdx.x & (PS_STRIPES - 1)][field],
                  (unsigned long long)tot);
}

// the same for 64-bit lane values, into the committed counters
__device__ __forceinline__ void ps_wave_add64(PsCounters* c, int field, unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, PS_WAVE);
    if ((threadIdx.x & (PS_WAVE - 1)) == 0 && v) atomicAdd(&c->stat[blockIdx.x & (PS_STRIPES - 1)][field], v);
}

// The tiles a span's pieces fall in: tiles g0 .. g1 (consecutive tile
// indices of one tile row), chunk range cx0 .. cx1 of the row.
struct PsSpanTiles {
    uint32_t g0, g1;
    int cx0, cx1, tx0;
};

// The chunk width, tile rows and tile columns are powers of two
// (psgpu_tiling_set) and y, x1, x2 are not negative, so the divisions are
// shifts (a division by a run-time divisor costs a dozen instructions).
__device__ __forceinline__ PsSpanTiles ps_span_tiles(int64_t tile_base, int n_tiles_x, int tile_rows, int tile_cols,
                                                     int chunk_w, int y, int x1, int x2)
{
    const int cs = __builtin_ctz((unsigned)chunk_w), ts = __builtin_ctz((unsigned)tile_cols);
    const int rs = __builtin_ctz((unsigned)tile_rows);
    PsSpanTiles s;
    s.cx0 = x1 >> cs;
    s.cx1 = x2 >> cs;
    s.tx0 = s.cx0 >> ts;
    s.g0 = (uint32_t)(tile_base + (int64_t)(y >> rs) * n_tiles_x + s.tx0);
    s.g1 = s.g0 + (uint32_t)((s.cx1 >> ts) - s.tx0);
    return s;
}

// the chunks [lo, hi] of the span inside tile k (g0 <= k <= g1)
__device__ __forceinline__ void ps_span_tile_chunks(const PsSpanTiles& s, int tile_cols, uint32_t k, int& lo, int& hi)
{
    int tx = s.tx0 + (int)(k - s.g0);
    lo = max(s.cx0, tx * tile_cols);
    hi = min(s.cx1, tx * tile_cols + tile_cols - 1);
}

// The distinct tiles of the wave's spans, in increasing order, up to 64 at a
// time: lane j receives the j-th tile (bk) and the wave's pieces in it (bn);
// returns how many.  cur is the lane's next tile (0xffffffff: none left) and
// moves on past the batch.  Every lane of the wave calls it.  A span's tiles
// are consecutive, so taking the smallest pending tile each round visits each
// (wave, tile) pair exactly once.
__device__ __forceinline__ int ps_wave_tile_batch(const PsSpanTiles& s, int tile_cols, uint32_t& cur, uint32_t& bk,
                                                  uint32_t& bn)
{
    const int lane = threadIdx.x & (PS_WAVE - 1);
    bk = 0xffffffffu;
    bn = 0;
    int j = 0;
    for (; j < PS_WAVE; j++) {
        const uint32_t k = ps_wave_min(cur);
        if (k == 0xffffffffu) break;
        const bool m = cur == k;
        int lo = 0, hi = -1;
        if (m) ps_span_tile_chunks(s, tile_cols, k, lo, hi);
        const uint32_t n = __builtin_amdgcn_readlane(ps_wave_incl_scan((uint32_t)(hi - lo + 1)), 63);
        if (lane == j) { bk = k; bn = n; }
        if (m) cur = k < s.g1 ? k + 1 : 0xffffffffu;
    }
    return j;
}

// ---------------------------------------------------------------------------
// fills (graph friendly: no hipMemset nodes)
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v, const PsCounters* c)
{
    if (c && ps_poisoned(c)) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// a clear command: fill f = blockIdx.y; clear4 (fbo.cpp:332-346) is a fill of
// rows 0 .. height-2, clear16 (fbo.cpp:348-371) one of the whole buffer.
// Rows without padding form one run, stored four values per thread; otherwise
// work items are pieces of 1024 values of one row (a 4 KiB stride keeps the
// row's alignment), and inside a piece the 16-byte aligned middle is stored
// four values per thread.
#define FILL_PIECE 1024
__global__ void __launch_bounds__(256) k_fills(const PsFill* __restrict__ fills, const PsCounters* c)
{
    if (c && ps_poisoned(c)) return;
    const PsFill fl = fills[blockIdx.y];
    const uint4 v4 = make_uint4(fl.value, fl.value, fl.value, fl.value);
    if (fl.scan == fl.width) {
        // rows without padding are one run: four values per thread, grid-stride
        const int64_t n = fl.rows * fl.width;
        uint32_t* p = fl.ptr;
        int64_t head = (int64_t)(((16u - ((uintptr_t)p & 15u)) & 15u) / 4);
        if (head > n) head = n;
        const int64_t nvec = (n - head) / 4, tail0 = head + nvec * 4;
        if (blockIdx.x == 0 && (int64_t)threadIdx.x < head) p[threadIdx.x] = fl.value;
        if (blockIdx.x == 0 && tail0 + (int64_t)threadIdx.x < n) p[tail0 + threadIdx.x] = fl.value;
        uint4* q = (uint4*)(p + head);
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nvec; i += (int64_t)gridDim.x * blockDim.x)
            q[i] = v4;
        return;
    }
    const int64_t npc = (fl.width + FILL_PIECE - 1) / FILL_PIECE;
    const int64_t total = fl.rows * npc;
    for (int64_t k = blockIdx.x; k < total; k += gridDim.x) {
        int64_t y = k / npc, x0 = (k - y * npc) * FILL_PIECE;
        int64_t n = fl.width - x0 < FILL_PIECE ? fl.width - x0 : FILL_PIECE;
        uint32_t* p = fl.ptr + y * fl.scan + x0;
        int head = (int)(((16u - ((uintptr_t)p & 15u)) & 15u) / 4);
        if (head > n) head = (int)n;
        int64_t nvec = (n - head) / 4;
        int64_t tail0 = head + nvec * 4;
        if ((int)threadIdx.x < head) p[threadIdx.x] = fl.value;
        if ((int64_t)threadIdx.x < nvec) ((uint4*)(p + head))[threadIdx.x] = v4;
        if (tail0 + threadIdx.x < n) p[tail0 + threadIdx.x] = fl.value;
    }
}


// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
// Draw of triangle t (the last draw with tri_start <= t; draws are sorted by
// tri_start and every t is covered).  The lanes' triangles lie in
// [t0, t0 + 63], t0 that of lane 0.  While more than 63 draws remain, a
// 64-way search (lane j loads pivot lo + j * step) narrows them to the
// interval holding t0's draw.  The last step loads the 63 draws from lo on
// into lanes 0..62 and the next one into lane 63: each lane's draw is then
// the last of those starting at or before its t, found from ballots and
// read lanes without another load, unless a lane's t reaches the draw after
// the 63 (then it steps on).  So one dependent load for up to 63 draws, two
// for up to 4,032.
__device__ __forceinline__ int ps_find_draw_wave(const PsDraw* __restrict__ draws, int nd, int64_t t)
{
    const int lane = threadIdx.x & 63;
    const int64_t t0 = (int64_t)__builtin_amdgcn_readfirstlane((int)(t >> 32)) << 32 |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)t);
    int lo = 0, hi = nd;
    while (hi - lo > 63) {
        const int step = (hi - lo + 63) / 64;
        const int p = lo + lane * step;
        const bool le = p < hi && draws[p].tri_start <= t0;
        const int cnt = __popcll(__ballot(le));   // pivot 0 (lo) is always <= t0
        // pivots below cnt are <= t0: t0's draw is in [lo + (cnt-1)*step, lo + cnt*step)
        const int nhi = lo + cnt * step;
        lo = lo + (cnt - 1) * step;
        hi = nhi < hi ? nhi : hi;
    }
    const int p = lo + lane;
    const int64_t sp = p < nd ? draws[p].tri_start : INT64_MAX;
    // draws lo .. lo+62 starting in (t0, t] for some lane
    const uint64_t later = __ballot(lane < 63 && sp > t0 && sp <= t0 + 63) ;
    int d = lo + __popcll(__ballot(lane < 63 && sp <= t0)) - 1;
    for (uint64_t m = later; m;) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        const int64_t sb = (int64_t)__builtin_amdgcn_readlane((int)(sp >> 32), b) << 32 |
                           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)sp, b);
        if (t >= sb) d = lo + b;
    }
    // the draw after the 63 starts at or before some lane's t: step on
    const int64_t s63 = (int64_t)__builtin_amdgcn_readlane((int)(sp >> 32), 63) << 32 |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)sp, 63);
    if (s63 <= t0 + 63)
        while (d + 1 < nd && draws[d + 1].tri_start <= t) d++;
    return d;
}

// row_part (flushes of one programme): the block's row total, for the scan
// of the row counts.  At least three waves per SIMD: a 1080p frame's 200k
// triangles make only about three per SIMD, and the bound of four left the
// Blinn-Phong instantiation at 128 VGPRs with a 12-byte spill.
template <int PROG, bool SHARE>
__global__ void __launch_bounds__(128, 3) k_setup(const PsDraw* __restrict__ draws, int nd, int64_t T, PsTri* __restrict__ tris,
                                               float* __restrict__ vary, int nvs, float* __restrict__ grads,
                                               uint32_t* __restrict__ nrows, uint32_t* __restrict__ row_part,
                                               PsCounters* cnt, const PsView* __restrict__ views)
{
    typedef PsProc<PROG> P;
    extern __shared__ float su[];   // per wave: the varyings of its 64 triangles, 3*nvs floats each
    __shared__ uint32_t wrows[2];
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool live = t < T;
    int di = ps_find_draw_wave(draws, nd, live ? t : T - 1);
    unsigned culled = 0, zrej = 0, dropped = 0;
    uint32_t myrows = 0;
    const int lane = threadIdx.x & 63, n3 = 3 * nvs;
    float* wl = su + (size_t)(threadIdx.x & ~63) * n3;
    bool stage = false;
    ps_with_draw(draws, di, [&](const PsDraw& d) {
    if (live && d.prog == PROG) {
        const int W = views[d.view].W, H = views[d.view].H;
        int64_t lt = t - d.tri_start;
        float pos[3][4];
        float u[3][P::NV > 0 ? P::NV : 1];
        float rw[3], z[3];
        bool pieces = false;
        const bool skip = SHARE && ps_shares_varyings(PROG) && d.vary_skip;
#pragma unroll
        for (int v = 0; v < 3; v++) {
            if (skip) ps_vp_pos(d, lt * 3 + v, pos[v], u[v], ps_grad_iu<P>());
            else P::vp(d, lt * 3 + v, pos[v], u[v]);
            // perspective division (vertthrd.cpp:37-47)
            float r = 1.0f / pos[v][3];
            pos[v][0] *= r; pos[v][1] *= r; pos[v][2] *= r; pos[v][3] *= r;
            z[v] = pos[v][2];
            rw[v] = r;
        }
        uint32_t n = 0;
        PsTri tr;
        bool keep = true;
        if ((d.behavior & PSGPU_FACE_CULLING) && ps_back_face(pos[0], pos[1], pos[2])) { keep = false; culled = 1; }
        if (keep) {
            // drawvao.cpp:51-56
#pragma unroll
            for (int v = 0; v < 3; v++)
                if (pos[v][2] < -1.0f || pos[v][2] > 1.0f) keep = false;
            if (!keep) zrej = 1;
        }
        if (keep) {
            // pushVertex (rasterizer.cpp:122-139)
            const int hw = W / 2, hh = H / 2;
            float vx[3], vy[3];
            bool bad = false;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                vy[v] = ((float)hh * pos[v][1]) + (float)hh;
                vx[v] = ((float)hw * pos[v][0]) + (float)hw;
                // rows outside the int range make the reference read rows it
                // never wrote (or crash); the triangle is dropped
                if (!(vy[v] >= -2147483648.0f && vy[v] < 2147483648.0f)) bad = true;
            }
            int first = ps_cvt_i32(vy[0]), last = first;
#pragma unroll
            for (int v = 1; v < 3; v++) {
                if (vy[v] > (float)last) last = ps_cvt_i32(vy[v]);
                else if (vy[v] < (float)first) first = ps_cvt_i32(vy[v]);
            }
            if (first >= H || last < 0) { first = 0; last = -1; }
            if (first < 0) first = 0;
            if (last >= H) last = H - 1;
            if (bad || first == last) { keep = false; dropped = 1; }
            if (keep) {
                uint32_t code;
                int split = H;
                if (vy[0] == vy[1]) {
                    code = (vx[0] < vx[1]) ? (0u | (1u << 2) | (2u << 4)) : (1u | (0u << 2) | (2u << 4));
                } else if (vy[0] == vy[2]) {
                    code = (vx[0] < vx[2]) ? (0u | (2u << 2) | (1u << 4)) : (2u | (0u << 2) | (1u << 4));
                } else if (vy[1] == vy[2]) {
                    code = (vx[1] < vx[2]) ? (1u | (2u << 2) | (0u << 4)) : (2u | (1u << 2) | (0u << 4));
                } else {
                    int top, bottom, third = 2;
                    if (vy[0] > vy[1]) { top = 0; bottom = 1; } else { top = 1; bottom = 0; }
                    if (ps_sel3(vy, top) < ps_sel3(vy, third)) { int s = top; top = third; third = s; }
                    else if (ps_sel3(vy, bottom) > ps_sel3(vy, third)) { int s = bottom; bottom = third; third = s; }
                    PsEdge tt = ps_edge(vx, vy, top, third), tb = ps_edge(vx, vy, top, bottom);
                    float y3 = ps_sel3(vy, third);
                    bool left = ps_edge_eval(tb, y3) > ps_edge_eval(tt, y3);
                    code = (uint32_t)top | ((uint32_t)bottom << 2) | ((uint32_t)third << 4) | 64u | (left ? 128u : 0u);
                    // last row of the lower half: processTriangle's iyMax for yMax = third.y
                    split = y3 > (float)(H - 1) ? H - 1 : ps_cvt_i32(y3);
                }
#pragma unroll
                for (int v = 0; v < 3; v++) { tr.vx[v] = vx[v]; tr.vy[v] = vy[v]; tr.rw[v] = rw[v]; tr.z[v] = z[v]; }
                tr.first = first;
                tr.nrows = last - first + 1;
                tr.split = split;
                tr.code = code | ((uint32_t)di << 8);
                tr.view = d.view;
                tr.vflags = (uint32_t)(d.behavior & 3) | (d.writes_colour ? 4u : 0u);
                tr.pad[0] = tr.pad[1] = 0;
                n = (uint32_t)(last - first + 1);
                // a triangle without rows has no spans: nothing reads its records
                // (rows can lie outside the vertices' y range, where edges
                // extrapolate, so no x range test is safe here)
                pieces = n > 0;
                if (pieces) tris[t] = tr;
                if (!SHARE && pieces) {   // staged: the wave stores whole triangles below
#pragma unroll
                    for (int v = 0; v < 3; v++)
#pragma unroll
                        for (int k = 0; k < P::NV; k++) wl[lane * n3 + v * nvs + k] = u[v][k];
                }
                if constexpr (P::GRAD) if (pieces) {   // screen gradients for the mip level, once per triangle
                    float g[6];
                    ps_tri_grad(tr, u[0], u[1], u[2], P::IU, g);
                    float4* go = (float4*)(grads + t * 8);
                    go[0] = make_float4(g[0], g[1], g[2], g[3]);
                    go[1] = make_float4(g[4], g[5], 0.0f, 0.0f);
                }
            }
        }
        nrows[t] = n;
        myrows = n;
        // varyings of triangles with rows, or of all when other views read them
        stage = SHARE ? !skip && (pieces || d.vary_all) : pieces;
        if (SHARE && stage) {
#pragma unroll
            for (int v = 0; v < 3; v++)
#pragma unroll
                for (int k = 0; k < P::NV; k++) wl[lane * n3 + v * nvs + k] = u[v][k];
        }
    }
    });
    // varyings of the wave's staged triangles: the wave's slots are one
    // contiguous block of 64 * 3*nvs floats, stored 16 bytes per lane and
    // instruction.  A float is written only when its triangle staged in this
    // launch (the slots of other programmes' triangles belong to their own
    // launches); floats past NV of a staged vertex receive what the LDS
    // holds, and nothing reads them.
    const uint64_t m = __ballot(stage);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (m) {
        const int64_t tw = t - lane;
        const int nf = (int)min((int64_t)PS_WAVE, T - tw) * n3;   // the wave's floats
        const float rn3 = 1.0f / (float)n3;   // (f + 0.5) * rn3 floors to f / n3 exactly for f < 64 * 60
        float* dst = vary + tw * n3;
        auto staged = [&](int f) { return (m >> (int)(((float)f + 0.5f) * rn3)) & 1ull; };
        for (int f = lane * 4; f < nf; f += 4 * PS_WAVE) {
            if (f + 4 <= nf && staged(f) && staged(f + 3)) {
                *(float4*)(dst + f) = *(const float4*)(wl + f);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (f + j < nf && staged(f + j)) dst[f + j] = wl[f + j];
            }
        }
    }
    ps_wave_add(cnt, PS_ST_CULLED, culled, true);
    ps_wave_add(cnt, PS_ST_ZREJ, zrej, true);
    ps_wave_add(cnt, PS_ST_DROPPED, dropped, true);
    if (row_part) {
        const uint32_t w = __builtin_amdgcn_readlane(ps_wave_incl_scan(myrows), 63);
        if (lane == 0) wrows[threadIdx.x >> 6] = w;
        __syncthreads();
        if (threadIdx.x == 0) row_part[blockIdx.x] = wrows[0] + wrows[1];
    }
}

// ---------------------------------------------------------------------------
// exclusive scan (uint32, or uint64 holding two 32-bit counts side by side)
// with a device-side length; out[n] = total
// ---------------------------------------------------------------------------
#define SCAN_ITEMS 16
#define SCAN_BLOCK (256 * SCAN_ITEMS)
#define PS_HEAVY_CAP 1024

// Pieces that make a tile heavy: three times the flush's mean per tile (at
// least 96).  The flush's pieces are k_span's 64-bit count in the pending
// statistics: ps_pend_pieces loads one stripe per lane (0 without counters),
// ps_wave_sum64 adds the wave's stripes.
__device__ __forceinline__ unsigned long long ps_pend_pieces(const PsCounters* c)
{
    static_assert(PS_STRIPES == 64, "one stripe per lane");
    return c ? c->pend[threadIdx.x & 63][PS_ST_SEGS] : 0ull;
}

__device__ __forceinline__ unsigned long long ps_wave_sum64(unsigned long long x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// the wave's pieces rounded down to multiples of 256 (plenty for a
// threshold): a 32-bit DPP scan instead of 64-bit shuffles
__device__ __forceinline__ unsigned long long ps_pend_approx(unsigned long long stripe)
{
    const uint32_t p256 = (uint32_t)min(stripe >> 8, (unsigned long long)0x3ffffff);
    return (unsigned long long)__builtin_amdgcn_readlane(ps_wave_incl_scan(p256), 63) << 8;
}

__device__ __forceinline__ uint32_t ps_heavy_min(unsigned long long pieces, int64_t n_tiles)
{
    const unsigned long long h = 3 * pieces / (unsigned long long)(n_tiles > 0 ? n_tiles : 1);
    return (uint32_t)min(max(h, 96ull), 0xffffffffull);
}

// inclusive wave scan of uint32, or of the two 32-bit halves of a uint64
// separately (the packed counts never carry from the low half)
__device__ __forceinline__ uint32_t ps_wave_incl_scan_t(uint32_t v) { return ps_wave_incl_scan(v); }
__device__ __forceinline__ unsigned long long ps_wave_incl_scan_t(unsigned long long v)
{
    return (unsigned long long)ps_wave_incl_scan((uint32_t)v) |
           ((unsigned long long)ps_wave_incl_scan((uint32_t)(v >> 32)) << 32);
}

__device__ __forceinline__ uint32_t ps_wave_bcast63(uint32_t v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ uint32_t ps_wave_bcast0(uint32_t v) { return __builtin_amdgcn_readlane(v, 0); }
__device__ __forceinline__ unsigned long long ps_wave_bcast0(unsigned long long v)
{
    return (unsigned long long)(uint32_t)__builtin_amdgcn_readlane((uint32_t)v, 0) |
           ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), 0) << 32);
}
__device__ __forceinline__ unsigned long long ps_wave_bcast63(unsigned long long v)
{
    return (unsigned long long)(uint32_t)__builtin_amdgcn_readlane((uint32_t)v, 63) |
           ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), 63) << 32);
}

// exclusive scan of one value per thread over a block of NW waves
template <typename T, int NW>
__device__ __forceinline__ T ps_block_excl_scan(T v, T* lds, T& total)
{
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = ps_wave_incl_scan_t(v);
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T wbase = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        T l = lds[w];
        wbase += (w < wave) ? l : (T)0;
        total += l;
    }
    __syncthreads();
    return wbase + x - v;
}

// Wave-striped ownership: the wave owns items [wbase, wbase + 64 * SCAN_ITEMS),
// in groups of 64 * PER (PER = the items in 16 bytes); in group g the lane
// owns the PER consecutive items wbase + g * 64 * PER + lane * PER + j, so
// each group is one coalesced 16-byte load or store per lane.
template <typename T>
__device__ __forceinline__ void ps_load_items(const T* in, int64_t wbase, int64_t n, T* v)
{
    constexpr int PER = 16 / sizeof(T);
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < SCAN_ITEMS / PER; g++) {
        const int64_t i = wbase + (int64_t)g * 64 * PER + lane * PER;
        if (i + PER <= n) {
            const uint4 u = *(const uint4*)(in + i);
            const T* t = (const T*)&u;
#pragma unroll
            for (int j = 0; j < PER; j++) v[g * PER + j] = t[j];
        } else {
#pragma unroll
            for (int j = 0; j < PER; j++) v[g * PER + j] = i + j < n ? in[i + j] : (T)0;
        }
    }
}

// Heavy tiles among the lane's items (wave-striped as above): those of at
// least heavy_min pieces.
__device__ __forceinline__ uint32_t ps_heavy_count(int64_t wbase, int64_t n, const unsigned long long* v, uint32_t heavy_min)
{
    constexpr int PER = 2;
    const int lane = threadIdx.x & 63;
    uint32_t nh = 0;
#pragma unroll
    for (int g = 0; g < SCAN_ITEMS / PER; g++)
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int64_t i = wbase + (int64_t)g * 64 * PER + lane * PER + j;
            nh += (i < n && (uint32_t)v[g * PER + j] >= heavy_min) ? 1u : 0u;
        }
    return nh;
}

// The lane's first slot in the heavy list for its nh heavy tiles, reserved
// with one atomic per wave on heavy[0] (0xffffffff: the wave has none).
__device__ __forceinline__ uint32_t ps_heavy_reserve(uint32_t nh, uint32_t* heavy)
{
    const int lane = threadIdx.x & 63;
    const uint32_t incl = ps_wave_incl_scan(nh);
    const uint32_t tot = __builtin_amdgcn_readlane(incl, 63);
    if (!tot) return 0xffffffffu;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(heavy, tot);
    return __builtin_amdgcn_readlane(base, 0) + incl - nh;
}

// The wave's exclusive prefix of its items (from `run` on) into out.  Tile
// scans (T = uint64, heavy given) also list the tiles of at least heavy_min
// pieces in heavy[1 ..], the lane's from slot hslot on (up to PS_HEAVY_CAP
// are listed), and mark the listed ones with bit 63 of their start.
template <typename T>
__device__ __forceinline__ void ps_store_scan(T* out, int64_t wbase, int64_t n, const T* v, T run,
                                              uint32_t heavy_min = 0xffffffffu, uint32_t* heavy = nullptr,
                                              uint32_t hslot = 0xffffffffu)
{
    constexpr int PER = 16 / sizeof(T);
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < SCAN_ITEMS / PER; g++) {
        T sum = 0;
#pragma unroll
        for (int j = 0; j < PER; j++) sum += v[g * PER + j];
        const T x = ps_wave_incl_scan_t(sum);
        T r = run + x - sum;
        const int64_t i = wbase + (int64_t)g * 64 * PER + lane * PER;
        T o[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            o[j] = r;
            if constexpr (sizeof(T) == 8) {
                if (hslot != 0xffffffffu && i + j < n && (uint32_t)v[g * PER + j] >= heavy_min) {
                    if (hslot < PS_HEAVY_CAP) {
                        heavy[1 + hslot] = (uint32_t)(i + j);
                        o[j] |= 1ull << 63;
                    }
                    hslot++;
                }
            }
            r += v[g * PER + j];
        }
        if (i + PER <= n) {
            uint4 u;
            T* t = (T*)&u;
#pragma unroll
            for (int j = 0; j < PER; j++) t[j] = o[j];
            *(uint4*)(out + i) = u;
        } else {
            for (int j = 0; j < PER; j++)
                if (i + j < n) out[i + j] = o[j];
        }
        run += ps_wave_bcast63(x);
    }
}

// The tile scan (T = uint64) adds 32-bit piece counts: when the flush's
// pieces, which k_span also counts in 64 bits into the pending statistics,
// do not fit, it reports the largest total and flags an overflow, so that
// the flush is split.  Called by the 64 lanes of one wave.
template <typename T>
__device__ __forceinline__ void ps_scan_total(T tot, int64_t n, int64_t cap, T* out, uint32_t* total_out,
                                              const PsCounters* c, unsigned long long pieces = 0)
{
    const int lane = threadIdx.x & 63;
    if (lane == 0 && n <= cap) {
        out[n] = tot;
        if (total_out) {
            *total_out = (uint32_t)tot;
            atomicMax(total_out + 2, (uint32_t)tot);
        }
    }
    if constexpr (sizeof(T) == 8) {
        if (lane == 0 && pieces != (uint32_t)tot) {
            if (total_out) atomicMax(total_out + 2, 0xffffffffu);
            atomicOr((int*)&c->overflow, 1);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_scan_reduce(const T* __restrict__ in, const uint32_t* n_ptr, int64_t n_host,
                                                     T* __restrict__ partial, const PsCounters* c)
{
    __shared__ T lds[4];
    int64_t n = n_ptr ? (int64_t)*n_ptr : n_host;
    int64_t wbase = blockIdx.x * (int64_t)SCAN_BLOCK + (int64_t)(threadIdx.x >> 6) * (64 * SCAN_ITEMS);
    T s = 0;
    if (!(c && ps_poisoned(c)) && wbase < n) {
        T v[SCAN_ITEMS];
        ps_load_items(in, wbase, n, v);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; i++) s += v[i];
    }
    T tot;
    ps_block_excl_scan<T, 4>(s, lds, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// sum of partial[0, end) over a 256-thread block
template <typename T>
__device__ __forceinline__ T ps_block_sum_prefix(const T* partial, int end, T* lds)
{
    T s = 0;
#pragma unroll 8
    for (int i = threadIdx.x; i < end; i += 256) s += partial[i];
    T tot;
    ps_block_excl_scan<T, 4>(s, lds, tot);
    return tot;
}

// block b: prefix of the partial sums before it (partial i sums items
// [i * pgran, (i + 1) * pgran), pgran dividing SCAN_BLOCK), then the block's
// own scan.  Block 0 also writes the total.  With a host-side length the
// overflow flag is tested only before the first store: the loads stay in
// bounds either way, so its round trip overlaps theirs.
template <typename T>
__global__ void __launch_bounds__(256) k_scan_final(const T* __restrict__ in, const uint32_t* n_ptr, int64_t n_host,
                                                    const T* __restrict__ partial, int nb, int pgran, T* __restrict__ out,
                                                    uint32_t* total_out, int64_t cap, const PsCounters* c,
                                                    uint32_t* heavy)
{
    __shared__ T lds[4];
    const bool poisoned = c && ps_poisoned(c);
    if (n_ptr && poisoned) return;
    const unsigned long long pend = sizeof(T) == 8 ? ps_pend_pieces(c) : 0ull;
    const int64_t n = n_ptr ? (int64_t)*n_ptr : n_host;
    const int64_t wbase = blockIdx.x * (int64_t)SCAN_BLOCK + (int64_t)(threadIdx.x >> 6) * (64 * SCAN_ITEMS);
    const bool mine = blockIdx.x * (int64_t)SCAN_BLOCK < n;
    T v[SCAN_ITEMS];
    if (mine) ps_load_items(in, wbase, n, v);
    T total = 0, before = 0;
    if (blockIdx.x == 0) total = ps_block_sum_prefix(partial, nb, lds);
    else if (mine) before = ps_block_sum_prefix(partial, blockIdx.x * (SCAN_BLOCK / pgran), lds);
    if (poisoned) return;
    const unsigned long long pieces = ps_wave_sum64(pend);
    if (blockIdx.x == 0 && threadIdx.x < 64) ps_scan_total(total, n, cap, out, total_out, c, pieces);
    if (!mine) return;
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) s += v[i];
    T tot;
    // the waves before this one: the block scan of the lane sums, taken at lane 0
    const T wpre = ps_block_excl_scan<T, 4>(s, lds, tot);
    uint32_t hm = 0xffffffffu, hslot = 0xffffffffu;
    if constexpr (sizeof(T) == 8) {
        if (heavy) {
            hm = ps_heavy_min(ps_pend_approx(pend), n);
            hslot = ps_heavy_reserve(ps_heavy_count(wbase, n, v, hm), heavy);
        }
    }
    ps_store_scan(out, wbase, n, v, before + ps_wave_bcast0(wpre), hm, heavy, hslot);
}

// One pass for tile counts up to SCAN_DIRECT_MAX.  Blocks of 16 waves; each
// lane owns two adjacent tiles (one 16-byte load), so a block covers
// DIRECT_BLOCK tiles and a wave needs one wave scan.  Block b adds the counts
// before its own tiles directly (up to 8 16-byte loads per thread in flight
// at a time, one round trip below 16k tiles) and counts the heavy tiles among
// them on the way, so each lane knows its slots in the heavy list without
// atomics (the list is in tile order).  The waves combine their totals in
// one LDS exchange; the last block writes the totals.
#define DIRECT_BLOCK 2048
#define SCAN_DIRECT_MAX (32 * DIRECT_BLOCK)
struct PsDirectWave {
    unsigned long long pre, own;   // counts before the block (the wave's share), the wave's own tiles
    uint32_t hpre, hown;           // heavy tiles among them
};
__global__ void __launch_bounds__(1024) k_scan_direct(const unsigned long long* __restrict__ in, int64_t n,
                                                      unsigned long long* __restrict__ out, uint32_t* total_out,
                                                      const PsCounters* c, uint32_t* heavy)
{
    typedef unsigned long long T;
    __shared__ PsDirectWave xw[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // tested before the first store: every load stays in bounds either way,
    // so the flag's round trip overlaps theirs
    const bool poisoned = ps_poisoned(c);
    const int64_t b0 = blockIdx.x * (int64_t)DIRECT_BLOCK;
    const int64_t i = b0 + 2 * (int64_t)threadIdx.x;   // the lane's tiles i, i + 1
    const T a = in[min(i, n - 1)], b = in[min(i + 1, n - 1)];
    // tiles before the block: 8 loads per chunk, addresses past b0 clamped
    // (b0 - 2, a multiple of DIRECT_BLOCK less 2) and their values dropped
    const int nk = (int)(b0 / DIRECT_BLOCK);   // 16-byte loads per thread in all
    ulonglong2 u[8];
    auto chunk_load = [&](int k0) {
#pragma unroll
        for (int k = 0; k < 8; k++)
            u[k] = *(const ulonglong2*)(in + min(((int64_t)(k0 + k) * 1024 + threadIdx.x) * 2, b0 - 2));
    };
    if (nk > 0) chunk_load(0);
    const unsigned long long pend = ps_pend_pieces(c);   // after the loads: one wait covers them all
    const uint32_t hm = heavy ? ps_heavy_min(ps_pend_approx(pend), n) : 0xffffffffu;
    T pre = 0;
    uint32_t hpre = 0;
    for (int k0 = 0; k0 < nk; k0 += 8) {
        if (k0 > 0) chunk_load(k0);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool ok = k0 + k < nk;
            const T x = ok ? u[k].x : 0ull, y = ok ? u[k].y : 0ull;
            pre += x + y;
            hpre += ((uint32_t)x >= hm ? 1u : 0u) + ((uint32_t)y >= hm ? 1u : 0u);   // hm >= 96: dropped values never count
        }
    }
    if (poisoned) return;
    const T va = i < n ? a : 0ull, vb = i + 1 < n ? b : 0ull;
    const uint32_t ha = (uint32_t)va >= hm ? 1u : 0u, hb = (uint32_t)vb >= hm ? 1u : 0u;
    // the wave's inclusive scans of its own tiles and heavy marks, and its
    // totals before the block
    const T own = ps_wave_incl_scan_t(va + vb);
    const uint32_t hown = ps_wave_incl_scan(ha + hb);
    const T pre_w = ps_wave_incl_scan_t(pre);
    const uint32_t hpre_w = ps_wave_incl_scan(hpre);
    if (lane == 63) xw[wave] = PsDirectWave{pre_w, own, hpre_w, hown};
    __syncthreads();
    T before = 0, woff = 0, btot = 0;
    uint32_t hbefore = 0, hwoff = 0, hbtot = 0;
#pragma unroll
    for (int w = 0; w < 16; w++) {
        const PsDirectWave e = xw[w];
        before += e.pre;
        hbefore += e.hpre;
        woff += w < wave ? e.own : 0ull;
        hwoff += w < wave ? e.hown : 0u;
        btot += e.own;
        hbtot += e.hown;
    }
    T r = before + woff + own - (va + vb);                 // tile i's start
    uint32_t hs = hbefore + hwoff + hown - (ha + hb);      // its first heavy slot
    T oa = r, ob = r + va;
    if (heavy) {
        if (ha && hs < PS_HEAVY_CAP) { heavy[1 + hs] = (uint32_t)i; oa |= 1ull << 63; }
        hs += ha;
        if (hb && hs < PS_HEAVY_CAP) { heavy[1 + hs] = (uint32_t)(i + 1); ob |= 1ull << 63; }
    }
    if (i + 2 <= n) {
        *(ulonglong2*)(out + i) = make_ulonglong2(oa, ob);
    } else if (i < n) {
        out[i] = oa;
    }
    if (blockIdx.x == gridDim.x - 1 && wave == 0) {
        ps_scan_total(before + btot, n, n, out, total_out, c, ps_wave_sum64(pend));
        if (heavy && lane == 0) heavy[0] = hbefore + hbtot;
    }
}

// Row scan from per-block partial sums (partial i sums items [i * pgran,
// (i + 1) * pgran), pgran dividing SCAN_BLOCK): blocks of 16 waves, each lane
// owning four adjacent items (one 16-byte load), so a block covers
// SCAN_BLOCK items and a wave needs one wave scan.  Block b adds the partials
// before it (two loads per thread below 2048 of them), the waves combine in
// one LDS exchange, and the last block writes the total.
struct PsRowWave {
    uint32_t pre, own;
};
__global__ void __launch_bounds__(1024) k_scan_parts(const uint32_t* __restrict__ in, int64_t n,
                                                     const uint32_t* __restrict__ partial, int pgran,
                                                     uint32_t* __restrict__ out, uint32_t* total_out,
                                                     uint32_t* __restrict__ bstart, int64_t bcap)
{
    __shared__ PsRowWave xw[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = blockIdx.x * (int64_t)SCAN_BLOCK + 4 * (int64_t)threadIdx.x;   // items i .. i + 3
    uint32_t v[4];
    if (i + 4 <= n) {
        const uint4 q = *(const uint4*)(in + i);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = in[min(i + j, n - 1)];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = i + j < n ? v[j] : 0u;
    }
    const int np = blockIdx.x * (SCAN_BLOCK / pgran);   // partials before the block
    uint32_t pre = 0;
#pragma unroll 4
    for (int k = threadIdx.x; k < np; k += 1024) pre += partial[k];
    const uint32_t sum = v[0] + v[1] + v[2] + v[3];
    const uint32_t own = ps_wave_incl_scan(sum);
    const uint32_t pre_w = ps_wave_incl_scan(pre);
    if (lane == 63) xw[wave] = PsRowWave{pre_w, own};
    __syncthreads();
    uint32_t before = 0, woff = 0, btot = 0;
#pragma unroll
    for (int w = 0; w < 16; w++) {
        const PsRowWave e = xw[w];
        before += e.pre;
        woff += w < wave ? e.own : 0u;
        btot += e.own;
    }
    uint32_t r = before + woff + own - sum;
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; j++) { o[j] = r; r += v[j]; }
    // the triangle of every 64th row item: item 64b lies in item i + j's rows
    if (bstart) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (i + j >= n) break;
            const int64_t b1 = min(((int64_t)o[j] + v[j] + 63) >> 6, bcap);
            for (int64_t b = ((int64_t)o[j] + 63) >> 6; b < b1; b++) bstart[b] = (uint32_t)(i + j);
        }
    }
    if (i + 4 <= n) {
        *(uint4*)(out + i) = make_uint4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (i + j < n) out[i + j] = o[j];
    }
    if (blockIdx.x == gridDim.x - 1 && wave == 0) ps_scan_total(before + btot, n, n, out, total_out, (const PsCounters*)nullptr);
}

// Triangle of row item i: the t with row_off[t] <= i < row_off[t + 1]
// (row_off is the exclusive scan of the row counts, row_off[T] = R).  The
// wave finds the triangle of its lowest unresolved item with a 64-way search
// over row_off (64 pivots per step, a few dependent loads), holds the next 64
// row starts one per lane, and every lane counts those at or below its own
// item with a 6-step search over the lanes.  Lanes whose item lies beyond
// that window go round again from the lowest of them, so every round
// resolves at least one lane and runs of triangles without rows are jumped.
// All lanes of the wave call it; i0 is the wave's first item.
// Flushes that search hold fewer than 2^31 triangles and R < 2^26 row items,
// so the arithmetic is 32-bit.  hint, when not negative, is the triangle of
// item i0 and hint_hi that of item i0 + 64 (the row scan's table of every
// 64th item's triangle): the first round then takes the 64 row starts after
// hint without a search, and a later one searches only up to hint_hi.
__device__ __forceinline__ int ps_row_tri(const uint32_t* __restrict__ row_off, int T, uint32_t R, uint32_t i0, int hint,
                                          int hint_hi)
{
    const int lane = threadIdx.x & 63;
    const uint32_t i = i0 + lane;
    int found = -1;
    uint64_t pend = __ballot(i < R);
    const int lo0 = hint > 0 ? hint : 0;
    const int hi0 = hint >= 0 && hint_hi >= 0 ? hint_hi + 1 : T;
    bool guess = hint >= 0;
    while (pend) {
        const uint32_t tgt = i0 + __builtin_ctzll(pend);
        int lo = lo0, hi = guess ? lo0 + 1 : hi0;   // row_off[lo] <= tgt; hi == T or row_off[hi] > tgt
        guess = false;
        while (hi - lo > 1) {
            const int step = (hi - lo + 63) >> 6;
            const int p = lo + lane * step;
            const bool le = p < hi && row_off[p] <= tgt;
            const int c = __popcll(__ballot(le));
            const int nhi = lo + c * step;
            lo += (c - 1) * step;
            hi = nhi < hi ? nhi : hi;
        }
        const int we = lo + 1 + lane;
        const uint32_t e = we <= T ? row_off[we] : 0xffffffffu;   // above every item (R < 2^26)
        int pos = 0;   // row starts of the window at or below i: 0..63, or 64 past it
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1) {
            const uint32_t v = (uint32_t)__shfl((int)e, pos + st - 1, 64);
            if (v <= i) pos += st;
        }
        if ((uint32_t)__builtin_amdgcn_readlane((int)e, 63) <= i) pos = 64;
        if (found < 0 && i < R && i >= tgt && pos < 64) found = lo + pos;
        pend = __ballot(found < 0 && i < R);
    }
    return found;
}

// ---------------------------------------------------------------------------
// span setup
// ---------------------------------------------------------------------------
#define SPAN_BINS 512
// Also counts, per tile, the span pieces in it and the waves (of 64 span
// items) that have pieces there: tcnt[t] += pieces | waves << 32.  Eight
// waves per SIMD (64 VGPRs, no spill): at seven the 256-view batch's span
// pass took 1.55 ms instead of 1.21.
__global__ void __launch_bounds__(256, 8) k_span(const PsTri* __restrict__ tris,
                                              const uint32_t* __restrict__ row_off, const uint32_t* __restrict__ bstart,
                                              int64_t T, int64_t cap,
                                              PsSpan* __restrict__ spans, unsigned long long* __restrict__ tcnt,
                                              const PsView* __restrict__ views, int n_views, int chunk_w, int tile_rows,
                                              int tile_cols, uint32_t* heavy, PsCounters* cnt)
{
    // per wave: piece counts of the tiles [kmin, kmin + SPAN_BINS) its spans touch
    __shared__ uint32_t bins_all[4][SPAN_BINS];
    uint32_t* bins = bins_all[threadIdx.x >> 6];
    const int lane = threadIdx.x & (PS_WAVE - 1);
    const bool poisoned = ps_poisoned(cnt);
    const int64_t R = row_off[T];
    // the only view of a one-view flush, read once (otherwise each triangle's)
    int64_t tb1 = 0;
    int ntx1 = 0, W1 = 0;
    if (n_views == 1) { tb1 = views[0].tile_base; ntx1 = views[0].n_tiles_x; W1 = views[0].W; }
    if (poisoned) return;
    if (heavy && blockIdx.x == 0 && threadIdx.x == 0) heavy[0] = 0;   // the two-pass tile scan lists heavy tiles atomically
    if (R > cap) {   // more rows than span records: flag it, write nothing
        if (threadIdx.x == 0 && blockIdx.x == 0) atomicOr(&cnt->overflow, 1);
        return;
    }
    uint32_t all_spans = 0, all_pieces = 0, all_frags = 0;   // the lane's statistics, added up once per wave
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < R; base += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = base + threadIdx.x;
        uint32_t nspan = 0, nfrag = 0;
        // the item's triangle, searched in row_off from the row scan's table
        const int64_t i0 = base + (threadIdx.x & ~63);
        const bool tab = i0 < R;
        const int tw = ps_row_tri(row_off, (int)T, (uint32_t)R, (uint32_t)i0, tab ? (int)bstart[i0 >> 6] : -1,
                                  tab && i0 + 64 < R ? (int)bstart[(i0 >> 6) + 1] : -1);
        PsSpanTiles st = {};
        bool has = false;
        if (i < R) {
            const int t = tw;
            const PsTri tr = tris[t];
            int64_t vtb = tb1;
            int vntx = ntx1, W = W1;
            if (n_views != 1) { const PsView& v = views[tr.view]; vtb = v.tile_base; vntx = v.n_tiles_x; W = v.W; }
            int row = tr.first + (int)(i - (int64_t)row_off[t]);
            PsEdge L, Rr;
            ps_row_edges(tr, row, L, Rr);
            int left = ps_cvt_i32(ps_edge_eval(L, (float)row) + 0.5f);
            int right = ps_cvt_i32(ps_edge_eval(Rr, (float)row) + 0.5f);
            PsSpan sp;
            sp.tri = -1;
            if (left != right) {   // drawvao.cpp:72-75
                nspan = 1;
                int x1 = left < 0 ? 0 : left;
                int x2 = right >= W ? W - 1 : right;
                if (x1 <= x2) {
                    PsSpanGeom g;
                    ps_span_geom(tr, row, left, right, L, Rr, g);
                    ps_span_vis(tr, g, sp.cf2, sp.cf2_step, sp.pz, sp.pz_step);
                    sp.tri = t;
                    sp.y = row | (int)((tr.vflags & 3u) << 24) | (int)(((tr.vflags >> 2) & 1u) << 28);
                    sp.x1 = x1;
                    sp.x2 = x2;
                    nfrag = (uint32_t)(x2 - x1 + 1);
                    st = ps_span_tiles(vtb, vntx, tile_rows, tile_cols, chunk_w, row, x1, x2);
                    has = true;
                }
            }
            spans[i] = sp;
        }
        // per tile of the wave's spans: its pieces there and one run.  When
        // the wave's tiles lie within SPAN_BINS of each other (about nine
        // tiles of one or two tile rows, typically), the lanes add their
        // pieces into LDS bins and one atomic per touched bin follows;
        // otherwise the wave visits its tiles in rounds (lane j adds the
        // j-th tile's counts, so a batch of tiles is one atomic).
        const uint32_t npieces = has ? (uint32_t)(st.cx1 - st.cx0 + 1) : 0u;
        const uint32_t kmin = ps_wave_min(has ? st.g0 : 0xffffffffu);
        const uint32_t kmax = ps_wave_max(has ? st.g1 : 0u);
        if (kmin != 0xffffffffu && kmax - kmin < (uint32_t)SPAN_BINS) {
            const uint32_t nbin = kmax - kmin + 1;
            for (uint32_t b = lane; b < nbin; b += PS_WAVE) bins[b] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (has)
                for (uint32_t k = st.g0; k <= st.g1; k++) {
                    int lo, hi;
                    ps_span_tile_chunks(st, tile_cols, k, lo, hi);
                    atomicAdd(&bins[k - kmin], (uint32_t)(hi - lo + 1));
                }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t b = lane; b < nbin; b += PS_WAVE) {
                const uint32_t c = bins[b];
                if (c) atomicAdd(&tcnt[kmin + b], (unsigned long long)c | (1ull << 32));
            }
            // the next iteration clears the bins only after every lane read them
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            uint32_t cur = has ? st.g0 : 0xffffffffu;
            for (;;) {
                uint32_t bk, bn;
                const int nb = ps_wave_tile_batch(st, tile_cols, cur, bk, bn);
                if (lane < nb) atomicAdd(&tcnt[bk], (unsigned long long)bn | (1ull << 32));
                if (nb < PS_WAVE) break;
            }
        }
        all_spans += nspan;
        all_pieces += npieces;
        all_frags += nfrag;
    }
    ps_wave_add(cnt, PS_ST_SPANS, all_spans, true);
    ps_wave_add(cnt, PS_ST_SEGS, all_pieces, true);
    ps_wave_add(cnt, PS_ST_FRAGS, all_frags, true);
}

// Spans at least this long keep varying checkpoints at each later piece, so
// that shading never replays more than about this many pixels; shorter spans
// are replayed from x1.
#define PS_CK_MIN 48
__device__ __forceinline__ bool ps_span_has_ck(const PsSpan& sp) { return sp.x2 - sp.x1 + 1 >= PS_CK_MIN; }

// ---------------------------------------------------------------------------
// segments: (bin key, span) pairs
// ---------------------------------------------------------------------------
// One thread per span: the span's pieces per bin chunk, in order.  The walk
// from x1 applies the serial `+= step` of the pixel loop (interp.cpp:82-92)
// pixel by pixel, so each piece starts with exactly the accumulators the
// reference holds at its first pixel; spans of colour-writing draws that
// cross chunks also keep the varying accumulators at each later piece.
// Placement: the wave visits its tiles in increasing order, SG_BATCH at a
// time.  Per tile, each lane's pieces go after those of the lanes before it
// (offsets kept in LDS), and lane j reserves the j-th tile's run of
// consecutive pieces with one atomic on the tile's counter (counting down
// from k_span's count), all lanes at once.  Then every lane walks its pieces
// of the batch's tiles into their slots; a span's tiles are consecutive in
// the visit order, so each walk goes on from tile to tile.
#define SG_BATCH 16
// NVC: room for the flush's widest varying checkpoint (at least its nvs
// floats); the accumulators live in registers across the walk, so the
// narrowest instantiation that fits keeps more waves resident (a 1080p
// Blinn-Phong flush: 117 VGPRs and four waves per SIMD at NVC 14 against
// 149 and three at 20, pieces 29.2 -> 24.8 us).  Up to NVC 14 the kernel
// fits 96 VGPRs without spilling, five waves per SIMD (23.6 -> 22.4 us);
// NVC 20 keeps its registers (a bound of four waves spilled).
template <int NVC>
__global__ void __launch_bounds__(256, NVC <= 14 ? 5 : 1) k_segments(const PsSpan* __restrict__ spans, const uint32_t* __restrict__ row_off,
                                                  int64_t T, PsFrame f, const PsTri* __restrict__ tris,
                                                  const PsDraw* __restrict__ draws, const float* __restrict__ vary, int nvs,
                                                  unsigned long long* __restrict__ tcnt, PsSeg* __restrict__ segs,
                                                  float* __restrict__ ck, int64_t cap, PsCounters* cnt)
{
    __shared__ uint32_t rel_all[4][SG_BATCH][PS_WAVE];   // per wave: [iteration][lane] offset in the run
    const bool poisoned = ps_poisoned(cnt);
    const int64_t R = row_off[T];
    const int64_t S = (uint32_t)f.tstart[f.n_tiles];
    if (poisoned) return;
    if (S > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&cnt->overflow, 1);
        return;
    }
    const int CW = f.chunk_w, TC = f.tile_cols;
    int64_t tb1 = 0;   // the only view of a one-view flush, read once
    int ntx1 = 0;
    if (f.n_views == 1) { tb1 = f.views[0].tile_base; ntx1 = f.views[0].n_tiles_x; }
    const int lane = threadIdx.x & (PS_WAVE - 1);
    uint32_t (*rel)[PS_WAVE] = rel_all[threadIdx.x >> 6];
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < R; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const uint32_t wave_id = (uint32_t)((base + (threadIdx.x & ~(PS_WAVE - 1))) >> 6);
        PsSpan sp;
        sp.tri = -1;
        if (i < R) sp = spans[i];
        const bool has = sp.tri >= 0;
        PsSpanTiles st = {};
        int nv = 0, y = 0;
        uint32_t flags = 0;
        bool need_ck = false;
        float vs[NVC], vt[NVC];
        if (has) {
            // the triangle only for a flush of several views (its view) or
            // for varying checkpoints
            PsTri tr;
            int64_t vtb = tb1;
            int vntx = ntx1;
            if (f.n_views != 1) {
                tr = tris[sp.tri];
                const PsView& v = f.views[tr.view];
                vtb = v.tile_base;
                vntx = v.n_tiles_x;
            }
            y = sp.y & 0xffffff;
            flags = ((uint32_t)(sp.y >> 24) & 3u) << 26 | ((uint32_t)(sp.y >> 28) & 1u) << 28;
            st = ps_span_tiles(vtb, vntx, f.tile_rows, TC, CW, y, sp.x1, sp.x2);
            need_ck = (st.cx1 > st.cx0) && ((sp.y >> 28) & 1) && ps_span_has_ck(sp);
            if (need_ck) {
                if (f.n_views == 1) tr = tris[sp.tri];
                const PsDraw& d = draws[tr.code >> 8];
                nv = d.nv;
                PsEdge L, Rr;
                ps_row_edges(tr, y, L, Rr);
                int left = ps_cvt_i32(ps_edge_eval(L, (float)y) + 0.5f);
                int right = ps_cvt_i32(ps_edge_eval(Rr, (float)y) + 0.5f);
                PsSpanGeom g;
                ps_span_geom(tr, y, left, right, L, Rr, g);
                const float* u0 = vary + (int64_t)ps_vary_index(d, sp.tri) * 3 * nvs;
                ps_span_varyings_n<NVC>(g, u0, u0 + nvs, u0 + 2 * nvs, vs, vt, nv);
            }
        }
        float cf2 = sp.cf2, pz = sp.pz;
        int x = sp.x1;
        int cx = st.cx0;                     // the lane's next chunk
        uint32_t cur = has ? st.g0 : 0xffffffffu;
        for (;;) {
            // the batch's tiles: lane j takes the j-th (bk) and its pieces (bn);
            // each lane notes its offset in the runs of its tiles, and which
            // iterations those are (consecutive: jt0 .. jt0 + ntb - 1)
            uint32_t bk = 0xffffffffu, bn = 0;
            int nb = 0, jt0 = -1, ntb = 0;
            for (; nb < SG_BATCH; nb++) {
                const uint32_t k = ps_wave_min(cur);
                if (k == 0xffffffffu) break;
                const bool m = cur == k;
                int lo = 0, hi = -1;
                if (m) ps_span_tile_chunks(st, TC, k, lo, hi);
                const uint32_t n = (uint32_t)(hi - lo + 1);
                const uint32_t incl = ps_wave_incl_scan(n);
                const uint32_t tot = __builtin_amdgcn_readlane(incl, 63);
                if (lane == nb) { bk = k; bn = tot; }
                if (m) {
                    rel[nb][lane] = incl - n;
                    if (jt0 < 0) jt0 = nb;
                    ntb++;
                    cur = k < st.g1 ? k + 1 : 0xffffffffu;
                }
            }
            if (nb == 0) break;
            uint32_t first = 0xffffffffu;
            if (lane < nb) {
                const unsigned long long take = (unsigned long long)bn | (1ull << 32);
                const unsigned long long left = atomicAdd(&tcnt[bk], (unsigned long long)0 - take);
                const unsigned long long end = PS_TS(f.tstart[bk + 1]);
                const uint32_t fp = (uint32_t)end - (uint32_t)left;
                const uint32_t slot = (uint32_t)(end >> 32) - (uint32_t)(left >> 32);
                // the counts k_span left must cover this run (they always do:
                // both kernels visit the same tiles); never write outside
                if ((uint32_t)left < bn || (uint32_t)(left >> 32) == 0u || (uint64_t)fp + bn > (uint64_t)cap) {
                    atomicOr(&cnt->overflow, 2);
                } else {
                    first = fp;
                    f.runs[2 * (size_t)slot] = fp;
                    f.runs[2 * (size_t)slot + 1] = (wave_id << 12) | (bn - 1u);
                }
            }
            // every lane walks its pieces of the batch's tiles
            const int maxtb = __builtin_amdgcn_readfirstlane(ps_wave_max((uint32_t)ntb));
            for (int q = 0; q < maxtb; q++) {
                const int jj = jt0 + q;
                const uint32_t run0 = __shfl(first, jj < 0 ? 0 : jj, PS_WAVE);
                if (q < ntb && run0 != 0xffffffffu) {
                    uint32_t s = run0 + rel[jj][lane];
                    const int hi = min(st.cx1, cx | (TC - 1));   // the tile's last chunk of the span (cx >= 0, TC a power of two)
                    for (; cx <= hi; cx++, s++) {
                        const int xa = max(sp.x1, cx * CW), xb = min(sp.x2, cx * CW + CW - 1);
                        for (; x < xa; x++) {
                            cf2 += sp.cf2_step;
                            pz += sp.pz_step;
                            if (need_ck) {
#pragma unroll
                                for (int k = 0; k < NVC; k++)
                                    if (k < nv) vs[k] = vs[k] + vt[k];
                            }
                        }
                        PsSeg sg;
                        sg.si = (uint32_t)i | flags;
                        sg.xab = (uint32_t)xa | ((uint32_t)xb << 16);
                        sg.cf2 = cf2;
                        sg.pz = pz;
                        sg.cst = sp.cf2_step;
                        sg.pst = sp.pz_step;
                        sg.y = y;
                        sg.win = need_ck && xa > sp.x1 ? (s | PS_SEG_CK) : (uint32_t)sp.tri;
                        segs[s] = sg;
                        if (need_ck && xa > sp.x1) {
                            float* c = ck + (int64_t)s * f.ck_stride;
#pragma unroll
                            for (int k = 0; k < NVC; k++)
                                if (k < nv) c[k] = vs[k];
                        }
                    }
                }
            }
            if (nb < SG_BATCH) break;
        }
    }
}

// ---------------------------------------------------------------------------
// tile lists: the runs of a tile in submission order
// ---------------------------------------------------------------------------
// bitonic sort of one (key, value) pair per lane over lanes [0, np), np a
// power of two up to 64, ascending by key
__device__ __forceinline__ void ps_sort_lanes(uint32_t& key, uint32_t& val, int np)
{
    const int lane = threadIdx.x & 63;
    for (int k = 2; k <= np; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t ok = __shfl_xor(key, j, 64), ov = __shfl_xor(val, j, 64);
            const bool up = (lane & k) == 0, low = (lane & j) == 0;
            if (low == up ? ok < key : ok > key) { key = ok; val = ov; }
        }
    }
}

// The runs of a tile are written in the order their atomics happened.  Ordered
// by wave they give the tile's pieces in submission order: wave order is span
// order, and inside a run the pieces are in span order.
//
// Window of up to 64 runs from run w0 on, one per lane: rbase = the run's
// first piece, rv = the list position of that piece (0xffffffff past the last
// run), vhi = the list position where the window ends, np = the smallest
// power of two at least the window's runs.  A tile of at most 64 runs is
// sorted here; longer lists come ordered by ps_order_runs, with each run's
// key replaced by its list position.
__device__ __forceinline__ void ps_run_window(const uint32_t* rl, uint32_t nr, uint32_t w0, uint32_t n_list,
                                              uint32_t& rbase, uint32_t& rv, uint32_t& vhi, int& np)
{
    const uint32_t j = w0 + (threadIdx.x & 63);
    const uint32_t nw = min(nr - w0, 64u);
    np = 1;
    while ((uint32_t)np < nw) np <<= 1;
    if (nr <= 64) {
        uint32_t key = j < nr ? rl[2 * j + 1] : 0xffffffffu;
        rbase = j < nr ? rl[2 * j] : 0u;
        ps_sort_lanes(key, rbase, np);
        const uint32_t c = j < nr ? (key & 0xfffu) + 1u : 0u;
        const uint32_t incl = ps_wave_incl_scan(c);
        rv = j < nr ? incl - c : 0xffffffffu;
        vhi = n_list;
    } else {
        rbase = j < nr ? rl[2 * j] : 0u;
        rv = j < nr ? rl[2 * j + 1] : 0xffffffffu;
        vhi = w0 + 64 < nr ? rl[2 * (w0 + 64) + 1] : n_list;
    }
}

// the piece at list position v of the window (rv[0] <= v): the last run that
// starts at or before v, found by a binary search over lanes [0, np)
__device__ __forceinline__ uint32_t ps_run_piece(uint32_t rv, uint32_t rbase, uint32_t v, int np)
{
    if (np == 1) return (uint32_t)__builtin_amdgcn_readlane(rbase, 0) + (v - (uint32_t)__builtin_amdgcn_readlane(rv, 0));
    int pos = 0;
    for (int st = np >> 1; st >= 1; st >>= 1)
        if (__shfl(rv, pos + st, 64) <= v) pos += st;
    return __shfl(rbase, pos, 64) + (v - __shfl(rv, pos, 64));
}

// A tile with more than 64 runs: one wave sorts them stably by wave (LSD
// radix, 7-bit digits of the wave index above the smallest, ping-ponging
// between the tile's parts a and b of the two run arrays), then replaces each
// run's key by the list position of its first piece.  cnt128: 128 words of
// the wave's LDS.  Returns the array that holds the result.
__device__ uint32_t* ps_order_runs(uint32_t* a, uint32_t* b, uint32_t nr, uint32_t* cnt128)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = ps_lt_mask();
    uint32_t mn = 0xffffffffu, mx = 0;
    for (uint32_t j = lane; j < nr; j += 64) {
        const uint32_t w = a[2 * j + 1] >> 12;
        mn = min(mn, w);
        mx = max(mx, w);
    }
    mn = ps_wave_min(mn);
    mx = ps_wave_max(mx);
    const int bits = mx > mn ? 32 - __clz(mx - mn) : 0;
    for (int shift = 0; shift < bits; shift += 7) {
        cnt128[lane] = 0;
        cnt128[lane + 64] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t j = lane; j < nr; j += 64)
            atomicAdd(&cnt128[(((a[2 * j + 1] >> 12) - mn) >> shift) & 127u], 1u);
        __builtin_amdgcn_wave_barrier();
        {
            const uint32_t c0 = cnt128[2 * lane], c1 = cnt128[2 * lane + 1];
            const uint32_t ex = ps_wave_incl_scan(c0 + c1) - c0 - c1;
            __builtin_amdgcn_wave_barrier();
            cnt128[2 * lane] = ex;
            cnt128[2 * lane + 1] = ex + c0;
            __builtin_amdgcn_wave_barrier();
        }
        for (uint32_t j0 = 0; j0 < nr; j0 += 64) {
            const uint32_t j = j0 + lane;
            const bool valid = j < nr;
            const uint32_t e0 = valid ? a[2 * j] : 0u, e1 = valid ? a[2 * j + 1] : 0u;
            const uint32_t dig = valid ? (((e1 >> 12) - mn) >> shift) & 127u : 0u;
            uint64_t same = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 7; bit++) {
                const uint64_t m = __ballot((dig >> bit) & 1u);
                same &= ((dig >> bit) & 1u) ? m : ~m;
            }
            const uint32_t before = (uint32_t)__popcll(same & lt);
            const uint32_t cur = cnt128[dig];
            __builtin_amdgcn_wave_barrier();
            if (valid) {
                b[2 * (cur + before)] = e0;
                b[2 * (cur + before) + 1] = e1;
                if (before == 0) cnt128[dig] = cur + (uint32_t)__popcll(same);
            }
            __builtin_amdgcn_wave_barrier();
        }
        __threadfence();   // the next pass reads what this one wrote
        uint32_t* t = a; a = b; b = t;
    }
    uint32_t carry = 0;
    for (uint32_t j0 = 0; j0 < nr; j0 += 64) {
        const uint32_t j = j0 + lane;
        const uint32_t c = j < nr ? (a[2 * j + 1] & 0xfffu) + 1u : 0u;
        const uint32_t incl = ps_wave_incl_scan(c);
        if (j < nr) a[2 * j + 1] = carry + incl - c;
        carry += __builtin_amdgcn_readlane(incl, 63);
    }
    __threadfence();
    return a;
}

// ---------------------------------------------------------------------------
// raster: ordered depth test per bin
// ---------------------------------------------------------------------------
#define RA_CHUNK 128    // tile-list entries staged in LDS per pass
#define RA_PER (RA_CHUNK / 64)

// LDS of one wave: depth and winner columns [c][lane] (bank = lane), then the
// staged pieces of the current chunk (SoA, in per-lane order), then per-lane
// counts and offsets
#define RA_WORDS(cw) ((cw) * 64 * 2 + RA_CHUNK * 7 + 128)

// KCW: the chunk width when it is 1, 2, 4 or 8, else 0.  With a known width
// each lane keeps its pixels' depth and winner in registers, so the walk has
// no LDS round trip per pixel; the LDS columns serve wider chunks.
template <bool DEBUG, int KCW>
__global__ void __launch_bounds__(512, 5) k_raster(const PsSpan* __restrict__ spans, const PsSeg* __restrict__ segs,
                         const PsTri* __restrict__ tris, const PsDraw* __restrict__ draws, PsFrame f, int64_t n_tiles,
                         PsCounters* cnt)
{
    extern __shared__ uint32_t lds[];
    const bool poisoned = ps_poisoned(cnt);
    const int CW = KCW ? KCW : f.chunk_w;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int waves = blockDim.x >> 6;
    // neighbouring tiles share an L2; with a heavy-tile list, the first
    // PS_HEAVY_CAP waves take the listed tiles (they start first, so the
    // longest walks do not end the kernel), the others every unlisted tile
    // (the list and the marks are in bounds, stale or not)
    int64_t tile = ps_xcd_remap<4>(blockIdx.x, gridDim.x) * (int64_t)waves + wave;
    bool skip = false;
    if (f.heavy) {
        if (tile < PS_HEAVY_CAP) {
            const uint32_t listed = min(f.heavy[0], (uint32_t)PS_HEAVY_CAP);
            const uint32_t t = f.heavy[1 + tile];
            skip = tile >= listed;
            tile = t;
        } else {
            tile -= PS_HEAVY_CAP;
            skip = tile >= n_tiles || (f.tstart[min(tile, n_tiles)] >> 63);
        }
    }
    if (poisoned) return;
    ps_commit_pending(cnt);
    if (skip || tile >= n_tiles) return;
    uint32_t* wl = lds + (size_t)wave * RA_WORDS(KCW ? 0 : CW);
    float* dep = (float*)wl;                       // LDS columns: chunks wider than 8 only
    int* win = (int*)(dep + (KCW ? 0 : CW) * 64);
    uint32_t* r_s = (uint32_t*)(win + (KCW ? 0 : CW) * 64);   // winner code (triangle, or checkpoint | piece)
    uint32_t* r_fl = r_s + RA_CHUNK;               // behaviour | writes colour (PsSeg.si high bits)
    uint32_t* r_x = r_fl + RA_CHUNK;               // xa | xb << 16
    float* r_cf2 = (float*)(r_x + RA_CHUNK);
    float* r_pz = r_cf2 + RA_CHUNK;
    float* r_cst = r_pz + RA_CHUNK;
    float* r_pst = r_cst + RA_CHUNK;
    uint32_t* lcnt = (uint32_t*)(r_pst + RA_CHUNK);
    uint32_t* loff = lcnt + 64;

    const int cw_shift = __builtin_ctz((unsigned)CW), tc_shift = __builtin_ctz((unsigned)f.tile_cols);
    const PsView& V = f.views[ps_view_of_tile(f, tile)];
    int64_t ltile = tile - V.tile_base;
    // a view has fewer than 2^31 tiles; tile_cols is a power of two
    const uint32_t lt32 = (uint32_t)ltile, ntx = (uint32_t)V.n_tiles_x;
    int ty = (int)(lt32 / ntx), tx = (int)(lt32 - (uint32_t)ty * ntx);
    int ry = lane >> tc_shift, rx = lane & (f.tile_cols - 1);
    int y = ty * f.tile_rows + ry;
    int cx = tx * f.tile_cols + rx;
    int X0 = cx * CW;
    const int VW = V.W;
    bool inside = y < V.H && X0 < VW;
    int ncol = inside ? min(CW, VW - X0) : 0;
    float* drow = V.depth + (int64_t)y * V.depth_scan;
    // a clear fused into the flush: the tile starts from its value, and the
    // whole tile is written back
    const bool dclear = V.dclear != 0;
    float dv[KCW ? KCW : 1];
    int wv[KCW ? KCW : 1];
    if constexpr (KCW > 0) {
#pragma unroll
        for (int k = 0; k < KCW; k++) {
            dv[k] = k < ncol ? (dclear ? V.dclear_value : drow[X0 + k]) : 1.0f;
            wv[k] = -1;
        }
    } else {
        for (int c = 0; c < ncol; c++) {
            dep[c * 64 + lane] = dclear ? V.dclear_value : drow[X0 + c];
            win[c * 64 + lane] = -1;
        }
    }
    unsigned long long passes = 0;
    bool dirty = dclear;
    const uint64_t lt = ps_lt_mask();
    // the tile's runs in submission order, up to 64 at a time (one per lane)
    const unsigned long long ta = PS_TS(f.tstart[tile]), tb = PS_TS(f.tstart[tile + 1]);
    const uint32_t n_list = (uint32_t)tb - (uint32_t)ta;
    const uint32_t r0 = (uint32_t)(ta >> 32), nr = (uint32_t)(tb >> 32) - r0;
    const uint32_t* rl = f.runs + 2 * (size_t)r0;
    if (nr > 64) rl = ps_order_runs(f.runs + 2 * (size_t)r0, f.runs2 + 2 * (size_t)r0, nr, lcnt);
    for (uint32_t w0 = 0; w0 < nr; w0 += 64) {
        uint32_t rbase, rv, vhi;
        int np;
        ps_run_window(rl, nr, w0, n_list, rbase, rv, vhi, np);
        const uint32_t vlo = __builtin_amdgcn_readlane(rv, 0);
        // software pipeline over chunks of the window: while chunk i is split
        // and walked, the records of chunk i+1 are in flight
        uint32_t vn[RA_PER];
        uint4 ran[RA_PER], rbn[RA_PER];
#pragma unroll
        for (int k = 0; k < RA_PER; k++) {
            uint32_t e = vlo + k * 64 + lane;
            uint32_t pp = ps_run_piece(rv, rbase, e, np);
            vn[k] = e < vhi ? pp : 0u;
            const uint4* p = (const uint4*)(segs + vn[k]);
            ran[k] = e < vhi ? p[0] : make_uint4(0, 0, 0, 0);
            rbn[k] = e < vhi ? p[1] : make_uint4(0, 0, 0, 0);
        }
        for (uint32_t cb = vlo; cb < vhi; cb += RA_CHUNK) {
            uint32_t ce = min(vhi, cb + RA_CHUNK);
            uint32_t v[RA_PER], key[RA_PER];
            uint4 ra[RA_PER], rb[RA_PER];
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                v[k] = vn[k];
                ra[k] = ran[k];
                rb[k] = rbn[k];
            }
            if (cb + RA_CHUNK < vhi) {
#pragma unroll
                for (int k = 0; k < RA_PER; k++) {
                    uint32_t e = cb + RA_CHUNK + k * 64 + lane;
                    uint32_t pp = ps_run_piece(rv, rbase, e, np);
                    vn[k] = e < vhi ? pp : 0u;
                    const uint4* p = (const uint4*)(segs + vn[k]);
                    ran[k] = e < vhi ? p[0] : make_uint4(0, 0, 0, 0);
                    rbn[k] = e < vhi ? p[1] : make_uint4(0, 0, 0, 0);
                }
            }
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                // the piece's lane: its row inside the tile and its chunk inside the row
                int py = (int)rb[k].z, pcx = (int)(ra[k].y & 0xffffu) >> cw_shift;
                key[k] = (uint32_t)(((py & (f.tile_rows - 1)) << tc_shift) | (pcx & (f.tile_cols - 1)));
            }
            // stable split by lane key: each entry's rank among the equal keys
            // before it (the same-key lanes of its group of 64 by ballots, plus
            // the count of earlier groups), then per-key offsets, then placement
            lcnt[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            uint32_t rank[RA_PER];
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                uint32_t e = cb + k * 64 + lane;
                bool valid = e < ce;
                uint32_t kk = key[k];
                uint64_t same = __ballot(valid);
#pragma unroll
                for (int bit = 0; bit < 6; bit++) {
                    uint64_t m = __ballot((kk >> bit) & 1u);
                    same &= ((kk >> bit) & 1u) ? m : ~m;
                }
                uint32_t before = (uint32_t)__popcll(same & lt);
                uint32_t cur = lcnt[kk];
                rank[k] = cur + before;
                __builtin_amdgcn_wave_barrier();
                if (valid && before == 0) lcnt[kk] = cur + (uint32_t)__popcll(same);
                __builtin_amdgcn_wave_barrier();
            }
            {
                uint32_t cntv = lcnt[lane];
                uint32_t x = cntv;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    uint32_t yv = __shfl_up(x, o, 64);
                    if (lane >= o) x += yv;
                }
                loff[lane] = x - cntv;
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                uint32_t e = cb + k * 64 + lane;
                if (e < ce) {
                    uint32_t q = loff[key[k]] + rank[k];
                    r_s[q] = rb[k].w;
                    r_fl[q] = ra[k].x & ~PS_SEG_SPAN_MASK;
                    r_x[q] = ra[k].y;
                    r_cf2[q] = __uint_as_float(ra[k].z);
                    r_pz[q] = __uint_as_float(ra[k].w);
                    r_cst[q] = __uint_as_float(rb[k].x);
                    r_pst[q] = __uint_as_float(rb[k].y);
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (inside) {
                uint32_t l0 = loff[lane], l1 = l0 + lcnt[lane];
                for (uint32_t q = l0; q < l1; q++) {
                    const uint32_t wcode = r_s[q];
                    uint32_t fl = r_fl[q];
                    int beh = (fl >> 26) & 3;
                    bool wc = (fl >> 28) & 1;
                    uint32_t xab = r_x[q];
                    int a = (int)(xab & 0xffffu), b = (int)(xab >> 16);
                    float cf2 = r_cf2[q], pz = r_pz[q];
                    const float cst = r_cst[q], pst = r_pst[q];
                    uint32_t th = 0;
                    int tid = 0;
                    if (DEBUG) {
                        int tri = (wcode & PS_SEG_CK) ? spans[segs[wcode & ~PS_SEG_CK].si & PS_SEG_SPAN_MASK].tri : (int)wcode;
                        const PsTri tr = tris[tri];
                        const PsDraw& d = draws[tr.code >> 8];
                        tid = (int)(d.id_base + (tri - d.tri_start));
                        uint32_t h = (uint32_t)tid;
                        h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
                        th = h;
                    }
                    if constexpr (KCW > 0) {
                        // the chunk's pixels in order, those of [a, b] stepping the accumulators
#pragma unroll
                        for (int k = 0; k < KCW; k++) {
                            const int x = X0 + k;
                            if (x < a || x > b) continue;
                            // interpolateNextStep (interp.cpp:82-92)
                            float r = 1.0f / cf2;
                            cf2 += cst;
                            float z = pz * r;
                            pz += pst;
                            float cur = (beh & PSGPU_TEST_DEPTH) ? dv[k] : 1.0f;
                            int64_t pix = (int64_t)y * VW + x;
                            if (DEBUG) { f.dbg_rcount[pix] += 1; f.dbg_rhash[pix] ^= th; }
                            // fragthrd.cpp:227
                            if (-1.0f < z && (z - cur < -0.0001f)) {
                                passes++;
                                if (beh & PSGPU_UPDATE_DEPTH) { dv[k] = z; dirty = true; }
                                if (wc) wv[k] = (int)wcode;
                                if (DEBUG) { f.dbg_pcount[pix] += 1; f.dbg_phash[pix] ^= th; f.dbg_last[pix] = tid; }
                            }
                        }
                    } else {
                        for (int x = a; x <= b; x++) {
                            // interpolateNextStep (interp.cpp:82-92)
                            float r = 1.0f / cf2;
                            cf2 += cst;
                            float z = pz * r;
                            pz += pst;
                            int c = x - X0;
                            float cur = (beh & PSGPU_TEST_DEPTH) ? dep[c * 64 + lane] : 1.0f;
                            int64_t pix = (int64_t)y * VW + x;
                            if (DEBUG) { f.dbg_rcount[pix] += 1; f.dbg_rhash[pix] ^= th; }
                            // fragthrd.cpp:227
                            if (-1.0f < z && (z - cur < -0.0001f)) {
                                passes++;
                                if (beh & PSGPU_UPDATE_DEPTH) { dep[c * 64 + lane] = z; dirty = true; }
                                if (wc) win[c * 64 + lane] = (int)wcode;
                                if (DEBUG) { f.dbg_pcount[pix] += 1; f.dbg_phash[pix] ^= th; f.dbg_last[pix] = tid; }
                            }
                        }
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    int* wrow = V.winner + (int64_t)y * VW;
    if constexpr (KCW > 0) {
#pragma unroll
        for (int k = 0; k < KCW; k++)
            if (k < ncol) {
                if (dirty) drow[X0 + k] = dv[k];
                wrow[X0 + k] = wv[k];
            }
    } else {
        for (int c = 0; c < ncol; c++) {
            if (dirty) drow[X0 + c] = dep[c * 64 + lane];
            wrow[X0 + c] = win[c * 64 + lane];
        }
    }
    ps_wave_add64(cnt, PS_ST_SHADED, passes);
}

// ---------------------------------------------------------------------------
// raster with the FP in the ordered walk: programmes that blend or discard
// ---------------------------------------------------------------------------
// A discarded fragment leaves the depth unchanged and a blended one depends
// on the colour beneath it, so neither can be deferred to one shading pass
// per pixel.  The same tile walk runs the programme's FP on every depth-test
// pass, in submission order, and writes (or blends) the colour at once.
template <int PROG, bool DEBUG>
__global__ void __launch_bounds__(512) k_raster_imm(const PsSpan* __restrict__ spans, const PsSeg* __restrict__ segs,
                         const PsTri* __restrict__ tris, const PsDraw* __restrict__ draws, PsFrame f, int64_t n_tiles,
                         PsCounters* cnt, const float* __restrict__ vary, int nvs, const float* __restrict__ ck)
{
    typedef PsProc<PROG> P;
    extern __shared__ uint32_t lds[];
    if (ps_poisoned(cnt)) return;
    ps_commit_pending(cnt);
    const int CW = f.chunk_w;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int waves = blockDim.x >> 6;
    int64_t tile = ps_xcd_remap<4>(blockIdx.x, gridDim.x) * (int64_t)waves + wave;   // neighbouring tiles share an L2
    if (tile >= n_tiles) return;
    uint32_t* wl = lds + (size_t)wave * RA_WORDS(CW);
    float* dep = (float*)wl;
    int* win = (int*)(dep + CW * 64);
    uint32_t* r_s = (uint32_t*)(win + CW * 64);   // segment index
    uint32_t* r_fl = r_s + RA_CHUNK;               // behaviour | writes colour (PsSeg.si high bits)
    uint32_t* r_x = r_fl + RA_CHUNK;               // xa | xb << 16
    float* r_cf2 = (float*)(r_x + RA_CHUNK);
    float* r_pz = r_cf2 + RA_CHUNK;
    float* r_cst = r_pz + RA_CHUNK;
    float* r_pst = r_cst + RA_CHUNK;
    uint32_t* lcnt = (uint32_t*)(r_pst + RA_CHUNK);
    uint32_t* loff = lcnt + 64;

    const int cw_shift = __builtin_ctz((unsigned)CW), tc_shift = __builtin_ctz((unsigned)f.tile_cols);
    const PsView& V = f.views[ps_view_of_tile(f, tile)];
    int64_t ltile = tile - V.tile_base;
    // a view has fewer than 2^31 tiles; tile_cols is a power of two
    const uint32_t lt32 = (uint32_t)ltile, ntx = (uint32_t)V.n_tiles_x;
    int ty = (int)(lt32 / ntx), tx = (int)(lt32 - (uint32_t)ty * ntx);
    int ry = lane >> tc_shift, rx = lane & (f.tile_cols - 1);
    int y = ty * f.tile_rows + ry;
    int cx = tx * f.tile_cols + rx;
    int X0 = cx * CW;
    const int VW = V.W;
    bool inside = y < V.H && X0 < VW;
    int ncol = inside ? min(CW, VW - X0) : 0;
    float* drow = V.depth + (int64_t)y * V.depth_scan;
    // a clear fused into the flush: the tile starts from its value, and the
    // whole tile is written back
    const bool dclear = V.dclear != 0;
    for (int c = 0; c < ncol; c++) {
        dep[c * 64 + lane] = dclear ? V.dclear_value : drow[X0 + c];
        win[c * 64 + lane] = -1;
    }
    unsigned long long passes = 0;
    bool dirty = dclear;
    const uint64_t lt = ps_lt_mask();
    // the tile's runs in submission order, up to 64 at a time (one per lane)
    const unsigned long long ta = PS_TS(f.tstart[tile]), tb = PS_TS(f.tstart[tile + 1]);
    const uint32_t n_list = (uint32_t)tb - (uint32_t)ta;
    const uint32_t r0 = (uint32_t)(ta >> 32), nr = (uint32_t)(tb >> 32) - r0;
    const uint32_t* rl = f.runs + 2 * (size_t)r0;
    if (nr > 64) rl = ps_order_runs(f.runs + 2 * (size_t)r0, f.runs2 + 2 * (size_t)r0, nr, lcnt);
    for (uint32_t w0 = 0; w0 < nr; w0 += 64) {
        uint32_t rbase, rv, vhi;
        int np;
        ps_run_window(rl, nr, w0, n_list, rbase, rv, vhi, np);
        const uint32_t vlo = __builtin_amdgcn_readlane(rv, 0);
        // software pipeline over chunks of the window: while chunk i is split
        // and walked, the records of chunk i+1 are in flight
        uint32_t vn[RA_PER];
        uint4 ran[RA_PER], rbn[RA_PER];
#pragma unroll
        for (int k = 0; k < RA_PER; k++) {
            uint32_t e = vlo + k * 64 + lane;
            uint32_t pp = ps_run_piece(rv, rbase, e, np);
            vn[k] = e < vhi ? pp : 0u;
            const uint4* p = (const uint4*)(segs + vn[k]);
            ran[k] = e < vhi ? p[0] : make_uint4(0, 0, 0, 0);
            rbn[k] = e < vhi ? p[1] : make_uint4(0, 0, 0, 0);
        }
        for (uint32_t cb = vlo; cb < vhi; cb += RA_CHUNK) {
            uint32_t ce = min(vhi, cb + RA_CHUNK);
            uint32_t v[RA_PER], key[RA_PER];
            uint4 ra[RA_PER], rb[RA_PER];
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                v[k] = vn[k];
                ra[k] = ran[k];
                rb[k] = rbn[k];
            }
            if (cb + RA_CHUNK < vhi) {
#pragma unroll
                for (int k = 0; k < RA_PER; k++) {
                    uint32_t e = cb + RA_CHUNK + k * 64 + lane;
                    uint32_t pp = ps_run_piece(rv, rbase, e, np);
                    vn[k] = e < vhi ? pp : 0u;
                    const uint4* p = (const uint4*)(segs + vn[k]);
                    ran[k] = e < vhi ? p[0] : make_uint4(0, 0, 0, 0);
                    rbn[k] = e < vhi ? p[1] : make_uint4(0, 0, 0, 0);
                }
            }
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                // the piece's lane: its row inside the tile and its chunk inside the row
                int py = (int)rb[k].z, pcx = (int)(ra[k].y & 0xffffu) >> cw_shift;
                key[k] = (uint32_t)(((py & (f.tile_rows - 1)) << tc_shift) | (pcx & (f.tile_cols - 1)));
            }
            // stable split by lane key: each entry's rank among the equal keys
            // before it (the same-key lanes of its group of 64 by ballots, plus
            // the count of earlier groups), then per-key offsets, then placement
            lcnt[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            uint32_t rank[RA_PER];
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                uint32_t e = cb + k * 64 + lane;
                bool valid = e < ce;
                uint32_t kk = key[k];
                uint64_t same = __ballot(valid);
#pragma unroll
                for (int bit = 0; bit < 6; bit++) {
                    uint64_t m = __ballot((kk >> bit) & 1u);
                    same &= ((kk >> bit) & 1u) ? m : ~m;
                }
                uint32_t before = (uint32_t)__popcll(same & lt);
                uint32_t cur = lcnt[kk];
                rank[k] = cur + before;
                __builtin_amdgcn_wave_barrier();
                if (valid && before == 0) lcnt[kk] = cur + (uint32_t)__popcll(same);
                __builtin_amdgcn_wave_barrier();
            }
            {
                uint32_t cntv = lcnt[lane];
                uint32_t x = cntv;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    uint32_t yv = __shfl_up(x, o, 64);
                    if (lane >= o) x += yv;
                }
                loff[lane] = x - cntv;
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int k = 0; k < RA_PER; k++) {
                uint32_t e = cb + k * 64 + lane;
                if (e < ce) {
                    uint32_t q = loff[key[k]] + rank[k];
                    r_s[q] = v[k];
                    r_fl[q] = ra[k].x & ~PS_SEG_SPAN_MASK;
                    r_x[q] = ra[k].y;
                    r_cf2[q] = __uint_as_float(ra[k].z);
                    r_pz[q] = __uint_as_float(ra[k].w);
                    r_cst[q] = __uint_as_float(rb[k].x);
                    r_pst[q] = __uint_as_float(rb[k].y);
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (inside) {
                uint32_t l0 = loff[lane], l1 = l0 + lcnt[lane];
                for (uint32_t q = l0; q < l1; q++) {
                    int sidx = (int)r_s[q];
                    uint32_t fl = r_fl[q];
                    int beh = (fl >> 26) & 3;
                    bool wc = (fl >> 28) & 1;
                    uint32_t xab = r_x[q];
                    int a = (int)(xab & 0xffffu), b = (int)(xab >> 16);
                    float cf2 = r_cf2[q], pz = r_pz[q];
                    const float cst = r_cst[q], pst = r_pst[q];
                    uint32_t th = 0;
                    int tid = 0;
                    if (DEBUG) {
                        int si = (int)(segs[sidx].si & PS_SEG_SPAN_MASK);
                        int tri = spans[si].tri;
                        const PsTri tr = tris[tri];
                        const PsDraw& d = draws[tr.code >> 8];
                        tid = (int)(d.id_base + (tri - d.tri_start));
                        uint32_t h = (uint32_t)tid;
                        h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
                        th = h;
                    }
                    // the span's varyings at pixel a: from the piece's checkpoint or
                    // by the serial walk from x1 (interp.cpp:82-92), as the deferred
                    // shading does
                    const int si = (int)(segs[sidx].si & PS_SEG_SPAN_MASK);
                    const PsSpan sp = spans[si];
                    const PsTri tr = tris[sp.tri];
                    const PsDraw& d = draws[tr.code >> 8];
                    float vs[P::NV > 0 ? P::NV : 1], vt[P::NV > 0 ? P::NV : 1];
                    {
                        PsEdge Le, Re;
                        ps_row_edges(tr, y, Le, Re);
                        int left = ps_cvt_i32(ps_edge_eval(Le, (float)y) + 0.5f);
                        int right = ps_cvt_i32(ps_edge_eval(Re, (float)y) + 0.5f);
                        PsSpanGeom g;
                        ps_span_geom(tr, y, left, right, Le, Re, g);
                        const float* u0 = vary + (int64_t)ps_vary_index(d, sp.tri) * 3 * nvs;
                        ps_span_varyings<P::NV>(g, u0, u0 + nvs, u0 + 2 * nvs, vs, vt);
                        int xs = sp.x1;
                        if (a > sp.x1 && ps_span_has_ck(sp) && wc) {
                            const float* cp = ck + (int64_t)sidx * f.ck_stride;
#pragma unroll
                            for (int i = 0; i < P::NV; i++) vs[i] = cp[i];
                            xs = a;
                        }
                        for (int xx = xs; xx < a; xx++) {
#pragma unroll
                            for (int i = 0; i < P::NV; i++) vs[i] = vs[i] + vt[i];
                        }
                    }
                    const bool blend = (d.behavior & PSGPU_ALPHABLEND) != 0;
                    const int64_t crow = (V.colour_top_down ? (int64_t)(V.colour_h - 1 - y) : (int64_t)y) * V.colour_scan;
                    for (int x = a; x <= b; x++) {
                        // interpolateNextStep (interp.cpp:82-92)
                        float r = 1.0f / cf2;
                        cf2 += cst;
                        float z = pz * r;
                        pz += pst;
                        int c = x - X0;
                        float cur = (beh & PSGPU_TEST_DEPTH) ? dep[c * 64 + lane] : 1.0f;
                        int64_t pix = (int64_t)y * VW + x;
                        if (DEBUG) { f.dbg_rcount[pix] += 1; f.dbg_rhash[pix] ^= th; }
                        // fragthrd.cpp:227-236: the FP runs on a pass; a discarded
                        // fragment leaves depth as it was
                        if (-1.0f < z && (z - cur < -0.0001f)) {
                            passes++;
                            float in[P::NV > 0 ? P::NV : 1];
#pragma unroll
                            for (int i = 0; i < P::NV; i++) in[i] = vs[i] * r;
                            uint32_t col = 0;
                            int fr = P::fpi(d, in, col);
                            if (fr != 2 && (beh & PSGPU_UPDATE_DEPTH)) { dep[c * 64 + lane] = z; dirty = true; }
                            if (fr == 1 && wc) {   // FBOBridge::write4 (fragthrd.cpp:70-82)
                                uint32_t* cp = V.colour + crow + x;
                                *cp = blend ? ps_blend4(*cp, col) : col;
                            }
                            if (DEBUG) { f.dbg_pcount[pix] += 1; f.dbg_phash[pix] ^= th; f.dbg_last[pix] = tid; }
                        }
#pragma unroll
                        for (int i = 0; i < P::NV; i++) vs[i] = vs[i] + vt[i];
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    for (int c = 0; c < ncol; c++)
        if (dirty) drow[X0 + c] = dep[c * 64 + lane];
    ps_wave_add64(cnt, PS_ST_SHADED, passes);
}

// ---------------------------------------------------------------------------
// shade: deferred FP of the last colour-writing pass per pixel
// ---------------------------------------------------------------------------
// The winner plane names the triangle, so most pixels go winner -> triangle
// -> (draw, varyings) and re-derive the span's start (x1, cf2) from the
// triangle exactly as k_span did; only pixels of a piece with a varying
// checkpoint read that piece and its span first.
template <int PROG>
__global__ void __launch_bounds__(256, 7) k_shade(const PsSpan* __restrict__ spans, const PsSeg* __restrict__ segs,
                                               const float* __restrict__ ck, const PsTri* __restrict__ tris,
                                               const float* __restrict__ vary, int nvs, const float* __restrict__ grads,
                                               const PsDraw* __restrict__ draws, PsFrame f, const PsCounters* cnt, int clears)
{
    typedef PsProc<PROG> P;
    const bool poisoned = ps_poisoned(cnt);
    const int64_t lb = ps_xcd_remap<4>(blockIdx.x, gridDim.x);
    const PsView& V = f.views[ps_view_of_stile(f, lb)];
    int64_t st = lb - V.stile_base;
    // a view has fewer than 2^31 shading blocks
    const uint32_t st32 = (uint32_t)st, nsx = (uint32_t)V.n_stiles_x;
    int sy = (int)(st32 / nsx), sx = (int)(st32 - (uint32_t)sy * nsx);
    int x = sx * 16 + (threadIdx.x & 15);
    int y = sy * 16 + (threadIdx.x >> 4);
    if (x >= V.W || y >= V.H) return;
    const int64_t crow = V.colour_top_down ? (int64_t)(V.colour_h - 1 - y) : (int64_t)y;
    int w = V.winner[(int64_t)y * V.W + x];
    if (poisoned) return;
    if (w == -1) {   // no fragment colours the pixel: a fused clear of its row, if any, is all it gets
        if (clears && crow < V.cclear_rows) V.colour[crow * V.colour_scan + x] = V.cclear_value;
        return;
    }
    // a piece with a checkpoint: its first pixel xa carries the accumulators
    // as the serial walk from x1 leaves them
    const bool from_ck = (uint32_t)w & PS_SEG_CK;
    const uint32_t wi = (uint32_t)w & ~PS_SEG_CK;
    int tri = (int)wi, xa = 0;
    float cfa = 0.0f;
    if (from_ck) {
        const PsSeg sg = segs[wi];
        tri = spans[sg.si & PS_SEG_SPAN_MASK].tri;
        xa = (int)(sg.xab & 0xffffu);
        cfa = sg.cf2;
    }
    const PsTri tr = tris[tri];
    ps_with_draw(draws, (int)(tr.code >> 8), [&](const PsDraw& d) {
    if (d.prog != PROG) return;
    PsEdge L, R;
    ps_row_edges(tr, y, L, R);
    int left = ps_cvt_i32(ps_edge_eval(L, (float)y) + 0.5f);
    int right = ps_cvt_i32(ps_edge_eval(R, (float)y) + 0.5f);
    PsSpanGeom g;
    ps_span_geom(tr, y, left, right, L, R, g);
    float cf2, cf2_step, pz, pz_step;
    ps_span_vis(tr, g, cf2, cf2_step, pz, pz_step);
    const float* u0 = vary + (int64_t)ps_vary_index(d, tri) * 3 * nvs;
    float vs[P::NV > 0 ? P::NV : 1], vt[P::NV > 0 ? P::NV : 1];
    ps_span_varyings<P::NV>(g, u0, u0 + nvs, u0 + 2 * nvs, vs, vt);
    // from the span's first pixel (leftClamped), or from the checkpoint,
    // the serial walk on to the pixel
    int xs = left < 0 ? 0 : left;
    if (from_ck) {
        const float* c = ck + (int64_t)wi * f.ck_stride;
#pragma unroll
        for (int i = 0; i < P::NV; i++) vs[i] = c[i];
        xs = xa;
        cf2 = cfa;
    }
    for (int k = xs; k < x; k++) {
        cf2 += cf2_step;
#pragma unroll
        for (int i = 0; i < P::NV; i++) vs[i] = vs[i] + vt[i];
    }
    float r = 1.0f / cf2;
    float in[P::NV > 0 ? P::NV : 1];
#pragma unroll
    for (int i = 0; i < P::NV; i++) in[i] = vs[i] * r;
    if constexpr (PROG == PSGPU_PROC_BLINN) {
        if (V.gbuf) {   // geometry buffer of the visible fragment (rows bottom-up)
            float4* gp = (float4*)(V.gbuf + ((int64_t)y * V.W + x) * 12);
            gp[0] = make_float4(in[6], in[7], in[8], 1.0f);
            gp[1] = make_float4(in[9], in[10], in[11], 0.0f);
            gp[2] = make_float4(in[12], in[13], 0.0f, 0.0f);
        }
    }
    float grad[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (P::GRAD) {
        const float4* gi = (const float4*)(grads + (int64_t)tri * 8);
        float4 g0 = gi[0], g1 = gi[1];
        grad[0] = g0.x; grad[1] = g0.y; grad[2] = g0.z; grad[3] = g0.w; grad[4] = g1.x; grad[5] = g1.y;
    }
    uint32_t c;
    if (P::fp(d, in, r, grad, c)) V.colour[crow * V.colour_scan + x] = c;
    else if (crow < V.cclear_rows) V.colour[crow * V.colour_scan + x] = V.cclear_value;
    });
}

// ---------------------------------------------------------------------------
// mip chain: rounded 2x2 mean, level l-1 -> l
// ---------------------------------------------------------------------------
__global__ void k_mip(const uint8_t* __restrict__ src, int ws, int hs, int ss, uint8_t* __restrict__ dst, int wd, int hd,
                      int elem)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= wd || y >= hd) return;
    int y0 = 2 * y, y1 = min(2 * y + 1, hs - 1);
    int x0 = 2 * x, x1 = min(2 * x + 1, ws - 1);
    for (int c = 0; c < elem; c++) {
        int s = src[(int64_t)y0 * ss + x0 * elem + c] + src[(int64_t)y0 * ss + x1 * elem + c] +
                src[(int64_t)y1 * ss + x0 * elem + c] + src[(int64_t)y1 * ss + x1 * elem + c];
        dst[(int64_t)y * wd * elem + x * elem + c] = (uint8_t)((s + 2) >> 2);
    }
}

__global__ void k_selftest(int op, const float* a, const float* b, float* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r = 0.0f;
    switch (op) {
    case 0: r = a[i] / b[i]; break;
    case 1: r = sqrtf(a[i]); break;
    case 2: r = 1.0f / a[i]; break;
    case 3: r = (float)ps_cvt_i32(a[i]); break;
    case 4: r = floorf(a[i]); break;
    case 5: r = (float)(int)ps_cvt_u32(a[i]); break;
    case 6: r = ps_log2(a[i]); break;
    case 7: r = a[i]; break;   // plain copy: 4 bytes read + 4 written per element
    case 8: case 9: {           // wrapped texel pair (v, v + 1) modulo m; int bits in and out
        int v = __float_as_int(a[i]), m = __float_as_int(b[i]), c0, c1;
        ps_clamp_pair(1, m, v, c0, c1);
        r = __int_as_float(op == 8 ? c0 : c1);
        break;
    }
    case 10: r = (float)ps_sse_pack_byte(a[i]); break;
    }
    out[i] = r;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline int gsz(int64_t n, int b) { int64_t g = (n + b - 1) / b; return (int)(g < 1 ? 1 : g); }

void psk_fill_u32(hipStream_t s, uint32_t* p, int64_t n, uint32_t v, const PsCounters* c)
{
    if (n <= 0) return;
    int g = gsz(n, 256); if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, p, n, v, c);
}

int64_t psk_fill_pieces(int64_t rows, int64_t width) { return rows * ((width + FILL_PIECE - 1) / FILL_PIECE); }

void psk_fills(hipStream_t s, const PsFill* fills, int n, int64_t max_pieces, const PsCounters* c)
{
    if (n <= 0) return;
    int64_t g = max_pieces < 1 ? 1 : max_pieces; if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fills, dim3((unsigned)g, n), dim3(256), 0, s, fills, c);
}


int psk_setup_block() { return 128; }

void psk_setup(hipStream_t s, int prog, bool share, const PsDraw* draws, int nd, int64_t T, PsTri* tris, float* vary, int nvs,
               float* grads, uint32_t* nrows, uint32_t* row_part, PsCounters* cnt, const PsView* views)
{
    dim3 g(gsz(T, 128)), b(128);
    size_t lds = (size_t)128 * 3 * nvs * sizeof(float);
#define SU_CASE(K)                                                                                                  \
    case K:                                                                                                         \
        if (share && ps_shares_varyings(K))                                                                         \
            hipLaunchKernelGGL((k_setup<K, true>), g, b, lds, s, draws, nd, T, tris, vary, nvs, grads, nrows, row_part, cnt, views); \
        else                                                                                                        \
            hipLaunchKernelGGL((k_setup<K, false>), g, b, lds, s, draws, nd, T, tris, vary, nvs, grads, nrows, row_part, cnt, views); \
        break;
    switch (prog) {
    SU_CASE(PSGPU_PROC_DEF01) SU_CASE(PSGPU_PROC_DEF02) SU_CASE(PSGPU_PROC_DEF03) SU_CASE(PSGPU_PROC_DEF04)
    SU_CASE(PSGPU_PROC_DEF05) SU_CASE(PSGPU_PROC_LAMBERT) SU_CASE(PSGPU_PROC_BLINN) SU_CASE(PSGPU_PROC_SHADOWED)
    SU_CASE(PSGPU_PROC_CLOUD) SU_CASE(PSGPU_PROC_CLOUD_SHADOW)
    }
#undef SU_CASE
}

int psk_scan_blocks(int64_t cap) { return gsz(cap, SCAN_BLOCK); }


void psk_scan_parts(hipStream_t s, const uint32_t* in, int64_t n, uint32_t* out, const uint32_t* parts, int pgran,
                    uint32_t* total_out, uint32_t* bstart, int64_t bcap)
{
    hipLaunchKernelGGL(k_scan_parts, dim3(psk_scan_blocks(n)), dim3(1024), 0, s, in, n, parts, pgran, out, total_out,
                       bstart, bcap);
}

int psk_heavy_cap() { return PS_HEAVY_CAP; }

void psk_scan_tiles(hipStream_t s, const unsigned long long* tcnt, int64_t n, unsigned long long* tstart,
                    unsigned long long* partial, uint32_t* total_out, const PsCounters* c, uint32_t* heavy)
{
    if (n <= SCAN_DIRECT_MAX) {
        hipLaunchKernelGGL(k_scan_direct, dim3(gsz(n, DIRECT_BLOCK)), dim3(1024), 0, s, tcnt, n, tstart,
                           total_out, c, heavy);
        return;
    }
    int nb = psk_scan_blocks(n);
    hipLaunchKernelGGL(k_scan_reduce<unsigned long long>, dim3(nb), dim3(256), 0, s, tcnt, nullptr, n, partial, c);
    hipLaunchKernelGGL(k_scan_final<unsigned long long>, dim3(nb), dim3(256), 0, s, tcnt, nullptr, n, partial, nb,
                       SCAN_BLOCK, tstart, total_out, n, c, heavy);
}


void psk_span(hipStream_t s, const PsTri* tris, const uint32_t* row_off, const uint32_t* bstart, int64_t T,
              PsSpan* spans, unsigned long long* tcnt, const PsView* views, int n_views, int chunk_w, int tile_rows,
              int64_t cap, uint32_t* heavy, PsCounters* cnt)
{
    int g = gsz(cap, 256); if (g > 2048) g = 2048;
    const int tile_cols = 64 / tile_rows;
    hipLaunchKernelGGL(k_span, dim3(g), dim3(256), 0, s, tris, row_off, bstart, T, cap, spans, tcnt, views, n_views,
                       chunk_w, tile_rows, tile_cols, heavy, cnt);
}

void psk_scan_rows(hipStream_t s, const uint32_t* in, int64_t n, uint32_t* out, uint32_t* partial, uint32_t* total_out,
                   uint32_t* bstart, int64_t bcap)
{
    hipLaunchKernelGGL(k_scan_reduce<uint32_t>, dim3(psk_scan_blocks(n)), dim3(256), 0, s, in, (const uint32_t*)nullptr, n,
                       partial, (const PsCounters*)nullptr);
    psk_scan_parts(s, in, n, out, partial, SCAN_BLOCK, total_out, bstart, bcap);
}

void psk_segments(hipStream_t s, const PsSpan* spans, const uint32_t* row_off, int64_t T, const PsFrame& f,
                  const PsTri* tris, const PsDraw* draws, const float* vary, int nvs, unsigned long long* tcnt, PsSeg* segs,
                  float* ck, int64_t span_cap, int64_t seg_cap, PsCounters* cnt)
{
    int g = gsz(span_cap, 256); if (g > 2048) g = 2048;
#define SG_LAUNCH(NVC)                                                                                                \
    hipLaunchKernelGGL(k_segments<NVC>, dim3(g), dim3(256), 0, s, spans, row_off, T, f, tris, draws, vary, nvs, tcnt, \
                       segs, ck, seg_cap, cnt)
    if (nvs <= 4) SG_LAUNCH(4);
    else if (nvs <= 8) SG_LAUNCH(8);
    else if (nvs <= 14) SG_LAUNCH(14);
    else SG_LAUNCH(PS_MAX_NV);
#undef SG_LAUNCH
}

size_t psk_raster_lds(int chunk_w, int waves) { return (size_t)waves * RA_WORDS(chunk_w) * 4; }

void psk_raster(hipStream_t s, bool debug, const PsSpan* spans, const PsSeg* segs, const PsTri* tris, const PsDraw* draws,
                const PsFrame& f, int waves_per_group, PsCounters* cnt)
{
    int64_t n_tiles = f.n_tiles;
    const int cw = f.chunk_w;
    const bool known = cw == 1 || cw == 2 || cw == 4 || cw == 8;
    size_t lds = psk_raster_lds(known ? 0 : cw, waves_per_group);
    dim3 g(gsz(n_tiles + (f.heavy ? PS_HEAVY_CAP : 0), waves_per_group)), b(64 * waves_per_group);
#define RA_LAUNCH(D, K) hipLaunchKernelGGL((k_raster<D, K>), g, b, lds, s, spans, segs, tris, draws, f, n_tiles, cnt)
    switch (known ? cw : 0) {
    case 1: if (debug) RA_LAUNCH(true, 1); else RA_LAUNCH(false, 1); break;
    case 2: if (debug) RA_LAUNCH(true, 2); else RA_LAUNCH(false, 2); break;
    case 4: if (debug) RA_LAUNCH(true, 4); else RA_LAUNCH(false, 4); break;
    case 8: if (debug) RA_LAUNCH(true, 8); else RA_LAUNCH(false, 8); break;
    default: if (debug) RA_LAUNCH(true, 0); else RA_LAUNCH(false, 0); break;
    }
#undef RA_LAUNCH
}

void psk_raster_imm(hipStream_t s, int prog, bool debug, const PsSpan* spans, const PsSeg* segs, const PsTri* tris,
                    const PsDraw* draws, const PsFrame& f, int waves_per_group, PsCounters* cnt, const float* vary, int nvs,
                    const float* ck)
{
    int64_t n_tiles = f.n_tiles;
    size_t lds = psk_raster_lds(f.chunk_w, waves_per_group);
    dim3 g(gsz(n_tiles, waves_per_group)), b(64 * waves_per_group);
#define RI_CASE(K)                                                                                                   \
    case K:                                                                                                          \
        if (debug)                                                                                                   \
            hipLaunchKernelGGL((k_raster_imm<K, true>), g, b, lds, s, spans, segs, tris, draws, f, n_tiles, cnt,     \
                               vary, nvs, ck);                                                                       \
        else                                                                                                         \
            hipLaunchKernelGGL((k_raster_imm<K, false>), g, b, lds, s, spans, segs, tris, draws, f, n_tiles, cnt,    \
                               vary, nvs, ck);                                                                       \
        break;
    switch (prog) {
    RI_CASE(PSGPU_PROC_CLOUD) RI_CASE(PSGPU_PROC_CLOUD_SHADOW)
    default: break;
    }
#undef RI_CASE
}

void psk_shade(hipStream_t s, int prog, const PsSpan* spans, const PsSeg* segs, const float* ck, const PsTri* tris,
               const float* vary, int nvs, const float* grads, const PsDraw* draws, const PsFrame& f, const PsCounters* cnt,
               bool clears)
{
    dim3 g((unsigned)f.n_stiles), b(256);
#define SH_CASE(K) case K: hipLaunchKernelGGL(k_shade<K>, g, b, 0, s, spans, segs, ck, tris, vary, nvs, grads, draws, f, cnt, \
                                              clears ? 1 : 0); break;
    switch (prog) {
    SH_CASE(PSGPU_PROC_DEF01) SH_CASE(PSGPU_PROC_DEF02) SH_CASE(PSGPU_PROC_DEF03) SH_CASE(PSGPU_PROC_DEF04)
    SH_CASE(PSGPU_PROC_LAMBERT) SH_CASE(PSGPU_PROC_BLINN) SH_CASE(PSGPU_PROC_SHADOWED)
    default: break;
    }
#undef SH_CASE
}

void psk_mip(hipStream_t s, const uint8_t* src, int ws, int hs, int ss, uint8_t* dst, int wd, int hd, int elem)
{
    dim3 g((wd + 255) / 256, hd), b(256);
    hipLaunchKernelGGL(k_mip, g, b, 0, s, src, ws, hs, ss, dst, wd, hd, elem);
}

void psk_selftest(hipStream_t s, int op, const float* a, const float* b, float* out, int n)
{
    hipLaunchKernelGGL(k_selftest, dim3(gsz(n, 256)), dim3(256), 0, s, op, a, b, out, n);
}
