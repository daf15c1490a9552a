// ps_types.h -- device-side records shared by the raster kernels and the host
// orchestration of libpsgpu.  Plain structs: the host fills them and copies
// them to HBM; kernels read them through global pointers.
#pragma once
#include <stdint.h>

#define PS_MAX_NV      20   // widest varying block (DEF03: 5 float4 slots)
#define PS_N_PROGS     11   // programme kinds are 1 .. PS_N_PROGS-1 (PSGPU_PROC_*)
#define PS_MAX_LEVELS  16
#define PS_MAX_LAYERS  6

// programmes (PSGPU_PROC_* values) whose clip position is PV (M p) and whose
// varyings depend on the vertex streams, M, Mrot and shadowPV only, so that
// views of one scene can share them
constexpr bool ps_shares_varyings(int prog)
{
    return prog == 1 || prog == 2 || prog == 3 || prog == 6 || prog == 7 || prog == 8 || prog == 9;
}

// One texture (a PuresoftFBO used as a texture, fbo.h:6-100) in HBM.
// Layer-major; inside a layer level 0 (rows of `scanline` bytes) then the
// tightly packed mip levels.
struct PsTex {
    const uint8_t* data;
    int width, height, scanline, elem_len;
    int layers, wrap, levels, pad;
    int64_t layer_bytes;
    int64_t level_off[PS_MAX_LEVELS];
};

// Per-draw record: the snapshot of programme, behaviour, vertex streams and
// uniforms that drawVAO would see (drawvao.cpp:12-20).
struct PsDraw {
    int view;          // index of the view (render target + viewport) in the flush
    int prog;          // PSGPU_PROC_*
    int behavior;      // PSGPU_* behaviour bits
    int nv;            // varying floats of the programme
    int writes_colour; // FP writes FBO slot 0
    int64_t tri_start; // first global triangle id of this draw (flush-local)
    int64_t tri_count;
    int64_t id_base;   // debug triangle-id base
    // views of one scene in one flush share its varyings (they do not depend
    // on the view): triangle t of the draw reads those of triangle t + vary_delta;
    // vary_all: write them for every triangle (another draw reads them);
    // vary_skip: write none, compute positions only
    int vary_delta, vary_all, vary_skip, pad0;
    const uint8_t* slot[5];
    int unit[5];
    int spec_exp;
    float PV[16], M[16], Mrot[16];
    float invView[16];  // DEF04
    float PVM[16];      // DEF05
    float lightPos[4], cameraPos[4], lightDir[4];
    float shadowPV[16];  // PSGPU_PROC_SHADOWED: world -> shadow-map clip space
    PsTex tex0, tex1;
    PsTex tex2;          // PSGPU_PROC_SHADOWED: the FP32 shadow map
};

// Triangle after vertex processing and setup (vertthrd.cpp:14-65,
// rasterizer.cpp:122-283).  `code` packs the row->edge mapping:
//   bits 0-1 a, 2-3 b, 4-5 c (vertex indices), bit 6 standing, bit 7 points-left
//   flat   : left edge (a,c), right edge (b,c) for every row
//   standing (a=top, b=bottom, c=third): rows <= split use the lower half
struct PsTri {
    float vx[3], vy[3];
    float rw[3], z[3];
    int first;     // first drawn row
    int nrows;     // drawn rows (0 = no rows)
    int split;     // last row of the lower half (standing triangles)
    uint32_t code; // edge code | draw index << 8
    // what the span and piece kernels need of the draw, without reading it:
    // its view, and its depth behaviour bits | writes colour << 2
    int view;
    uint32_t vflags;
    int pad[2];
};

// One scanline span with at least one pixel (drawvao.cpp:66-110, interp.cpp:26-80).
struct PsSpan {
    int tri;        // flush-local triangle index
    int y;          // row
    int x1, x2;     // leftClamped, rightClamped (x1 <= x2)
    float cf2, cf2_step, pz, pz_step;
};

// The piece of a span inside one bin chunk, with the serial accumulators
// (correction factor, depth numerator) as they stand at its first pixel xa.
// si packs the span index (bits 0-25), the depth behaviour bits (26-27) and
// whether the draw writes colour (28).  win is what the winner plane keeps
// for the piece: the span's triangle, or PS_SEG_CK | the piece's own index
// when the piece keeps a varying checkpoint.
#define PS_SEG_SPAN_MASK 0x3ffffffu
#define PS_SEG_CK 0x80000000u
struct PsSeg {
    uint32_t si;
    uint32_t xab;   // xa | xb << 16 (first and last pixel of the piece)
    float cf2, pz, cst, pst;
    int y;          // row
    uint32_t win;
};

// One view of a flush: a viewport and its colour/depth targets.  A flush
// renders any number of views with the same kernels; tiles and shading
// blocks of all views are numbered consecutively.
struct PsView {
    int W, H;
    int n_chunks, n_tiles_x, n_tiles_y;
    int n_stiles_x;                 // 16x16 shading blocks per row
    int64_t tile_base;              // first raster tile of the view
    int64_t stile_base;             // first shading block of the view
    uint32_t* colour; int colour_scan; int colour_top_down; int colour_h;  // scan in uint32
    float* depth; int depth_scan;   // scan in floats
    // W*H: what shading needs of the last colour-writing pass: its triangle,
    // or PS_SEG_CK | its piece when the piece keeps a varying checkpoint;
    // -1 for none
    int* winner;
    float* gbuf;                    // optional W*H*12 floats: N.xyz 1 | wp.xyz 0 | u v 0 0 of the winner
    // clears that precede the flush's draws, applied by the raster and shade
    // kernels instead of a fill: the whole depth buffer (W x H, no padding)
    // to dclear_value when dclear, and memory rows [0, cclear_rows) of the
    // colour target (W x H) to cclear_value
    int dclear; float dclear_value;
    int cclear_rows; uint32_t cclear_value;
};

// Flush-wide parameters passed by value to kernels.
struct PsFrame {
    int chunk_w, tile_rows, tile_cols;   // raster tiling
    int ck_stride;                       // floats per varying checkpoint (the flush's widest NV)
    int n_views;
    int64_t n_tiles, n_stiles;
    const PsView* views;
    // tile lists.  Each wave of 64 span items writes its pieces of a tile as
    // one run of consecutive pieces.  tstart[t] holds the first piece (low 32
    // bits) and the first run (bits 32-62) of tile t (n_tiles + 1 entries);
    // run r is runs[2r] = its first piece, runs[2r+1] = wave << 12 | pieces - 1.
    // runs2 is scratch of the same size for ordering long run lists.
    const unsigned long long* tstart;
    uint32_t* runs;
    uint32_t* runs2;
    // heavy tiles, rastered first (null: every tile in order): [0] = how
    // many, then the listed tiles, whose tstart entries carry bit 63
    const uint32_t* heavy;
    // optional debug planes (single-view flushes), index y*W+x
    int32_t* dbg_last; int32_t* dbg_rcount; uint32_t* dbg_rhash; int32_t* dbg_pcount; uint32_t* dbg_phash;
};

// One fill of a clear command: rows x width elements from ptr, rows `scan` apart.
struct PsFill {
    uint32_t* ptr;
    int64_t rows, width, scan;
    uint32_t value;
    int pad;
};

// Device-side counters.  Statistics are striped over 64 slots (by block
// index) so that waves of many blocks do not serialise on one address.  The
// kernels before a flush's raster pass count into `pend`; the raster pass,
// which runs only once the flush is past every point where it can overflow
// its scratch, moves `pend` into `stat`.  So `stat` holds the statistics of
// completed flushes only, and a flush re-run after an overflow counts once.
// `overflow`, which every block of the later kernels reads, has a 256-byte
// line to itself: with the first pending stripe on its line, that stripe's
// atomics slowed the span and piece kernels by 10-25 %.
enum { PS_ST_CULLED = 0, PS_ST_ZREJ, PS_ST_DROPPED, PS_ST_SPANS, PS_ST_FRAGS, PS_ST_SHADED, PS_ST_SEGS, PS_ST_N = 8 };
#define PS_STRIPES 64
struct PsCounters {
    unsigned long long stat[PS_STRIPES][PS_ST_N];
    int overflow;
    int pad[63];
    unsigned long long pend[PS_STRIPES][PS_ST_N];
};
