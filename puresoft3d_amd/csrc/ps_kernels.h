// ps_kernels.h -- host launchers of the raster kernels (ps_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "ps_types.h"

void psk_fill_u32(hipStream_t s, uint32_t* p, int64_t n, uint32_t v, const PsCounters* c);
// work pieces of a fill (1024 values of a row); psk_fills takes the largest
int64_t psk_fill_pieces(int64_t rows, int64_t width);
void psk_fills(hipStream_t s, const PsFill* fills, int n, int64_t max_pieces, const PsCounters* c);
// row_part (optional, flushes of one programme): each block of psk_setup_block()
// triangles writes its row total there
int psk_setup_block();
void psk_setup(hipStream_t s, int prog, bool share, const PsDraw* draws, int nd, int64_t T, PsTri* tris, float* vary, int nvs,
               float* grads, uint32_t* nrows, uint32_t* row_part, PsCounters* cnt, const PsView* views);
int psk_scan_blocks(int64_t cap);
// exclusive scan of the row counts from given partial sums of pgran items
// each (pgran divides 4096); out[n] and *total_out receive the total, and
// bstart (bcap entries) the triangle of every 64th row item
void psk_scan_parts(hipStream_t s, const uint32_t* in, int64_t n, uint32_t* out, const uint32_t* parts, int pgran,
                    uint32_t* total_out, uint32_t* bstart, int64_t bcap);
// the same from the counts alone (block totals into partial first)
void psk_scan_rows(hipStream_t s, const uint32_t* in, int64_t n, uint32_t* out, uint32_t* partial, uint32_t* total_out,
                   uint32_t* bstart, int64_t bcap);
// exclusive scan of the per-tile counts (pieces | runs << 32) into tstart
// (n + 1 entries); the piece total into *total_out, its running maximum at
// total_out[2]
// heavy (optional, 1 + psk_heavy_cap() words): the heavy-tile list of PsFrame.heavy
int psk_heavy_cap();
void psk_scan_tiles(hipStream_t s, const unsigned long long* tcnt, int64_t n, unsigned long long* tstart,
                    unsigned long long* partial, uint32_t* total_out, const PsCounters* c, uint32_t* heavy);
// spans of every row item; tcnt[t] += the pieces in tile t | the waves with pieces there << 32
void psk_span(hipStream_t s, const PsTri* tris, const uint32_t* row_off, const uint32_t* bstart, int64_t T,
              PsSpan* spans, unsigned long long* tcnt, const PsView* views, int n_views, int chunk_w, int tile_rows,
              int64_t cap, uint32_t* heavy, PsCounters* cnt);
// the pieces, into runs of the tile lists f.tstart / f.runs; counts tcnt back down to zero.
void psk_segments(hipStream_t s, const PsSpan* spans, const uint32_t* row_off, int64_t T, const PsFrame& f,
                  const PsTri* tris, const PsDraw* draws, const float* vary, int nvs, unsigned long long* tcnt, PsSeg* segs,
                  float* ck, int64_t span_cap, int64_t seg_cap, PsCounters* cnt);
size_t psk_raster_lds(int chunk_w, int waves);
void psk_raster(hipStream_t s, bool debug, const PsSpan* spans, const PsSeg* segs, const PsTri* tris, const PsDraw* draws,
                const PsFrame& f, int waves_per_group, PsCounters* cnt);
// the raster walk with the FP of `prog` on every pass (blending / discarding
// programmes), writing colour at once; no winner buffer, no shading pass
void psk_raster_imm(hipStream_t s, int prog, bool debug, const PsSpan* spans, const PsSeg* segs, const PsTri* tris,
                    const PsDraw* draws, const PsFrame& f, int waves_per_group, PsCounters* cnt, const float* vary, int nvs,
                    const float* ck);
// clears: this launch also writes the views' fused colour clears (one launch of a flush does)
void psk_shade(hipStream_t s, int prog, const PsSpan* spans, const PsSeg* segs, const float* ck, const PsTri* tris,
               const float* vary, int nvs, const float* grads, const PsDraw* draws, const PsFrame& f,
               const PsCounters* cnt, bool clears);
void psk_mip(hipStream_t s, const uint8_t* src, int ws, int hs, int ss, uint8_t* dst, int wd, int hd, int elem);
void psk_selftest(hipStream_t s, int op, const float* a, const float* b, float* out, int n);
