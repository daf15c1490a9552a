"""Autotuner of the raster tiling.

A raster tile is one wavefront: 64 lanes, each owning a run of `chunk_width`
pixels of one row; `tile_rows` rows per tile, so a tile is
tile_rows x (64 / tile_rows * chunk_width) pixels.  The search covers tiles from
8x8 to 64x64 pixels and 1-8 wavefronts per workgroup (LDS permitting), timing
the caller's workload (typically a graph replay) for each candidate.
"""
import time

import numpy as np

LDS_LIMIT = 65536


def lds_bytes(chunk_width, waves):
    """LDS of the raster kernel: per wave, depth and winner columns plus 128
    staged pieces of 7 words and 128 words of counts (ps_kernels.hip RA_WORDS)"""
    return waves * (chunk_width * 64 * 2 + 128 * 7 + 128) * 4


def candidates(min_side=8, max_side=64, waves=(1, 2, 4, 8)):
    out = []
    for cw in (1, 2, 4, 8, 16, 32, 64):
        for tr in (1, 2, 4, 8, 16, 32, 64):
            tw = 64 // tr * cw
            if not (min_side <= tr <= max_side and min_side <= tw <= max_side):
                continue
            for w in waves:
                if lds_bytes(cw, w) <= LDS_LIMIT:
                    out.append((cw, tr, w))
    return out


def autotune(ctx, run, reps=5, cands=None, log=None):
    """run(): submits the workload once (asynchronously).  Returns
    (best (chunk_width, tile_rows, waves), [(tiling, tile_h, tile_w, ms), ...])."""
    cands = cands or candidates()
    table = []
    for cw, tr, w in cands:
        try:
            ctx.tiling(cw, tr, w)
        except ValueError:    # refused by the library (LDS)
            continue
        run()                 # capture / warm up
        ctx.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            ctx.sync()
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        table.append(((cw, tr, w), tr, 64 // tr * cw, ms))
        if log:
            log("tile %2dx%-2d chunk %2d waves %d: %.3f ms" % (tr, 64 // tr * cw, cw, w, ms))
    best = min(table, key=lambda r: r[3])
    ctx.tiling(*best[0])
    return best[0], table
