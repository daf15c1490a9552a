"""Mint tests/golden/raster_golden.npz from the reference's own rasterizer.

Runs oracle/_ref/libref_raster.so -- src/puresoft3d/rasterizer.cpp compiled
from the reference tree by `make -C oracle ref` -- on seeded triangles and
stores inputs and outputs.  Categories: ordinary, pixel-grid snapped (flat
top/bottom by exact equality), near-flat (|dy| < 1e-6 edges), huge x (x86
INT_MIN conversions), tiny, degenerate, far below the screen.
Run from the repo root:  python tests/golden/make_raster_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

SIZES = [(64, 48), (33, 17), (7, 5), (160, 120)]
PER_CATEGORY = 100


def gen(rng, cat, W, H):
    if cat == 0:
        v = rng.uniform(-1.5, 1.5, (3, 2))
    elif cat == 1:
        v = rng.integers(-4, max(W, H) + 4, (3, 2)) / np.array([W / 2, H / 2]) - 1.0
        v = np.round(v * 8) / 8
    elif cat == 2:
        v = rng.uniform(-1, 1, (3, 2))
        v[1, 1] = v[0, 1] + rng.uniform(-1e-7, 1e-7)
    elif cat == 3:
        v = rng.uniform(-1, 1, (3, 2))
        v[rng.integers(0, 3), 0] = rng.choice([1e5, -1e5, 3e7, -3e9, 5e9])
    elif cat == 4:
        c = rng.uniform(-1, 1, 2)
        v = c + rng.uniform(-0.03, 0.03, (3, 2))
    elif cat == 5:
        v = rng.uniform(-1, 1, (3, 2))
        v[2] = v[rng.integers(0, 2)]
    else:
        v = rng.uniform(-1, 1, (3, 2))
        v[:, 1] -= rng.uniform(0, 3)
    v4 = np.zeros((3, 4), np.float32)
    v4[:, :2] = v
    v4[:, 2] = rng.uniform(-1, 1, 3)
    v4[:, 3] = 1
    return v4


def main():
    R = po.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref_raster.so missing: run `make -C oracle ref` where the reference tree exists")
    rng = np.random.default_rng(20261202)
    sizes, verts, ok, first, last, vout, row_off, rows = [], [], [], [], [], [], [0], []
    for W, H in SIZES:
        for cat in range(7):
            for _ in range(PER_CATEGORY):
                v = gen(rng, cat, W, H)
                o, f, l, rr, vv = po.push_triangle(R.ref_push_triangle, W, H, *v)
                written = np.where(rr[:, 0] != 0x7FFFFFFF)[0]
                if len(written):
                    assert o and written.min() >= f and written.max() <= l, "reference wrote outside [first,last]"
                sizes.append((W, H)); verts.append(v); ok.append(o); first.append(f); last.append(l); vout.append(vv)
                blk = rr[f: l + 1] if (o and f <= l) else rr[:0]
                rows.append(blk)
                row_off.append(row_off[-1] + len(blk))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "raster_golden.npz")
    np.savez_compressed(out, sizes=np.array(sizes, np.int32), verts=np.array(verts, np.float32),
                        ok=np.array(ok, np.int32), first=np.array(first, np.int32), last=np.array(last, np.int32),
                        verts_out=np.array(vout, np.float32), row_off=np.array(row_off, np.int64),
                        rows=np.concatenate(rows, 0).astype(np.int32))
    print("wrote", out, len(sizes), "cases", row_off[-1], "rows")


if __name__ == "__main__":
    main()
