"""The batching and serving workloads at their full size.

* 256 views of 16 resident scenes (10k-500k triangles, views 256^2-1024^2,
  1024^2 mipmapped textures) in one graph replay: every view equals its
  one-view-per-flush render bit for bit (colour and depth), and eight of
  them (the smallest, the largest and six seeded others) equal the oracle.
* 2,000 Poisson requests over 32 resident scenes (1k-1M triangles, views
  128^2-2048^2, 512^2-4096^2 textures) through the continuous-batching
  server: every served image equals the same view rendered alone, and the
  three smallest equal the oracle.
"""
import hashlib

import numpy as np
import pytest

from conftest import announce, oracle_view

pytestmark = pytest.mark.gpu

N = pytest.importorskip("puresoft3d_amd.native")
from puresoft3d_amd import scenes as S  # noqa: E402

CLEAR = 0x00302010


def _digest(*arrays):
    h = hashlib.sha1()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.timeout(400)
def test_batch_256_views_full_size(oracle):
    assets = S.batch_assets(n_scenes=16, seed=11, tri_lo=10_000, tri_hi=500_000, tex_size=1024)
    views = S.batch_views(assets, n_views=256, seed=12, side_lo=256, side_hi=1024)
    announce("batch of 256 views, %d scenes, %.1f M px" % (len(assets), sum(v.width * v.height for v in views) / 1e6))
    ctx = N.Context(64, 64)
    try:
        b = S.GpuBatch(ctx, assets, views)
        ctx.capture_begin()
        b.submit()
        g = ctx.capture_end()        # runs the recording once
        ctx.graph_launch(g, 1)       # the replay is what is checked
        ctx.sync()
        batched = []
        keep = {}
        areas = np.array([v.width * v.height for v in views])
        rng = np.random.default_rng(3)
        check = sorted({int(np.argmin(areas)), int(np.argmax(areas))} |
                       set(int(i) for i in rng.choice(len(views), 6, replace=False)))
        for i in range(len(views)):
            col, dep = b.read(i)
            batched.append(_digest(col, dep.view(np.uint32)))
            if i in check:
                keep[i] = col
        announce("batch: one view per flush, 256 views")
        for i in range(len(views)):
            b.submit([i])
            ctx.sync()
            col, dep = b.read(i)
            assert _digest(col, dep.view(np.uint32)) == batched[i], "view %d differs from its one-view render" % i
        ctx.graph_destroy(g)
    finally:
        ctx.close()
    for i in check:
        announce("batch: view %d (%dx%d) against the oracle" % (i, views[i].width, views[i].height))
        sc = S.view_as_scene(assets, views[i], "view%d" % i)
        f = oracle.Frame(sc.width, sc.height, depth_scanline=sc.width, top_down=False)
        f.colour.view(np.uint32)[:] = 0x00302010
        f.clear_depth(1.0)
        oracle.render_scene(sc, frame=f, clear=False, n_threads=8)
        np.testing.assert_array_equal(keep[i], f.colour, err_msg="view %d" % i)


@pytest.mark.timeout(600)
def test_serving_2000_requests_full_size(oracle):
    from puresoft3d_amd import serving
    assets = S.serving_assets(n_scenes=32, seed=21)
    views, arr = S.poisson_requests(assets, n=2000, rate=8000.0, seed=23)
    announce("serve 2000 Poisson requests at 8000/s over %d scenes (%.2f M tris)" % (
        len(assets), sum(a.triangles() for a in assets) / 1e6))
    ctx = N.Context(64, 64)
    try:
        srv = serving.RenderServer(ctx, assets)
        res, st, imgs = srv.serve(views, arr, max_batch_views=64, max_batch_pixels=16 << 20, max_inflight=2,
                                  keep=range(len(views)))
        assert st["views"] == len(views) and st["batches"] < len(views) and st["max_batch_views_seen"] > 1
        assert (res["done_s"] >= res["launched_s"]).all() and (res["launched_s"] >= res["arrival_s"]).all()
        announce("serve: every image against its render alone")
        for i, v in enumerate(views):
            alone = srv.render_alone(v, CLEAR)
            assert np.array_equal(imgs[i], alone), "request %d (%dx%d) differs from its render alone" % (
                i, v.width, v.height)
        srv.close()
    finally:
        ctx.close()
    order = np.argsort([v.width * v.height for v in views])
    for i in order[:3]:
        announce("serve: request %d (%dx%d) against the oracle" % (i, views[i].width, views[i].height))
        f, _ = oracle_view(oracle, S.view_as_scene(assets, views[i]), CLEAR)
        np.testing.assert_array_equal(imgs[i], f.colour)
