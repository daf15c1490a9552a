"""The serving loop (include/psserve.h): continuous batching over resident
scenes with HBM-budget admission.  Each served image must equal the same view
rendered alone, and the oracle's render of it."""
import numpy as np
import pytest

from conftest import announce, oracle_view

from puresoft3d_amd import native as N
from puresoft3d_amd import scenes as S

CLEAR = 0x00302010


def test_poisson_requests_shape():
    assets = S.serving_assets(n_scenes=3, seed=2, tri_lo=200, tri_hi=2000, tex_lo=32, tex_hi=64)
    views, arr = S.poisson_requests(assets, n=4000, rate=1000.0, seed=5, side_lo=16, side_hi=256)
    assert (np.diff(arr) >= 0).all() and arr[0] == 0.0
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 1e-3) < 1e-4                      # exponential with mean 1/rate
    assert abs(np.median(gaps) - np.log(2) * 1e-3) < 1e-4
    area = np.array([v.width * v.height for v in views], float)
    assert area.min() >= 16 * 16 * 0.7 and area.max() <= 256 * 256 * 1.3
    la = np.log(area)
    assert abs(la.mean() - (np.log(256.0) + np.log(65536.0)) / 2) < 0.15   # log-uniform
    assert {v.scene for v in views} == {0, 1, 2}


def test_serve_rejects_bad_arguments():
    from puresoft3d_amd import serving
    import ctypes as C
    L = serving.lib()
    cfg = serving.Config(0, 1 << 30, 16, 8, 1 << 20, 2, 0, 1)
    h = C.c_void_p()
    assert L.psserve_create(None, None, 0, C.byref(cfg), C.byref(h)) == N.ERR_INVALID and not h.value
    assert L.psserve_run(None, None, 0, None, None, None) == N.ERR_INVALID
    bad = serving.Config(0, 1 << 30, 16, 0, 1 << 20, 2, 0, 1)      # no views per batch
    assert L.psserve_set_config(None, C.byref(bad)) == N.ERR_INVALID
    assert L.psserve_destroy(None) == 0
    ctxs = (C.c_void_p * 2)(None, None)
    assert L.psserve_create_multi(ctxs, 2, None, 1, C.byref(cfg), C.byref(h)) == N.ERR_INVALID and not h.value
    assert L.psserve_create_multi(None, 0, None, 1, C.byref(cfg), C.byref(h)) == N.ERR_INVALID
    neg = serving.Config(0, 1 << 30, 16, 8, 1 << 20, 2, 0, 1, -1)   # negative stream count
    assert L.psserve_set_config(None, C.byref(neg)) == N.ERR_INVALID


def _small_server(ctx):
    from puresoft3d_amd import serving
    assets = S.serving_assets(n_scenes=4, seed=31, tri_lo=500, tri_hi=20000, tex_lo=64, tex_hi=256)
    return assets, serving.RenderServer(ctx, assets)


@pytest.mark.gpu
def test_served_images_equal_alone_and_oracle(oracle):
    from puresoft3d_amd import serving
    ctx = N.Context(64, 64)
    try:
        assets, srv = _small_server(ctx)
        views, arr = S.poisson_requests(assets, n=48, rate=20000.0, seed=7, side_lo=24, side_hi=320)
        announce("serve 48 views, batches <= 16 views")
        res, st, imgs = srv.serve(views, arr, max_batch_views=16, max_inflight=2, keep=range(len(views)))
        assert st["views"] == len(views) and st["batches"] > 1 and st["max_batch_views_seen"] > 1
        assert (res["done_s"] >= res["launched_s"]).all() and (res["launched_s"] >= res["arrival_s"]).all()
        alone = [srv.render_alone(v) for v in views]
        for i in range(len(views)):
            np.testing.assert_array_equal(imgs[i], alone[i], err_msg="view %d" % i)
        summ = serving.latency_summary(res)
        assert summ["p99_ms"] >= summ["p50_ms"] > 0
        res_d, st_d, _ = srv.serve(views, arr, max_batch_views=16, to_host=False)
        assert st_d["views"] == len(views) and (res_d["done_s"] > 0).all()
        srv.close()
    finally:
        ctx.close()
    order = np.argsort([v.width * v.height for v in views])
    for i in order[:3]:
        f, _ = oracle_view(oracle, S.view_as_scene(assets, views[i]), CLEAR)
        np.testing.assert_array_equal(imgs[i], f.colour)


@pytest.mark.gpu
def test_two_render_streams_serve_identical_images():
    """batches spread over two contexts (render streams) give the same images
    as one stream and as each view rendered alone"""
    from puresoft3d_amd import serving
    ctx, ctx2 = N.Context(64, 64), N.Context(64, 64)
    try:
        assets = S.serving_assets(n_scenes=4, seed=31, tri_lo=500, tri_hi=20000, tex_lo=64, tex_hi=256)
        srv = serving.RenderServer(ctx, assets, extra_ctxs=[ctx2])
        views, arr = S.poisson_requests(assets, n=48, rate=20000.0, seed=13, side_lo=24, side_hi=320)
        announce("serve 48 views over two render streams")
        res, st, imgs = srv.serve(views, arr, max_batch_views=8, max_inflight=2, keep=range(len(views)))
        assert st["views"] == len(views) and st["batches"] > 2
        assert (res["done_s"] >= res["launched_s"]).all() and (res["launched_s"] >= res["arrival_s"]).all()
        res1, st1, imgs1 = srv.serve(views, arr, max_batch_views=8, max_inflight=2, keep=range(len(views)),
                                     streams=1)
        alone = [srv.render_alone(v) for v in views]
        for i in range(len(views)):
            np.testing.assert_array_equal(imgs[i], alone[i], err_msg="view %d" % i)
            np.testing.assert_array_equal(imgs1[i], alone[i], err_msg="view %d, one stream" % i)
        srv.close()
    finally:
        ctx2.close()
        ctx.close()


@pytest.mark.gpu
def test_budget_admission_and_plain_mode():
    ctx = N.Context(64, 64)
    try:
        assets, srv = _small_server(ctx)
        views, arr = S.poisson_requests(assets, n=30, rate=50000.0, seed=9, side_lo=64, side_hi=256)
        announce("serve 30 views, tight budget")
        budget = 3 * 256 * 256 * 8          # about three 256^2 targets
        res, st, imgs = srv.serve(views, arr, hbm_budget=budget, scratch_bytes_per_px=0, max_batch_views=8,
                                  keep=range(len(views)))
        assert st["views"] == 30 and st["peak_bytes"] <= budget
        assert st["target_evictions"] > 0 or st["target_allocs"] <= 3
        announce("serve 30 views, one view per launch")
        res1, st1, imgs1 = srv.serve(views, arr, max_batch_views=1, max_inflight=1, keep=range(len(views)))
        assert st1["batches"] == 30 and st1["max_batch_views_seen"] == 1
        for i in range(len(views)):
            np.testing.assert_array_equal(imgs[i], imgs1[i])
        srv.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_async_read_after_overflow_is_replayed():
    """a context that starts with tiny scratch overflows on its first flush; a
    fence after an asynchronous read must only report done with the re-run
    (correct) pixels in host memory"""
    ctx = N.Context(64, 64, flags=1)
    try:
        assets, srv = _small_server(ctx)
        views, _ = S.poisson_requests(assets, n=4, rate=1000.0, seed=11, side_lo=96, side_hi=200)
        ref = [srv.render_alone(v) for v in views]
        srv.close()
        ctx2 = N.Context(64, 64, flags=1)
        try:
            from puresoft3d_amd import serving
            srv2 = serving.RenderServer(ctx2, assets)
            announce("async read after overflow")
            bufs = []
            targets = []
            for v in views:
                c = ctx2.texture_create(v.width, v.height, 4, wrap=N.CLAMP)
                d = ctx2.texture_create(v.width, v.height, 4, wrap=N.CLAMP)
                targets.append((c, d))
                sc = srv2.scenes[v.scene]
                ctx2.program_use(srv2.prog)
                ctx2.target(c)
                ctx2.depth_target(d)
                ctx2.viewport(v.width, v.height)
                ctx2.clear_depth(1.0)
                ctx2.clear_target(CLEAR)
                u = v.uniforms()
                u[9] = np.array([sc.albedo_tex], np.int32)
                u[10] = np.array([sc.normal_tex], np.int32)
                for k, val in u.items():
                    ctx2.uniform(k, val)
                ctx2.draw(sc.vao)
            for (c, d), v in zip(targets, views):
                buf = ctx2.host_alloc(v.width * v.height * 4)
                buf[:] = 0xEE
                ctx2.read_async(c, buf, w=v.width, h=v.height)
                bufs.append(buf)
            f = ctx2.fence()
            while not ctx2.fence_done(f):
                pass
            assert ctx2.stats()["retries"] >= 1
            for buf, v, r in zip(bufs, views, ref):
                np.testing.assert_array_equal(buf.reshape(v.height, v.width, 4), r)
            for buf in bufs:
                ctx2.host_free(buf)
            srv2.close()
        finally:
            ctx2.close()
    finally:
        ctx.close()
