"""psgpu_sync_noalloc, the synchronising entry point that allocates nothing:
its frames against psgpu_sync's and the oracle's, bit for bit, over calls in
a row that reuse the fixed scratch array with a smaller input after a larger
one, and its refusals."""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

N = pytest.importorskip("puresoft3d_amd.native")
from puresoft3d_amd import scenes as S  # noqa: E402


def _frame(sc, noalloc):
    """one frame of the scene on a fresh context, synchronised by either entry point"""
    ctx = N.Context(sc.width, sc.height)
    try:
        gs = S.GpuScene(ctx, sc)
        ctx.debug_enable(True)
        ctx.debug_reset()
        ctx.sync()                       # set-up work (uploads, mip chains) done before the frame
        gs.frame(flush=False)
        if noalloc:
            ctx.sync_noalloc()
        else:
            ctx.sync()
        return (ctx.read_colour(), ctx.read_depth(), ctx.debug_read(sc.width, sc.height), ctx.stats())
    finally:
        ctx.close()


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    for k in ("last_pass", "pass_count", "pass_hash", "raster_count", "raster_hash"):
        np.testing.assert_array_equal(a[2][k], b[2][k])
    for k in ("frags_shaded", "frags_rasterised", "tris_culled", "spans"):
        assert a[3][k] == b[3][k], k


def test_noalloc_matches_sync_then_smaller(oracle):
    """a 640x360 terrain frame, then a smaller 320x240 sphere frame: each
    equals the same frame synchronised by psgpu_sync and the oracle's"""
    big = S.terrain_scene(width=640, height=360, cells=80, n_spheres=8, tex_size=256)
    small = S.sphere_scene(width=320, height=240)
    for sc in (big, small):
        announce("sync_noalloc " + sc.name)
        got = _frame(sc, True)
        want = _frame(sc, False)
        _same(got, want)
        f, ost = oracle.render_scene(sc, debug=True)
        np.testing.assert_array_equal(got[0], f.colour)
        np.testing.assert_array_equal(got[1].view(np.uint32), f.depth.view(np.uint32))
        np.testing.assert_array_equal(got[2]["pass_hash"], f.pass_hash)
        assert got[3]["frags_shaded"] == ost.frags_shaded > 0


def test_noalloc_repeated_frames_on_one_context(oracle):
    """frames in a row on one context, the second after a smaller viewport:
    the array keeps the last call's contents, and nothing in it needs to
    start at zero"""
    sc = S.terrain_scene(width=480, height=270, cells=60, n_spheres=4, tex_size=128)
    announce("sync_noalloc repeated " + sc.name)
    f, _ = oracle.render_scene(sc)
    ctx = N.Context(sc.width, sc.height)
    try:
        gs = S.GpuScene(ctx, sc)
        ctx.sync()
        for _ in range(2):
            gs.frame(flush=False)
            ctx.sync_noalloc()
            np.testing.assert_array_equal(ctx.read_colour(), f.colour)
        ctx.viewport(200, 120)            # a smaller input: a corner of the same scene
        gs.frame(flush=False)
        ctx.sync_noalloc()
        part = ctx.read_colour()
        ctx.viewport(200, 120)
        gs.frame(flush=False)
        ctx.sync()
        np.testing.assert_array_equal(part, ctx.read_colour())
    finally:
        ctx.close()


def test_noalloc_refusals():
    """a flush whose pixels alone exceed the array is refused before anything
    runs and stays queued for psgpu_sync; work already submitted must be
    synchronised first"""
    announce("sync_noalloc refusals")
    sc = S.sphere_scene(width=320, height=240)
    ctx = N.Context(8192, 8192)            # a 256 MiB winner plane alone
    try:
        sc.width, sc.height = 8192, 8192
        gs = S.GpuScene(ctx, sc)
        ctx.sync()
        gs.frame(flush=False)
        with pytest.raises(MemoryError):
            ctx.sync_noalloc()
        ctx.sync()                          # the queue was left intact
        assert ctx.stats()["frags_shaded"] > 0
        gs.frame(flush=True)
        with pytest.raises(N.PsGpuStateError):
            ctx.sync_noalloc()
        ctx.sync()
    finally:
        ctx.close()
