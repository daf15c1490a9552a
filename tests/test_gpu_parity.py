"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact: colour (BGRA8), depth (FP32 bits), and per pixel the last
depth-test-passing triangle, the number of rasterised fragments, the number of
depth-test passes, and xor-hashes of the rasterised and passing triangle ids.
"""
import os

import numpy as np
import pytest

from conftest import announce, oracle_view
import scenegen

pytestmark = pytest.mark.gpu

N = pytest.importorskip("puresoft3d_amd.native")
from puresoft3d_amd import scenes as S  # noqa: E402


def gpu_render(scene, tiling=None, debug=True, frames=1):
    ctx = N.Context(scene.width, scene.height)
    try:
        if tiling:
            ctx.tiling(*tiling)
        gs = S.GpuScene(ctx, scene)
        if debug:
            ctx.debug_enable(True)
        for _ in range(frames):
            if debug:
                ctx.debug_reset()
            gs.frame()
        col = ctx.read_colour()
        dep = ctx.read_depth()
        dbg = ctx.debug_read(scene.width, scene.height) if debug else None
        st = ctx.stats()
        return col, dep, dbg, st
    finally:
        ctx.close()


def assert_same(scene, oracle, tiling=None, frames=1):
    announce("%s tiling=%s" % (scene.name, tiling))
    col, dep, dbg, st = gpu_render(scene, tiling, frames=frames)
    f, ost = oracle.render_scene(scene, debug=True)
    if getattr(scene, "shadow", None) is None:   # the planes record the display pass only
        assert dbg["raster_count"].sum() == ost.frags_rasterised
    np.testing.assert_array_equal(dbg["raster_count"], f.raster_count)
    np.testing.assert_array_equal(dbg["raster_hash"], f.raster_hash)
    np.testing.assert_array_equal(dbg["pass_count"], f.pass_count)
    np.testing.assert_array_equal(dbg["pass_hash"], f.pass_hash)
    np.testing.assert_array_equal(dbg["last_pass"], f.last_pass)
    np.testing.assert_array_equal(dep.view(np.uint32), f.depth.view(np.uint32))
    np.testing.assert_array_equal(col, f.colour)
    assert st["frags_shaded"] == ost.frags_shaded * frames
    assert st["frags_rasterised"] == ost.frags_rasterised * frames
    assert st["tris_culled"] == ost.tris_culled * frames
    return st


def test_math_selftest():
    announce("selftest_math")
    rng = np.random.default_rng(0)
    ctx = N.Context(8, 8)
    try:
        a = np.concatenate([rng.uniform(-1e3, 1e3, 4000), rng.uniform(-1, 1, 2000) * 1e-38,
                            [0.0, -0.0, 1e-45, 3e38, np.inf, -np.inf, np.nan]]).astype(np.float32)
        b = np.concatenate([rng.uniform(-10, 10, 4000), rng.uniform(-1, 1, 2000),
                            [1.0, -3.0, 7.0, 1e-38, 2.0, np.inf, 1.0]]).astype(np.float32)
        with np.errstate(all="ignore"):
            np.testing.assert_array_equal(ctx.selftest_math(0, a, b).view(np.uint32), (a / b).view(np.uint32))
            sq = np.abs(a)
            np.testing.assert_array_equal(ctx.selftest_math(1, sq).view(np.uint32), np.sqrt(sq).view(np.uint32))
            np.testing.assert_array_equal(ctx.selftest_math(2, b).view(np.uint32),
                                          (np.float32(1.0) / b).view(np.uint32))
            x = np.concatenate([a, np.array([2.2e9, -2.2e9, 2147483520.0, -2147483648.0, 1e20], np.float32)])
            ref = np.where((x >= -2147483648.0) & (x < 2147483648.0), np.trunc(np.nan_to_num(x)), -2147483648.0)
            np.testing.assert_array_equal(ctx.selftest_math(3, x), ref.astype(np.float32))
            np.testing.assert_array_equal(ctx.selftest_math(4, a).view(np.uint32), np.floor(a).view(np.uint32))
    finally:
        ctx.close()


def test_wrapped_texel_pairs():
    """clampCoord's wrap (fbo.cpp:552-594) of a texel pair (v, v + 1): the
    kernels' reciprocal-based floor modulo equals the integer remainder, plus m
    when negative, for every divisor up to 8192 and dividends around the fast
    path's edges (+-2^21), the multiples of m and the int32 extremes"""
    announce("selftest wrapped texel pairs")
    rng = np.random.default_rng(3)
    ms, vs = [], []
    for m in list(range(1, 8193)) + [16383, 32767, 65535]:
        q = rng.integers(-(1 << 21) // m - 2, (1 << 21) // m + 2, 8)
        v = np.concatenate([q * m, q * m - 1, q * m + 1, rng.integers(-(1 << 22), 1 << 22, 8),
                            [0, -1, (1 << 21) - 1, -(1 << 21), 1 << 21, -(1 << 21) - 1,
                             2**31 - 1, -2**31, 2**31 - 2]])
        ms.append(np.full(v.size, m)); vs.append(v)
    m = np.concatenate(ms).astype(np.int64)
    v = np.clip(np.concatenate(vs), -2**31, 2**31 - 1).astype(np.int64)
    v1 = ((v + 1 + 2**31) % 2**32) - 2**31
    want0 = np.mod(v, m).astype(np.int32)      # Python's mod is the floor modulo
    want1 = np.mod(v1, m).astype(np.int32)
    a = v.astype(np.int32).view(np.float32)
    b = m.astype(np.int32).view(np.float32)
    ctx = N.Context(8, 8)
    try:
        got0 = ctx.selftest_math(8, a, b).view(np.int32)
        got1 = ctx.selftest_math(9, a, b).view(np.int32)
    finally:
        ctx.close()
    np.testing.assert_array_equal(got0, want0)
    np.testing.assert_array_equal(got1, want1)


def test_log2_matches_oracle(oracle):
    announce("selftest log2")
    rng = np.random.default_rng(1)
    x = np.concatenate([10.0 ** rng.uniform(-44, 38, 5000), [1.0, 2.0, 0.5, 1e-45, 0.0, np.inf]]).astype(np.float32)
    ctx = N.Context(8, 8)
    try:
        g = ctx.selftest_math(6, x)
    finally:
        ctx.close()
    o = np.array([oracle.lib().or_log2_approx(float(v)) for v in x], np.float32)
    np.testing.assert_array_equal(g.view(np.uint32), o.view(np.uint32))


def test_sphere_640x480_bitexact(oracle):
    sc = S.sphere_scene()
    st = assert_same(sc, oracle)
    assert st["frags_shaded"] > 50000


@pytest.mark.parametrize("tiling", [(16, 16, 1), (8, 8, 4), (32, 64, 2), (4, 1, 8), (64, 32, 1)])
def test_sphere_tilings_identical(oracle, tiling):
    sc = S.sphere_scene(width=320, height=240)
    assert_same(sc, oracle, tiling=tiling)


@pytest.mark.parametrize("prog", [N.PROC_DEF01, N.PROC_DEF02, N.PROC_DEF03, N.PROC_DEF04, N.PROC_DEF05,
                                  N.PROC_LAMBERT, N.PROC_BLINN, N.PROC_SHADOWED, N.PROC_CLOUD,
                                  N.PROC_CLOUD_SHADOW])
@pytest.mark.parametrize("seed", [1, 2])
def test_micro_scenes(oracle, prog, seed):
    sc = scenegen.micro_scene(prog, seed)
    assert_same(sc, oracle)


@pytest.mark.parametrize("prog", [N.PROC_DEF01, N.PROC_DEF02, N.PROC_DEF03, N.PROC_DEF04, N.PROC_DEF05,
                                  N.PROC_LAMBERT, N.PROC_BLINN, N.PROC_SHADOWED, N.PROC_CLOUD,
                                  N.PROC_CLOUD_SHADOW])
@pytest.mark.parametrize("extra", [0, 1, 2])
def test_hazard_scenes(oracle, prog, extra):
    """vertices with w at or below zero and non-finite coordinates, which the
    reference never clips (drawvao.cpp:39-63), among ordinary triangles, and
    vertex streams of 3k + 1 and 3k + 2 entries (vertthrd.cpp:25-29)"""
    sc = scenegen.hazard_scene(prog, 40 + prog, extra)
    assert_same(sc, oracle)


def test_sse_pack_selftest(oracle):
    """FP_Cloud's cvtps2dq/packusdw/packuswb tail on the device against the
    oracle's (which runs cvtps2dq itself): ties, the word and byte
    saturations, NaN and out-of-range values"""
    announce("selftest sse pack")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-300, 600, 20000), rng.uniform(-1e5, 1e5, 5000),
                        np.arange(-4, 70000, 0.5), [np.nan, np.inf, -np.inf, 3e9, -3e9, 2147483520.0, -2147483648.0,
                                                   32767.5, 32768.5, 65535.5, 255.5, 254.5]]).astype(np.float32)
    ctx = N.Context(8, 8)
    try:
        got = ctx.selftest_math(10, x)
    finally:
        ctx.close()
    want = np.array([oracle.sse_pack_byte(v) for v in x], np.float32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("seed", [1, 2])
def test_blend_scenes_bitexact(oracle, seed):
    """blended and discarding programmes between deferred Blinn-Phong draws,
    under every depth behaviour: each programme change starts a new flush and
    the raster walk keeps submission order per pixel"""
    assert_same(scenegen.blend_scene(seed), oracle)


@pytest.mark.parametrize("case", [(1e30, 1.0), (-1e30, 1.0), (1e30, 100.0)])
def test_cloud_quads_bitexact(oracle, case):
    assert_same(scenegen.cloud_quad_scene(case[0], case[1], red=170), oracle)


@pytest.mark.parametrize("red", [149, 150])
def test_cloud_shadow_quads_bitexact(oracle, red):
    assert_same(scenegen.cloud_shadow_quad(red), oracle)


def test_cloud_layer_bitexact(oracle):
    """the demo's cloud: its shadow programme in the shadow pass (discarding
    thin cloud), then the blended, lit and shadowed cloud over the shadowed
    terrain"""
    sc = S.cloud_terrain_scene()
    assert_same(sc, oracle)


def test_extra_fbo_slots_are_recorded_only(oracle):
    """setFBO on slots 1..15 binds as the reference binds (pipeline.cpp:260-275)
    and nothing writes there: the texture keeps its pixels, the frame is the
    oracle's; a bad texture index or slot 16 is refused"""
    sc = scenegen.micro_scene(N.PROC_BLINN, 3)
    announce("extra FBO slots " + sc.name)
    ctx = N.Context(sc.width, sc.height)
    try:
        gs = S.GpuScene(ctx, sc)
        px = np.arange(16 * 16 * 4, dtype=np.uint32).astype(np.uint8).reshape(16, 16, 4)
        t = ctx.texture_create(16, 16, 4, pixels=px)
        for slot in (1, 7, 15):
            ctx.target(t, slot)
        ctx.target(N.DISPLAY, 7)   # NULL: unbound again
        for bad in ((t, 16), (t, -1), (t + 100, 3)):
            with pytest.raises(IndexError):   # PSGPU_ERR_RANGE
                ctx.target(*bad)
        # texel addressing limits: 2^22 texels a side, rows below 2^24 bytes
        for w, h, scan in (((1 << 22) + 1, 1, None), (1, (1 << 22) + 1, None), (4, 4, 1 << 24)):
            with pytest.raises(IndexError):   # PSGPU_ERR_RANGE, refused before any allocation
                ctx.texture_create(w, h, 4, scanline=scan)
        gs.frame()
        col = ctx.read_colour()
        np.testing.assert_array_equal(ctx.texture_download(t), px.reshape(16, 64))
    finally:
        ctx.close()
    f, _ = oracle.render_scene(sc)
    np.testing.assert_array_equal(col, f.colour)


_SWEEP_PROGS = [N.PROC_DEF01, N.PROC_DEF02, N.PROC_DEF03, N.PROC_DEF04, N.PROC_DEF05, N.PROC_LAMBERT, N.PROC_BLINN,
                N.PROC_SHADOWED, N.PROC_CLOUD, N.PROC_CLOUD_SHADOW]


@pytest.mark.parametrize("seed", range(int(os.environ.get("PS_SWEEP_SEEDS", "30"))))
def test_random_scene_sweep(oracle, seed):
    """every programme over seeded random viewports (odd sizes included),
    triangle counts and raster tilings (PS_SWEEP_SEEDS widens the sweep)"""
    rng = np.random.default_rng(1000 + seed)
    prog = _SWEEP_PROGS[seed % len(_SWEEP_PROGS)]
    w, h = int(rng.integers(17, 230)), int(rng.integers(9, 170))
    if w % 4 == 1:   # the reference's default depth scanline would be shorter than the row
        w += 1
    n = int(rng.integers(20, 400))
    if seed % 3 == 2:    # a third of the sweep carries the unclipped-vertex hazards and a partial triangle
        sc = scenegen.hazard_scene(prog, 100 + seed, extra_vertices=seed % 5 % 3, width=w, height=h, n=n)
    else:
        sc = scenegen.micro_scene(prog, 100 + seed, width=w, height=h, n=n)
    cw, tr, wv = int(2 ** rng.integers(0, 7)), int(2 ** rng.integers(0, 7)), int(2 ** rng.integers(0, 4))
    while wv > 1 and wv * (cw * 128 + 128 * 7 + 128) * 4 > 65536:   # the raster kernel's LDS per group
        wv //= 2
    assert_same(sc, oracle, tiling=(cw, tr, wv))


def test_large_flush_of_rowless_triangles(oracle):
    """a flush of more than 2^20 triangles, most of them sub-pixel and without
    rows (1.1 M triangles of a dense sphere at 320x240): k_span's search
    starts from the row scan's table of every 64th row item's triangle and
    stops at the next entry, however long the runs of row-less triangles"""
    sc = S.sphere_scene(width=320, height=240, tex_size=64)
    sc.meshes["sphere"] = S.uv_sphere(600, 920, 1.0, seed=1)
    sc.name = "uv_sphere_1.1M_320x240"
    assert sc.triangles() > (1 << 20)
    assert_same(sc, oracle)


def test_terrain_reduced_bitexact(oracle):
    sc = S.terrain_scene(width=480, height=270, cells=60, n_spheres=8, tex_size=256)
    assert_same(sc, oracle)


def test_terrain_1080p_bitexact(oracle):
    """the benchmark workload itself: 1920x1080, 200,480 triangles, 2048^2
    trilinear albedo and normal maps, default 8x32 tiles (8,100 tiles: 13-bit
    keys, two sort passes and the tile-start search)"""
    sc = S.terrain_scene()
    st = assert_same(sc, oracle)
    assert st["frags_shaded"] > 1_000_000


def test_shadow_map_two_passes_bitexact(oracle):
    """a DEF05 pass renders the terrain and spheres from the light into a
    256^2 FP32 depth texture; the display pass samples it through the
    projective sampler (PSGPU_PROC_SHADOWED).  The draw that samples the map
    flushes the pass that renders it first."""
    sc = S.shadow_terrain_scene()
    st = assert_same(sc, oracle)
    assert st["frags_shaded"] > 50000


@pytest.mark.parametrize("which", ["sphere", "plane"])
def test_objx_assets_bitexact(oracle, which):
    """the reference's own meshes (src/test/sphere.objx: 528 triangles;
    src/test2/plane.objx: 13 meshes, 10,414 triangles, a third of the jar's
    tangents NaN as exported), loaded as the demo loader prepares them"""
    sc = scenegen.objx_scene(which)
    st = assert_same(sc, oracle)
    assert st["frags_shaded"] > 10000


@pytest.mark.parametrize("which", ["plane", "sphere"])
def test_objx_assets_with_pictures_bitexact(oracle, which):
    """the reference's own meshes with its own pictures (decoded by PIL as
    its loader would leave them): plane.objx with marmite.jpg, top.jpg and
    penhold.jpg (tex1light1 and colr1light1), sphere.objx as the demo's
    Earth (earth.jpg and earth.dot3.png, tex1bump1light1)"""
    pytest.importorskip("PIL")
    assert_same(scenegen.objx_scene(which, 640, 480, pictures=True), oracle)


def test_skybox_pictures_bitexact(oracle):
    """the demo's purplenebula skybox through DEF04"""
    pytest.importorskip("PIL")
    assert_same(scenegen.skybox_scene(), oracle)


def test_two_frames_accumulate(oracle):
    sc = S.sphere_scene(width=200, height=150)
    assert_same(sc, oracle, frames=2)


def test_mips_match_oracle(oracle):
    announce("mips")
    spec = S.TextureSpec([S.noise_texture(40, 3)], levels=S.mip_levels(40, 40))
    spec2 = S.TextureSpec([np.random.default_rng(5).integers(0, 256, (37, 23, 4), dtype=np.uint8)],
                          levels=S.mip_levels(23, 37))
    ctx = N.Context(8, 8)
    try:
        for sp in (spec, spec2):
            t = ctx.texture_create(sp.width, sp.height, 4, mip_levels=sp.levels, pixels=sp.images[0])
            ctx.texture_generate_mips(t)
            ot = oracle.OracleTexture(sp)
            for l in range(sp.levels):
                np.testing.assert_array_equal(ctx.texture_download(t, 0, l).ravel(), ot.level(0, l))
    finally:
        ctx.close()


def test_graph_replay_matches_eager(oracle):
    """a recorded frame replayed as a hipGraph gives the same image, depth and
    fragment planes as the eager path, replay after replay"""
    sc = S.terrain_scene(width=320, height=180, cells=40, n_spheres=4, tex_size=128)
    announce("graph replay " + sc.name)
    f, ost = oracle.render_scene(sc, debug=True)
    ctx = N.Context(sc.width, sc.height)
    try:
        gs = S.GpuScene(ctx, sc)
        ctx.debug_enable(True)
        ctx.capture_begin()
        gs.frame()
        g = ctx.capture_end()          # runs the recording once
        for rep in range(3):
            ctx.debug_reset()
            ctx.graph_launch(g, 1)
            col, dep = ctx.read_colour(), ctx.read_depth()
            dbg = ctx.debug_read(sc.width, sc.height)
            np.testing.assert_array_equal(col, f.colour)
            np.testing.assert_array_equal(dep.view(np.uint32), f.depth.view(np.uint32))
            np.testing.assert_array_equal(dbg["pass_hash"], f.pass_hash)
            np.testing.assert_array_equal(dbg["last_pass"], f.last_pass)
        ctx.reset_stats()
        ctx.graph_launch(g, 5)
        ctx.sync()
        assert ctx.stats()["frags_shaded"] == 5 * ost.frags_shaded
    finally:
        ctx.close()


def test_capacity_growth_replays_log(oracle):
    """tiny initial scratch buffers overflow; the sync replays the log and the
    result is unchanged"""
    sc = S.terrain_scene(width=640, height=360, cells=120, n_spheres=8, tex_size=128)
    announce("capacity growth " + sc.name)
    f, ost = oracle.render_scene(sc)
    ctx = N.Context(sc.width, sc.height, flags=1)
    try:
        gs = S.GpuScene(ctx, sc)
        ctx.clear_colour(0x11223344)
        gs.frame()
        gs.frame()
        col, dep = ctx.read_colour(), ctx.read_depth()
        np.testing.assert_array_equal(col, f.colour)
        np.testing.assert_array_equal(dep.view(np.uint32), f.depth.view(np.uint32))
        st = ctx.stats()
        assert st["retries"] >= 1
        assert st["frags_shaded"] == 2 * ost.frags_shaded
    finally:
        ctx.close()


def test_batched_views_match_oracle_and_single(oracle):
    """views of several sizes over two scenes, rendered in one flush, equal
    the oracle's per-view render and the GPU's one-view-per-flush render"""
    assets = S.batch_assets(n_scenes=2, seed=3, tri_lo=2000, tri_hi=20000, tex_size=128)
    views = S.batch_views(assets, n_views=7, seed=4, side_lo=48, side_hi=200)
    announce("batched views %s" % [(v.width, v.height, v.scene) for v in views])
    ctx = N.Context(64, 64)
    try:
        b = S.GpuBatch(ctx, assets, views)
        b.submit()
        ctx.sync()
        st = ctx.stats()
        assert st["flushes"] == 1
        got = [b.read(i) for i in range(len(views))]
        for i in range(len(views)):       # one view per flush
            b.submit([i])
            ctx.sync()
            col, dep = b.read(i)
            np.testing.assert_array_equal(col, got[i][0])
            np.testing.assert_array_equal(dep.view(np.uint32), got[i][1].view(np.uint32))
    finally:
        ctx.close()
    for i in range(len(views)):
        sc = b.view_scene(i)
        f, _ = oracle_view(oracle, sc, 0x00302010)
        np.testing.assert_array_equal(got[i][0], f.colour)
        np.testing.assert_array_equal(got[i][1].view(np.uint32), f.depth.view(np.uint32))


def test_batched_views_graph_replay(oracle):
    assets = S.batch_assets(n_scenes=3, seed=5, tri_lo=1000, tri_hi=8000, tex_size=64)
    views = S.batch_views(assets, n_views=12, seed=6, side_lo=32, side_hi=128)
    announce("batched graph %d views" % len(views))
    ctx = N.Context(64, 64)
    try:
        b = S.GpuBatch(ctx, assets, views)
        ctx.capture_begin()
        b.submit()
        g = ctx.capture_end()
        ref = [b.read(i) for i in range(len(views))]
        for k in range(len(views)):           # dirty every target, then replay
            ctx.target(b.targets[k][0])
            ctx.clear_target(0xFFFFFFFF)
        ctx.target(N.DISPLAY)
        ctx.graph_launch(g, 2)
        ctx.sync()
        for i in range(len(views)):
            col, dep = b.read(i)
            np.testing.assert_array_equal(col, ref[i][0])
            np.testing.assert_array_equal(dep.view(np.uint32), ref[i][1].view(np.uint32))
        # another tiling (8x8 tiles: many more bin segments) recaptures the graph
        for tiling in [(1, 8, 2), (64, 64, 1), (4, 16, 8)]:
            ctx.tiling(*tiling)
            for k in range(len(views)):
                ctx.target(b.targets[k][0])
                ctx.clear_target(0xFFFFFFFF)
            ctx.target(N.DISPLAY)
            ctx.graph_launch(g, 1)
            ctx.sync()
            for i in range(len(views)):
                col, dep = b.read(i)
                np.testing.assert_array_equal(col, ref[i][0])
                np.testing.assert_array_equal(dep.view(np.uint32), ref[i][1].view(np.uint32))
    finally:
        ctx.close()
    for i in (0, len(views) - 1):
        f, _ = oracle_view(oracle, b.view_scene(i), 0x00302010)
        np.testing.assert_array_equal(ref[i][0], f.colour)


@pytest.mark.parametrize("seed", [1, 2])
def test_behaviors_bitexact(oracle, seed):
    assert_same(scenegen.behavior_scene(seed), oracle)


@pytest.mark.parametrize("prog", [N.PROC_DEF02, N.PROC_DEF03, N.PROC_BLINN])
@pytest.mark.parametrize("tiling", [None, (4, 16, 2), (64, 64, 1)])
def test_long_spans_bitexact(oracle, prog, tiling):
    """spans of hundreds of pixels: the per-chunk pieces start from the
    accumulators of the serial walk, for depth and for every varying"""
    sc = scenegen.big_triangle_scene(prog, 5)
    st = assert_same(sc, oracle, tiling=tiling)
    assert st["segments"] > 4 * st["spans"] // 3


@pytest.mark.parametrize("prog", [N.PROC_DEF02, N.PROC_BLINN])
def test_sliver_extrapolated_spans_bitexact(oracle, prog):
    sc = scenegen.sliver_scene(prog, 9)
    assert_same(sc, oracle)


def _hip_runtime():
    """the HIP runtime libpsgpu is linked against, for a post-processor's own work"""
    import ctypes as C
    L = C.CDLL("libamdhip64.so.7")
    L.hipMemset2DAsync.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_size_t, C.c_size_t, C.c_void_p]
    L.hipMemset2DAsync.restype = C.c_int
    return L


@pytest.mark.parametrize("tiny", [False, True])
def test_post_process_runs_in_stream_order(oracle, tiny):
    """PuresoftPipeline::postProcess (post.cpp:3-18): the post-processor sees
    the frame after the draws before it, its work lands before later commands,
    and with scratch buffers too small at first (tiny) the replayed flush is
    followed by the post-processor again"""
    sc = S.terrain_scene(width=320, height=180, cells=60, n_spheres=4, tex_size=128) if tiny else \
        S.sphere_scene(width=200, height=150)
    announce("post-process %s tiny=%s" % (sc.name, tiny))
    hip = _hip_runtime()
    calls = []

    def band(frame, stream):    # rows 40..59, columns 0..99 of the colour target := 0x7f bytes
        calls.append((frame.width, frame.height, frame.colour_top_down))
        return hip.hipMemset2DAsync(frame.colour + 40 * frame.colour_pitch, frame.colour_pitch, 0x7F, 100 * 4, 20,
                                    stream)

    f, _ = oracle.render_scene(sc)
    want = f.colour.reshape(sc.height, sc.width, 4).copy()
    want[40:60, :100] = 0x7F
    ctx = N.Context(sc.width, sc.height, flags=1 if tiny else 0)
    try:
        gs = S.GpuScene(ctx, sc)
        gs.frame()
        ctx.post_process(band)
        col = ctx.read_colour().reshape(sc.height, sc.width, 4)
        np.testing.assert_array_equal(col, want)
        assert calls[0] == (sc.width, sc.height, 1)
        if tiny:
            assert ctx.stats()["retries"] >= 1
        assert len(calls) == 1   # after an overflow it waits for the replay instead of running twice
        # a clear after it overwrites the band: later commands come after its work
        ctx.clear_colour(0x00112233)
        col = ctx.read_colour().reshape(sc.height, sc.width, 4)
        assert np.all(col[40:60, :100].view(np.uint32) == 0x00112233)
    finally:
        ctx.close()


class _ShiftUp:
    """A post-processor that is not idempotent: every memory row of the colour
    target but the last takes the row after it (through a device copy of the
    rows), so running it twice moves the frame by two rows."""

    def __init__(self, hip, width, height):
        import ctypes as C
        self.C, self.hip, self.calls = C, hip, 0
        hip.hipMemcpy2DAsync.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.c_int, C.c_void_p]
        hip.hipMemcpy2DAsync.restype = C.c_int
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipMalloc.restype = C.c_int
        hip.hipFree.argtypes = [C.c_void_p]
        self.tmp = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.tmp), width * 4 * height) == 0

    def __call__(self, frame, stream):
        self.calls += 1
        w, p, h = frame.width * 4, frame.colour_pitch, frame.height
        d2d = 3   # hipMemcpyDeviceToDevice
        rc = self.hip.hipMemcpy2DAsync(self.tmp, w, frame.colour + p, p, w, h - 1, d2d, stream)
        return rc or self.hip.hipMemcpy2DAsync(frame.colour, p, self.tmp, w, w, h - 1, d2d, stream)

    def close(self):
        self.hip.hipFree(self.tmp)


def _shifted(img, times=1):
    out = img.copy()
    for _ in range(times):
        out[:-1] = out[1:].copy()
    return out


@pytest.mark.parametrize("tiny", [False, True])
def test_post_process_in_place_runs_once(oracle, tiny):
    """an in-place post-processor (like the reference's PP_DepthofField, which
    adds to every byte, test2/testpost.cpp:9-44) is applied exactly once, also
    when the flush before it first overflowed its scratch buffers (tiny) and
    the log replay re-ran that flush"""
    sc = S.terrain_scene(width=320, height=180, cells=60, n_spheres=4, tex_size=128)
    announce("post-process in place %s tiny=%s" % (sc.name, tiny))
    f, _ = oracle.render_scene(sc)
    want = _shifted(f.colour.reshape(sc.height, sc.width, 4))
    ctx = N.Context(sc.width, sc.height, flags=1 if tiny else 0)
    shift = _ShiftUp(_hip_runtime(), sc.width, sc.height)
    try:
        gs = S.GpuScene(ctx, sc)
        gs.frame()
        ctx.post_process(shift)
        col = ctx.read_colour().reshape(sc.height, sc.width, 4)
        if tiny:
            assert ctx.stats()["retries"] >= 1
        assert shift.calls == 1
        np.testing.assert_array_equal(col, want)
    finally:
        ctx.close()
        shift.close()


def test_stale_graph_launch_runs_once(oracle):
    """a graph of draws and an in-place post-processor, launched once after a
    VBO it draws was updated: the eager run that re-sizes and re-captures it
    is that launch, so the frame is drawn and shifted once, not twice"""
    sc = S.sphere_scene(width=200, height=150)
    announce("stale graph launch " + sc.name)
    f, _ = oracle.render_scene(sc)
    want = _shifted(f.colour.reshape(sc.height, sc.width, 4))
    ctx = N.Context(sc.width, sc.height)
    shift = _ShiftUp(_hip_runtime(), sc.width, sc.height)
    try:
        gs = S.GpuScene(ctx, sc)
        ctx.clear_depth(sc.clear_depth)
        ctx.clear_colour(sc.clear_colour)
        ctx.flush()
        ctx.capture_begin()
        gs.frame(clear=False)
        ctx.post_process(shift)
        g = ctx.capture_end()                 # runs the recording once
        np.testing.assert_array_equal(ctx.read_colour().reshape(sc.height, sc.width, 4), want)
        ctx.clear_depth(sc.clear_depth)
        ctx.clear_colour(sc.clear_colour)
        vbo = ctx.vao_get(gs.vao["sphere"], 0)
        ctx.vbo_update(vbo, sc.meshes["sphere"][0])   # same data: the graph is stale all the same
        ctx.graph_launch(g, 1)
        np.testing.assert_array_equal(ctx.read_colour().reshape(sc.height, sc.width, 4), want)
        # a current graph replays `count` times: the second replay draws nothing
        # new (the depth test fails on equal depth) and shifts once more
        ctx.graph_launch(g, 1)
        np.testing.assert_array_equal(ctx.read_colour().reshape(sc.height, sc.width, 4), _shifted(want))
    finally:
        ctx.close()
        shift.close()


def test_graph_after_vbo_update_grows_scratch(oracle):
    """a graph captured for small geometry, launched after its VBOs were
    updated to geometry with many more rows (scratch buffers tiny at first):
    it is re-sized and captured again, and renders the new geometry exactly"""
    small = S.sphere_scene(width=320, height=240)
    small.meshes["sphere"] = S.uv_sphere(32, 32, 0.04, seed=1)
    big = S.sphere_scene(width=320, height=240)
    announce("graph after vbo update " + big.name)
    f, ost = oracle.render_scene(big)
    ctx = N.Context(big.width, big.height, flags=1)
    try:
        gs = S.GpuScene(ctx, small)
        ctx.capture_begin()
        gs.frame()
        g = ctx.capture_end()
        ctx.sync()
        for slot, arr in big.meshes["sphere"].items():
            ctx.vbo_update(ctx.vao_get(gs.vao["sphere"], slot), arr)
        ctx.reset_stats()
        ctx.graph_launch(g, 2)
        col, dep = ctx.read_colour(), ctx.read_depth()
        np.testing.assert_array_equal(col, f.colour)
        np.testing.assert_array_equal(dep.view(np.uint32), f.depth.view(np.uint32))
        st = ctx.stats()
        assert st["retries"] >= 1
        assert st["frags_shaded"] == 2 * ost.frags_shaded
        ctx.graph_launch(g, 1)                 # current again: a plain replay
        np.testing.assert_array_equal(ctx.read_colour(), f.colour)
    finally:
        ctx.close()


def test_view_size_limits():
    """viewports wider than 65,535 pixels or taller than 2^24 - 1 rows are
    refused: intermediate records keep a column in 16 bits and a row in 24"""
    with pytest.raises(N.PsGpuError):
        N.Context(65536, 1)
    ctx = N.Context(64, 48)
    try:
        with pytest.raises(IndexError):       # PSGPU_ERR_RANGE
            ctx.viewport(65536, 8)
        with pytest.raises(IndexError):
            ctx.viewport(8, 1 << 24)
        ctx.viewport(65535, 8)
        ctx.viewport(64, 48)
    finally:
        ctx.close()


@pytest.mark.parametrize("tiny", [False, True])
def test_flushes_beyond_the_span_limit_are_split(oracle, tiny):
    """a flush that needs more spans than one flush may hold is split into
    flushes of about half its triangles (cutting a draw where needed) and
    re-run in order, with the same colour, depth and fragment planes
    (PSGPU_CREATE_SMALL_FLUSHES lowers the limit from 2^26 to 4,096 spans)"""
    sc = S.terrain_scene(width=640, height=360, cells=60, n_spheres=6, tex_size=128)
    announce("split flushes %s tiny=%s" % (sc.name, tiny))
    f, ost = oracle.render_scene(sc, debug=True)
    assert ost.spans > 3 * 4096
    ctx = N.Context(sc.width, sc.height, flags=2 | (1 if tiny else 0))
    try:
        gs = S.GpuScene(ctx, sc)
        ctx.debug_enable(True)
        ctx.debug_reset()
        gs.frame()
        col, dep = ctx.read_colour(), ctx.read_depth()
        dbg = ctx.debug_read(sc.width, sc.height)
        st = ctx.stats()
        assert st["splits"] >= 2
        np.testing.assert_array_equal(col, f.colour)
        np.testing.assert_array_equal(dep.view(np.uint32), f.depth.view(np.uint32))
        for k in ("last_pass", "pass_count", "pass_hash", "raster_count", "raster_hash"):
            np.testing.assert_array_equal(dbg[k], getattr(f, k))
        assert st["frags_shaded"] == ost.frags_shaded and st["frags_rasterised"] == ost.frags_rasterised
        # a recorded command list whose flush would need splitting is refused
        ctx.capture_begin()
        gs.frame()
        with pytest.raises(MemoryError):
            ctx.capture_end()
    finally:
        ctx.close()
