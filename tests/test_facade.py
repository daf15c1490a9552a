"""The C++ PuresoftPipeline facade (puresoft3d_amd/host) and its psrender
program: scenes rendered through the reference-shaped API must equal the
oracle's frames bit for bit, as the C-ABI path does."""
import os
import subprocess

import numpy as np
import pytest

from conftest import announce
import scenegen

from puresoft3d_amd import native as N
from puresoft3d_amd import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "puresoft3d_amd")
PSRENDER = os.path.join(PKG, "host", "psrender")
FACADE = os.path.join(PKG, "libpuresoft3d.so")

# PuresoftPipeline members the facade keeps from pipeline.h:24-156
FACADE_METHODS = ["createTexture", "destroyTexture", "getTexture", "addProcessor", "destroyProcessor",
                  "createProgramme", "destroyProgramme", "useProgramme", "createVAO", "attachVBO", "detachVBO",
                  "getVBO", "destroyVAO", "setViewport", "setDepth", "setFBO", "setUniform", "drawVAO",
                  "swapBuffers", "enable", "disable", "clearDepth", "clearColour", "postProcess", "setRenderer",
                  "saveTexture", "saveTextureAsRaw", "getTaskQCounters"]


def _built():
    if not (os.path.exists(PSRENDER) and os.path.exists(FACADE)):
        pytest.skip("facade not built (python -m puresoft3d_amd.build)")


def test_facade_exports_pipeline_api():
    _built()
    out = subprocess.run(["nm", "-DC", "--defined-only", FACADE], capture_output=True, text=True, check=True).stdout
    for m in FACADE_METHODS:
        assert "PuresoftPipeline::%s(" % m in out, m


def test_psrender_usage():
    _built()
    r = subprocess.run([PSRENDER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def _read_scene_header(path):
    raw = open(path, "rb").read()
    magic, ver, w, h, cd, cc = np.frombuffer(raw[:24], "<u4").tolist()
    return magic, ver, w, h, raw


def test_scene_file_layout(tmp_path):
    sc = scenegen.micro_scene(N.PROC_BLINN, 1)
    p = str(tmp_path / "s.pssc")
    S.save_scene(sc, p)
    magic, ver, w, h, raw = _read_scene_header(p)
    assert (magic, ver, w, h) == (S.SCENE_MAGIC, 1, sc.width, sc.height)
    tex = sum(20 + t.layers * t.width * t.height * 4 for t in sc.textures)
    mesh = sum(4 + sum(12 + a.size * 4 for a in m.values()) for m in sc.meshes.values())
    uni = 4 + sum(8 + np.asarray(v).nbytes for v in sc.uniforms.values())
    draws = 4 + sum(16 + sum(8 + np.asarray(v).nbytes for v in d.uniforms.values()) for d in sc.draws)
    assert len(raw) == 24 + 4 + tex + 4 + mesh + uni + draws


def _psrender(scene, tmp_path, frames=1, dump=None):
    p = str(tmp_path / (scene.name + ".pssc"))
    o = str(tmp_path / (scene.name + ".bgra"))
    S.save_scene(scene, p)
    announce("psrender %s frames=%d" % (scene.name, frames))
    r = subprocess.run([PSRENDER, p, o, str(frames)] + ([dump] if dump else []), capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stderr
    if dump:
        assert "task queue counters 0 0 0 0" in r.stdout
    return np.fromfile(o, np.uint8).reshape(scene.height, scene.width, 4)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["sphere", "blinn", "cube", "behaviors", "blends"])
def test_psrender_matches_oracle(oracle, tmp_path, which):
    _built()
    sc = {"sphere": lambda: S.sphere_scene(width=320, height=240),
          "blinn": lambda: scenegen.micro_scene(N.PROC_BLINN, 3),
          "cube": lambda: scenegen.micro_scene(N.PROC_DEF04, 3),
          "behaviors": lambda: scenegen.behavior_scene(4),
          "blends": lambda: scenegen.blend_scene(5)}[which]()
    got = _psrender(sc, tmp_path, frames=2)
    f, _ = oracle.render_scene(sc)
    np.testing.assert_array_equal(got, f.colour)


@pytest.mark.gpu
def test_debug_api_dumps(oracle, tmp_path):
    """saveTextureAsRaw (fbo.cpp:542-550) writes the display and the depth
    buffer as they lie in memory; saveTexture writes them as BMP (32-bit
    BGRA bottom row first; FP32 depth as 8-bit grey of value * 255);
    getTaskQCounters reports zeros, as the reference does without PROFILING"""
    _built()
    sc = scenegen.micro_scene(N.PROC_BLINN, 4)
    pre = str(tmp_path / "dump")
    got = _psrender(sc, tmp_path, dump=pre)
    f, _ = oracle.render_scene(sc)
    W, H = sc.width, sc.height
    raw = np.fromfile(pre + ".display.raw", np.uint8).reshape(H, W, 4)
    np.testing.assert_array_equal(raw, got)
    dep = np.fromfile(pre + ".depth.raw", np.float32).reshape(H, -1)
    np.testing.assert_array_equal(dep.view(np.uint32), f.depth.view(np.uint32))
    bmp = np.fromfile(pre + ".display.bmp", np.uint8)
    assert bytes(bmp[:2]) == b"BM" and bmp.size == 54 + W * H * 4
    np.testing.assert_array_equal(bmp[54:].reshape(H, W, 4), f.colour[::-1])
    gb = np.fromfile(pre + ".depth.bmp", np.uint8)
    line = (W + 3) // 4 * 4
    assert gb.size == 54 + 1024 + line * H
    grey = gb[54 + 1024:].reshape(H, line)[:, :W]
    want = (f.depth[:, :W] * np.float32(255.0)).astype(np.int32).astype(np.uint8)
    np.testing.assert_array_equal(grey, want)


PSDEMO = os.path.join(PKG, "host", "psdemo")


def _stub_picture(name, w, h):
    """the pixels psdemo's stub picture loader writes for a file name"""
    s = 0
    for ch in name:
        s = (s * 31 + ord(ch)) & 255
    y, x = np.mgrid[0:h, 0:w].astype(np.int64)
    img = np.empty((h, w, 4), np.uint8)
    img[..., 0] = (7 * x + 3 * y + s) & 255
    img[..., 1] = (5 * x + 11 * y + 2 * s) & 255
    img[..., 2] = (13 * x + y + 3 * s) & 255
    img[..., 3] = 255
    return img


def _psdemo_scene(W, H):
    """psdemo's frame as a scene: the skybox (DEF04, six stub pictures) and
    the textured DEF01 quad, with the same floats"""
    f32 = np.float32
    sc = S.Scene(W, H)
    sc.name = "psdemo_%dx%d" % (W, H)
    z = f32(0.999)
    sky = np.array([[-1, -1, z, 1], [1, -1, z, 1], [1, 1, z, 1], [-1, -1, z, 1], [1, 1, z, 1], [-1, 1, z, 1]], np.float32)
    a = f32(0.6)
    pos = np.array([[-a, -a, 0.2, 1], [a, -a, 0.2, 1], [a, a, 0.2, 1], [-a, -a, 0.2, 1], [a, a, 0.2, 1], [-a, a, 0.2, 1]],
                   np.float32)
    uv = ((pos[:, :2] + a) / (f32(2) * a)).astype(np.float32)
    nrm = np.tile(np.array([0, 0, 1, 0], np.float32), (6, 1))
    sc.meshes["sky"] = {0: sky}
    sc.meshes["quad"] = {0: pos, 3: nrm, 4: uv}
    faces = ["sky_rt.png", "sky_lf.png", "sky_up.png", "sky_dn.png", "sky_ft.png", "sky_bk.png"]
    sc.textures.append(S.TextureSpec([_stub_picture(n, 32, 32) for n in faces], wrap=N.CLAMP))
    sc.textures.append(S.TextureSpec([_stub_picture("quad.png", 64, 32)], wrap=N.WRAP))
    I = S.colmajor(np.eye(4))
    c, s = f32(0.87758256), f32(0.47942554)
    V = np.array([c, 0, -s, 0, 0, 1, 0, 0, s, 0, c, 0, 0, 0, 0, 1], np.float32)
    sc.draws.append(S.DrawSpec("sky", N.PROC_DEF04, uniforms={0: I, 1: V, 2: np.array([0], np.int32)}))
    sc.draws.append(S.DrawSpec("quad", N.PROC_DEF01, uniforms={3: I, 4: I, 5: I, 7: S.vec4(-3, 4, 6, 0),
                                                               8: S.vec4(0, 0, 5, 0), 9: np.array([1], np.int32)}))
    sc.clear_colour = 0x00203040
    return sc


@pytest.mark.gpu
def test_reference_resource_idioms(oracle, tmp_path):
    """psdemo: textures filled through getTexture's pointer, once for a 2D
    picture created without pixels and once per layer of a skybox created
    from a NULL image (test/scenobj.cpp:165-219), then a post-processor of
    the reference's own form (test2/testpost.h: paddb +50 over every row):
    the presented frame equals the oracle's frame plus 50 per byte"""
    if not os.path.exists(PSDEMO):
        pytest.skip("psdemo not built (python -m puresoft3d_amd.build)")
    W, H = 160, 120
    out = str(tmp_path / "demo.bgra")
    announce("psdemo %dx%d" % (W, H))
    fbo = str(tmp_path / "fbo.bgra")
    r = subprocess.run([PSDEMO, out, str(W), str(H), fbo], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.uint8).reshape(H, W, 4)
    sc = _psdemo_scene(W, H)
    f, st = oracle.render_scene(sc)
    assert st.frags_shaded > W * H // 2
    np.testing.assert_array_equal(got, (f.colour.astype(np.int64) + 50).astype(np.uint8))
    # the quad drawn again into a host FBO bound with setFBO(0, fbo): its
    # memory holds the draw over the 0x11 bytes it was filled with
    sc.draws = sc.draws[1:]
    g = oracle.Frame(W, H, depth_scanline=W, top_down=False)
    g.colour[:] = 0x11
    g.clear_depth(1.0)
    oracle.render_scene(sc, frame=g, clear=False)
    np.testing.assert_array_equal(np.fromfile(fbo, np.uint8).reshape(H, W, 4), g.colour)
