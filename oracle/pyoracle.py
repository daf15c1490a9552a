"""ctypes binding of the CPU oracle (oracle/libpsoracle.so) and, when it has
been built, of the reference rasterizer harness (oracle/_ref/libref_raster.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsoracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_raster.so")

MAX_SLOTS = 16
MAX_UNIFORMS = 1024

UPDATE_DEPTH, TEST_DEPTH, FACE_CULLING, ALPHABLEND = 1, 2, 4, 8
DEFAULT_BEHAVIOR = UPDATE_DEPTH | TEST_DEPTH | FACE_CULLING


class Texture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("scanline", C.c_int),
                ("elem_len", C.c_int), ("layers", C.c_int), ("wrap", C.c_int),
                ("levels", C.c_int), ("data", C.c_void_p)]


class Target(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int),
                ("colour", C.c_void_p), ("colour_scanline", C.c_int),
                ("colour_top_down", C.c_int),
                ("depth", C.c_void_p), ("depth_scanline", C.c_int),
                ("dbg_last_pass", C.c_void_p), ("dbg_raster_count", C.c_void_p),
                ("dbg_raster_hash", C.c_void_p), ("dbg_pass_count", C.c_void_p),
                ("dbg_pass_hash", C.c_void_p)]


class Draw(C.Structure):
    _fields_ = [("prog", C.c_int), ("behavior", C.c_int),
                ("slot_data", C.c_void_p * MAX_SLOTS),
                ("slot_unit_bytes", C.c_int * MAX_SLOTS),
                ("slot_count", C.c_int * MAX_SLOTS),
                ("uniforms", C.c_void_p * MAX_UNIFORMS),
                ("textures", C.c_void_p), ("n_textures", C.c_int),
                ("tri_base", C.c_int), ("shade_fp64", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("tris_in", C.c_int64), ("tris_culled", C.c_int64),
                ("tris_zrejected", C.c_int64), ("tris_dropped", C.c_int64),
                ("spans", C.c_int64), ("frags_rasterised", C.c_int64),
                ("frags_shaded", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        f = C.POINTER(C.c_float)
        i = C.POINTER(C.c_int)
        L.or_draw_vao.argtypes = [C.POINTER(Draw), C.POINTER(Target), C.POINTER(Stats), C.c_int]
        L.or_draw_vao.restype = C.c_int
        L.or_clear_depth.argtypes = [C.POINTER(Target), C.c_float]
        L.or_clear_colour.argtypes = [C.POINTER(Target), C.c_uint32]
        L.or_push_triangle.argtypes = [C.c_int, C.c_int, f, f, f, i, i, i, f]
        L.or_push_triangle.restype = C.c_int
        L.or_interp_setup.argtypes = [f, f, f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, f, f, f, f]
        L.or_is_back_face.argtypes = [f, f, f]
        L.or_is_back_face.restype = C.c_int
        L.or_build_mips.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_sample_nearest.argtypes = [C.POINTER(Texture), C.c_int, C.c_float, C.c_float, C.c_void_p]
        L.or_sample_cube.argtypes = [C.POINTER(Texture), f, C.c_void_p]
        L.or_sample_bilinear.argtypes = [C.POINTER(Texture), C.c_int, C.c_int, C.c_float, C.c_float, f]
        L.or_log2_approx.argtypes = [C.c_float]
        L.or_log2_approx.restype = C.c_float
        L.or_sample_projection.argtypes = [C.POINTER(Texture), f]
        L.or_sample_projection.restype = C.c_float
        L.or_blend4.argtypes = [C.c_void_p, C.c_void_p]
        L.or_blend4.restype = None
        L.or_sse_pack_byte.argtypes = [C.c_float]
        L.or_sse_pack_byte.restype = C.c_uint8
        L.or_tri_hash.argtypes = [C.c_uint32]
        L.or_tri_hash.restype = C.c_uint32
        L.or_pool_shutdown.argtypes = []
        L.or_pool_shutdown.restype = None
        vp = C.c_void_p
        L.or_diff_forward_f64.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, C.c_double, C.c_double,
                                          C.c_double, vp, vp, vp]
        L.or_diff_forward_f64.restype = None
        L.or_diff_backward_f64.argtypes = [C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, C.c_double, C.c_double,
                                           C.c_double, vp, vp, vp, vp, vp]
        L.or_diff_backward_f64.restype = None
        _lib = L
    return _lib


def ref():
    """The reference rasterizer built from the reference sources, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        f = C.POINTER(C.c_float)
        i = C.POINTER(C.c_int)
        L.ref_push_triangle.argtypes = [C.c_int, C.c_int, f, f, f, i, i, i, f]
        L.ref_push_triangle.restype = C.c_int
        _ref = L
    return _ref


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def push_triangle(fn, width, height, v0, v1, v2, sentinel=True):
    """Run a pushTriangle implementation; returns (ok, first, last, rows[H,8], verts[6])."""
    v = [np.ascontiguousarray(np.asarray(x, np.float32).reshape(4)) for x in (v0, v1, v2)]
    rows = np.full((height, 8), 0x7FFFFFFF, np.int32)
    rows[:, 2:4] = -7
    rows[:, 6:8] = -7
    first = C.c_int()
    last = C.c_int()
    verts = np.zeros(6, np.float32)
    ok = fn(width, height, _fp(v[0]), _fp(v[1]), _fp(v[2]), C.byref(first), C.byref(last), _ip(rows), _fp(verts))
    return int(ok), first.value, last.value, rows, verts


def texture_struct(arr, wrap=0, levels=1, layers=1, width=None, height=None, scanline=None):
    """arr: uint8 buffer holding all layers and levels."""
    t = Texture()
    t.width = width
    t.height = height
    t.scanline = scanline if scanline is not None else width * 4
    t.elem_len = 4
    t.layers = layers
    t.wrap = wrap
    t.levels = levels
    t.data = arr.ctypes.data
    return t


class Frame:
    """Host-side render target for the oracle: BGRA8 display (top-down rows)
    plus the default FP32 depth buffer (pipeline.cpp:31) and debug planes."""

    def __init__(self, width, height, debug=False, depth_scanline=None, top_down=True):
        self.width, self.height = width, height
        if depth_scanline is None:
            depth_scanline = int(width / 4.0 + 0.5) * 4
        if depth_scanline < width:
            raise ValueError("default depth scanline shorter than the width (width % 4 == 1)")
        self.colour = np.zeros((height, width, 4), np.uint8)
        self.depth = np.zeros((height, depth_scanline), np.float32)
        self.t = Target()
        self.t.width, self.t.height = width, height
        self.t.colour = self.colour.ctypes.data
        self.t.colour_scanline = width * 4
        self.t.colour_top_down = 1 if top_down else 0
        self.t.depth = self.depth.ctypes.data
        self.t.depth_scanline = depth_scanline
        self.debug = debug
        if debug:
            n = width * height
            self.last_pass = np.full(n, -1, np.int32)
            self.raster_count = np.zeros(n, np.int32)
            self.raster_hash = np.zeros(n, np.uint32)
            self.pass_count = np.zeros(n, np.int32)
            self.pass_hash = np.zeros(n, np.uint32)
            self.t.dbg_last_pass = self.last_pass.ctypes.data
            self.t.dbg_raster_count = self.raster_count.ctypes.data
            self.t.dbg_raster_hash = self.raster_hash.ctypes.data
            self.t.dbg_pass_count = self.pass_count.ctypes.data
            self.t.dbg_pass_hash = self.pass_hash.ctypes.data

    def clear_depth(self, v=1.0):
        lib().or_clear_depth(C.byref(self.t), v)

    def clear_colour(self, bgra=0):
        lib().or_clear_colour(C.byref(self.t), bgra)

    def depth_image(self):
        return self.depth[:, : self.width]


class DrawCall:
    """Keeps numpy buffers alive while the oracle reads them."""

    def __init__(self, prog, slots, uniforms, textures=(), behavior=DEFAULT_BEHAVIOR, tri_base=0, shade_fp64=0):
        self.keep = []
        d = Draw()
        d.prog = prog
        d.behavior = behavior
        for s, arr in slots.items():
            arr = np.ascontiguousarray(arr, np.float32)
            self.keep.append(arr)
            n = arr.shape[0]
            unit = arr.nbytes // n
            d.slot_data[s] = arr.ctypes.data
            d.slot_unit_bytes[s] = unit
            d.slot_count[s] = n
        for k, v in uniforms.items():
            a = np.ascontiguousarray(v)
            self.keep.append(a)
            d.uniforms[k] = a.ctypes.data
        self.tex_structs = (Texture * max(1, len(textures)))()
        for i, t in enumerate(textures):
            self.tex_structs[i] = t.struct()
            self.keep.append(t)
        d.textures = C.cast(self.tex_structs, C.c_void_p)
        d.n_textures = len(textures)
        d.tri_base = tri_base
        d.shade_fp64 = shade_fp64
        self.d = d

    def run(self, frame, n_threads=1, stats=None):
        st = stats if stats is not None else Stats()
        rc = lib().or_draw_vao(C.byref(self.d), C.byref(frame.t), C.byref(st), n_threads)
        if rc != 0:
            raise RuntimeError("oracle draw failed: %d" % rc)
        return st


# ---------------------------------------------------------------- scenes
class OracleTexture:
    """Host texture in the oracle's layout: layer-major, level 0 then the tightly
    packed mip levels (built by or_build_mips)."""

    def __init__(self, spec):
        w, h, layers, levels = spec.width, spec.height, spec.layers, spec.levels
        sizes = [max(1, w >> l) * max(1, h >> l) * 4 for l in range(levels)]
        per = sum(sizes)
        self.data = np.zeros(per * layers, np.uint8)
        for L, im in enumerate(spec.images):
            self.data[L * per: L * per + sizes[0]] = np.ascontiguousarray(im, np.uint8).ravel()
        if levels > 1:
            lib().or_build_mips(self.data.ctypes.data, w, h, 4, layers, levels)
        self.w, self.h, self.layers, self.levels, self.wrap = w, h, layers, levels, spec.wrap
        self.sizes, self.per = sizes, per

    def level(self, layer, l):
        off = layer * self.per + sum(self.sizes[:l])
        return self.data[off: off + self.sizes[l]]

    def struct(self):
        return texture_struct(self.data, wrap=self.wrap, levels=self.levels, layers=self.layers,
                              width=self.w, height=self.h)


class _TexSpec:
    """one-layer, one-level, clamped texture of raw 4-byte texels"""

    def __init__(self, img):
        self.images = [img]
        self.height, self.width = img.shape[:2]
        self.layers, self.levels, self.wrap = 1, 1, 0


def sample_projection(tex, proj):
    """PuresoftSamplerProjection::get of a projected coordinate (4 floats)"""
    p = np.ascontiguousarray(proj, np.float32)
    s = tex.struct()
    return float(lib().or_sample_projection(C.byref(s), _fp(p)))


def blend4(dst, src):
    """PuresoftFBO::blend4 of BGRA bytes src onto dst; returns the new dst"""
    d = np.array(dst, np.uint8).copy()
    sv = np.ascontiguousarray(src, np.uint8)
    lib().or_blend4(d.ctypes.data, sv.ctypes.data)
    return d


def sse_pack_byte(f):
    """one lane of FP_Cloud's cvtps2dq/packusdw/packuswb tail"""
    return int(lib().or_sse_pack_byte(float(np.float32(f))))


def mesh_triangles(mesh):
    return min(a.shape[0] for a in mesh.values()) // 3


APPROX_MODES = {"x86": 2, "bound_hi": 3, "bound_lo": 4}


def render_scene(scene, n_threads=1, debug=False, fp64=False, frame=None, clear=True, approx=None):
    """Run a puresoft3d_amd.scenes.Scene through the oracle: clearDepth +
    clearColour + the draws in order.  Returns (Frame, Stats).  fp64: shade
    in FP64.  approx: FP32 with the reference's approximate reciprocals where
    its DEF01-DEF04 call div_3_4 and norm_3_4: "x86" as this CPU's rcpps /
    rsqrtss compute them, "bound_hi" / "bound_lo" at either end of their
    1.5 * 2^-12 relative error bound."""
    f = frame if frame is not None else Frame(scene.width, scene.height, debug=debug)
    st = Stats()
    texs = [OracleTexture(t) for t in scene.textures]
    sp = getattr(scene, "shadow", None)
    base = 0   # triangle ids count every draw in submission order
    if sp is not None:
        # the shadow map first: DEF05 into a size x size FP32 depth texture
        sf = Frame(sp.size, sp.size, depth_scanline=sp.size, top_down=False)
        sf.clear_depth(1.0)
        for d in sp.draws:
            u = {3: sp.light_pv}
            u.update(d.uniforms)
            DrawCall(d.prog, scene.meshes[d.mesh], u, texs, behavior=DEFAULT_BEHAVIOR, tri_base=base).run(
                sf, n_threads, st)
            base += mesh_triangles(scene.meshes[d.mesh])
        img = np.ascontiguousarray(sf.depth[:, : sp.size]).view(np.uint8).reshape(sp.size, sp.size, 4)
        texs.append(OracleTexture(_TexSpec(img)))
    if clear:
        f.clear_depth(scene.clear_depth)
        f.clear_colour(scene.clear_colour)
    for d in scene.draws:
        u = dict(scene.uniforms)
        u.update(d.uniforms)
        beh = DEFAULT_BEHAVIOR if d.behavior is None else d.behavior
        mode = APPROX_MODES[approx] if approx else (1 if fp64 else 0)
        dc = DrawCall(d.prog, scene.meshes[d.mesh], u, texs, behavior=beh, tri_base=base, shade_fp64=mode)
        dc.run(f, n_threads, st)
        base += mesh_triangles(scene.meshes[d.mesh])
    return f, st


def pool_shutdown():
    """stop the oracle's worker threads (they spin while idle, like the reference's)"""
    lib().or_pool_shutdown()


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def diff_forward_f64(gbuf, tex, albedo_scale, spec_exp, ks, light, camera):
    """FP64 differentiable shading of a geometry buffer (H, W, 12) with an RGB
    texture (th, tw, 3); returns (H, W, 3) float64"""
    g, t = _f32(gbuf), _f32(tex)
    H, W = g.shape[:2]
    th, tw = t.shape[:2]
    out = np.zeros((H, W, 3), np.float64)
    lt, cm = _f32(light), _f32(camera)
    lib().or_diff_forward_f64(W, H, g.ctypes.data, tw, th, t.ctypes.data, albedo_scale, spec_exp, ks,
                              lt.ctypes.data, cm.ctypes.data, out.ctypes.data)
    return out


def diff_backward_f64(gbuf, tex, albedo_scale, spec_exp, ks, light, camera, gout):
    """FP64 adjoint: (dLoss/dtex (th, tw, 3), dLoss/d[albedo_scale, spec_exp]) for gout = dLoss/dout"""
    g, t = _f32(gbuf), _f32(tex)
    H, W = g.shape[:2]
    th, tw = t.shape[:2]
    go = np.ascontiguousarray(gout, np.float64)
    gt = np.zeros((th, tw, 3), np.float64)
    gp = np.zeros(2, np.float64)
    lt, cm = _f32(light), _f32(camera)
    lib().or_diff_backward_f64(W, H, g.ctypes.data, tw, th, t.ctypes.data, albedo_scale, spec_exp, ks,
                               lt.ctypes.data, cm.ctypes.data, go.ctypes.data, gt.ctypes.data, gp.ctypes.data)
    return gt, gp
