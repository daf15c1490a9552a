"""ctypes binding of libpsgpu.so, the C ABI declared in include/psgpu.h.

The HIP library is the only compute path: if it is missing this module raises
instead of falling back to anything else.  Status codes map to the exception
types the reference pipeline throws (pipeline.cpp, fbo.cpp):
std::out_of_range -> IndexError, std::invalid_argument -> ValueError,
std::bad_alloc -> MemoryError.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsgpu.so")

OK, ERR_INVALID, ERR_RANGE, ERR_NOMEM, ERR_DEVICE, ERR_STATE, ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
UPDATE_DEPTH, TEST_DEPTH, FACE_CULLING, ALPHABLEND = 1, 2, 4, 8
DEFAULT_BEHAVIOR = UPDATE_DEPTH | TEST_DEPTH | FACE_CULLING   # pipeline.cpp:34
PROC_DEF01, PROC_DEF02, PROC_DEF03, PROC_DEF04, PROC_DEF05, PROC_LAMBERT, PROC_BLINN = 1, 2, 3, 4, 5, 6, 7
PROC_SHADOWED = 8   # Blinn-Phong lit through a projective shadow map (uniforms 12, 13)
PROC_CLOUD = 9      # the demo's cloud layer: lit, shadowed, alpha from the texel's red; blended under
                    # ALPHABLEND (uniforms 15, 16: shadow map, light PV; shaded in the raster walk)
PROC_CLOUD_SHADOW = 10   # depth only, discarded where the texel's red < 150 (shaded in the raster walk)
CLAMP, WRAP = 0, 1
DISPLAY = -1
DEFAULT_DEPTH = -1


class PsGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("psgpu error %d: %s" % (code, msg))
        self.code = code


class PsGpuStateError(PsGpuError):
    pass


class PsGpuUnsupported(PsGpuError):
    pass


class Stats(C.Structure):
    _fields_ = [("tris_in", C.c_int64), ("tris_culled", C.c_int64), ("tris_zrejected", C.c_int64),
                ("tris_dropped", C.c_int64), ("spans", C.c_int64), ("frags_rasterised", C.c_int64),
                ("frags_shaded", C.c_int64), ("segments", C.c_int64), ("flushes", C.c_int64),
                ("retries", C.c_int64), ("splits", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class PostFrame(C.Structure):
    """psgpu_post_frame: device addresses of the colour target and depth buffer"""
    _fields_ = [("colour", C.c_void_p), ("colour_pitch", C.c_int), ("colour_top_down", C.c_int),
                ("depth", C.c_void_p), ("depth_pitch", C.c_int), ("width", C.c_int), ("height", C.c_int)]


POST_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(PostFrame), C.c_void_p)


class Tiling(C.Structure):
    _fields_ = [("chunk_width", C.c_int), ("tile_rows", C.c_int), ("waves_per_group", C.c_int)]


_lib = None

_SIGS = {
    "psgpu_create": ([C.c_int, C.c_int, C.c_int, C.c_uint, C.POINTER(C.c_void_p)], C.c_int),
    "psgpu_destroy": ([C.c_void_p], C.c_int),
    "psgpu_last_error": ([C.c_void_p], C.c_char_p),
    "psgpu_stream": ([C.c_void_p], C.c_void_p),
    "psgpu_texture_create": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "psgpu_texture_upload": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_texture_download": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_texture_generate_mips": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_texture_destroy": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_texture_info": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                            C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "psgpu_vbo_create": ([C.c_void_p, C.c_size_t, C.c_size_t, C.POINTER(C.c_int)], C.c_int),
    "psgpu_vbo_update": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "psgpu_vbo_destroy": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_vao_create": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "psgpu_vao_attach": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "psgpu_vao_detach": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "psgpu_vao_get": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "psgpu_vao_destroy": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_program_create": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "psgpu_program_destroy": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_program_use": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_uniform_set": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_behavior_enable": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_behavior_disable": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_behavior_get": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "psgpu_viewport": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "psgpu_target_set": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "psgpu_depth_set": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_clear_depth": ([C.c_void_p, C.c_float], C.c_int),
    "psgpu_clear_colour": ([C.c_void_p, C.c_uint32], C.c_int),
    "psgpu_clear_target": ([C.c_void_p, C.c_uint32], C.c_int),
    "psgpu_draw": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_flush": ([C.c_void_p], C.c_int),
    "psgpu_sync": ([C.c_void_p], C.c_int),
    "psgpu_sync_noalloc": ([C.c_void_p], C.c_int),
    "psgpu_read_colour": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_read_depth": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_write_colour": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_write_depth": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_display_device_ptr": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "psgpu_depth_device_ptr": ([C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int)], C.c_int),
    "psgpu_get_stats": ([C.c_void_p, C.POINTER(Stats)], C.c_int),
    "psgpu_reset_stats": ([C.c_void_p], C.c_int),
    "psgpu_swap_buffers": ([C.c_void_p], C.c_int),
    "psgpu_post_process": ([C.c_void_p, POST_FN, C.c_void_p], C.c_int),
    "psgpu_debug_enable": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_debug_reset": ([C.c_void_p], C.c_int),
    "psgpu_debug_read": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "psgpu_tiling_set": ([C.c_void_p, C.POINTER(Tiling)], C.c_int),
    "psgpu_tiling_get": ([C.c_void_p, C.POINTER(Tiling)], C.c_int),
    "psgpu_profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_profile_read": ([C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                            C.POINTER(C.c_int64), C.POINTER(C.c_int)], C.c_int),
    "psgpu_profile_reset": ([C.c_void_p], C.c_int),
    "psgpu_capture_begin": ([C.c_void_p], C.c_int),
    "psgpu_capture_end": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "psgpu_graph_launch": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "psgpu_graph_destroy": ([C.c_void_p, C.c_int], C.c_int),
    "psgpu_selftest_math": ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "psgpu_host_alloc": ([C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "psgpu_target_gbuffer": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "psgpu_diff_forward": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                           C.c_int),
    "psgpu_diff_backward": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p], C.c_int),
    "psgpu_host_free": ([C.c_void_p, C.c_void_p], C.c_int),
    "psgpu_texture_read_async": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                  C.c_size_t], C.c_int),
    "psgpu_fence": ([C.c_void_p, C.POINTER(C.c_uint64)], C.c_int),
    "psgpu_fence_query": ([C.c_void_p, C.c_uint64, C.POINTER(C.c_int)], C.c_int),
    "psgpu_fence_wait": ([C.c_void_p, C.c_uint64], C.c_int),
}


def declared_symbols():
    """Every entry point declared in include/psgpu.h (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "psgpu.h")
    names = []
    with open(hdr) as fh:
        for line in fh:
            m = re.match(r"^\s*(?:int|const char\*|void\*)\s+(psgpu_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names


_runtime = None


def runtime():
    """which HIP runtime libpsgpu bound to: "torch" (the one torch loaded
    before it), "rocm" (the system one), or None before the first load"""
    return _runtime


def lib():
    global _lib, _runtime
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpsgpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        # libpsgpu binds by soname to a HIP runtime torch has already loaded;
        # importing torch after this point would load a second runtime
        _runtime = "torch" if "torch" in sys.modules else "rocm"
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One psgpu context (a PuresoftPipeline bound to one GPU and one stream)."""

    def __init__(self, width, height, device=0, flags=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.psgpu_create(device, width, height, flags, C.byref(h))
        if rc != OK:
            raise PsGpuError(rc, "psgpu_create(%d, %d, %d) failed" % (device, width, height))
        self.h = h
        self.width, self.height = width, height
        self.depth_scanline = int(width / 4.0 + 0.5) * 4

    def close(self):
        if self.h:
            self.L.psgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            msg = self.L.psgpu_last_error(self.h).decode()
            cls = {ERR_RANGE: IndexError, ERR_INVALID: ValueError, ERR_NOMEM: MemoryError}.get(rc)
            if cls is not None:
                raise cls(msg)
            if rc == ERR_STATE:
                raise PsGpuStateError(rc, msg)
            if rc == ERR_UNSUPPORTED:
                raise PsGpuUnsupported(rc, msg)
            raise PsGpuError(rc, msg)

    # textures
    def texture_create(self, width, height, elem_len=4, scanline=None, extra_layers=0, wrap=CLAMP, mip_levels=1,
                       pixels=None):
        out = C.c_int()
        if scanline is None:
            scanline = width * elem_len
        buf = None if pixels is None else np.ascontiguousarray(pixels)
        self._chk(self.L.psgpu_texture_create(self.h, width, scanline, height, elem_len, extra_layers, wrap,
                                              mip_levels, _ptr(buf), C.byref(out)))
        return out.value

    def texture_upload(self, tex, pixels, layer=0, level=0):
        buf = np.ascontiguousarray(pixels)
        self._chk(self.L.psgpu_texture_upload(self.h, tex, layer, level, _ptr(buf), buf.nbytes))

    def texture_download(self, tex, layer=0, level=0):
        w, s, h, e = (C.c_int() for _ in range(4))
        self._chk(self.L.psgpu_texture_info(self.h, tex, level, C.byref(w), C.byref(s), C.byref(h), C.byref(e)))
        out = np.zeros(s.value * h.value, np.uint8)
        self._chk(self.L.psgpu_texture_download(self.h, tex, layer, level, _ptr(out), out.nbytes))
        return out.reshape(h.value, s.value)

    def texture_generate_mips(self, tex):
        self._chk(self.L.psgpu_texture_generate_mips(self.h, tex))

    def texture_destroy(self, tex):
        self._chk(self.L.psgpu_texture_destroy(self.h, tex))

    # vertex streams
    def vbo_create(self, data):
        data = np.ascontiguousarray(data, np.float32)
        n = data.shape[0]
        out = C.c_int()
        self._chk(self.L.psgpu_vbo_create(self.h, data.nbytes // n, n, C.byref(out)))
        self._chk(self.L.psgpu_vbo_update(self.h, out.value, _ptr(data)))
        return out.value

    def vbo_update(self, vbo, data):
        data = np.ascontiguousarray(data, np.float32)
        self._chk(self.L.psgpu_vbo_update(self.h, vbo, _ptr(data)))

    def vbo_destroy(self, vbo):
        self._chk(self.L.psgpu_vbo_destroy(self.h, vbo))

    def vao_create(self):
        out = C.c_int()
        self._chk(self.L.psgpu_vao_create(self.h, C.byref(out)))
        return out.value

    def vao_attach(self, vao, slot, vbo):
        prev = C.c_int()
        self._chk(self.L.psgpu_vao_attach(self.h, vao, slot, vbo, C.byref(prev)))
        return prev.value

    def vao_detach(self, vao, slot):
        prev = C.c_int()
        self._chk(self.L.psgpu_vao_detach(self.h, vao, slot, C.byref(prev)))
        return prev.value

    def vao_get(self, vao, slot):
        out = C.c_int()
        self._chk(self.L.psgpu_vao_get(self.h, vao, slot, C.byref(out)))
        return out.value

    def vao_destroy(self, vao):
        self._chk(self.L.psgpu_vao_destroy(self.h, vao))

    # programmes and state
    def program_create(self, vp, ip=None, fp=None):
        out = C.c_int()
        self._chk(self.L.psgpu_program_create(self.h, vp, vp if ip is None else ip, vp if fp is None else fp,
                                              C.byref(out)))
        return out.value

    def program_destroy(self, prog):
        self._chk(self.L.psgpu_program_destroy(self.h, prog))

    def program_use(self, prog):
        self._chk(self.L.psgpu_program_use(self.h, prog))

    def uniform(self, slot, value):
        if value is None:
            self._chk(self.L.psgpu_uniform_set(self.h, slot, None, 0))
            return
        a = np.ascontiguousarray(value)
        self._chk(self.L.psgpu_uniform_set(self.h, slot, _ptr(a), a.nbytes))

    def enable(self, bits):
        self._chk(self.L.psgpu_behavior_enable(self.h, bits))

    def disable(self, bits):
        self._chk(self.L.psgpu_behavior_disable(self.h, bits))

    def viewport(self, w, h):
        self._chk(self.L.psgpu_viewport(self.h, w, h))

    def target(self, tex, slot=0):
        self._chk(self.L.psgpu_target_set(self.h, slot, tex))

    def gbuffer_target(self, device_ptr, nbytes):
        """geometry buffer written by Blinn-Phong draws (device memory of the
        same HIP runtime, e.g. a torch tensor's data_ptr()); 0 stops it"""
        self._chk(self.L.psgpu_target_gbuffer(self.h, C.c_void_p(device_ptr or None), nbytes))

    def depth_target(self, tex):
        self._chk(self.L.psgpu_depth_set(self.h, tex))

    def clear_depth(self, v=1.0):
        self._chk(self.L.psgpu_clear_depth(self.h, v))

    def clear_colour(self, bgra=0):
        self._chk(self.L.psgpu_clear_colour(self.h, bgra))

    def clear_target(self, bgra=0):
        self._chk(self.L.psgpu_clear_target(self.h, bgra))

    def draw(self, vao):
        self._chk(self.L.psgpu_draw(self.h, vao))

    def flush(self):
        self._chk(self.L.psgpu_flush(self.h))

    def sync(self):
        self._chk(self.L.psgpu_sync(self.h))

    def sync_noalloc(self):
        """psgpu_sync_noalloc: run the queued clears and draws with the fixed
        device scratch array (no allocation) and wait"""
        self._chk(self.L.psgpu_sync_noalloc(self.h))

    def read_colour(self, out=None):
        """the display (top-down BGRA8 rows), into `out` when given"""
        if out is None:
            out = np.zeros((self.height, self.width, 4), np.uint8)
        elif out.nbytes != self.width * self.height * 4 or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous %d x %d x 4 byte array" % (self.height, self.width))
        self._chk(self.L.psgpu_read_colour(self.h, _ptr(out), out.nbytes))
        return out

    def read_depth(self):
        out = np.zeros((self.height, self.depth_scanline), np.float32)
        self._chk(self.L.psgpu_read_depth(self.h, _ptr(out), out.nbytes))
        return out

    def stats(self):
        s = Stats()
        self._chk(self.L.psgpu_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._chk(self.L.psgpu_reset_stats(self.h))

    def debug_enable(self, on=True):
        self._chk(self.L.psgpu_debug_enable(self.h, 1 if on else 0))

    def debug_reset(self):
        self._chk(self.L.psgpu_debug_reset(self.h))

    def debug_read(self, w, h):
        n = w * h
        planes = [np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.uint32), np.zeros(n, np.int32),
                  np.zeros(n, np.uint32)]
        self._chk(self.L.psgpu_debug_read(self.h, *[_ptr(p) for p in planes]))
        return dict(zip(["last_pass", "raster_count", "raster_hash", "pass_count", "pass_hash"], planes))

    def tiling(self, chunk_width=None, tile_rows=None, waves_per_group=None):
        t = Tiling()
        self._chk(self.L.psgpu_tiling_get(self.h, C.byref(t)))
        if chunk_width is None and tile_rows is None and waves_per_group is None:
            return (t.chunk_width, t.tile_rows, t.waves_per_group)
        if chunk_width is not None:
            t.chunk_width = chunk_width
        if tile_rows is not None:
            t.tile_rows = tile_rows
        if waves_per_group is not None:
            t.waves_per_group = waves_per_group
        self._chk(self.L.psgpu_tiling_set(self.h, C.byref(t)))
        return (t.chunk_width, t.tile_rows, t.waves_per_group)

    def profile(self, on=True):
        self._chk(self.L.psgpu_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        self._chk(self.L.psgpu_profile_reset(self.h))

    def profile_read(self):
        n = 64
        names = (C.c_char_p * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        k = C.c_int()
        self._chk(self.L.psgpu_profile_read(self.h, n, names, ms, cnt, C.byref(k)))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k.value)}

    def capture_begin(self):
        self._chk(self.L.psgpu_capture_begin(self.h))

    def capture_end(self):
        out = C.c_int()
        self._chk(self.L.psgpu_capture_end(self.h, C.byref(out)))
        return out.value

    def graph_launch(self, graph, count=1):
        self._chk(self.L.psgpu_graph_launch(self.h, graph, count))

    def graph_destroy(self, graph):
        self._chk(self.L.psgpu_graph_destroy(self.h, graph))

    # ---- asynchronous readback (serving)
    def host_alloc(self, nbytes):
        """pinned host buffer as a uint8 numpy array; free with host_free"""
        p = C.c_void_p()
        self._chk(self.L.psgpu_host_alloc(self.h, nbytes, C.byref(p)))
        return np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(p.value))

    def host_free(self, arr):
        self._chk(self.L.psgpu_host_free(self.h, arr.ctypes.data))

    def read_async(self, tex, out, x=0, y=0, w=None, h=None, layer=0):
        """queue a copy of rows y..y+h-1, columns x..x+w-1 (memory order) into
        `out` (uint8, h rows of at least w*4 bytes)"""
        if w is None or h is None:
            raise ValueError("region size required")
        pitch = out.strides[0] if out.ndim > 1 else w * 4
        if out.nbytes < pitch * (h - 1) + w * 4:
            raise ValueError("host buffer too small")
        self._chk(self.L.psgpu_texture_read_async(self.h, tex, layer, x, y, w, h, out.ctypes.data, pitch))

    def fence(self):
        f = C.c_uint64()
        self._chk(self.L.psgpu_fence(self.h, C.byref(f)))
        return f.value

    def fence_done(self, f):
        d = C.c_int()
        self._chk(self.L.psgpu_fence_query(self.h, f, C.byref(d)))
        return bool(d.value)

    def fence_wait(self, f):
        self._chk(self.L.psgpu_fence_wait(self.h, f))

    def post_process(self, fn):
        """PuresoftPipeline::postProcess: fn(frame: PostFrame, stream: int) -> int
        runs in stream order after the draws queued so far and enqueues its own
        device work on `stream`; it sees the finished frame and runs once per
        call.  The callback object is kept alive with the context."""
        cb = POST_FN(lambda user, frame, stream: int(fn(frame.contents, stream or 0)))
        self._post_cbs = getattr(self, "_post_cbs", []) + [cb]
        self._chk(self.L.psgpu_post_process(self.h, cb, None))

    def selftest_math(self, op, a, b=None):
        a = np.ascontiguousarray(a, np.float32)
        b = a if b is None else np.ascontiguousarray(b, np.float32)
        out = np.zeros_like(a)
        self._chk(self.L.psgpu_selftest_math(self.h, op, _ptr(a), _ptr(b), _ptr(out), a.size))
        return out

    def stream(self):
        return self.L.psgpu_stream(self.h)
