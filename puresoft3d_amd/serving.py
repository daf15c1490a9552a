"""Render serving over libpsgpu: resident scenes, a Poisson stream of view
requests, continuous batching with HBM-budget admission (the loop itself is
native: include/psserve.h, puresoft3d_amd/host/psserve.cpp)."""
import ctypes as C
import os

import numpy as np

from . import native as N

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libpuresoft3d.so")


class Scene(C.Structure):
    _fields_ = [("vao", C.c_int), ("albedo_tex", C.c_int), ("normal_tex", C.c_int)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int), ("width", C.c_int), ("height", C.c_int), ("pv", C.c_float * 16),
                ("eye", C.c_float * 4), ("light", C.c_float * 4), ("spec_exp", C.c_int), ("arrival_s", C.c_double)]


class Config(C.Structure):
    _fields_ = [("program", C.c_int), ("hbm_budget", C.c_size_t), ("scratch_bytes_per_px", C.c_int),
                ("max_batch_views", C.c_int), ("max_batch_pixels", C.c_int64), ("max_inflight", C.c_int),
                ("clear_bgra", C.c_uint32), ("deliver_to_host", C.c_int), ("render_streams", C.c_int)]


class Result(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admitted_s", C.c_double), ("launched_s", C.c_double),
                ("done_s", C.c_double), ("batch", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("batches", C.c_int64), ("views", C.c_int64), ("pixels", C.c_int64), ("target_allocs", C.c_int64),
                ("target_evictions", C.c_int64), ("peak_bytes", C.c_int64), ("max_batch_views_seen", C.c_int),
                ("wall_s", C.c_double), ("host_submit_s", C.c_double), ("host_flush_s", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        N.lib()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("%s is not built (python -m puresoft3d_amd.build)" % HOST_LIB)
        L = C.CDLL(HOST_LIB)
        L.psserve_create.argtypes = [C.c_void_p, C.POINTER(Scene), C.c_int, C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.psserve_create.restype = C.c_int
        L.psserve_create_multi.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(Scene), C.c_int, C.POINTER(Config),
                                           C.POINTER(C.c_void_p)]
        L.psserve_create_multi.restype = C.c_int
        L.psserve_set_config.argtypes = [C.c_void_p, C.POINTER(Config)]
        L.psserve_set_config.restype = C.c_int
        L.psserve_destroy.argtypes = [C.c_void_p]
        L.psserve_destroy.restype = C.c_int
        L.psserve_run.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.POINTER(Result), C.POINTER(C.c_void_p),
                                  C.POINTER(Stats)]
        L.psserve_run.restype = C.c_int
        L.psserve_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class RenderServer:
    """Scenes resident on one context, or on each of several contexts of the
    device (render streams: ctx plus extra_ctxs); serve() runs a request
    stream.  The native server keeps its pooled render targets between runs;
    close() it before closing the contexts."""

    def __init__(self, ctx, assets, gen_mips=True, extra_ctxs=()):
        self.ctx, self.assets = ctx, assets
        self.ctxs = [ctx] + list(extra_ctxs)
        self.all_scenes = (Scene * (len(assets) * len(self.ctxs)))()
        progs = []
        for c, cx in enumerate(self.ctxs):
            for k, a in enumerate(assets):
                ids = []
                for t in a.textures:
                    idx = cx.texture_create(t.width, t.height, 4, wrap=t.wrap, mip_levels=t.levels, pixels=t.images[0])
                    if t.levels > 1 and gen_mips:
                        cx.texture_generate_mips(idx)
                    ids.append(idx)
                v = cx.vao_create()
                for s, arr in a.mesh.items():
                    cx.vao_attach(v, s, cx.vbo_create(arr))
                self.all_scenes[c * len(assets) + k] = Scene(v, ids[0], ids[1])
            progs.append(cx.program_create(N.PROC_BLINN))
        if len(set(progs)) != 1:
            raise RuntimeError("the Blinn-Phong programme has different ids on the contexts: %s" % progs)
        self.prog = progs[0]
        self.scenes = self.all_scenes   # scene k of the first context at [k]
        L = lib()
        h = C.c_void_p()
        ctxs = (C.c_void_p * len(self.ctxs))(*[cx.h for cx in self.ctxs])
        self._check(L.psserve_create_multi(ctxs, len(self.ctxs), self.all_scenes, len(assets),
                                           C.byref(self._config()), C.byref(h)))
        self.h = h

    @staticmethod
    def _check(rc):
        if rc:
            raise RuntimeError("psserve failed (%d): %s" % (rc, lib().psserve_last_error().decode()))

    def _config(self, hbm_budget=8 << 30, max_batch_views=64, max_batch_pixels=16 << 20, max_inflight=2,
                scratch_bytes_per_px=48, clear_bgra=0x00302010, to_host=True, streams=0):
        return Config(self.prog, hbm_budget, scratch_bytes_per_px, max_batch_views, max_batch_pixels, max_inflight,
                      clear_bgra, 1 if to_host else 0, streams)

    def close(self):
        """frees the server's pooled targets (the scenes stay on the context)"""
        if getattr(self, "h", None):
            self._check(lib().psserve_destroy(self.h))
            self.h = None

    def requests(self, views, arrivals):
        reqs = (Request * len(views))()
        for i, (vw, t) in enumerate(zip(views, arrivals)):
            u = vw.uniforms()
            r = reqs[i]
            r.scene, r.width, r.height = vw.scene, vw.width, vw.height
            r.pv[:] = [float(x) for x in u[3]]
            r.eye[:] = [float(x) for x in u[8]]
            r.light[:] = [float(x) for x in u[7]]
            r.spec_exp = int(u[11][0])
            r.arrival_s = float(t)
        return reqs

    def serve(self, views, arrivals, hbm_budget=8 << 30, max_batch_views=64, max_batch_pixels=16 << 20,
              max_inflight=2, scratch_bytes_per_px=48, clear_bgra=0x00302010, keep=None, to_host=True, streams=0):
        """Runs the stream in real time.  keep: indices whose images to return
        (needs to_host).  to_host=False leaves images in HBM: a request is done
        when its view is rendered.  streams: render streams (contexts) used,
        0 for all.  Returns (per-request result dict of arrays, stats dict,
        images dict)."""
        if keep is not None and not to_host:
            raise ValueError("images are returned only when delivered to the host")
        L = lib()
        n = len(views)
        reqs = self.requests(views, arrivals)
        cfg = self._config(hbm_budget, max_batch_views, max_batch_pixels, max_inflight, scratch_bytes_per_px,
                           clear_bgra, to_host, streams)
        self._check(L.psserve_set_config(self.h, C.byref(cfg)))
        res = (Result * n)()
        st = Stats()
        imgs, ptrs = {}, None
        if keep is not None:
            ptrs = (C.c_void_p * n)()
            for i in keep:
                imgs[i] = np.zeros((views[i].height, views[i].width, 4), np.uint8)
                ptrs[i] = imgs[i].ctypes.data
        self._check(L.psserve_run(self.h, reqs, n, res, ptrs, C.byref(st)))
        out = {k: np.array([getattr(r, k) for r in res]) for k, _ in Result._fields_}
        stats = {k: getattr(st, k) for k, _ in Stats._fields_}
        return out, stats, imgs

    def render_alone(self, view, clear_bgra=0x00302010):
        """one view through the plain path (targets of exactly its size), synchronously"""
        ctx = self.ctx
        c = ctx.texture_create(view.width, view.height, 4, wrap=N.CLAMP)
        d = ctx.texture_create(view.width, view.height, 4, wrap=N.CLAMP)
        try:
            sc = self.scenes[view.scene]
            ctx.program_use(self.prog)
            ctx.target(c)
            ctx.depth_target(d)
            ctx.viewport(view.width, view.height)
            ctx.clear_depth(1.0)
            ctx.clear_target(clear_bgra)
            u = view.uniforms()
            u[9] = np.array([sc.albedo_tex], np.int32)
            u[10] = np.array([sc.normal_tex], np.int32)
            for k, val in u.items():
                ctx.uniform(k, val)
            ctx.draw(sc.vao)
            ctx.target(N.DISPLAY)
            ctx.depth_target(N.DEFAULT_DEPTH)
            ctx.viewport(ctx.width, ctx.height)
            return ctx.texture_download(c).reshape(view.height, view.width, 4)
        finally:
            ctx.texture_destroy(c)
            ctx.texture_destroy(d)


def latency_summary(res):
    lat = res["done_s"] - res["arrival_s"]
    span = res["done_s"].max() - res["arrival_s"].min()
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3), "views_per_s": float(len(lat) / span) if span > 0 else None}
