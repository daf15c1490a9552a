"""Differentiable texture sampling and shading behind torch.autograd.Function.

The raster pipeline renders a view once and leaves a geometry buffer of the
visible fragments (psgpu_target_gbuffer): interpolated normal, world position
and texture coordinates per pixel.  That fixes visibility.  Shading is then a
differentiable function of a learnable FP32 RGB texture and two material
parameters (albedo scale, specular exponent):

    c = albedo_scale * bilinear(tex, u, v) * max(n.L, 0) + ks * max(n.H, 0)^spec_exp

with a hand-written adjoint (psgpu_diff_backward): texel gradients are
scattered with atomics, parameter gradients are reduced per wavefront.

torch must be imported before libpsgpu is first loaded, so that both use the
one HIP runtime torch brings (this module imports torch first).

    python -m puresoft3d_amd.adjoint --bench     # optimisation steps/s on the 512^2 workload
"""
import ctypes as C

import numpy as np
import torch

from . import native as N
from . import scenes as S


class DiffDesc(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("tex_w", C.c_int), ("tex_h", C.c_int),
                ("light", C.c_float * 4), ("camera", C.c_float * 4), ("ks", C.c_float)]


def _require_torch_runtime():
    if N.runtime() == "rocm":
        raise RuntimeError("libpsgpu was loaded before torch: it bound to ROCm's HIP runtime, which torch tensors "
                           "do not share; import puresoft3d_amd.adjoint (or torch) before creating a Context")


class GeometryBuffer:
    """The visible fragments of one view, rendered once by the raster pipeline
    (Blinn-Phong programme over slots 0-4 of `mesh`) into a (H, W, 12) float32
    CUDA tensor.  Rows are bottom-up, as in every texture target."""

    def __init__(self, ctx, mesh, width, height, eye, target=(0.0, 0.0, 0.0), fov_deg=45.0, light=(4.0, 5.0, 6.0)):
        _require_torch_runtime()
        self.ctx, self.width, self.height = ctx, width, height
        self.eye = np.asarray(eye, np.float64)
        self.light = np.asarray(light, np.float64)
        self.gbuf = torch.zeros((height, width, 12), device="cuda", dtype=torch.float32)
        torch.cuda.synchronize()
        P = S.perspective(np.radians(fov_deg), width / height, 0.5, 50.0)
        V = S.look_at(self.eye, target)
        ph = np.full((4, 4, 4), 128, np.uint8)       # the colour pass needs textures; its colour is unused
        t0 = ctx.texture_create(4, 4, 4, pixels=ph)
        t1 = ctx.texture_create(4, 4, 4, pixels=ph)
        col = ctx.texture_create(width, height, 4, wrap=N.CLAMP)
        dep = ctx.texture_create(width, height, 4, wrap=N.CLAMP)
        vao = ctx.vao_create()
        for s, arr in mesh.items():
            ctx.vao_attach(vao, s, ctx.vbo_create(arr))
        prog = ctx.program_create(N.PROC_BLINN)
        try:
            ctx.program_use(prog)
            ctx.target(col)
            ctx.depth_target(dep)
            ctx.viewport(width, height)
            ctx.clear_depth(1.0)
            ctx.clear_target(0)
            ctx.gbuffer_target(self.gbuf.data_ptr(), self.gbuf.numel() * 4)
            for k, v in {3: S.colmajor(P @ V), 4: S.colmajor(np.eye(4)), 5: S.colmajor(np.eye(4)),
                         7: S.vec4(*self.light), 8: S.vec4(*self.eye), 9: np.array([t0], np.int32),
                         10: np.array([t1], np.int32), 11: np.array([8], np.int32)}.items():
                ctx.uniform(k, v)
            ctx.draw(vao)
            ctx.sync()
        finally:
            ctx.gbuffer_target(0, 0)
            ctx.target(N.DISPLAY)
            ctx.depth_target(N.DEFAULT_DEPTH)
            ctx.viewport(ctx.width, ctx.height)
            ctx.vao_destroy(vao)
            for t in (t0, t1, col, dep):
                ctx.texture_destroy(t)
            ctx.program_destroy(prog)
        self.covered = int((self.gbuf[..., 3] != 0).sum().item())

    def desc(self, tex, ks):
        d = DiffDesc()
        d.width, d.height = self.width, self.height
        d.tex_h, d.tex_w = int(tex.shape[0]), int(tex.shape[1])
        d.light[:] = [float(x) for x in self.light] + [0.0]
        d.camera[:] = [float(x) for x in self.eye] + [0.0]
        d.ks = float(ks)
        return d


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_inputs(tex, params):
    if tex.dtype != torch.float32 or tex.dim() != 3 or tex.shape[2] != 3 or not tex.is_contiguous() or not tex.is_cuda:
        raise ValueError("tex must be a contiguous CUDA float32 tensor of shape (th, tw, 3)")
    if params.dtype != torch.float32 or params.numel() != 2 or not params.is_contiguous() or not params.is_cuda:
        raise ValueError("params must be a contiguous CUDA float32 tensor [albedo_scale, spec_exp]")


class ShadeFunction(torch.autograd.Function):
    """image (H, W, 3) = shade(tex (th, tw, 3), params [albedo_scale, spec_exp]) on a fixed geometry buffer"""

    @staticmethod
    def forward(fctx, tex, params, gb, ks):
        _check_inputs(tex, params)
        L = N.lib()
        out = torch.empty((gb.height, gb.width, 3), device=tex.device, dtype=torch.float32)
        d = gb.desc(tex, ks)
        rc = L.psgpu_diff_forward(gb.ctx.h, C.byref(d), C.c_void_p(gb.gbuf.data_ptr()), C.c_void_p(tex.data_ptr()),
                                  C.c_void_p(params.data_ptr()), C.c_void_p(out.data_ptr()), _stream())
        if rc:
            raise RuntimeError("psgpu_diff_forward failed (%d)" % rc)
        fctx.save_for_backward(tex, params)
        fctx.gb, fctx.ks = gb, ks
        return out

    @staticmethod
    def backward(fctx, gout):
        tex, params = fctx.saved_tensors
        gb = fctx.gb
        gout = gout.contiguous()
        gtex = torch.zeros_like(tex)
        gpar = torch.zeros_like(params)
        d = gb.desc(tex, fctx.ks)
        rc = N.lib().psgpu_diff_backward(gb.ctx.h, C.byref(d), C.c_void_p(gb.gbuf.data_ptr()),
                                         C.c_void_p(tex.data_ptr()), C.c_void_p(params.data_ptr()),
                                         C.c_void_p(gout.data_ptr()), C.c_void_p(gtex.data_ptr()),
                                         C.c_void_p(gpar.data_ptr()), _stream())
        if rc:
            raise RuntimeError("psgpu_diff_backward failed (%d)" % rc)
        return gtex, gpar, None, None


def shade(tex, params, gb, ks=0.5):
    return ShadeFunction.apply(tex, params, gb, ks)


# ---------------------------------------------------------------- the 512^2 workload
def workload(ctx, width=512, height=512, tex_size=1024, seed=41):
    """50k-triangle displaced sphere at 512^2; a target image shaded from a
    seeded noise texture (albedo 1.2, exponent 24); the learnable texture
    starts at 0.5 everywhere, albedo scale 1.0, exponent 12"""
    mesh = S.displaced_sphere(125, 200, seed=seed)
    gb = GeometryBuffer(ctx, mesh, width, height, eye=(0.0, 0.4, 3.0))
    noise = S.tiled_texture(lambda t, sd: S.noise_texture(t, sd, cells=16, octaves=3), tex_size, seed + 1)
    tex_true = torch.from_numpy(noise[..., :3].astype(np.float32) / 255.0).cuda().contiguous()
    p_true = torch.tensor([1.2, 24.0], device="cuda")
    with torch.no_grad():
        target = shade(tex_true, p_true, gb)
    tex = torch.full((tex_size, tex_size, 3), 0.5, device="cuda", requires_grad=True)
    params = torch.tensor([1.0, 12.0], device="cuda", requires_grad=True)
    return gb, target, tex, params


def bench(steps=200, warmup=20):
    """optimisation steps (forward, L2 loss, backward, Adam update) per second"""
    import json
    import time
    ctx = N.Context(64, 64)
    gb, target, tex, params = workload(ctx)
    opt = torch.optim.Adam([tex, params], lr=1e-2)

    def step():
        opt.zero_grad(set_to_none=False)
        loss = ((shade(tex, params, gb) - target) ** 2).sum()
        loss.backward()
        opt.step()
        return loss

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # kernel-only timing of one forward + backward
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    gout = torch.ones_like(target)
    ev[0].record()
    for _ in range(50):
        out = ShadeFunction.forward(_Ctx(), tex.detach(), params.detach(), gb, 0.5)
    ev[1].record()
    c = _Ctx()
    ShadeFunction.forward(c, tex.detach(), params.detach(), gb, 0.5)
    for _ in range(50):
        ShadeFunction.backward(c, gout)
    ev[2].record()
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / 50
    bwd_ms = ev[1].elapsed_time(ev[2]) / 50
    res = {"workload": "adjoint 512x512, 50k-tri displaced sphere, 1024^2 FP32 RGB texture",
           "steps_per_s": round(steps / dt, 1), "ms_per_step": round(dt * 1e3 / steps, 4),
           "forward_ms": round(fwd_ms, 4), "backward_ms": round(bwd_ms, 4), "covered_px": gb.covered,
           "final_loss": float(loss.item()), "albedo_scale": float(params[0].item()),
           "spec_exp": float(params[1].item())}
    del out
    ctx.close()
    print(json.dumps(res), flush=True)
    return res


class _Ctx:
    """stand-in for autograd's context when the Function's methods are timed directly"""

    def save_for_backward(self, *t):
        self.saved_tensors = t


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--bench", action="store_true")
    ap.add_argument("--steps", type=int, default=200)
    a = ap.parse_args()
    if a.bench:
        bench(a.steps)
