"""Checks of the adjoint (puresoft3d_amd.adjoint) on the GPU, run in a process
of their own by tests/test_adjoint.py: torch must be imported before libpsgpu
loads, which a pytest process that already loaded it cannot undo.

Prints one JSON line of measured errors:
  forward vs the FP64 oracle (max abs), texel and parameter gradients vs the
  FP64 oracle adjoint (relative), central finite differences (h = 1e-3) on
  random texels and both parameters, and the loss falling under Adam.
"""
import json
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from puresoft3d_amd import adjoint as A  # noqa: E402
from puresoft3d_amd import native as N  # noqa: E402
import pyoracle  # noqa: E402


def main():
    full = "--full" in sys.argv
    n_fd = 1000 if full else 200
    W = H = 512 if full else 192
    ts = 1024 if full else 256
    print("[adjoint-check] %dx%d tex %d fd %d" % (W, H, ts, n_fd), flush=True)
    ctx = N.Context(64, 64)
    res = {"runtime": N.runtime(), "width": W, "tex": ts}
    gb, target, tex, params = A.workload(ctx, width=W, height=H, tex_size=ts)
    res["covered_px"] = gb.covered
    ks = 0.5
    # a texture with structure, so that every term is exercised
    rng = np.random.default_rng(3)
    tex0 = torch.from_numpy(rng.uniform(0.1, 0.9, (ts, ts, 3)).astype(np.float32)).cuda()
    p0 = torch.tensor([0.9, 16.0], device="cuda")
    g_np = gb.gbuf.cpu().numpy()
    t_np = tex0.cpu().numpy()
    light, cam = gb.light.astype(np.float32), gb.eye.astype(np.float32)

    # forward vs FP64
    out = A.shade(tex0, p0, gb, ks)
    ref = pyoracle.diff_forward_f64(g_np, t_np, 0.9, 16.0, ks, light, cam)
    res["forward_max_abs_err"] = float(np.abs(out.cpu().numpy().astype(np.float64) - ref).max())

    # backward vs FP64 adjoint, loss = sum((out - target)^2)
    t_leaf = tex0.clone().requires_grad_(True)
    p_leaf = p0.clone().requires_grad_(True)
    out = A.shade(t_leaf, p_leaf, gb, ks)
    loss = ((out - target) ** 2).sum()
    loss.backward()
    gout = 2.0 * (ref - target.cpu().numpy().astype(np.float64))
    gt_ref, gp_ref = pyoracle.diff_backward_f64(g_np, t_np, 0.9, 16.0, ks, light, cam, gout)
    gt = t_leaf.grad.cpu().numpy().astype(np.float64)
    gp = p_leaf.grad.cpu().numpy().astype(np.float64)
    res["grad_tex_rel_err"] = float(np.linalg.norm(gt - gt_ref) / np.linalg.norm(gt_ref))
    res["grad_tex_max_err_over_max"] = float(np.abs(gt - gt_ref).max() / np.abs(gt_ref).max())
    res["grad_params_rel_err"] = [float(abs(gp[k] - gp_ref[k]) / abs(gp_ref[k])) for k in range(2)]

    # central differences, h = 1e-3, loss summed in FP64
    def L(tx, pp):
        with torch.no_grad():
            o = A.shade(tx, pp, gb, ks)
            return float(((o.double() - target.double()) ** 2).sum().item())

    h = 1e-3
    flat = tex0.view(-1)
    idx = rng.choice(flat.numel(), n_fd, replace=False)
    fd = np.zeros(n_fd)
    for k, i in enumerate(idx):
        v = flat[i].item()
        flat[i] = v + h
        lp = L(tex0, p0)
        flat[i] = v - h
        lm = L(tex0, p0)
        flat[i] = v
        fd[k] = (lp - lm) / (2 * h)
    an = gt.reshape(-1)[idx]
    scale = np.abs(gt).max()
    res["fd_texels"] = int(n_fd)
    res["fd_texels_nonzero"] = int((np.abs(an) > 0).sum())
    res["fd_max_abs_err_over_max_grad"] = float(np.abs(fd - an).max() / scale)
    nz = np.abs(an) > 1e-3 * scale
    res["fd_rel_err_nonzero_max"] = float((np.abs(fd - an)[nz] / np.abs(an[nz])).max()) if nz.any() else 0.0
    fdp = []
    for k in range(2):
        pp = p0.clone()
        pp[k] += h
        lp = L(tex0, pp)
        pp[k] -= 2 * h
        lm = L(tex0, pp)
        fdp.append((lp - lm) / (2 * h))
    res["fd_params_rel_err"] = [float(abs(fdp[k] - gp[k]) / abs(gp[k])) for k in range(2)]

    # a few Adam steps on the workload reduce the loss
    opt = torch.optim.Adam([tex, params], lr=2e-2)
    losses = []
    for _ in range(30):
        opt.zero_grad()
        lo = ((A.shade(tex, params, gb) - target) ** 2).sum()
        lo.backward()
        opt.step()
        losses.append(float(lo.item()))
    res["loss_first"], res["loss_last"] = losses[0], losses[-1]
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
