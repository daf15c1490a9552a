"""The adjoint of texture sampling and shading: the FP64 oracle pinned by
closed forms and its own finite differences (CPU), then the GPU adjoint
against that oracle and against finite differences (GPU, own process)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KS = 0.5


def _gbuf(n, rng, normal=None, pos=None, uv=None):
    g = np.zeros((1, n, 12), np.float32)
    g[0, :, 0:3] = normal if normal is not None else rng.normal(size=(n, 3))
    g[0, :, 3] = 1.0
    g[0, :, 4:7] = pos if pos is not None else rng.uniform(-1, 1, (n, 3))
    g[0, :, 8:10] = uv if uv is not None else rng.uniform(-1, 2, (n, 2))
    return g


def test_forward_closed_forms(oracle):
    """light and camera straight along the normal: c = a*T + ks at a texel
    corner; light at 60 deg and camera on the normal: lam = cos 60 and
    s = cos 30; half-way between texels the sample is their mean"""
    rng = np.random.default_rng(0)
    tex = rng.uniform(0, 1, (5, 8, 3)).astype(np.float32)
    t64 = tex.astype(np.float64)
    u, v = 3 / 8, 2 / 5
    g = _gbuf(1, rng, normal=[[0, 0, 1]], pos=[[0.2, -0.1, 0.0]], uv=[[u, v]])
    out = oracle.diff_forward_f64(g, tex, 1.3, 7.0, KS, np.float32([0.2, -0.1, 10.0]), np.float32([0.2, -0.1, 4.0]))
    np.testing.assert_allclose(out[0, 0], 1.3 * t64[2, 3] + KS, rtol=1e-12)
    s60, c60 = np.sin(np.pi / 3), np.cos(np.pi / 3)
    light = np.float32([0.2 + 10 * s60, -0.1, 10 * c60])
    out = oracle.diff_forward_f64(g, tex, 1.3, 7.0, KS, light, np.float32([0.2, -0.1, 4.0]))
    np.testing.assert_allclose(out[0, 0], 1.3 * t64[2, 3] * c60 + KS * np.cos(np.pi / 6) ** 7, rtol=2e-6)
    g2 = _gbuf(1, rng, normal=[[0, 0, 1]], pos=[[0, 0, 0]], uv=[[(3 + 0.5) / 8, v]])
    out = oracle.diff_forward_f64(g2, tex, 1.0, 7.0, 0.0, np.float32([0, 0, 5]), np.float32([0, 0, 5]))
    np.testing.assert_allclose(out[0, 0], 0.5 * (t64[2, 3] + t64[2, 4]), rtol=1e-7)
    # wrapping: u = 1 + j/tw samples texel j
    g3 = _gbuf(1, rng, normal=[[0, 0, 1]], pos=[[0, 0, 0]], uv=[[1 + 3 / 8, v - 1]])
    out3 = oracle.diff_forward_f64(g3, tex, 1.0, 7.0, 0.0, np.float32([0, 0, 5]), np.float32([0, 0, 5]))
    np.testing.assert_allclose(out3[0, 0], t64[2, 3], rtol=1e-7)


def test_forward_is_affine_in_texture(oracle):
    rng = np.random.default_rng(1)
    g = _gbuf(50, rng)
    t1 = rng.uniform(0, 1, (6, 7, 3)).astype(np.float32)
    t2 = rng.uniform(0, 1, (6, 7, 3)).astype(np.float32)
    args = (1.1, 9.0, KS, np.float32([1, 2, 3]), np.float32([-1, 0.5, 2]))
    zero = oracle.diff_forward_f64(g, np.zeros_like(t1), *args)   # specular part alone
    a = oracle.diff_forward_f64(g, t1, *args) - zero
    b = oracle.diff_forward_f64(g, t2, *args) - zero
    ab = oracle.diff_forward_f64(g, (t1 + t2).astype(np.float32), *args) - zero
    np.testing.assert_allclose(ab, a + b, atol=1e-12)


def test_backward_matches_fp64_finite_differences(oracle):
    """L = sum(gout * out): dL/dtex and dL/d(a, e) by the adjoint equal
    central differences of the forward (h = 1e-6, FP64)"""
    rng = np.random.default_rng(2)
    g = _gbuf(40, rng)
    tex = rng.uniform(0, 1, (4, 5, 3)).astype(np.float64)
    gout = rng.normal(size=(1, 40, 3))
    light, cam = np.float32([1, 2, 3]), np.float32([-1, 0.5, 2])
    a, e = 1.1, 9.0
    gt, gp = oracle.diff_backward_f64(g, tex.astype(np.float32), a, e, KS, light, cam, gout)

    # texels enter linearly, so differences of a float32 texture are exact enough
    # only through the float32 cast: use the exact linear map instead
    def Lf(t, aa=a, ee=e):
        return float((oracle.diff_forward_f64(g, t.astype(np.float32), aa, ee, KS, light, cam) * gout).sum())

    base = Lf(np.zeros_like(tex))
    fd = np.zeros_like(tex)
    for idx in np.ndindex(tex.shape):
        t = np.zeros_like(tex)
        t[idx] = 1.0
        fd[idx] = Lf(t) - base          # exact: out is affine in tex
    np.testing.assert_allclose(gt, fd, rtol=1e-9, atol=1e-12)
    h = 1e-6
    tf = tex.astype(np.float32)
    da = (Lf(tf, a + h) - Lf(tf, a - h)) / (2 * h)
    de = (Lf(tf, a, e + h) - Lf(tf, a, e - h)) / (2 * h)
    np.testing.assert_allclose(gp, [da, de], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("size", ["reduced", "full"])
def test_gpu_adjoint(size):
    args = [sys.executable, "-u", os.path.join(HERE, "adjoint_check.py")] + (["--full"] if size == "full" else [])
    print("\n[gpu-case] adjoint check %s" % size, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res)
    assert res["runtime"] == "torch"
    assert res["covered_px"] > 0.2 * res["width"] ** 2
    assert res["forward_max_abs_err"] <= 1e-5
    assert res["grad_tex_rel_err"] <= 1e-5
    assert max(res["grad_params_rel_err"]) <= 1e-4
    assert res["fd_texels_nonzero"] > 0
    assert res["fd_max_abs_err_over_max_grad"] <= 1e-3
    assert max(res["fd_params_rel_err"]) <= 1e-2
    assert res["loss_last"] < 0.5 * res["loss_first"]
