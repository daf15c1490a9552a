#!/usr/bin/env python3
"""Benchmark of the Puresoft3D raster hot path on one MI355X.

Workload (one step = one frame): 1920x1080, seeded terrain of 300x300 cells
(180,000 triangles) plus 64 instanced spheres of 320 triangles, 2048^2 albedo
and normal maps with full mip chains, trilinear sampling, per-pixel
Blinn-Phong with normal mapping, one point light, FP32 depth.  A frame is
clearDepth + clearColour + 65 drawVAO calls, with scenes and textures resident
in HBM when the timed region starts.

Prints one JSON line (see the keys below).  The CPU baseline is the oracle
(a C restatement of the reference pipeline with the reference's threading:
one producer thread plus row-interleaved workers), run on a bounded number
of frames of the same workload on the host cores.

Further sections of the same line (each can be skipped):
  batched  256 views of 16 resident scenes in one graph-replayed launch,
           batch latency p50/p99 and views/s against one view per launch
  served   2,000 Poisson view requests over 32 resident scenes, continuous
           batching: latency p50/p99 and views/s against the plain path
  adjoint  optimisation steps/s of the differentiable texture workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fragments shaded/s + 1080p frames/s vs host CPU; p50/p99 view latency when batched"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(scene):
    """B = 3*T*S_v + W*H*(4 colour + 4 depth) + texture bytes (all mip levels)"""
    vb = scene.vertex_bytes()
    fb = scene.width * scene.height * 8
    tb = 0
    for t in scene.textures:
        for l in range(t.levels):
            tb += max(1, t.width >> l) * max(1, t.height >> l) * 4 * t.layers
    return vb, fb, tb


def kernel_alg_bytes(name, scene):
    vb, fb, tb = algorithmic_bytes(scene)
    px = scene.width * scene.height
    return {"setup": vb, "raster": px * 4, "shade": px * 4 + tb}.get(name)


PROFILE_NAMES = {"setup": "k_setup<7, false>", "raster": "k_raster<false, 2>", "shade": "k_shade<7>",
                 "span": "k_span", "segments": "k_segments<14>"}


def measured_traffic(kernel, path=os.path.join(ROOT, "profiles", "traffic_terrain1080p.md")):
    """HBM bytes per launch of a kernel of this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration it states), or None"""
    name = PROFILE_NAMES.get(kernel)
    if not name or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 7 and cells[0] == name:
            try:
                return int(round(float(cells[6]) * 1e6))
            except ValueError:
                return None
    return None


def measured_valu_share(kernel, path=os.path.join(ROOT, "profiles", "valu_terrain1080p.md")):
    """VALU issue share of a kernel of this workload (SQ_INSTS_VALU x 4 cycles
    over the SIMD-cycles of its duration at peak clock), from the committed
    rocprofv3 SQ-counter summary, or None"""
    name = PROFILE_NAMES.get(kernel)
    if not name or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 9 and cells[0] == name and cells[8].endswith("%"):
            try:
                return round(float(cells[8][:-1]) / 100.0, 3)
            except ValueError:
                return None
    return None


def host_cpus():
    """logical CPUs of the host, those this process may run on, and the model"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return os.cpu_count() or 1, usable, model


def cpu_baseline(scene, seconds, threads, warmup=3, max_frames=30):
    """the oracle (1 producer + threads-1 row-interleaved spinning workers, the
    reference's architecture) on `warmup` untimed and up to `max_frames` timed
    frames, bounded by `seconds`"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    frames, t_frames = 0, []
    frag = 0
    for _ in range(warmup):   # page-in, thread pool
        pyoracle.render_scene(scene, n_threads=threads)
    t_start = time.perf_counter()
    while frames < max_frames and (frames < 3 or time.perf_counter() - t_start < seconds):
        t0 = time.perf_counter()
        _, st = pyoracle.render_scene(scene, n_threads=threads)
        t_frames.append(time.perf_counter() - t0)
        frag += st.frags_shaded
        frames += 1
    pyoracle.pool_shutdown()
    tot = sum(t_frames)
    return {"value": frames / tot, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "%d timed frames (after %d untimed) of the same 1080p workload (clear + 65 draws each), oracle "
                      "with 1 producer + %d row-interleaved worker threads" % (frames, warmup, threads - 1),
            "ms_per_frame_p50": round(float(np.median(t_frames) * 1e3), 2),
            "ms_per_frame_p99": round(float(np.percentile(t_frames, 99) * 1e3), 2),
            "fragments_shaded_per_s": frag / tot}


def batch_section(args, N, S, log):
    """256 views (area log-uniform 256^2..1024^2) over 16 resident scenes of
    10k-500k triangles, 1024^2 mipmapped textures, Blinn-Phong trilinear;
    the whole batch is one recorded command list replayed as a hipGraph."""
    from puresoft3d_amd import autotune as AT
    t0 = time.perf_counter()
    assets = S.batch_assets(n_scenes=16, seed=11, tri_lo=10_000, tri_hi=500_000, tex_size=1024)
    views = S.batch_views(assets, n_views=256, seed=12, side_lo=256, side_hi=1024)
    log("[bench] batch: %d scenes (%d..%d tris), %d views, generated in %.1f s" % (
        len(assets), min(a.triangles() for a in assets), max(a.triangles() for a in assets), len(views),
        time.perf_counter() - t0))
    ctx = N.Context(64, 64)
    b = S.GpuBatch(ctx, assets, views)
    ctx.capture_begin()
    b.submit()
    g = ctx.capture_end()
    run = lambda: ctx.graph_launch(g, 1)  # noqa: E731
    tuned = None
    if not args.no_tune:
        best, table = AT.autotune(ctx, run, reps=3)
        tuned = {"best": {"chunk_width": best[0], "tile_rows": best[1], "waves_per_group": best[2]},
                 "searched": len(table),
                 "top_ms": {"%dx%d/c%d/w%d" % (r[1], r[2], r[0][0], r[0][2]): round(r[3], 3)
                            for r in sorted(table, key=lambda r: r[3])[:5]}}
    for _ in range(3):
        run()
    ctx.sync()
    ctx.reset_stats()
    lat = []
    for _ in range(args.batch_steps):
        t1 = time.perf_counter()
        run()
        ctx.sync()
        lat.append(time.perf_counter() - t1)
    st = ctx.stats()
    tot = sum(lat)
    px = sum(v.width * v.height for v in views)
    # per-kernel time of one batch (eager, HIP events on the context stream)
    ctx.profile(True)
    ctx.profile_reset()
    b.submit()
    ctx.sync()
    kprof = {k: round(v[0], 4) for k, v in ctx.profile_read().items()}
    ctx.profile(False)
    # one view per launch, same views: the plain path's per-view latency
    single = []
    for i in range(0, len(views), 8):
        t1 = time.perf_counter()
        b.submit([i])
        ctx.sync()
        single.append(time.perf_counter() - t1)
    ctx.close()
    out = {"views": len(views), "scenes": len(assets), "pixels_per_batch": px,
           "triangles_per_batch": int(sum(assets[v.scene].triangles() for v in views)),
           "batch_ms_p50": round(float(np.median(lat)) * 1e3, 3),
           "batch_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "views_per_s": round(len(views) * args.batch_steps / tot, 1),
           "fragments_shaded_per_s": round(st["frags_shaded"] / tot, 1),
           "single_view_ms_p50": round(float(np.median(single)) * 1e3, 3),
           "single_view_views_per_s": round(1.0 / float(np.mean(single)), 1),
           "kernel_ms_per_batch": kprof}
    if tuned:
        out["autotune"] = tuned
    return out, assets, views


def serve_section(args, N, S, log):
    """2,000 Poisson view requests (area log-uniform 128^2..2048^2) over 32
    resident scenes (1k-1M triangles, 512^2-4096^2 mipmapped textures),
    continuous batching with HBM-budget admission on one render stream,
    against the plain one-view-per-launch path on the same GPU; then the same
    batching over two render streams (contexts) at the lower loads."""
    from puresoft3d_amd import serving as SV
    t0 = time.perf_counter()
    assets = S.serving_assets(n_scenes=32, seed=21)
    n = args.serve_requests
    views, _ = S.poisson_requests(assets, n=n, rate=1.0, seed=23)
    log("[bench] serve: %d scenes (%.2f M tris), %d requests, generated in %.1f s" % (
        len(assets), sum(a.triangles() for a in assets) / 1e6, n, time.perf_counter() - t0))
    ctxs = [N.Context(64, 64) for _ in range(max(1, args.serve_streams))]
    ctx = ctxs[0]
    srv = SV.RenderServer(ctx, assets, extra_ctxs=ctxs[1:])
    cfg = dict(max_batch_views=args.serve_batch_views, max_batch_pixels=args.serve_batch_pixels << 20,
               max_inflight=2, hbm_budget=args.serve_budget_gib << 30, streams=1)
    plain_cfg = dict(max_batch_views=1, max_inflight=1, streams=1)
    # two batches in flight in all: queued on one stream, or one on each of two
    cfg2 = dict(cfg, max_inflight=1, streams=len(ctxs))
    zeros = np.zeros(n)
    srv.serve(views, zeros, **cfg)                       # warm-up: target pools and scratch sizes
    if len(ctxs) > 1:
        srv.serve(views, zeros, **cfg2)
    srv.serve(views, zeros, **plain_cfg)
    out = {"requests": n, "scenes": len(assets), "config": {k: int(v) for k, v in cfg.items()}}
    if not args.no_tune:   # raster tiling for this mix of views, on a sample of 64 requests
        from puresoft3d_amd import autotune as AT
        sample = views[:64]
        best, _ = AT.autotune(ctx, lambda: srv.serve(sample, zeros[:64], to_host=False, **cfg), reps=2,
                              cands=AT.candidates(waves=(1, 2)))
        for c in ctxs[1:]:
            c.tiling(*best)
        out["tiling"] = {"chunk_width": best[0], "tile_rows": best[1], "waves_per_group": best[2]}
    keep = list(range(0, n, max(1, n // 8)))
    for mode in ("host", "device"):
        to_host = mode == "host"
        # plain path: one view per launch, back to back
        n_plain = min(n, 600)
        res_p, st_p, _ = srv.serve(views[:n_plain], zeros[:n_plain], to_host=to_host, **plain_cfg)
        single = res_p["done_s"] - res_p["launched_s"]
        plain_vps = n_plain / st_p["wall_s"]
        sec = {"plain": {"views_per_s": round(plain_vps, 1),
                         "host_submit_ms_per_view": round(st_p["host_submit_s"] * 1e3 / n_plain, 4),
                         "single_view_ms_p50": round(float(np.percentile(single, 50)) * 1e3, 3),
                         "single_view_ms_p99": round(float(np.percentile(single, 99)) * 1e3, 3)},
               "runs": []}
        for load in args.serve_loads:
            rate = load * plain_vps
            _, arr = S.poisson_requests(assets, n=n, rate=rate, seed=23)
            res, st, imgs = srv.serve(views, arr, keep=keep if to_host else None, to_host=to_host, **cfg)
            summ = SV.latency_summary(res)
            lat = res["done_s"] - res["arrival_s"]
            run = {"offered_x_plain": load, "offered_views_per_s": round(rate, 1),
                   "views_per_s": round(summ["views_per_s"], 1),
                   "speedup_vs_plain": round(summ["views_per_s"] / plain_vps, 2),
                   "latency_ms_p50": round(summ["p50_ms"], 3), "latency_ms_p99": round(summ["p99_ms"], 3),
                   "p99_vs_single_view_p99": round(float(np.percentile(lat, 99)) /
                                                   float(np.percentile(single, 99)), 2),
                   "batches": st["batches"], "mean_batch_views": round(st["views"] / max(1, st["batches"]), 1),
                   "host_submit_ms_per_batch": round(st["host_submit_s"] * 1e3 / max(1, st["batches"]), 4),
                   "peak_target_bytes": st["peak_bytes"], "target_allocs": st["target_allocs"],
                   "target_evictions": st["target_evictions"]}
            if to_host:
                same = sum(int(np.array_equal(imgs[i], srv.render_alone(views[i]))) for i in keep)
                run["images_identical_to_alone"] = "%d/%d" % (same, len(keep))
            sec["runs"].append(run)
            log("[bench] serve (%s) load %.1fx: %s" % (mode, load, json.dumps(run)))
        out["deliver_" + mode] = sec
    # latency against the batch size cap, device delivery at twice the plain rate
    sweep = []
    plain_dev = out["deliver_device"]["plain"]["views_per_s"]
    _, arr = S.poisson_requests(assets, n=n, rate=2.0 * plain_dev, seed=23)
    for cap in args.serve_caps:
        c2 = dict(cfg, max_batch_pixels=cap << 20)
        res, st, _ = srv.serve(views, arr, to_host=False, **c2)
        summ = SV.latency_summary(res)
        sweep.append({"max_batch_mpix": cap, "views_per_s": round(summ["views_per_s"], 1),
                      "latency_ms_p50": round(summ["p50_ms"], 3), "latency_ms_p99": round(summ["p99_ms"], 3),
                      "mean_batch_views": round(st["views"] / max(1, st["batches"]), 1)})
        log("[bench] serve (device, 2x) cap %d MPix: %s" % (cap, json.dumps(sweep[-1])))
    out["batch_cap_sweep_device_2x"] = sweep
    # the same batching over two render streams, one batch in flight on each,
    # at the loads below saturation (above it the streams' separate target
    # pools churn inside the budget)
    if len(ctxs) > 1:
        for mode in ("host", "device"):
            to_host = mode == "host"
            plain_vps = out["deliver_" + mode]["plain"]["views_per_s"]
            two = []
            for load in [x for x in args.serve_loads if x <= 2.0]:
                _, arr = S.poisson_requests(assets, n=n, rate=load * plain_vps, seed=23)
                res, st, _ = srv.serve(views, arr, to_host=to_host, **cfg2)
                summ = SV.latency_summary(res)
                two.append({"offered_x_plain": load, "views_per_s": round(summ["views_per_s"], 1),
                            "speedup_vs_plain": round(summ["views_per_s"] / plain_vps, 2),
                            "latency_ms_p50": round(summ["p50_ms"], 3), "latency_ms_p99": round(summ["p99_ms"], 3),
                            "mean_batch_views": round(st["views"] / max(1, st["batches"]), 1),
                            "target_allocs": st["target_allocs"], "target_evictions": st["target_evictions"]})
                log("[bench] serve (%s, two streams) load %.1fx: %s" % (mode, load, json.dumps(two[-1])))
            out["two_streams_" + mode] = two
    srv.close()
    for c in ctxs:
        c.close()
    return out


def adjoint_section(args, log):
    """optimisation steps of the adjoint workload (512^2, 50k-tri displaced
    sphere, 1024^2 FP32 RGB texture), in a process of its own: it uses torch
    on the GPU, and torch must load its HIP runtime before libpsgpu does"""
    import subprocess
    r = subprocess.run([sys.executable, "-u", "-m", "puresoft3d_amd.adjoint", "--bench", "--steps",
                        str(args.adjoint_steps)], capture_output=True, text=True, timeout=600, cwd=ROOT)
    if r.returncode != 0:
        raise RuntimeError("adjoint bench failed: " + r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    log("[bench] adjoint: %s" % json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="reduced scene for quick checks")
    ap.add_argument("--mode", choices=["graph", "eager"], default="graph")
    ap.add_argument("--no-tune", action="store_true", help="keep the default raster tiling (frame and batch)")
    ap.add_argument("--tiling", type=str, default="", metavar="CW,TR,W",
                    help="fixed raster tiling of the 1080p frame (chunk width, tile rows, waves per group); implies "
                         "no tuning of it")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-view batch workload")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving workload")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-loads", type=float, nargs="+", default=[1.5, 2.0, 3.0, 4.0],
                    help="offered request rates as multiples of the plain path's views/s")
    ap.add_argument("--serve-batch-views", type=int, default=64)
    ap.add_argument("--serve-caps", type=int, nargs="*", default=[2, 4, 8], metavar="MPIX",
                    help="batch pixel caps swept at twice the plain rate (device delivery)")
    ap.add_argument("--serve-batch-pixels", type=int, default=16, metavar="MPIX")
    ap.add_argument("--serve-budget-gib", type=int, default=8)
    ap.add_argument("--serve-streams", type=int, default=2, help="render streams (contexts) of the batched server")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--frame-only", action="store_true", help="only the 1080p frame (and the CPU baseline)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="also run one plain copy kernel of MIB MiB (k_selftest op 7) to calibrate HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (views would be sharded "
              "across GPUs with scenes replicated). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # the adjoint runs in a child process, started before this one touches the GPU
    if args.frame_only:
        args.no_batch = args.no_serve = args.no_adjoint = True
    adj = adjoint_section(args, lambda m: print(m, flush=True)) if not args.no_adjoint else None

    from puresoft3d_amd import native as N
    from puresoft3d_amd import scenes as S

    if args.serve_only:
        log = lambda m: print(m, flush=True)  # noqa: E731
        print(json.dumps({"served": serve_section(args, N, S, log)}), flush=True)
        return

    if args.small:
        scene = S.terrain_scene(width=960, height=540, cells=150, n_spheres=16, tex_size=1024)
    else:
        scene = S.terrain_scene()
    T = scene.triangles()
    print("[bench] scene %s: %d triangles" % (scene.name, T), flush=True)
    ctx = N.Context(scene.width, scene.height)
    gs = S.GpuScene(ctx, scene)
    print("[bench] resident; warming up", flush=True)

    def frame():
        gs.frame()

    graph = None
    if args.mode == "graph":
        ctx.capture_begin()
        gs.frame()
        graph = ctx.capture_end()

        def frame():  # noqa: F811
            ctx.graph_launch(graph, 1)

    tuned = None
    if args.tiling:
        ctx.tiling(*[int(v) for v in args.tiling.split(",")])
    elif not args.no_tune:
        from puresoft3d_amd import autotune as AT
        best, table = AT.autotune(ctx, frame, reps=5)
        top = sorted(table, key=lambda r: r[3])[:5]
        tuned = {"best": {"chunk_width": best[0], "tile_rows": best[1], "waves_per_group": best[2]},
                 "searched": len(table),
                 "top_ms": {"%dx%d/c%d/w%d" % (r[1], r[2], r[0][0], r[0][2]): round(r[3], 4) for r in top}}
        print("[bench] tiling autotuned: %s" % json.dumps(tuned), flush=True)
    for _ in range(args.warmup):
        frame()
    ctx.sync()
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.sync()
    dt = time.perf_counter() - t0
    st = ctx.stats()
    fps = args.steps / dt
    print("[bench] timed %d frames in %.3f s" % (args.steps, dt), flush=True)

    # end to end: the same frames, each read back to host memory (8.3 MB of
    # BGRA8 over PCIe) before the next starts
    host_img = np.empty((scene.height, scene.width, 4), np.uint8)
    n_e2e = max(10, min(args.steps // 4, 50))
    t0 = time.perf_counter()
    for _ in range(n_e2e):
        frame()
        ctx.read_colour(out=host_img)
    e2e_fps = n_e2e / (time.perf_counter() - t0)
    print("[bench] end to end (readback included): %.1f frames/s" % e2e_fps, flush=True)

    # per-kernel durations with HIP events on the context stream
    ctx.profile(True)
    ctx.profile_reset()
    nprof = max(10, min(args.steps, 50))
    for _ in range(nprof):
        gs.frame()
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    per_kernel = {k: v[0] / max(1, v[1]) for k, v in prof.items()}
    dom = max(per_kernel, key=per_kernel.get)
    alg = kernel_alg_bytes(dom, scene)
    vb, fb, tb = algorithmic_bytes(scene)
    B = vb + fb + tb
    roof = {"bound": "hbm", "kernel": dom, "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
            "traffic": measured_traffic(dom)}
    if roof["traffic"] is not None:
        roof["traffic_source"] = "profiles/traffic_terrain1080p.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)"
    valu = measured_valu_share(dom)
    if valu is not None:   # the exact-FP32 shading is bound by vector ALU issue, not HBM
        roof["valu_issue_share"] = valu
        roof["valu_source"] = "profiles/valu_terrain1080p.md (rocprofv3 SQ_INSTS_VALU per wave, peak clock)"
    if alg:
        achieved = alg / (per_kernel[dom] * 1e-3) / 1e9
        roof.update(achieved=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4), alg_bytes=alg,
                    avg_ms=round(per_kernel[dom], 4))

    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural terrain, spheres and noise textures)",
        "config": {"workload": scene.name, "width": scene.width, "height": scene.height, "triangles": T,
                   "shading": "Blinn-Phong + normal map, trilinear 2048^2 mipmapped", "mode": args.mode},
        "e2e_frames_per_s": round(e2e_fps, 1),
        "e2e_note": "frames each followed by the 8.3 MB colour readback to host memory (PCIe-inclusive); not value",
        "fragments_shaded_per_s": round(st["frags_shaded"] / dt, 1),
        "fragments_rasterised_per_s": round(st["frags_rasterised"] / dt, 1),
        "frame_alg_bytes": B,
        "frame_alg_GBs": round(B * fps / 1e9, 1),
        "kernel_ms": {k: round(v, 4) for k, v in sorted(per_kernel.items())},
        "tiling": tuned,
        "roofline": roof,
    }
    if args.calib_copy:
        n = args.calib_copy * (1 << 20) // 4
        ctx.selftest_math(7, np.ones(n, np.float32))
        out["calib_copy_bytes"] = {"read": n * 4, "written": n * 4}
    ctx.close()
    if not args.no_batch:
        log = lambda m: print(m, flush=True)  # noqa: E731
        out["batched"], _, _ = batch_section(args, N, S, log)
    if not args.no_serve:
        log = lambda m: print(m, flush=True)  # noqa: E731
        out["served"] = serve_section(args, N, S, log)
    if adj is not None:
        out["adjoint"] = adj
    if not args.no_cpu:
        nproc, usable, model = host_cpus()
        # every CPU this process may use, up to the 16 threads one GPU's share
        # of the host holds; the reference's minimum (2: a producer and one
        # worker) as a second point
        thr = args.cpu_threads or max(2, min(usable, 16))
        print("[bench] cpu baseline: oracle with %d threads for <= %.0f s" % (thr, args.cpu_seconds), flush=True)
        cb = cpu_baseline(scene, args.cpu_seconds, thr)
        cb.update(nproc=nproc, usable_cpus=usable, cpu_model=model)
        if thr != 2:
            print("[bench] cpu baseline: oracle with 2 threads for <= %.0f s" % (args.cpu_seconds / 2), flush=True)
            c2 = cpu_baseline(scene, args.cpu_seconds / 2, 2, warmup=1, max_frames=10)
            cb["n2"] = {k: c2[k] for k in ("value", "cores", "sample", "ms_per_frame_p50", "fragments_shaded_per_s")}
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = round(fps / cb["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
