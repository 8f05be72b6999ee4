"""Benchmark: LBVH build + thin-lens render of a 1M-triangle random soup on one MI355X.

Headline step ("soup1M"): a 1,000,000-triangle random soup (centroids
uniform in the unit cube, lognormal edges with median 0.002 and sigma 0.5,
uniformly random orientation, fp32), 63-bit Morton LBVH build (triangle
AABBs + scene bounds -> Morton keys -> onesweep radix sort -> fused
hierarchy/refit climb), then a 1920x1080 thin-lens depth-of-field render
(aperture 0.05, focus distance 1.5, fov 40) at 16 spp with one shadow ray
per hit. Inputs are resident in HBM before the timed region.

Prints one JSON line. value = primary (+ bounce) rays per second of the
whole step (build + render), in Mrays/s. The default run then measures the
rest of the metric on its own workloads and adds them as extra keys:
refit-only frames of a 10M-triangle displaced sphere (refit Mtris/s), 64
batched scenes against the per-scene launch loop (speed-up), and the
differentiable render at 200K triangles (backward / forward time).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cuda-rt_amd"))

HBM_PEAK_GBS = 8000.0
# FP64 vector issue rate of MI355X: 256 CU x 64 lanes x 2.4 GHz = 39.3 T lane-op/s
# (the published 78.6 TFLOP/s counts an FMA as two flops).
FP64_LANE_OPS_PEAK_T = 39.3
# Algorithmic work per unit (fp32 canonical counts; FP64 evaluation counts twice).
OPS_PER_BOX_TEST = 16       # one slab test (a binary node entered = 2, a 4-wide node = 4)
OPS_PER_TRI_TEST = 42       # Moller-Trumbore, evaluated in FP64
FP64_FACTOR = 2
# The default traversal layout (4-wide nodes) tests child boxes in fp32.
BOX_TEST_FACTOR = 1
# Algorithmic HBM bytes per triangle of a 63-bit build: vertices read twice
# (72 B), key + id written (12 B), 8 sort passes of 24 B + 8 B of histogram,
# sorted keys read (4 B), topology written (16 B), refit gathers (40 B) and
# 2N boxes written (48 B).
BUILD_BYTES_PER_TRI_63 = 396


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--workload", choices=["soup1M", "batch64", "diff200k", "sphere10M", "build1M", "refit10M"],
                    default="soup1M",
                    help="soup1M: the headline single-scene step; batch64: 64 scenes, batched vs per-scene loop; "
                         "diff200k: differentiable render forward + adjoint backward; "
                         "sphere10M: refit-only frames of a 10M-triangle displaced sphere with diffuse bounces; "
                         "build1M / refit10M: the soup's 63-bit rebuild / the sphere's refit alone (for counters)")
    ap.add_argument("--scenes", type=int, default=64, help="batch64: number of scenes")
    ap.add_argument("--mode", choices=["wavefront", "megakernel"], default="wavefront",
                    help="render kernel structure (identical images)")
    ap.add_argument("--one-stream", action="store_true",
                    help="soup1M: every render pass on one stream (for per-kernel profiles; same image)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="headline: keep the default tuning (by default rt_render_autotune picks the single-scene "
                         "tuning before the timed steps)")
    ap.add_argument("--headline-only", action="store_true",
                    help="soup1M: skip the refit, batch and differentiable legs of the metric")
    ap.add_argument("--cpu-pixel-stride", type=int, default=1,
                    help="CPU baseline renders every k-th pixel (launch order)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run device copies of this many MiB (HBM counter calibration) and exit")
    return ap.parse_args()


def committed_traffic(kind, args):
    """HBM bytes per traversal launch of the default workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/traffic_soup1M.json),
    scaled by the copy-kernel calibration recorded there; None for another
    workload."""
    path = os.path.join(ROOT, "profiles", "traffic_soup1M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    w = rec["workload"]
    if (w["tris"], w["width"], w["height"]) != (args.tris, args.width, args.height) or args.mode != "wavefront":
        return None
    k = rec["kernels"][kind]
    return {"bytes": int(k["bytes_per_launch"]), "fetch_bytes": int(k["fetch_bytes"]),
            "write_bytes": int(k["write_bytes"]), "source": "profiles/traffic_soup1M.json"}


def committed_build_traffic(name, tris):
    """HBM bytes of one build or refit (the sum over its kernels), from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes in
    profiles/traffic_build.json, scaled by the copy calibration recorded
    there; None if that workload was not profiled at this size."""
    path = os.path.join(ROOT, "profiles", "traffic_build.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    w = rec.get("workloads", {}).get(name)
    if not w or w["tris"] != tris:
        return None
    return {"bytes": int(w["bytes"]), "fetch_bytes": int(w["fetch_bytes"]), "write_bytes": int(w["write_bytes"]),
            "bytes_per_tri": round(w["bytes"] / tris, 1), "source": "profiles/traffic_build.json"}


def calibrate_copy(mib):
    import torch
    a = torch.empty(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "bytes_read": a.numel() * 4, "bytes_written": a.numel() * 4,
                      "copies": 3}))


def host_info():
    """What the CPU baseline ran on: logical CPUs visible to this process and
    the CPU model (from /proc/cpuinfo, as lscpu reports it)."""
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
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"nproc": affinity or os.cpu_count(), "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(hm, mats, light, cam, opts, tris, stride_all=4, stride_one=64):
    """Time the FP64 CPU oracle (a restatement of the reference algorithm) on
    bounded samples of the same workload: the full 63-bit build, then the
    render of every stride-th launch pixel, once on all host threads (at
    most 16, the box's CPU share) and once on one thread."""
    import numpy as np
    from oracle import oracle as O
    info = host_info()
    threads = max(1, min(info["nproc"] or 1, 16))

    def run(nthreads, stride):
        t0 = time.perf_counter()
        b = O.build(hm, 63, nthreads=nthreads)
        t_build = time.perf_counter() - t0
        pixels = np.arange(0, cam.width * cam.height, stride, dtype=np.int64)
        t0 = time.perf_counter()
        _, counts = O.render(hm, b, mats, light, cam, opts, pixels=pixels, nthreads=nthreads)
        t_render = time.perf_counter() - t0
        return t_build, t_render, pixels.size, counts

    tb, tr, npx, c = run(threads, stride_all)
    tb1, tr1, npx1, c1 = run(1, stride_one)
    return {
        "value": round((c[0] + c[2]) / tr / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle FP64 build of all {tris} tris ({tb:.2f} s) + render of every {stride_all}th launch "
                  f"pixel ({npx} px x {opts.spp} spp = {int(c[0])} primary, {int(c[1])} shadow rays, "
                  f"{tr:.2f} s) on {threads} threads",
        "build_mtris_s": round(tris / tb / 1e6, 4),
        "single_thread": {
            "value": round((c1[0] + c1[2]) / tr1 / 1e6, 4), "unit": "Mrays/s", "cores": 1,
            "build_mtris_s": round(tris / tb1 / 1e6, 4),
            "sample": f"build of all {tris} tris ({tb1:.2f} s) + render of every {stride_one}th launch pixel "
                      f"({npx1} px, {int(c1[0])} primary rays, {tr1:.2f} s)"},
        "host": info,
    }


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def run_batch(ctx, steps, warmup, count=64):
    """C4: 64 random soups (log-uniform 1K..2M triangles), cameras of 256^2,
    512^2 or 1024^2 at 1..8 spp. Step = rebuild every BVH + render every
    scene. The per-scene launch loop against the batched form (autotuned,
    replayed from a hipGraph); the images must be identical."""
    import torch

    import rtamd
    from rtamd import workloads
    specs = workloads.batch_scenes(count)
    sc = []
    for sp in specs:
        dm = rtamd.DeviceMesh(sp["hm"])
        sp["opts"].flags |= rtamd.RT_RENDER_WAVEFRONT
        sc.append(dict(mesh=dm, bvh=rtamd.BVH(ctx, dm, 63), mats=rtamd.materials_tensor(sp["mats"]),
                       nmat=len(sp["mats"]), light=sp["light"], cam=sp["cam"], opts=sp["opts"],
                       rgb=torch.zeros(sp["cam"].width * sp["cam"].height * 3, dtype=torch.float32, device="cuda")))
    tris = sum(s["mesh"].n for s in sc)
    pixels = sum(s["cam"].width * s["cam"].height for s in sc)
    primary = sum(s["cam"].width * s["cam"].height * s["opts"].spp for s in sc)
    loop_rgb = [torch.zeros_like(s["rgb"]) for s in sc]

    def loop_step():
        for s, out in zip(sc, loop_rgb):
            s["bvh"].rebuild(s["mesh"])
            rtamd.render(ctx, s["bvh"], s["mesh"], s["mats"], s["nmat"], s["light"], s["cam"], s["opts"], rgb=out)

    loop_ms = _timed(loop_step, steps, warmup)
    batch = rtamd.Batch(ctx, sc)
    best, tuned_ms = batch.autotune()
    batch.rebuild()
    batch.render()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            batch.rebuild()
            batch.render()
    torch.cuda.current_stream().wait_stream(side)
    batch_ms = _timed(g.replay, steps, warmup)
    same = all(torch.equal(a, s["rgb"]) for a, s in zip(loop_rgb, sc))
    res = {
        "workload": f"batch{len(sc)}_soups_1K-2M_tris_256-1024px_1-8spp", "scenes": len(sc), "tris": tris,
        "pixels": pixels, "primary_rays": primary, "steps": steps, "warmup": warmup,
        "speedup": round(loop_ms / batch_ms, 3),
        "loop_ms": round(loop_ms, 3), "batch_graph_ms": round(batch_ms, 3),
        "batch_mrays_s": round(primary / (batch_ms * 1e-3) / 1e6, 2),
        "loop_mrays_s": round(primary / (loop_ms * 1e-3) / 1e6, 2),
        "identical_images": bool(same),
        "tuning": best.as_dict(), "autotune_best_render_ms": round(tuned_ms, 3),
        "stack_overflows": batch.diagnostics(),
    }
    batch.close()
    for s in sc:
        s["bvh"].close()
    return res


# Algorithmic HBM bytes per triangle of a refit-only frame: 12 B indices +
# 36 B gathered vertices + 4 B ids + 48 B box writes + 48 B child-box reads
# + 4 B arrival flags.
REFIT_BYTES_PER_TRI = 152


def run_sphere(ctx, steps, warmup):
    """C3: refit-only frames of a 9,996,980-triangle icosphere (frequency 707,
    radius 1, 3-octave value-noise displacement of amplitude 0.05), 30-bit
    build once; each frame moves every vertex by Gaussian jitter (sigma
    1e-4, precomputed frames resident in HBM), refits the BVH and renders
    1024x1024 at 4 spp with cosine-weighted diffuse bounces to depth 4 and
    one shadow ray per hit. Step = refit + render."""
    import numpy as np
    import torch

    import rtamd
    from rtamd import workloads
    hm = workloads.sphere10m_mesh()
    frames = []
    for k in range(3):
        j = workloads.sphere10m_frame(hm, k)
        frames.append([torch.from_numpy(a).cuda() for a in (j.vx, j.vy, j.vz)])
    dm = rtamd.DeviceMesh(hm)
    mats, light, cam, opts = workloads.sphere10m_setup()
    opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    md = rtamd.materials_tensor(mats)
    bvh = rtamd.BVH(ctx, dm, 30)
    rgb = torch.empty(cam.width * cam.height * 3, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    state = {"k": 0}

    def step(ev=None, cnt=None):
        f = frames[state["k"] % len(frames)]
        state["k"] += 1
        dm.vx, dm.vy, dm.vz = f
        if ev:
            ev[0].record(stream)
        bvh.refit(dm)
        if ev:
            ev[1].record(stream)
        rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, opts, rgb=rgb, counters=cnt)
        if ev:
            ev[2].record(stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    refit_ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
    render_ms = float(np.median([e[1].elapsed_time(e[2]) for e in evs]))
    # ray counts (exact) from one more, untimed frame
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    step(cnt=cnt)
    primary, shadow, bounce = (int(x) for x in cnt.cpu().numpy()[:3])
    ms = wall * 1e3 / steps
    refit_gbs = REFIT_BYTES_PER_TRI * hm.n / (refit_ms * 1e-3) / 1e9
    res = {
        "workload": "sphere10M_refit_1024x1024_4spp_depth4", "tris": hm.n, "steps": steps, "warmup": warmup,
        "ms_per_step": round(ms, 3),
        "mrays_s": round((primary + bounce) / (ms * 1e-3) / 1e6, 2),
        "refit_ms": round(refit_ms, 4), "refit_mtris_s": round(hm.n / (refit_ms * 1e-3) / 1e6, 1),
        "roofline_refit": {"kernels": "k_init..k_wide_fill (one refit)", "bound": "hbm",
                           "achieved": round(refit_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(refit_gbs / HBM_PEAK_GBS, 4), "bytes_per_tri": REFIT_BYTES_PER_TRI,
                           "traffic": committed_build_traffic("refit_sphere10M", hm.n)},
        "render_ms": round(render_ms, 3),
        "render_mrays_s": round((primary + bounce) / (render_ms * 1e-3) / 1e6, 2),
        "primary_rays": primary, "bounce_rays": bounce, "shadow_rays": shadow,
        "diagnostics": dict(zip(("stack_overflows", "build_errors"), bvh.diagnostics())),
    }
    bvh.close()
    return res


def run_diff(ctx, steps, warmup):
    """C5: differentiable render step on a 200K-triangle displaced icosphere
    (frequency 100, displacement 0.05): 512x512, 16 spp, 2 diffuse bounces,
    per-triangle albedo U[0.2, 0.9], light intensity 10. The target image is
    rendered at the true albedo; the parameters are the albedo perturbed by
    N(0, 0.05). Step = forward (render + path record) + backward (adjoint
    kernel) of the L2 image loss."""
    import numpy as np
    import torch

    import rtamd
    from rtamd import diff, workloads
    hm = workloads.diff200k_mesh()
    dm = rtamd.DeviceMesh(hm)
    light, cam = workloads.diff200k_setup()
    spp, depth = 16, 2
    R = diff.DiffRenderer(ctx, dm, light, cam, spp, depth, seed=1)
    true_alb, pert_alb = workloads.diff200k_albedo(hm.n)
    alb = torch.tensor(pert_alb, dtype=torch.float64, device="cuda")
    target = R.forward(torch.tensor(true_alb, dtype=torch.float64, device="cuda"), 10.0).clone()
    s = 10.0

    def step(ev=None):
        if ev:
            ev[0].record()
        img = R.forward(alb, s)
        if ev:
            ev[1].record()
        g = R.backward(alb, s, 2.0 * (img - target))
        if ev:
            ev[2].record()
        return g

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        g_alb, g_light, g_pos = step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
    bwd_ms = float(np.median([e[1].elapsed_time(e[2]) for e in evs]))
    ms = wall * 1e3 / steps
    primary = cam.width * cam.height * spp
    return {
        "workload": "diff_icosphere200K_512x512_16spp_depth2", "tris": hm.n, "steps": steps, "warmup": warmup,
        "ms_per_step": round(ms, 3), "mpaths_s": round(primary / (ms * 1e-3) / 1e6, 2),
        "forward_ms": round(fwd_ms, 3), "backward_ms": round(bwd_ms, 3),
        "bwd_over_fwd": round(bwd_ms / fwd_ms, 3),
        "loss": float(((target - R.forward(alb, s)) ** 2).sum()),
        "grad_norms": {"albedo": float(g_alb.norm()), "light": float(g_light), "positions": float(g_pos.norm())},
        "record_mib": round(R.records.numel() * 4 / 2**20, 1),
    }


def run_build_only(ctx, steps, warmup, refit):
    """Steps of the build alone: a 63-bit rebuild of the 1M soup (refit =
    False) or a refit of the 10M sphere on alternating jittered positions
    (refit = True). Reports the median step time from HIP events."""
    import numpy as np
    import torch

    import rtamd
    from rtamd import workloads
    if refit:
        hm = workloads.sphere10m_mesh()
        frames = [[torch.from_numpy(a).cuda() for a in (f.vx, f.vy, f.vz)]
                  for f in (workloads.sphere10m_frame(hm, k) for k in range(2))]
        bits, bytes_per_tri = 30, REFIT_BYTES_PER_TRI
    else:
        hm = workloads.soup_mesh(1_000_000)
        frames = None
        bits, bytes_per_tri = 63, BUILD_BYTES_PER_TRI_63
    dm = rtamd.DeviceMesh(hm)
    bvh = rtamd.BVH(ctx, dm, bits)
    stream = torch.cuda.current_stream()
    state = {"k": 0}

    def step(ev=None):
        if frames:
            dm.vx, dm.vy, dm.vz = frames[state["k"] % 2]
            state["k"] += 1
        if ev:
            ev[0].record(stream)
        if refit:
            bvh.refit(dm)
        else:
            bvh.rebuild(dm)
        if ev:
            ev[1].record(stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    for k in range(steps):
        step(evs[k])
    torch.cuda.synchronize()
    ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
    gbs = bytes_per_tri * hm.n / (ms * 1e-3) / 1e9
    name = "refit_sphere10M" if refit else "build_soup1M"
    res = {"workload": name, "tris": hm.n, "ms_per_step": round(ms, 4), "mtris_s": round(hm.n / (ms * 1e-3) / 1e6, 1),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_tri": bytes_per_tri,
                        "traffic": committed_build_traffic(name, hm.n)},
           "diagnostics": dict(zip(("stack_overflows", "build_errors"), bvh.diagnostics()))}
    bvh.close()
    return res


def _standalone(args, metric, value_key, unit, res, workload):
    out = {"metric": metric, "value": res[value_key], "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res.get("ms_per_step", res.get("batch_graph_ms")),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": workload}, "cpu_baseline": None}
    out.update({k: v for k, v in res.items() if k not in out})
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    import numpy as np
    import torch

    import rtamd
    from rtamd import workloads

    if args.workload == "batch64":
        res = run_batch(rtamd.Context(0), args.steps, args.warmup, args.scenes)
        _standalone(args, "batched hipGraph step vs per-scene launch loop (64 scenes, build + render)", "speedup",
                    "x speed-up", res, res["workload"])
        return
    if args.workload == "diff200k":
        res = run_diff(rtamd.Context(0), args.steps, args.warmup)
        _standalone(args, "differentiable render step (forward + adjoint backward)", "mpaths_s", "Mpaths/s", res,
                    res["workload"])
        return
    if args.workload in ("build1M", "refit10M"):
        refit = args.workload == "refit10M"
        res = run_build_only(rtamd.Context(0), args.steps, args.warmup, refit)
        _standalone(args, "refit Mtris/s @10M tris" if refit else "LBVH build Mtris/s @1M tris", "mtris_s",
                    "Mtris/s", res, res["workload"])
        return
    if args.workload == "sphere10M":
        res = run_sphere(rtamd.Context(0), args.steps, args.warmup)
        _standalone(args, "Mrays/s (primary+bounce), refit-only frames @10M tris", "mrays_s", "Mrays/s", res,
                    res["workload"])
        return

    dev = torch.device("cuda:0")
    ctx = rtamd.Context(0)
    hm = workloads.soup_mesh(args.tris)
    dm = rtamd.DeviceMesh(hm)
    mats, light, cam, opts = workloads.c2_setup(args.width, args.height, args.spp)
    if args.mode == "wavefront":
        opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    if args.one_stream:
        opts.flags |= rtamd.RT_RENDER_ONE_STREAM
    md = rtamd.materials_tensor(mats)
    stream = torch.cuda.current_stream()
    bvh = rtamd.BVH(ctx, dm, 63)
    rgb = torch.empty(args.width * args.height * 3, dtype=torch.float32, device=dev)
    tuned = None
    if args.autotune and args.mode == "wavefront":
        # setup, not timed: the context keeps the pick for this scene's size
        best, best_ms, _ = rtamd.render_autotune(ctx, bvh, dm, md, len(mats), light, cam, opts, rgb=rgb)
        tuned = {"tuning": best.as_dict(), "autotune_render_ms": round(best_ms, 3)}

    def step(events=None, cnt=None):
        if events:
            events[0].record(stream)
        bvh.rebuild(dm)
        if events:
            events[1].record(stream)
        rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, opts, rgb=rgb, counters=cnt)
        if events:
            events[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    # HIP events around every traversal launch, on the stream the library
    # launches on (the dominant kernel's live duration for the roofline).
    # The timed steps count no work; the counts come from an untimed render.
    ctx.set_kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kstats = ctx.kernel_stats()
    # one more step, untimed, with the work counters: ray counts (exact) and
    # the child-box and triangle tests of each traversal kernel
    cnt = torch.zeros(5, dtype=torch.int64, device=dev)
    step(cnt=cnt)
    kwork = ctx.kernel_stats()
    # and one render with every pass on one stream: each traversal launch
    # then runs alone on the GPU (in the timed steps a sample's any-hit
    # launch overlaps the next sample's closest-hit launch, so their spans
    # include time shared with the other kernel)
    iso = None
    if args.mode == "wavefront":
        o1 = rtamd.render_opts(opts.spp, opts.depth, opts.seed, opts.flags | rtamd.RT_RENDER_ONE_STREAM)
        rtamd.render(ctx, bvh, dm, md, len(mats), light, cam, o1, rgb=rgb)
        iso = ctx.kernel_stats()
    ctx.set_kernel_timing(False)
    diag = bvh.diagnostics()
    build_ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
    render_ms = float(np.median([e[1].elapsed_time(e[2]) for e in evs]))
    primary, shadow, bounce, box_tests, tests = (int(x) for x in cnt.cpu().numpy())
    ms_per_step = wall * 1e3 / args.steps
    value = (primary + bounce) / (ms_per_step * 1e-3) / 1e6

    def kernel_roofline(name, ks, work):
        # algorithmic lane-ops per launch (from the counting render) / the
        # launch's average duration (from the timed steps)
        if ks["launches"] == 0 or work["launches"] == 0:
            return None
        avg_ms = ks["ms"] / ks["launches"]
        box = work["box_tests"] // work["launches"]
        tri = work["tri_tests"] // work["launches"]
        ops = box * OPS_PER_BOX_TEST * BOX_TEST_FACTOR + tri * OPS_PER_TRI_TEST * FP64_FACTOR
        ach = ops / (avg_ms * 1e-3) / 1e12
        return {"kernel": name, "bound": "valu", "achieved": round(ach, 3), "peak": FP64_LANE_OPS_PEAK_T,
                "unit": "T lane-op/s", "frac": round(ach / FP64_LANE_OPS_PEAK_T, 4),
                "avg_launch_ms": round(avg_ms, 4), "launches": ks["launches"],
                "work_per_launch": {"box_tests": box, "tri_tests": tri, "lane_ops": ops,
                                    "ops_per_box_test": OPS_PER_BOX_TEST * BOX_TEST_FACTOR,
                                    "ops_per_tri_test": OPS_PER_TRI_TEST * FP64_FACTOR,
                                    "counted": "child-box tests of filled node slots; triangle tests"}}

    roof = kernel_roofline("k_trace_persistent<closest hit, 4-wide quantized> (one launch = one sample's primary rays)",
                           kstats["closest"], kwork["closest"])
    roof_any = kernel_roofline("k_trace_persistent<any hit, 4-wide quantized> (one launch = two samples' shadow rays)",
                               kstats["any"], kwork["any"])
    if roof is None:
        # megakernel: one k_render launch per step does all traversal and shading
        ops = box_tests * OPS_PER_BOX_TEST * BOX_TEST_FACTOR + tests * OPS_PER_TRI_TEST * FP64_FACTOR
        ach = ops / (render_ms * 1e-3) / 1e12
        roof = {"kernel": "k_render (megakernel, one launch per step)", "bound": "valu", "achieved": round(ach, 3),
                "peak": FP64_LANE_OPS_PEAK_T, "unit": "T lane-op/s", "frac": round(ach / FP64_LANE_OPS_PEAK_T, 4),
                "avg_launch_ms": round(render_ms, 4), "traffic": None}
    else:
        roof["traffic"] = committed_traffic("closest", args)
        roof_any["traffic"] = committed_traffic("any", args)
        roof["note"] = ("avg_launch_ms is the launch's span in the timed steps, where it can share the GPU with the "
                        "any-hit launch of the two samples before it; roofline_isolated times the same kernel alone")
    roof_iso = None
    if iso is not None:
        roof_iso = {"closest": kernel_roofline("k_trace_persistent<closest hit> alone", iso["closest"],
                                               kwork["closest"]),
                    "any": kernel_roofline("k_trace_persistent<any hit> alone", iso["any"], kwork["any"])}
    build_bytes = BUILD_BYTES_PER_TRI_63 * args.tris
    build_gbs = build_bytes / (build_ms * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s (primary+bounce) & LBVH build Mtris/s @1M tris; bwd/fwd time ratio",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"soup{args.tris}_63bit_{args.width}x{args.height}_{args.spp}spp_thinlens",
                   "tris": args.tris, "width": args.width, "height": args.height, "spp": args.spp,
                   "morton_bits": 63, "aperture": 0.05, "focus": 1.5},
        "render_mode": args.mode,
        "render_tuning": tuned or {"tuning": "default"},
        "roofline": roof,
        "roofline_any_hit": roof_any,
        "roofline_isolated": roof_iso,
        "render_work": {"box_tests": box_tests, "tri_tests": tests},
        "roofline_build": {
            "kernels": "k_init..k_wide_fill (one rebuild)",
            "bound": "hbm",
            "achieved": round(build_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(build_gbs / HBM_PEAK_GBS, 4),
            "traffic": committed_build_traffic("build_soup1M", args.tris),
            "bytes_per_tri": BUILD_BYTES_PER_TRI_63,
        },
        "build_ms": round(build_ms, 4),
        "build_mtris_s": round(args.tris / (build_ms * 1e-3) / 1e6, 1),
        "render_ms": round(render_ms, 3),
        "render_mrays_s": round((primary + bounce) / (render_ms * 1e-3) / 1e6, 2),
        "primary_rays": primary,
        "shadow_rays": shadow,
        "shadow_rays_traced": kwork["any"]["rays"] if kwork["any"]["launches"] else None,
        "total_rays_incl_shadow_mrays_s": round((primary + bounce + shadow) / (render_ms * 1e-3) / 1e6, 2),
        "diagnostics": {"stack_overflows": diag[0], "build_errors": diag[1]},
        "cpu_baseline": None,
    }
    bvh.close()
    if not args.headline_only:
        # the rest of the metric, each on its own workload (fixed step counts)
        print("bench: C3 refit-only frames, 10M triangles", file=sys.stderr, flush=True)
        c3 = run_sphere(ctx, 5, 2)
        print("bench: C4 batch of 64 scenes", file=sys.stderr, flush=True)
        c4 = run_batch(ctx, 3, 1)
        print("bench: C5 differentiable render, 200K triangles", file=sys.stderr, flush=True)
        c5 = run_diff(ctx, 5, 2)
        out["refit_mtris_s"] = c3["refit_mtris_s"]
        out["batch_speedup"] = c4["speedup"]
        out["bwd_over_fwd"] = c5["bwd_over_fwd"]
        out["c3_refit"] = c3
        out["c4_batch"] = c4
        out["c5_diff"] = c5
    if not args.no_cpu:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(hm, mats, light, cam, opts, args.tris)
        out["gpu_vs_cpu"] = round(out["render_mrays_s"] / max(out["cpu_baseline"]["value"], 1e-9), 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
