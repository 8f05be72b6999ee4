"""Benchmark: LBVH build + thin-lens render of a 1M-triangle random soup on one MI355X.

Workload (one step, "soup1M"): 1,000,000-triangle random soup (centroids
uniform in the unit cube, lognormal edges with median 0.002 and sigma 0.5,
uniformly random orientation, fp32), 63-bit Morton LBVH build (triangle
AABBs + scene bounds -> Morton keys -> onesweep radix sort -> fused
hierarchy/refit climb), then a 1920x1080 thin-lens depth-of-field render
(aperture 0.05, focus distance 1.5, fov 40) at 16 spp with one shadow ray
per hit. Inputs are resident in HBM before the timed region.

Prints one JSON line. value = primary rays per second of the whole step
(build + render), in Mrays/s.
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
# Algorithmic HBM bytes per triangle of a 63-bit build.
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
    ap.add_argument("--workload", choices=["soup1M", "batch64", "diff200k", "sphere10M"], default="soup1M",
                    help="soup1M: the headline single-scene step; batch64: 64 scenes, batched vs per-scene loop; "
                         "diff200k: differentiable render forward + adjoint backward; "
                         "sphere10M: refit-only frames of a 10M-triangle displaced sphere with diffuse bounces")
    ap.add_argument("--scenes", type=int, default=64, help="batch64: number of scenes")
    ap.add_argument("--mode", choices=["wavefront", "megakernel"], default="wavefront",
                    help="render kernel structure (identical images)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
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


def calibrate_copy(mib):
    import torch
    a = torch.empty(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "bytes_read": a.numel() * 4, "bytes_written": a.numel() * 4,
                      "copies": 3}))


def cpu_baseline(hm, mats, light, cam, opts, stride, tris):
    """Time the FP64 CPU oracle (a restatement of the reference algorithm) on a
    bounded sample: the full 63-bit build, then every stride-th launch pixel."""
    import numpy as np
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    b = O.build(hm, 63, nthreads=threads)
    t_build = time.perf_counter() - t0
    pixels = np.arange(0, cam.width * cam.height, stride, dtype=np.int64)
    t0 = time.perf_counter()
    _, counts = O.render(hm, b, mats, light, cam, opts, pixels=pixels, nthreads=threads)
    t_render = time.perf_counter() - t0
    return {
        "value": round(counts[0] / t_render / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle FP64 build of all {tris} tris ({t_build:.2f} s, "
                  f"{tris / t_build / 1e6:.3f} Mtris/s) + render of "
                  f"{'every pixel' if stride == 1 else f'every {stride}th pixel'} "
                  f"({pixels.size} px x {opts.spp} spp = {int(counts[0])} primary, {int(counts[1])} shadow rays, "
                  f"{t_render:.2f} s)",
        "build_mtris_s": round(tris / t_build / 1e6, 4),
    }


def bench_batch(args):
    """64 random soups (log-uniform 1K..2M triangles), cameras of 256^2, 512^2
    or 1024^2 at 1..8 spp. Step = rebuild every BVH + render every scene.
    Compares the per-scene launch loop with the batched form replayed from a
    hipGraph, and checks that the images are identical."""
    import numpy as np
    import torch

    import rtamd
    from tests import scenes
    ctx = rtamd.Context(0)
    specs = scenes.batch_scenes(args.scenes, 0x5EED0004)
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

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    loop_ms = timed(loop_step, args.steps, args.warmup)
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
    batch_ms = timed(g.replay, args.steps, args.warmup)
    same = all(torch.equal(a, s["rgb"]) for a, s in zip(loop_rgb, sc))
    out = {
        "metric": "batched hipGraph step vs per-scene launch loop (64 scenes, build + render)",
        "value": round(loop_ms / batch_ms, 3),
        "unit": "x speed-up",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(batch_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batch{len(sc)}_soups_1K-2M_tris_256-1024px_1-8spp", "scenes": len(sc),
                   "tris": tris, "pixels": pixels, "primary_rays": primary},
        "loop_ms": round(loop_ms, 3), "batch_graph_ms": round(batch_ms, 3),
        "batch_mrays_s": round(primary / (batch_ms * 1e-3) / 1e6, 2),
        "loop_mrays_s": round(primary / (loop_ms * 1e-3) / 1e6, 2),
        "identical_images": bool(same),
        "tuning": best.as_dict(), "autotune_best_render_ms": round(tuned_ms, 3),
        "stack_overflows": batch.diagnostics(),
    }
    print(json.dumps(out), flush=True)


REFIT_BYTES_PER_TRI = 152  # refit-only frame: see DESIGN.md


def bench_sphere(args):
    """Refit-only frames: a 9,996,980-triangle icosphere (frequency 707,
    radius 1, 3-octave value-noise displacement of amplitude 0.05), 30-bit
    build once; each frame moves every vertex by Gaussian jitter (sigma
    1e-4, precomputed frames resident in HBM), refits the BVH and renders
    1024x1024 at 4 spp with cosine-weighted diffuse bounces to depth 4 and
    one shadow ray per hit. Step = refit + render."""
    import numpy as np
    import torch

    import rtamd
    ctx = rtamd.Context(0)
    hm = rtamd.gen_icosphere(707, 1.0, (0.0, 0.0, -3.0), 0.05, 0x5EED0003)
    frames = []
    for k in range(3):
        j = rtamd.gen_jitter(hm, 1e-4, 0x5EED1000 + k)
        frames.append([torch.from_numpy(a).cuda() for a in (j.vx, j.vy, j.vz)])
    dm = rtamd.DeviceMesh(hm)
    mats = [rtamd.material(kd=(0.8, 0.8, 0.8))]
    md = rtamd.materials_tensor(mats)
    light = rtamd.light(position=(2.0, 4.0, 0.0))
    cam = rtamd.camera(1024, 1024, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    opts = rtamd.render_opts(4, depth=4, flags=rtamd.RT_RENDER_CLAMP | rtamd.RT_RENDER_WAVEFRONT)
    bvh = rtamd.BVH(ctx, dm, 30)
    rgb = torch.empty(1024 * 1024 * 3, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(5, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    state = {"k": 0}

    def step(ev=None):
        f = frames[state["k"] % len(frames)]
        state["k"] += 1
        dm.vx, dm.vy, dm.vz = f
        if ev:
            ev[0].record(stream)
        bvh.refit(dm)
        if ev:
            ev[1].record(stream)
        rtamd.render(ctx, bvh, dm, md, 1, light, cam, opts, rgb=rgb, counters=cnt)
        if ev:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    cnt.zero_()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    refit_ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
    render_ms = float(np.median([e[1].elapsed_time(e[2]) for e in evs]))
    c = cnt.cpu().numpy() // args.steps
    primary, shadow, bounce = int(c[0]), int(c[1]), int(c[2])
    ms = wall * 1e3 / args.steps
    refit_gbs = REFIT_BYTES_PER_TRI * hm.n / (refit_ms * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s (primary+bounce), refit-only frames @10M tris",
        "value": round((primary + bounce) / (ms * 1e-3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "sphere10M_refit_1024x1024_4spp_depth4", "tris": hm.n, "width": 1024, "height": 1024,
                   "spp": 4, "depth": 4, "jitter_sigma": 1e-4},
        "refit_ms": round(refit_ms, 4), "refit_mtris_s": round(hm.n / (refit_ms * 1e-3) / 1e6, 1),
        "roofline_refit": {"bound": "hbm", "achieved": round(refit_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(refit_gbs / HBM_PEAK_GBS, 4), "bytes_per_tri": REFIT_BYTES_PER_TRI},
        "render_ms": round(render_ms, 3),
        "render_mrays_s": round((primary + bounce) / (render_ms * 1e-3) / 1e6, 2),
        "primary_rays": primary, "bounce_rays": bounce, "shadow_rays": shadow,
        "diagnostics": dict(zip(("stack_overflows", "build_errors"), bvh.diagnostics())),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_diff(args):
    """Differentiable render step on a 200K-triangle displaced icosphere
    (frequency 100, displacement 0.05): 512x512, 16 spp, 2 diffuse bounces,
    per-triangle albedo U[0.2, 0.9], light intensity 10. The target image is
    rendered at the true albedo; the parameters are the albedo perturbed by
    N(0, 0.05). Step = forward (render + path record) + backward (adjoint
    kernel) of the L2 image loss."""
    import numpy as np
    import torch

    import rtamd
    from rtamd import diff
    ctx = rtamd.Context(0)
    hm = rtamd.gen_icosphere(100, 1.0, (0.0, 0.0, -3.0), 0.05, 0x5EED0006)
    dm = rtamd.DeviceMesh(hm)
    light = rtamd.light(position=(2.0, 3.0, 0.0), color=(1.0, 1.0, 1.0), radius=0.5, c0=0.5, c1=0.2, c2=0.05)
    cam = rtamd.camera(512, 512, fov_deg=45.0, fstop=float("inf"), focal_point=1.0)
    spp, depth = 16, 2
    R = diff.DiffRenderer(ctx, dm, light, cam, spp, depth, seed=1)
    rng = np.random.default_rng(0x5EED0007)
    true_alb = rng.uniform(0.2, 0.9, (hm.n, 3))
    alb = torch.tensor(np.clip(true_alb + rng.normal(0, 0.05, true_alb.shape), 0.0, 1.0), dtype=torch.float64,
                       device="cuda")
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

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        g_alb, g_light, g_pos = step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
    bwd_ms = float(np.median([e[1].elapsed_time(e[2]) for e in evs]))
    ms = wall * 1e3 / args.steps
    primary = cam.width * cam.height * spp
    out = {
        "metric": "differentiable render step (forward + adjoint backward)",
        "value": round(primary / (ms * 1e-3) / 1e6, 2),
        "unit": "Mpaths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "diff_icosphere200K_512x512_16spp_depth2", "tris": hm.n, "width": 512,
                   "height": 512, "spp": spp, "depth": depth},
        "forward_ms": round(fwd_ms, 3), "backward_ms": round(bwd_ms, 3),
        "backward_over_forward": round(bwd_ms / fwd_ms, 3),
        "loss": float(((target - R.forward(alb, s)) ** 2).sum()),
        "grad_norms": {"albedo": float(g_alb.norm()), "light": float(g_light), "positions": float(g_pos.norm())},
        "record_mib": round(R.records.numel() * 4 / 2**20, 1),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.workload == "batch64":
        bench_batch(args)
        return
    if args.workload == "diff200k":
        bench_diff(args)
        return
    if args.workload == "sphere10M":
        bench_sphere(args)
        return
    import numpy as np
    import torch

    import rtamd
    from tests import scenes

    dev = torch.device("cuda:0")
    ctx = rtamd.Context(0)
    hm = scenes.soup_mesh(args.tris)
    dm = rtamd.DeviceMesh(hm)
    mats, light, cam, opts = scenes.c2_setup(args.width, args.height, args.spp)
    if args.mode == "wavefront":
        opts.flags |= rtamd.RT_RENDER_WAVEFRONT
    md = rtamd.materials_tensor(mats)
    stream = torch.cuda.current_stream()
    bvh = rtamd.BVH(ctx, dm, 63)
    rgb = torch.empty(args.width * args.height * 3, dtype=torch.float32, device=dev)
    cnt = torch.zeros(5, dtype=torch.int64, device=dev)

    def step(events=None):
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
    cnt.zero_()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    # HIP events around every traversal launch, on the stream the library
    # launches on (the dominant kernel's live duration for the roofline)
    ctx.set_kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kstats = ctx.kernel_stats()
    # one more render, untimed, with every pass on one stream: each traversal
    # launch then runs alone on the GPU (in the timed steps a sample's any-hit
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
    c = cnt.cpu().numpy() // args.steps
    primary, shadow, bounce, box_tests, tests = (int(x) for x in c)
    ms_per_step = wall * 1e3 / args.steps
    value = (primary + bounce) / (ms_per_step * 1e-3) / 1e6

    def kernel_roofline(name, ks):
        # algorithmic lane-ops per launch / the launch's average duration
        if ks["launches"] == 0:
            return None
        launches = ks["launches"]
        avg_ms = ks["ms"] / launches
        ops = (ks["box_tests"] * OPS_PER_BOX_TEST * BOX_TEST_FACTOR + ks["tri_tests"] * OPS_PER_TRI_TEST * FP64_FACTOR)
        ach = ops / launches / (avg_ms * 1e-3) / 1e12
        return {"kernel": name, "bound": "valu", "achieved": round(ach, 3), "peak": FP64_LANE_OPS_PEAK_T,
                "unit": "T lane-op/s", "frac": round(ach / FP64_LANE_OPS_PEAK_T, 4),
                "avg_launch_ms": round(avg_ms, 4), "launches": ks["launches"],
                "work_per_launch": {"box_tests": ks["box_tests"] // launches, "tri_tests": ks["tri_tests"] // launches,
                                    "lane_ops": ops // launches, "ops_per_box_test": OPS_PER_BOX_TEST * BOX_TEST_FACTOR,
                                    "ops_per_tri_test": OPS_PER_TRI_TEST * FP64_FACTOR}}

    roof = kernel_roofline("k_trace_persistent<closest hit, 4-wide quantized> (one launch = one sample's primary rays)",
                           kstats["closest"])
    roof_any = kernel_roofline("k_trace_persistent<any hit, 4-wide quantized> (one launch = one sample's shadow rays)",
                               kstats["any"])
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
        roof["note"] = ("avg_launch_ms is the launch's span in the timed steps, where it shares the GPU with the "
                        "previous sample's any-hit launch; roofline_isolated times the same kernel alone")
    roof_iso = None
    if iso is not None:
        roof_iso = {"closest": kernel_roofline("k_trace_persistent<closest hit> alone", iso["closest"]),
                    "any": kernel_roofline("k_trace_persistent<any hit> alone", iso["any"])}
    build_bytes = BUILD_BYTES_PER_TRI_63 * args.tris
    build_gbs = build_bytes / (build_ms * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s (primary+bounce) & LBVH build Mtris/s @1M tris",
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
        "roofline": roof,
        "roofline_any_hit": roof_any,
        "roofline_isolated": roof_iso,
        "render_work": {"box_tests": box_tests, "tri_tests": tests},
        "roofline_build": {
            "kernels": "k_init..k_climb_top (one rebuild)",
            "bound": "hbm",
            "achieved": round(build_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(build_gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_tri": BUILD_BYTES_PER_TRI_63,
        },
        "build_ms": round(build_ms, 4),
        "build_mtris_s": round(args.tris / (build_ms * 1e-3) / 1e6, 1),
        "render_ms": round(render_ms, 3),
        "render_mrays_s": round((primary + bounce) / (render_ms * 1e-3) / 1e6, 2),
        "primary_rays": primary,
        "shadow_rays": shadow,
        "total_rays_incl_shadow_mrays_s": round((primary + bounce + shadow) / (render_ms * 1e-3) / 1e6, 2),
        "diagnostics": {"stack_overflows": diag[0], "build_errors": diag[1]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(hm, mats, light, cam, opts, args.cpu_pixel_stride, args.tris)
        out["gpu_vs_cpu"] = round(out["render_mrays_s"] / max(out["cpu_baseline"]["value"], 1e-9), 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
