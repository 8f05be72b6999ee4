"""Benchmark: LBVH build + thin-lens render of a 1M-triangle random soup on one MI355X.

Workload (one step): 1,000,000-triangle random soup (centroids uniform in the
unit cube, lognormal edges, random orientation, fp32), 63-bit Morton LBVH
build (AABBs -> Morton -> onesweep sort -> Karras -> atomic refit), then a
1920x1080 thin-lens depth-of-field render (aperture 0.05, focus 1.5) at 16
spp with one shadow ray per hit. Inputs are resident in HBM before timing.

Prints one JSON line. value = primary rays per second of the whole step
(build + render) in Mrays/s.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cuda-rt_amd"))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import rtamd
    from tests import scenes

    dev = torch.device("cuda:0")
    ctx = rtamd.Context(0)
    hm = scenes.soup_mesh(args.tris)
    dm = rtamd.DeviceMesh(hm)
    mats, light, cam, opts = scenes.c2_setup(args.width, args.height, args.spp)
    md = rtamd.materials_tensor(mats)
    stream = torch.cuda.current_stream()
    bvh = rtamd.BVH(ctx, dm, 63)
    rgb = torch.empty(args.width * args.height * 3, dtype=torch.float32, device=dev)
    cnt = torch.zeros(3, dtype=torch.int64, device=dev)

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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    build_ms = [e[0].elapsed_time(e[1]) for e in evs]
    render_ms = [e[1].elapsed_time(e[2]) for e in evs]
    counts = cnt.cpu().numpy() // args.steps
    primary = int(counts[0])
    ms_per_step = wall * 1e3 / args.steps
    value = primary / (ms_per_step * 1e-3) / 1e6
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
                   "morton_bits": 63},
        "build_ms": round(float(np.median(build_ms)), 4),
        "build_mtris_s": round(args.tris / (np.median(build_ms) * 1e-3) / 1e6, 1),
        "render_ms": round(float(np.median(render_ms)), 3),
        "render_mrays_s": round(primary / (np.median(render_ms) * 1e-3) / 1e6, 2),
        "shadow_rays": int(counts[1]),
        "primary_rays": primary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
